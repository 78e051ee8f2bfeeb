"""The 4e superconductor chain (4e.cpp) on the GPU: Z2 x Z2 x Z2 sectors,
graded fermionic MPOs of bond 9 with two sites per tensor, open and periodic.
Pinned by the exact low spectrum of the eight-site chain recorded in
tests/golden/ref_vectors_s4e.json (the oracle's exact two-site operator, whose
builders, environments and task lists tests/test_ref_vectors_s4e.py checks
against the compiled reference): at m=256 nothing is truncated on four
tensors, so the step energies are exact."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import dmrg_amd
from dmrg_amd import driver

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
FIX = json.loads((ROOT / "tests" / "golden" / "ref_vectors_s4e.json").read_text())
CHAIN = FIX["chain"]
P = tuple(CHAIN["params"][k] for k in ("t1", "t3", "f0", "f1", "frac"))
EXE = Path(dmrg_amd.__file__).resolve().parent / "dmrgx"


@pytest.mark.parametrize("label", ["open_s0", "open_s3", "pbc_s0"])
@pytest.mark.parametrize("schedule", ["dmrg2", "dmrg"])
def test_4e_eight_sites_exact(gpu, label, schedule):
    sp = CHAIN["spectra"][label]
    print(f"case 4e L=8 {label} {schedule}", flush=True)
    res = driver.run_dmrg("4e", 8, 256, 1, P, sector=sp["sector"], schedule=schedule, pbc=sp["pbc"])
    assert abs(res.energy - sp["lowest"][0]) < 1e-9, (res.energy, sp["lowest"])
    if schedule == "dmrg2":
        # the two-site step on tensors (1, 2) is the whole chain: its three
        # lowest eigenvalues are the chain's
        last = res.steps[-1]
        assert (last[0], last[1]) == (1, 2)
        assert np.max(np.abs(last[2:5] - np.array(sp["lowest"]))) < 1e-8, (last[2:5], sp["lowest"])


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_4e_mixed_precision(gpu, prec):
    sp = CHAIN["spectra"]["open_s0"]
    print(f"case 4e L=8 {prec}", flush=True)
    res = driver.run_dmrg("4e", 8, 256, 1, P, prec=prec, tol_lo=1e-5)
    assert abs(res.energy - sp["lowest"][0]) < 1e-8, (res.energy, sp["lowest"][0])


def test_4e_longer_chain_and_correlations(gpu):
    """Sixteen sites (eight tensors), two sweeps at m=64: the state is
    normalised and the main's correlation lines come in its order (13 end
    and normalisation lines, then two Cooper-pair lines per interior tensor,
    then the pair-pair lines).  The correlation values themselves go through
    calc(mps, mpo) with graded signs, which has no independent check here."""
    print("case 4e L=16 m=64 correlations", flush=True)
    a = driver.run_dmrg("4e", 16, 64, 2, P, keep_state=True)
    b = driver.run_dmrg("4e", 16, 64, 2, P, keep_state=True, prec="bf16x3", tol_lo=1e-5)
    assert abs(a.energy - b.energy) < 1e-8
    # both states are truncated at about 4e-8: they agree to that level
    trunc = max(a.stats["trunc_err"], b.stats["trunc_err"])
    assert abs(a.state.fidelity(b.state) - 1) < max(1e-9, 10 * trunc)
    rows = a.state.correlations_4e()
    n = 8
    assert len(rows) == 13 + 2 * (n - 2) + (n - 2) + (n - 3)
    assert np.all(rows[:13, 0] == n)
    # calc(mps): the truncation keeps the state unnormalised (src/qtensor.cpp:1122-1262
    # never rescales the kept singular values), so the norm is 1 - (discarded weight)
    nrm = a.state.norm2()
    assert abs(rows[12, 1] - nrm.real) < 1e-12 and abs(rows[12, 2]) < 1e-10
    assert abs(rows[12, 1] - 1) < max(1e-10, 10 * a.stats["trunc_err"])
    assert np.all(np.isfinite(rows))
    cp = rows[13:13 + 2 * (n - 2), 0]
    assert list(cp) == [x for q in range(1, n - 1) for x in (2 * q + 1, 2 * q + 2)]
    a.state.close()
    b.state.close()


def test_cli_4e_data_file(gpu, tmp_path):
    assert EXE.exists(), "dmrg_amd/dmrgx is not built (python -m dmrg_amd.build)"
    print("case cli 4e L=8", flush=True)
    args = ["4e", "l=8", "m=256", "n=1", "t1=1", "t3=0.5", "f0=0.3", "f1=0.2", "frac=0.25"]
    r = subprocess.run([str(EXE), *args], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    e = float(r.stdout.split("Energy:")[1].split()[0])
    assert abs(e - CHAIN["spectra"]["open_s0"]["lowest"][0]) < 1e-8
    data = sorted(tmp_path.glob("data*"))
    assert len(data) == 1
    rows = np.loadtxt(data[0], ndmin=2)
    assert rows.shape == (13 + 2 * 2 + 2 + 1, 3)
    assert abs(rows[12, 1] - 1) < 1e-10
