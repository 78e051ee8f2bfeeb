"""The command-line driver (dmrg_amd/dmrgx, the reference mains' interface)
and the reference's energy/singular output files: a Kitaev run against exact
diagonalisation, the files' shapes, and the file rows against the step log of
the same run through the library."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import dmrg_amd
from dmrg_amd import driver
from oracle import exact

pytestmark = pytest.mark.gpu

EXE = Path(dmrg_amd.__file__).resolve().parent / "dmrgx"


def test_cli_kitaev_files(gpu, tmp_path):
    assert EXE.exists(), "dmrg_amd/dmrgx is not built (python -m dmrg_amd.build)"
    print("case cli kitaev L=12", flush=True)
    args = ["kitaev", "l=12", "m=32", "n=2", "t=-1", "p=-1", "u=0.25"]
    r = subprocess.run([str(EXE), *args], cwd=tmp_path, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    suffix = "_L_12_m_32_n_2_s_0_t_-1_p_-1_u_0.25_.txt"
    E = np.loadtxt(tmp_path / ("energy" + suffix), ndmin=2)
    S = np.loadtxt(tmp_path / ("singular" + suffix), ndmin=2)
    assert E.shape[1] == 2 + 6 and S.shape[1] == 3 + 20 + 1 and len(E) == len(S)
    e_ed = exact.kitaev_ed(12, -1.0, -1.0, 0.25)[0]
    assert abs(E[-1, 2] - e_ed) < 1e-9, (E[-1, 2], e_ed)
    norm = [l for l in r.stdout.splitlines() if l.startswith("Normalization:")]
    assert norm and abs(float(norm[0].split()[1]) - 1) < 1e-9
    # the same schedule through the library: identical step sequence and values
    res = driver.run_dmrg("kitaev", 12, 32, 2, (-1.0, -1.0, 0.25), schedule="dmrg")
    assert len(res.steps) == len(E)
    assert np.array_equal(res.steps[:, :2], E[:, :2])
    assert np.max(np.abs(res.steps[:, 2] - E[:, 2])) < 1e-9
    assert np.max(np.abs(res.steps[:, 9] - S[:, 2])) < 1e-9    # entropy
    assert np.max(np.abs(res.steps[:, 8] - S[:, -1])) < 1e-9   # truncation error
