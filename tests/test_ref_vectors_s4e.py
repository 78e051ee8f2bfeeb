"""The Z2 x Z2 x Z2 group, the 4e model and the periodic edge environments
against vectors of the reference's own code (tests/golden/ref_vectors_s4e.json,
written by tests/golden/make_ref_vectors_s4e.py with oracle/_ref/ref_harness_s4e,
compiled from the reference's sources and 4e.cpp under its SYMMETRY 8 / S4E /
PBC switches).  Pinned here:
  * the 4e MPO builders H_4e1 / H_4e2, their edges and the main's bond-1
    operators (4e.cpp:397-779), entry by entry;
  * update_two's environments and the periodic edge environments through
    qtensor::id(bulk) (src/core.cpp:188-199, src/qtensor.cpp:157-182), by
    per-block norms and fixed random projections;
  * generate_dim_sym and both get_map task lists under the S4E fusion table
    and fermion parities (src/useful.cpp:4-27), bit for bit through SHA-256
    digests, in the oracle and in the C++ mirror.
CPU only."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from dmrg_amd import driver
from oracle import models
from oracle import qtensor as qt
from oracle.qtensor import QTensor, generate_dim_sym

ROOT = Path(__file__).resolve().parents[1]
FIX = json.loads((ROOT / "tests" / "golden" / "ref_vectors_s4e.json").read_text())
CASES = {c["name"]: c for c in FIX["cases"]}


def digest(obj):
    return hashlib.sha256(json.dumps(obj, separators=(",", ":"), default=int).encode()).hexdigest()


def projections(flat, seed=FIX["projection_seed"]):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(8):
        r = rng.uniform(-1.0, 1.0, flat.size)
        p = complex(np.dot(r, flat))
        out.append([p.real, p.imag])
    return np.array(out)


def same_small(t: QTensor, ref):
    assert t.dim == ref["dim"] and t.sym == ref["sym"] and list(t.dir) == ref["dir"]
    flat = t.flat()
    assert flat.size == ref["size"]
    want = np.zeros(ref["size"], dtype=complex)
    for j, re, im in ref["nz"]:
        want[j] = re + 1j * im
    assert np.max(np.abs(flat - want), initial=0.0) <= 1e-15 * max(1.0, np.max(np.abs(want)))


def same_large(t: QTensor, ref):
    assert t.dim == ref["dim"] and list(t.dir) == ref["dir"]
    assert len(t.sym) == ref["n_blocks"] and digest(t.sym) == ref["sym_sha"]
    flat = t.flat()
    norms = [float(np.linalg.norm(v)) for v in t.val]
    assert np.allclose(norms, ref["block_norms"], rtol=1e-13, atol=1e-13)
    p, want = projections(flat), np.array(ref["proj"])
    assert np.max(np.abs(p - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


def oracle_case(c):
    qt.set_symmetry(qt.S4E)
    t1, t3, f0, f1, p1, frac = c["params"]
    W = models.H_4e(c["second"], t1, t3, f0, f1, p1, frac)
    ls = models.H_4e(False, t1, t3, f0, f1, p1, 0.0).left()
    rs = models.H_4e(True, t1, t3, f0, f1, p1, frac).right()
    lenv = QTensor().contract_index(ls, 1, W, 0).exchange(3, 4)
    renv = QTensor().contract_index(rs, 1, W.exchange(0, 2), 0, True).exchange(0, 1)
    ope = ls.id(W)
    ledge = QTensor().contract_index(ls, 1, ope, 0).exchange(3, 4)
    redge = QTensor().contract_index(rs, 1, ope, 0, True).exchange(0, 1)
    return W, lenv, renv, ledge, redge


def test_s4e_table_is_xor_and_parities():
    """The fusion table the oracle restates is bitwise XOR (the C++ mirror
    computes XOR); ff marks the sectors 2..5 odd."""
    qt.set_symmetry(qt.S4E)
    for a in range(8):
        for b in range(8):
            assert qt.ss(a, b) == a ^ b == qt.ss(a, -b)
    assert [qt.ff(a) for a in range(8)] == [0, 0, 1, 1, 1, 1, 0, 0]
    qt.set_symmetry(2)


@pytest.mark.parametrize("name", list(CASES))
def test_4e_builders_match_reference(name):
    c = CASES[name]
    W, *_ = oracle_case(c)
    ref = c["reference"]
    same_small(W, ref["W"])
    same_small(W.left(), ref["Wleft"])
    same_small(W.right(), ref["Wright"])


def test_4e_operators_match_reference():
    qt.set_symmetry(qt.S4E)
    ref = next(iter(CASES.values()))["reference"]
    ops = {"op_up1": ("lcu", 0.5, False), "op_down1": ("lcd", 0.5, False), "op_up2": ("rcu", 0.5, False),
           "op_down2": ("rcd", 0.5, False), "op_pair1": ("lp", 0.5, False), "op_pair2": ("rp", 0.5, False),
           "op_pairpair": ("pp", 0.5, False), "op_pair_dagger1": ("lp", -0.5, True),
           "op_pair_dagger2": ("rp", -0.5, True)}
    for key, (tab, coef, flip) in ops.items():
        same_small(models.op_4e(tab, coef, flip), ref[key])
    same_small(models.op_4e("lp").id(), ref["op_id"])


@pytest.mark.parametrize("name", list(CASES))
def test_environments_and_edge_environments_match_reference(name):
    c = CASES[name]
    _, lenv, renv, ledge, redge = oracle_case(c)
    ref = c["reference"]
    same_large(lenv, ref["lenv"])
    same_large(renv, ref["renv"])
    same_large(ledge, ref["ledge"])
    same_large(redge, ref["redge"])


def _check_maps(ref, dim, sym, rmap, dim_ret, sym_ret, lmap, ldim):
    assert dim == ref["dim"] and len(sym) == ref["sym"]["n"] and digest(sym) == ref["sym"]["sha"]
    assert len(rmap) == ref["rmap"]["n"] and rmap[:16] == ref["rmap"]["head"] and digest(rmap) == ref["rmap"]["sha"]
    assert dim_ret == ref["dim_ret"] and digest(sym_ret) == ref["sym_ret"]["sha"]
    assert len(lmap) == ref["lmap"]["n"] and lmap[:16] == ref["lmap"]["head"] and digest(lmap) == ref["lmap"]["sha"]
    assert ldim == ref["ldim"]


def test_get_map_oracle_matches_reference_s4e():
    c = CASES["h4e1_s0"]
    _, lenv, renv, _, _ = oracle_case(c)
    dim, sym = generate_dim_sym(lenv.dimension_all(), 2, renv.dimension_all(), 2, c["sector"])
    rmap, dim_ret, sym_ret = renv.get_map(dim, sym, "N", "T", 2, 0, False, fermion=True)
    lmap, ldim, _ = lenv.get_map(dim_ret, sym_ret, "N", "T", 3, 0, True, sym_ret=sym, fermion=True)
    _check_maps(c["reference"], dim, sym, rmap, dim_ret, sym_ret, lmap, ldim)


@pytest.mark.parametrize("name", list(CASES))
def test_get_map_mirror_matches_reference_s4e(name):
    """The C++ mirror (symmetry DMRGX_SYM_S4E) on the environments' structure,
    which same_large pins to the reference's."""
    c = CASES[name]
    _, lenv, renv, _, _ = oracle_case(c)
    S = -8  # DMRGX_SYM_S4E
    dim, sym = driver.generate_dim_sym(S, lenv.dim, lenv.sym, lenv.dir, 2, renv.dim, renv.sym, renv.dir, 2,
                                       c["sector"])
    rmap, dim_ret, sym_ret = driver.get_map(S, renv.dim, renv.sym, renv.dir, dim, sym, "N", "T", 2, 0, False,
                                            fermion=True)
    lmap, ldim, _ = driver.get_map(S, lenv.dim, lenv.sym, lenv.dir, dim_ret, sym_ret, "N", "T", 3, 0, True,
                                   sym_ret=sym, fermion=True)
    _check_maps(c["reference"], dim, sym, rmap, dim_ret, sym_ret, lmap, ldim)


def test_chain_spectrum_fixture_is_consistent():
    """The recorded exact spectra: sorted, the periodic chain differs from the
    open one, and sector 3 lies above sector 0."""
    sp = FIX["chain"]["spectra"]
    for v in sp.values():
        assert v["lowest"] == sorted(v["lowest"]) and v["n"] == 8192
    assert abs(sp["pbc_s0"]["lowest"][0] - sp["open_s0"]["lowest"][0]) > 0.1
    assert sp["open_s3"]["lowest"][0] > sp["open_s0"]["lowest"][0]
