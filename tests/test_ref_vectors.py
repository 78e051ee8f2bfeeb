"""The oracle and the library's integer bookkeeping against vectors computed by
the reference's own code (tests/golden/ref_vectors.json, written by
tests/golden/make_ref_vectors.py with oracle/_ref/ref_harness, which is
compiled from the reference's src/qtensor.cpp, src/tensor.cpp,
src/useful.cpp and src/operators.cpp with its shipped global.h: Z2, graded
signs on).  Pinned here, bit for bit where the quantity is integer:
  * the MPO builders and edges (src/operators.cpp, include/mpo.h:45-48);
  * update_two's environments from given stores: the env x MPO contraction
    with its graded signs and the leg exchanges (src/core.cpp:188-192,
    src/qtensor.cpp:902-983, src/tensor.cpp:652-695);
  * generate_dim_sym and both get_map calls (src/useful.cpp:151-194,
    src/qtensor.cpp:997-1119), in the oracle and in the C++ mirror.
CPU only.  When oracle/_ref/ref_harness is built (in the build container),
a fresh case is also run through it directly."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dmrg_amd import driver
from oracle import models
from oracle import qtensor as qt
from oracle.qtensor import QTensor, generate_dim_sym

ROOT = Path(__file__).resolve().parents[1]
FIX = json.loads((ROOT / "tests" / "golden" / "ref_vectors.json").read_text())
CASES = {c["name"]: c for c in FIX["cases"]}
HARNESS = ROOT / "oracle" / "_ref" / "ref_harness"


def to_qt(d, cplx):
    flat = np.asarray(d["re"], dtype=float)
    if cplx:
        flat = flat + 1j * np.asarray(d["im"], dtype=float)
    t = QTensor.from_flat(flat, d["dim"], d["sym"], dtype=np.complex128 if cplx else np.float64)
    t.dir = list(d["dir"])
    return t


def same_tensor(a: QTensor, ref: dict, cplx: bool, tol=1e-14):
    assert a.dim == ref["dim"]
    assert a.sym == ref["sym"]
    assert list(a.dir) == ref["dir"]
    want = np.asarray(ref["re"]) + (1j * np.asarray(ref["im"]) if cplx else 0)
    got = a.flat()
    assert got.shape == want.shape
    scale = max(1.0, float(np.max(np.abs(want)))) if want.size else 1.0
    assert np.max(np.abs(got - want), initial=0.0) <= tol * scale


def oracle_case(c):
    cplx = "im" in c["lstore"]
    qt.set_symmetry(2)
    mpo, n, ferm = models.make_model(c["model"], c["L"], **c["params"])
    ls, rs = to_qt(c["lstore"], cplx), to_qt(c["rstore"], cplx)
    W = mpo.ops[0]
    lenv = QTensor().contract_index(ls, 1, W, 0).exchange(3, 4)
    ope = W.exchange(0, 2)
    renv = QTensor().contract_index(rs, 1, ope, 0, True).exchange(0, 1)
    return cplx, mpo, W, lenv, renv


@pytest.mark.parametrize("name", list(CASES))
def test_mpo_builders_match_reference(name):
    c = CASES[name]
    cplx, mpo, W, _, _ = oracle_case(c)
    ref = c["reference"]
    same_tensor(W, ref["W"], cplx, tol=0)
    same_tensor(W.left(), ref["Wleft"], cplx, tol=0)
    same_tensor(W.right(), ref["Wright"], cplx, tol=0)


@pytest.mark.parametrize("name", list(CASES))
def test_environments_match_reference(name):
    c = CASES[name]
    cplx, _, _, lenv, renv = oracle_case(c)
    same_tensor(lenv, c["reference"]["lenv"], cplx)
    same_tensor(renv, c["reference"]["renv"], cplx)


@pytest.mark.parametrize("name", list(CASES))
def test_get_map_oracle_matches_reference(name):
    c = CASES[name]
    _, _, _, lenv, renv = oracle_case(c)
    ref = c["reference"]
    dim, sym = generate_dim_sym(lenv.dimension_all(), 2, renv.dimension_all(), 2, c["sector"])
    assert dim == ref["dim"] and sym == ref["sym"]
    rmap, dim_ret, sym_ret = renv.get_map(dim, sym, "N", "T", 2, 0, False, fermion=True)
    assert rmap == ref["rmap"]
    assert dim_ret == ref["dim_ret"] and sym_ret == ref["sym_ret"]
    lmap, ldim, _ = lenv.get_map(dim_ret, sym_ret, "N", "T", 3, 0, True, sym_ret=sym, fermion=True)
    assert lmap == ref["lmap"]
    assert ldim == ref["ldim"]


@pytest.mark.parametrize("name", list(CASES))
def test_get_map_mirror_matches_reference(name):
    """The C++ mirror (dmrgx_generate_dim_sym, dmrgx_get_map) on the
    reference's own environments: structure only, no device."""
    c = CASES[name]
    ref = c["reference"]
    le, re_ = ref["lenv"], ref["renv"]
    dim, sym = driver.generate_dim_sym(2, le["dim"], le["sym"], le["dir"], 2, re_["dim"], re_["sym"], re_["dir"], 2,
                                       c["sector"])
    assert dim == ref["dim"] and sym == ref["sym"]
    rmap, dim_ret, sym_ret = driver.get_map(2, re_["dim"], re_["sym"], re_["dir"], dim, sym, "N", "T", 2, 0, False,
                                            fermion=True)
    assert rmap == ref["rmap"]
    assert dim_ret == ref["dim_ret"] and sym_ret == ref["sym_ret"]
    lmap, ldim, _ = driver.get_map(2, le["dim"], le["sym"], le["dir"], dim_ret, sym_ret, "N", "T", 3, 0, True,
                                   sym_ret=sym, fermion=True)
    assert lmap == ref["lmap"]
    assert ldim == ref["ldim"]


def test_fixture_is_not_trivial():
    """The vectors exercise the graded sign: some tasks carry alpha = -1."""
    assert any(t[6] == -1 for c in FIX["cases"] for t in c["reference"]["rmap"])
    assert any(c["reference"]["lenv"]["dir"] != [0] * 5 for c in FIX["cases"])


@pytest.mark.skipif(not HARNESS.exists(), reason="oracle/_ref/ref_harness is built only in the build container")
def test_fresh_case_through_the_compiled_reference():
    """A case not in the fixture (another seed and coupling set) straight
    through the compiled reference: oracle and mirror agree with it."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", ROOT / "tests" / "golden" / "make_ref_vectors.py")
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    from oracle.dmrg import Runner
    mpo, n, ferm = models.make_model("kitaev", 10, t=-1.2, p=-0.6, u=0.3)
    R = Runner(mpo, n, ferm, seed=77)
    R.dmrg2(7, 1)
    ls, rs = R.mps.store[3], R.mps.store[6]
    text = (f"model kitaev -1.2 -0.6 0.3 0.0\nlstore {mk.tensor_text(ls)}\nrstore {mk.tensor_text(rs)}\n"
            f"sector 0\n")
    run = subprocess.run([str(HARNESS)], input=text, capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    ref = mk.parse_output(run.stdout, False)
    c = {"model": "kitaev", "L": 10, "params": dict(t=-1.2, p=-0.6, u=0.3), "lstore": mk.tensor_json(ls),
         "rstore": mk.tensor_json(rs)}
    _, _, _, lenv, renv = oracle_case(c)
    same_tensor(lenv, ref["lenv"], False)
    same_tensor(renv, ref["renv"], False)
    dim, sym = generate_dim_sym(lenv.dimension_all(), 2, renv.dimension_all(), 2, 0)
    rmap, dim_ret, sym_ret = renv.get_map(dim, sym, "N", "T", 2, 0, False, fermion=True)
    assert rmap == ref["rmap"]
    m2, _, _ = driver.get_map(2, ref["renv"]["dim"], ref["renv"]["sym"], ref["renv"]["dir"], dim, sym, "N", "T", 2,
                              0, False, fermion=True)
    assert m2 == ref["rmap"]
