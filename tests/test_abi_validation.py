"""Argument checks of the C ABI that run before the device is touched, so they
are exercised on a CPU-only host: each call must return an error code and a
message instead of launching work on bad input."""
import ctypes as C

import numpy as np
import pytest

import dmrg_amd
from dmrg_amd import driver


def _lib():
    return dmrg_amd.load_library()


def _err():
    return _lib().dmrgx_last_error().decode()


def _run(model, schedule, sign_rule, prec="fp64", L=8):
    cfg = driver._config(model, L, 16, 1, (1.0, 0.8, 0.5), 0, schedule, prec, sign_rule, True, 1, 1e-10, 1e-6, 8,
                         False)
    st = driver.RunStats()
    n = C.c_int()
    return _lib().dmrgx_run_dmrg_state(C.byref(cfg), C.byref(st), None, 0, C.byref(n), None)


@pytest.mark.parametrize("model", ["xyz", "potts3"])
def test_single_site_reference_rule_refused_for_bosonic_models(model):
    """Under the reference's sign rule the single-site move2right operator of a
    bosonic model is non-Hermitian (the reference solves it with complex
    Arnoldi, src/znaupd.cpp:78-197); the Hermitian solver refuses it."""
    rc = _run(model, "dmrg", "reference")
    assert rc != 0
    assert "non-Hermitian" in _err()


def test_unknown_schedule_and_precision_refused():
    cfg = driver._config("xyz", 8, 16, 1, (1.0, 0.8, 0.5), 0, "dmrg2", "fp64", "physical", True, 1, 1e-10, 1e-6, 8,
                         False)
    cfg.schedule = 7
    st = driver.RunStats()
    n = C.c_int()
    assert _lib().dmrgx_run_dmrg_state(C.byref(cfg), C.byref(st), None, 0, C.byref(n), None) != 0
    assert "schedule" in _err()
    cfg.schedule = 1
    cfg.prec = 9
    assert _lib().dmrgx_run_dmrg_state(C.byref(cfg), C.byref(st), None, 0, C.byref(n), None) != 0
    assert "precision" in _err()


def test_short_chain_refused():
    assert _run("xyz", "dmrg2", "physical", L=3) != 0
    assert "L must be" in _err()


def _heff_create(rmap, lmap, n_psi=8, n_store=8, nblk=1, blk=16):
    off = np.arange(nblk, dtype=np.int64) * blk
    size = np.full(nblk, blk, dtype=np.int64)
    rm = np.ascontiguousarray(rmap, dtype=np.int32)
    lm = np.ascontiguousarray(lmap, dtype=np.int32)
    h = C.c_void_p()
    i64p, i32p = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    return _lib().dmrgx_heff_create(C.byref(h), 0, 0, None, off.ctypes.data_as(i64p), size.ctypes.data_as(i64p),
                                    nblk, rm.ctypes.data_as(i32p), len(rm) // 7, None, off.ctypes.data_as(i64p),
                                    size.ctypes.data_as(i64p), nblk, lm.ctypes.data_as(i32p), len(lm) // 7, n_psi,
                                    n_store)


GOOD = [0, 0, 0, 2, 4, 0, 1]  # block 0 (4 x 4), in 0, out 0, row 2, col 4


@pytest.mark.parametrize("bad,msg", [
    ([1, 0, 0, 2, 4, 0, 1], "block index"),    # block past the end
    ([-1, 0, 0, 2, 4, 0, 1], "block index"),   # negative block
    ([0, -4, 0, 2, 4, 0, 1], "negative"),      # negative input offset
    ([0, 0, -8, 2, 4, 0, 1], "negative"),      # negative output offset
    ([0, 0, 0, 0, 4, 0, 1], "empty"),          # empty row
    ([0, 0, 0, 2, 3, 0, 1], "divide"),         # col does not divide the block
    ([0, 4, 0, 2, 4, 0, 1], "past the end"),   # reads past psi (4 + 2*4 > 8)
])
def test_heff_create_rejects_bad_task_tuples(bad, msg):
    assert _heff_create(bad, GOOD) != 0
    assert msg in _err(), _err()
    assert _heff_create(GOOD, bad) != 0


def test_task_gemm_rejects_bad_tuples():
    off = np.zeros(1, dtype=np.int64)
    size = np.full(1, 16, dtype=np.int64)
    m = np.asarray([0, 0, -3, 2, 4, 0, 1], dtype=np.int32)
    i64p, i32p = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    rc = _lib().dmrgx_task_gemm(0, 0, None, off.ctypes.data_as(i64p), size.ctypes.data_as(i64p), 1,
                                m.ctypes.data_as(i32p), 1, None, None, 64)
    assert rc != 0 and "negative" in _err()


def test_batched_warm_start_refused():
    cfg = driver._config("xyz", 8, 16, 1, (), 0, "dmrg2", "fp64", "physical", True, 1, 1e-10, 1e-6, 8, False,
                         warm_start=True)
    p = np.zeros((2, 4))
    e, tr = np.zeros(2), np.zeros(2)
    st = driver.RunStats()
    n = C.c_int()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = _lib().dmrgx_run_batch(C.byref(cfg), 2, dp(p), dp(e), dp(tr), C.byref(st), None, 0, C.byref(n), None)
    assert rc != 0
    assert "warm" in _err()


def test_periodic_refusals():
    """potts2 has no edge terms in the reference; periodic chains are not batched."""
    cfg = driver._config("potts2", 8, 16, 1, (1.0, 1.0), 0, "dmrg2", "fp64", "physical", True, 1, 1e-10, 1e-6, 8,
                         False, pbc=True)
    st = driver.RunStats()
    n = C.c_int()
    assert _lib().dmrgx_run_dmrg_state(C.byref(cfg), C.byref(st), None, 0, C.byref(n), None) != 0
    assert "periodic" in _err()
    cfg = driver._config("xyz", 8, 16, 1, (), 0, "dmrg2", "fp64", "physical", True, 1, 1e-10, 1e-6, 8, False,
                         pbc=True)
    p = np.ones((2, 4))
    assert _lib().dmrgx_run_batch(C.byref(cfg), 2, p.ctypes.data_as(C.POINTER(C.c_double)), None, None,
                                  C.byref(st), None, 0, C.byref(n), None) != 0
    assert "periodic" in _err()
