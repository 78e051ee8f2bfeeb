"""dmrgx_task_gemm_noalloc (the task-list GEMM that allocates nothing) against
dmrgx_task_gemm with FP64 arithmetic, through the C ABI: two calls in a row,
the second on a smaller task list, each compared with the allocating entry
point on the same inputs (FP64: 1e-12 of the largest output)."""
import ctypes as C

import numpy as np
import pytest

from tests.cases import synthetic_case

pytestmark = pytest.mark.gpu


def _call(gpu, fn, case, dt, x, y0):
    A = gpu.DeviceArray.from_numpy(case.rflat)
    xd = gpu.DeviceArray.from_numpy(x)
    yd = gpu.DeviceArray.from_numpy(y0)
    m = np.ascontiguousarray(np.asarray(case.rmap, dtype=np.int32))
    i64p, i32p = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    args = [A.ptr, case.roff.ctypes.data_as(i64p), case.rsize.ctypes.data_as(i64p), len(case.roff),
            m.ctypes.data_as(i32p), len(m), xd.ptr, yd.ptr, case.n_store]
    lib = gpu.lib()
    if fn == "alloc":
        gpu.check(lib.dmrgx_task_gemm(dt, gpu.FP64, *args), "dmrgx_task_gemm")
    else:
        gpu.check(lib.dmrgx_task_gemm_noalloc(dt, *args), "dmrgx_task_gemm_noalloc")
    return yd.numpy()


@pytest.mark.parametrize("dtype", [np.complex128, np.float64])
def test_task_gemm_noalloc_matches_task_gemm(gpu, dtype):
    dt = gpu.C64 if np.dtype(dtype).kind == "c" else gpu.R64
    # a large task list (split-K partial tiles in the scratch), then a smaller one
    for S, D, m, uneven in ((2, 5, 256, None), (3, 4, 60, [27, 20, 13])):
        print(f"case task_gemm_noalloc S={S} D={D} m={m} {np.dtype(dtype).name}", flush=True)
        case = synthetic_case(S, D, m, dtype=dtype, uneven=uneven)
        rng = np.random.default_rng(5)
        x = rng.standard_normal(case.n_psi)
        y0 = rng.standard_normal(case.n_store)  # outputs without a task stay as they are
        if dt == gpu.C64:
            x = x + 1j * rng.standard_normal(case.n_psi)
            y0 = y0 + 1j * rng.standard_normal(case.n_store)
        x, y0 = x.astype(dtype), y0.astype(dtype)
        ref = _call(gpu, "alloc", case, dt, x, y0)
        got = _call(gpu, "noalloc", case, dt, x, y0)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"  rel err {err:.2e}", flush=True)
        assert err < 1e-12, err
        # against the oracle's restatement of the same pass as well
        want = y0.copy()
        case.renv.contract_map(want, x, case.rmap, "N", "T", 2)
        assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-12
