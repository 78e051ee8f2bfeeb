"""dmrg_amd -- MI355X-native DMRG inner loop (two-site H_eff matvec, FP64
Lanczos, blockwise Jacobi SVD, environment updates) behind a C ABI.

The native library ``libdmrgx.so`` (built by ``python -m dmrg_amd.build``) is
loaded lazily through ctypes.  There is no CPU fallback: every call that
computes runs on the GPU, and loading fails loudly when the library or the
device is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libdmrgx.so"

R64, C64 = 0, 1
FP64, FP32, BF16X3 = 0, 1, 2
PREC_NAMES = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3}

# every extern "C" symbol declared in include/dmrgx.h and include/dmrgx_driver.h
EXPORTS = [
    "dmrgx_init", "dmrgx_finalize", "dmrgx_last_error", "dmrgx_device_info", "dmrgx_malloc",
    "dmrgx_free", "dmrgx_memcpy_h2d", "dmrgx_memcpy_d2h", "dmrgx_memset_zero", "dmrgx_synchronize",
    "dmrgx_stream", "dmrgx_task_gemm", "dmrgx_heff_create", "dmrgx_heff_apply", "dmrgx_heff_flops",
    "dmrgx_heff_destroy", "dmrgx_heff_adjoint", "dmrgx_heff_sizes", "dmrgx_eigsh", "dmrgx_eigsh_batched", "dmrgx_set_lanczos_graphs", "dmrgx_set_autotune", "dmrgx_autotune_keys", "dmrgx_block_svd", "dmrgx_block_svd_sweeps", "dmrgx_block_gemm", "dmrgx_leg_contract",
    "dmrgx_permute", "dmrgx_copy_boxes", "dmrgx_heff_time", "dmrgx_heff_time_adjoint", "dmrgx_heff_set_split", "dmrgx_heff_get_split",
    "dmrgx_task_gemm_noalloc", "dmrgx_set_skip_zero_slices", "dmrgx_heff_flops_exec", "dmrgx_calibrate_mfma",
    "dmrgx_heff_time_kernels",
]

_lib = None
TORCH_LOADED_FIRST = False


class DmrgxError(RuntimeError):
    pass


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libdmrgx.so without touching the GPU (safe on CPU-only hosts)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise DmrgxError(f"{p} is missing: build it with `python -m dmrg_amd.build`")
    lib = C.CDLL(str(p))
    _declare(lib)
    if path is None:
        _lib = lib
        # libdmrgx binds to whichever HIP runtime the process loaded first
        # (torch ships its own under the same soname); see dmrg_amd.autograd
        global TORCH_LOADED_FIRST
        TORCH_LOADED_FIRST = "torch" in sys.modules
    return lib


def _declare(lib):
    vp, i64p, i32p, dp = C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)
    sig = {
        "dmrgx_init": ([C.c_int], C.c_int),
        "dmrgx_finalize": ([], C.c_int),
        "dmrgx_last_error": ([], C.c_char_p),
        "dmrgx_device_info": ([C.POINTER(C.c_int), C.c_char_p, C.c_int], C.c_int),
        "dmrgx_malloc": ([C.c_size_t, C.POINTER(C.c_void_p)], C.c_int),
        "dmrgx_free": ([vp], C.c_int),
        "dmrgx_memcpy_h2d": ([vp, vp, C.c_size_t], C.c_int),
        "dmrgx_memcpy_d2h": ([vp, vp, C.c_size_t], C.c_int),
        "dmrgx_memset_zero": ([vp, C.c_size_t], C.c_int),
        "dmrgx_synchronize": ([], C.c_int),
        "dmrgx_stream": ([], vp),
        "dmrgx_task_gemm": ([C.c_int, C.c_int, vp, i64p, i64p, C.c_int, i32p, C.c_int, vp, vp, C.c_int64],
                            C.c_int),
        "dmrgx_task_gemm_noalloc": ([C.c_int, vp, i64p, i64p, C.c_int, i32p, C.c_int, vp, vp, C.c_int64],
                                    C.c_int),
        "dmrgx_heff_create": ([C.POINTER(vp), C.c_int, C.c_int, vp, i64p, i64p, C.c_int, i32p, C.c_int,
                               vp, i64p, i64p, C.c_int, i32p, C.c_int, C.c_int64, C.c_int64], C.c_int),
        "dmrgx_heff_apply": ([vp, vp, vp], C.c_int),
        "dmrgx_heff_flops": ([vp], C.c_double),
        "dmrgx_heff_flops_exec": ([vp], C.c_double),
        "dmrgx_set_skip_zero_slices": ([C.c_int], C.c_int),
        "dmrgx_heff_time": ([vp, vp, vp, C.c_int, dp], C.c_int),
        "dmrgx_heff_time_kernels": ([vp, vp, vp, C.c_int, dp], C.c_int),
        "dmrgx_heff_time_adjoint": ([vp, vp, vp, vp, vp, vp, C.c_int, dp], C.c_int),
        "dmrgx_heff_destroy": ([vp], C.c_int),
        "dmrgx_heff_adjoint": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "dmrgx_heff_sizes": ([vp, i64p], C.c_int),
        "dmrgx_eigsh": ([vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64, vp, dp, vp, vp, C.c_double,
                         C.POINTER(C.c_int), dp], C.c_int),
        "dmrgx_eigsh_batched": ([vp, C.c_int, i64p, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.POINTER(C.c_uint64), dp, vp, vp, C.c_double, C.POINTER(C.c_int), dp], C.c_int),
        "dmrgx_set_lanczos_graphs": ([C.c_int], C.c_int),
        "dmrgx_set_autotune": ([C.c_int], C.c_int),
        "dmrgx_autotune_keys": ([], C.c_int),
        "dmrgx_heff_set_split": ([vp, C.c_int, C.c_int], C.c_int),
        "dmrgx_heff_get_split": ([vp, C.POINTER(C.c_int)], C.c_int),
        "dmrgx_block_svd": ([C.c_int, C.c_int, i32p, i32p, vp, i64p, vp, i64p, vp, i64p, vp, i64p], C.c_int),
        "dmrgx_block_svd_sweeps": ([], C.c_int),
        "dmrgx_block_gemm": ([C.c_int, C.c_char, C.c_char, vp, vp, vp, i64p, C.c_int], C.c_int),
        "dmrgx_leg_contract": ([C.c_int, vp, vp, vp, i64p, C.c_int], C.c_int),
        "dmrgx_permute": ([C.c_int, vp, vp, i64p, C.c_int], C.c_int),
        "dmrgx_copy_boxes": ([C.c_int, vp, vp, i64p, C.c_int], C.c_int),
        "dmrgx_calibrate_mfma": ([C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    from . import driver as _driver
    _driver.declare(lib)


def lib() -> C.CDLL:
    return load_library()


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().dmrgx_last_error().decode(errors="replace")
        raise DmrgxError(f"{what} failed (rc={rc}): {msg}")


_initialised = False


def init(device: int = 0):
    global _initialised
    if not _initialised:
        check(lib().dmrgx_init(device), "dmrgx_init")
        _initialised = True


class DeviceArray:
    """A device allocation holding a numpy-typed vector (complex interleaved)."""

    def __init__(self, n: int, dtype=np.float64):
        init()
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(self.n * self.dtype.itemsize, 8)
        p = C.c_void_p()
        check(lib().dmrgx_malloc(self.nbytes, C.byref(p)), "dmrgx_malloc")
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib().dmrgx_memcpy_h2d(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes), "h2d")
        return d

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            check(lib().dmrgx_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes), "d2h")
        return out

    def zero(self):
        check(lib().dmrgx_memset_zero(self.ptr, self.nbytes), "memset")

    def free(self):
        if self.ptr:
            lib().dmrgx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _i64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
    return a, a.ctypes.data_as(C.POINTER(C.c_int64))


def _i32(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


class Heff:
    """dmrgx_heff plan: the two get_map passes of av() on device."""

    def __init__(self, dtype, prec, renv: DeviceArray, rblk_off, rblk_size, rmap,
                 lenv: DeviceArray, lblk_off, lblk_size, lmap, n_psi, n_store):
        init()
        self._keep = [renv, lenv]
        ro, rop = _i64(rblk_off)
        rs, rsp = _i64(rblk_size)
        lo, lop = _i64(lblk_off)
        ls, lsp = _i64(lblk_size)
        rm, rmp = _i32(np.asarray(rmap, dtype=np.int32).reshape(-1, 7))
        lm, lmp = _i32(np.asarray(lmap, dtype=np.int32).reshape(-1, 7))
        h = C.c_void_p()
        check(lib().dmrgx_heff_create(C.byref(h), dtype, prec, renv.ptr, rop, rsp, len(ro), rmp, len(rm),
                                      lenv.ptr, lop, lsp, len(lo), lmp, len(lm), int(n_psi), int(n_store)),
              "dmrgx_heff_create")
        self.h = h
        self.n_psi = int(n_psi)
        self.dtype = dtype

    @property
    def flops(self):
        return lib().dmrgx_heff_flops(self.h)

    def apply(self, x: DeviceArray, y: DeviceArray):
        check(lib().dmrgx_heff_apply(self.h, x.ptr, y.ptr), "dmrgx_heff_apply")

    def adjoint(self, psi: DeviceArray, g_out: DeviceArray, g_psi=None, g_renv=None, g_lenv=None):
        """Cotangents of psi and of the two environments for a cotangent of the output."""
        ptr = lambda a: a.ptr if a is not None else None
        check(lib().dmrgx_heff_adjoint(self.h, psi.ptr, g_out.ptr, ptr(g_psi), ptr(g_renv), ptr(g_lenv)),
              "dmrgx_heff_adjoint")

    @property
    def flops_exec(self):
        return lib().dmrgx_heff_flops_exec(self.h)

    def time(self, x: DeviceArray, y: DeviceArray, iters: int):
        """(ms per apply, ms R-pass kernel, ms L-pass kernel, R flops, L flops,
        R executed flops, L executed flops)"""
        ms = np.zeros(7)
        check(lib().dmrgx_heff_time(self.h, x.ptr, y.ptr, int(iters), ms.ctypes.data_as(C.POINTER(C.c_double))),
              "dmrgx_heff_time")
        return ms

    def time_kernels(self, x: DeviceArray, y: DeviceArray, iters: int):
        """(ms per launch of the R-pass GEMM kernel, of the L-pass kernel),
        each launched `iters` times back to back (dmrgx_heff_time_kernels)"""
        ms = np.zeros(2)
        check(lib().dmrgx_heff_time_kernels(self.h, x.ptr, y.ptr, int(iters),
                                            ms.ctypes.data_as(C.POINTER(C.c_double))), "dmrgx_heff_time_kernels")
        return ms

    def eigsh(self, nev=6, ncv=None, tol=1e-10, max_restart=1000, seed=1234, v0=None, refine=None,
              refine_tol=1e-10):
        """dmrgx_eigsh: (evals, lowest vector, info[6], FP64 residual of the returned pair)."""
        if ncv is None:
            ncv = (6 if self.dtype == C64 else 4) * nev
        npd = np.complex128 if self.dtype == C64 else np.float64
        vec = DeviceArray(self.n_psi, npd)
        ev = np.zeros(max(nev, 1))
        info = (C.c_int * 6)()
        res = np.zeros(1)
        check(lib().dmrgx_eigsh(self.h, nev, ncv, tol, max_restart, seed, v0.ptr if v0 is not None else None,
                                ev.ctypes.data_as(C.POINTER(C.c_double)), vec.ptr,
                                refine.h if refine is not None else None, refine_tol, info,
                                res.ctypes.data_as(C.POINTER(C.c_double))), "dmrgx_eigsh")
        return ev, vec, list(info), float(res[0])

    def eigsh_batched(self, lens, nev=6, ncv=None, tol=1e-10, max_restart=1000, seeds=None, refine=None,
                      refine_tol=1e-10):
        """dmrgx_eigsh_batched over chains of lengths `lens` concatenated in
        this plan: (evals n_chains x nev, the concatenated vectors, info
        n_chains x 6, FP64 residuals n_chains)."""
        if ncv is None:
            ncv = (6 if self.dtype == C64 else 4) * nev
        n = len(lens)
        ln = np.asarray(lens, dtype=np.int64)
        sd = np.asarray(seeds if seeds is not None else [1234] * n, dtype=np.uint64)
        npd = np.complex128 if self.dtype == C64 else np.float64
        vec = DeviceArray(self.n_psi, npd)
        ev = np.zeros((n, max(nev, 1)))
        info = np.zeros((n, 6), dtype=np.int32)
        res = np.zeros(n)
        check(lib().dmrgx_eigsh_batched(self.h, n, ln.ctypes.data_as(C.POINTER(C.c_int64)), nev, ncv, tol,
                                        max_restart, sd.ctypes.data_as(C.POINTER(C.c_uint64)),
                                        ev.ctypes.data_as(C.POINTER(C.c_double)), vec.ptr,
                                        refine.h if refine is not None else None, refine_tol,
                                        info.ctypes.data_as(C.POINTER(C.c_int)),
                                        res.ctypes.data_as(C.POINTER(C.c_double))), "dmrgx_eigsh_batched")
        return ev, vec, info, res

    def close(self):
        if self.h:
            lib().dmrgx_heff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
