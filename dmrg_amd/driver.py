"""ctypes bindings of include/dmrgx_driver.h: full DMRG runs on the GPU and the
structure-only bookkeeping entry points (get_map, generate_dim_sym)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

MODELS = {"xyz": 0, "kitaev": 1, "potts3": 2, "potts2": 3, "4e": 4}
COMPLEX_MODELS = ("xyz", "potts3", "4e")
STEP_LOG_WIDTH = 19
NPARAM = 8  # DMRGX_NPARAM
DRIVER_EXPORTS = ["dmrgx_run_dmrg", "dmrgx_synth_heff_create", "dmrgx_get_map", "dmrgx_generate_dim_sym",
                  "dmrgx_run_dmrg_state", "dmrgx_state_norm", "dmrgx_state_overlap", "dmrgx_state_expect",
                  "dmrgx_state_destroy", "dmrgx_state_derivative", "dmrgx_run_batch", "dmrgx_parse_args",
                  "dmrgx_state_heff_create", "dmrgx_state_position", "dmrgx_state_move_cut",
                  "dmrgx_state_inverse", "dmrgx_state_correlations_4e"]
BATCH_LOG_WIDTH = 11


class RunConfig(C.Structure):
    _fields_ = [("model", C.c_int), ("p", C.c_double * NPARAM), ("L", C.c_int), ("m", C.c_int), ("sweeps", C.c_int),
                ("sector", C.c_int), ("schedule", C.c_int), ("prec", C.c_int), ("sign_rule", C.c_int),
                ("ladder", C.c_int), ("seed", C.c_uint64), ("tol", C.c_double), ("tol_lo", C.c_double),
                ("refine_k", C.c_int), ("verbose", C.c_int), ("nev_lo", C.c_int), ("output", C.c_char_p),
                ("warm_start", C.c_int), ("pbc", C.c_int)]


class RunStats(C.Structure):
    _fields_ = [("energy", C.c_double), ("seconds", C.c_double), ("sweep_seconds", C.c_double),
                ("last_sweep_seconds", C.c_double), ("steps", C.c_int), ("matvecs", C.c_longlong),
                ("polish_matvecs", C.c_longlong), ("matvec_flops", C.c_double), ("trunc_err", C.c_double),
                ("entropy", C.c_double), ("t_eig", C.c_double), ("t_svd", C.c_double), ("t_env", C.c_double)]


def declare(lib):
    i32p = C.POINTER(C.c_int32)
    lib.dmrgx_run_dmrg.argtypes = [C.POINTER(RunConfig), C.POINTER(RunStats), C.POINTER(C.c_double), C.c_int,
                                   C.POINTER(C.c_int)]
    lib.dmrgx_run_dmrg.restype = C.c_int
    lib.dmrgx_synth_heff_create.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int,
                                            C.c_uint64, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]
    lib.dmrgx_synth_heff_create.restype = C.c_int
    lib.dmrgx_get_map.argtypes = [C.c_int, i32p, i32p, C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_int,
                                  i32p, i32p, C.c_int, C.POINTER(C.c_int), i32p, C.c_int]
    lib.dmrgx_get_map.restype = C.c_int
    lib.dmrgx_generate_dim_sym.argtypes = [C.c_int, i32p, C.c_int, i32p, C.c_int, C.c_int, i32p, C.c_int]
    lib.dmrgx_generate_dim_sym.restype = C.c_int
    dp = C.POINTER(C.c_double)
    lib.dmrgx_run_dmrg_state.argtypes = [C.POINTER(RunConfig), C.POINTER(RunStats), dp, C.c_int,
                                         C.POINTER(C.c_int), C.POINTER(C.c_void_p)]
    lib.dmrgx_run_dmrg_state.restype = C.c_int
    lib.dmrgx_state_norm.argtypes = [C.c_void_p, dp]
    lib.dmrgx_state_norm.restype = C.c_int
    lib.dmrgx_state_overlap.argtypes = [C.c_void_p, C.c_void_p, dp]
    lib.dmrgx_state_overlap.restype = C.c_int
    lib.dmrgx_state_expect.argtypes = [C.c_void_p, C.c_int, dp, dp]
    lib.dmrgx_state_expect.restype = C.c_int
    lib.dmrgx_state_derivative.argtypes = [C.c_void_p, dp, dp, C.c_int, dp]
    lib.dmrgx_state_derivative.restype = C.c_int
    lib.dmrgx_state_heff_create.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_void_p)]
    lib.dmrgx_state_heff_create.restype = C.c_int
    lib.dmrgx_run_batch.argtypes = [C.POINTER(RunConfig), C.c_int, dp, dp, dp, C.POINTER(RunStats), dp, C.c_int,
                                    C.POINTER(C.c_int), dp]
    lib.dmrgx_run_batch.restype = C.c_int
    lib.dmrgx_parse_args.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_char_p),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     dp, C.c_char_p, C.c_int]
    lib.dmrgx_parse_args.restype = C.c_int
    lib.dmrgx_state_position.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    lib.dmrgx_state_position.restype = C.c_int
    lib.dmrgx_state_move_cut.argtypes = [C.c_void_p, C.c_int]
    lib.dmrgx_state_move_cut.restype = C.c_int
    lib.dmrgx_state_inverse.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    lib.dmrgx_state_inverse.restype = C.c_int
    lib.dmrgx_state_correlations_4e.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int)]
    lib.dmrgx_state_correlations_4e.restype = C.c_int
    lib.dmrgx_state_destroy.argtypes = [C.c_void_p]
    lib.dmrgx_state_destroy.restype = C.c_int
    return lib


class State:
    """A finished run's MPS on the device (centre absorbed) and the
    reference's calc functions on it."""

    def __init__(self, handle, model):
        self.h = handle
        self.model = model

    def _call(self, fn, *args):
        from . import check
        out = np.zeros(2)
        check(fn(*args, out.ctypes.data_as(C.POINTER(C.c_double))), fn.__name__)
        return complex(out[0], out[1])

    def norm2(self) -> complex:
        """calc(mps) = <a|a>"""
        from . import lib
        return self._call(lib().dmrgx_state_norm, self.h)

    def overlap(self, other: "State") -> complex:
        """calc(mps, mps) = <other|self>"""
        from . import lib
        return self._call(lib().dmrgx_state_overlap, self.h, other.h)

    def expect(self, params, model=None) -> complex:
        """calc(mps, mpo) = <a|H(model, params)|a>"""
        from . import lib
        p = (C.c_double * NPARAM)(*([float(x) for x in params] + [0.0] * (NPARAM - len(params))))
        return self._call(lib().dmrgx_state_expect, self.h, MODELS[model or self.model], p)

    def derivative(self, params, dparams, prec="fp64"):
        """(dE0/dt along dparams at params, E0 of the centre) from the adjoint
        H_eff kernels with recomputed environment tangents."""
        from . import PREC_NAMES, check, lib
        pad = lambda v: (C.c_double * NPARAM)(*([float(x) for x in v] + [0.0] * (NPARAM - len(v))))
        out = np.zeros(2)
        check(lib().dmrgx_state_derivative(self.h, pad(params), pad(dparams), PREC_NAMES[prec],
                                           out.ctypes.data_as(C.POINTER(C.c_double))), "dmrgx_state_derivative")
        return float(out[0]), float(out[1])

    def centre_heff(self, prec="fp64"):
        """The H_eff plan at the centre bond of this state (environments
        recomputed from the MPS); its .psi_ptr is the two-site state."""
        from . import C64, PREC_NAMES, R64, check, lib
        h, n, psi = C.c_void_p(), C.c_int64(), C.c_void_p()
        check(lib().dmrgx_state_heff_create(self.h, PREC_NAMES[prec], C.byref(h), C.byref(n), C.byref(psi)),
              "dmrgx_state_heff_create")
        dt = C64 if self.model in COMPLEX_MODELS else R64
        return SynthHeff.from_handle(h, n.value, dt, psi.value)

    def position(self) -> int:
        """mps::position: the bond that holds the singular values"""
        from . import check, lib
        c = C.c_int()
        check(lib().dmrgx_state_position(self.h, C.byref(c)), "dmrgx_state_position")
        return c.value

    def move_cut(self, steps: int):
        """mps::move_cut_right (steps > 0) or move_cut_left, |steps| times"""
        from . import check, lib
        check(lib().dmrgx_state_move_cut(self.h, int(steps)), "dmrgx_state_move_cut")

    def inverse(self) -> "State":
        """mps::inverse: the chain read from the other end, as a new state"""
        from . import check, lib
        h = C.c_void_p()
        check(lib().dmrgx_state_inverse(self.h, C.byref(h)), "dmrgx_state_inverse")
        return State(h, self.model)

    def correlations_4e(self) -> np.ndarray:
        """The 4e main's data-file lines (x, re, im), dmrgx_state_correlations_4e"""
        from . import check, lib
        rows = np.zeros((4096, 3))
        n = C.c_int()
        check(lib().dmrgx_state_correlations_4e(self.h, rows.ctypes.data_as(C.POINTER(C.c_double)), len(rows),
                                                C.byref(n)), "dmrgx_state_correlations_4e")
        return rows[:n.value].copy()

    def fidelity(self, other: "State") -> float:
        """|<b|a>| / sqrt(<a|a> <b|b>)"""
        return abs(self.overlap(other)) / np.sqrt(abs(self.norm2()) * abs(other.norm2()))

    def close(self):
        from . import lib
        if self.h:
            lib().dmrgx_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class RunResult:
    energy: float
    stats: dict
    steps: np.ndarray  # (n_steps, STEP_LOG_WIDTH)
    state: State | None = None


def _config(model, L, m, sweeps, params, sector, schedule, prec, sign_rule, ladder, seed, tol, tol_lo, refine_k,
            verbose, nev_lo=0, output=None, warm_start=False, pbc=False):
    from . import PREC_NAMES
    cfg = RunConfig()
    cfg.model = MODELS[model]
    for i, v in enumerate(list(params)[:NPARAM]):
        cfg.p[i] = float(v)
    cfg.L, cfg.m, cfg.sweeps, cfg.sector = int(L), int(m), int(sweeps), int(sector)
    cfg.schedule = 1 if schedule == "dmrg2" else 0
    cfg.prec = PREC_NAMES[prec]
    cfg.sign_rule = 0 if sign_rule == "physical" else 1
    cfg.ladder = 1 if ladder else 0
    cfg.seed = seed
    cfg.tol, cfg.tol_lo, cfg.refine_k = tol, tol_lo, refine_k
    cfg.verbose = 2 if verbose == "stderr" else (1 if verbose else 0)
    cfg.nev_lo = int(nev_lo)
    cfg.output = output.encode() if output else None
    cfg.warm_start = 1 if warm_start else 0
    cfg.pbc = 1 if pbc else 0
    return cfg


def run_dmrg(model: str, L: int, m: int, sweeps: int, params=(), sector=0, schedule="dmrg2", prec="fp64",
             sign_rule="physical", ladder=True, seed=1234, tol=1e-10, tol_lo=1e-6, refine_k=8,
             verbose=False, keep_state=False, nev_lo=0, output=None, warm_start=False, pbc=False) -> RunResult:
    from . import check, init, lib
    init()
    cfg = _config(model, L, m, sweeps, params, sector, schedule, prec, sign_rule, ladder, seed, tol, tol_lo,
                  refine_k, verbose, nev_lo, output, warm_start, pbc)
    st = RunStats()
    cap = 20000
    log = np.zeros((cap, STEP_LOG_WIDTH))
    n = C.c_int()
    h = C.c_void_p()
    check(lib().dmrgx_run_dmrg_state(C.byref(cfg), C.byref(st), log.ctypes.data_as(C.POINTER(C.c_double)), cap,
                                     C.byref(n), C.byref(h) if keep_state else None), "dmrgx_run_dmrg")
    stats = {k: getattr(st, k) for k, _ in RunStats._fields_}
    state = State(h, model) if keep_state else None
    return RunResult(st.energy, stats, log[:n.value].copy(), state)


def parse_args(argv, names):
    """dmrgx_parse_args (the reference's set_para_val): (L, m, sweeps, sector,
    parameter values, output file suffix)."""
    from . import check, lib
    a = (C.c_char_p * max(len(argv), 1))(*[x.encode() for x in argv])
    nm = (C.c_char_p * max(len(names), 1))(*[x.encode() for x in names])
    L, m, n, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    vals = np.zeros(max(len(names), 1))
    buf = C.create_string_buffer(4096)
    check(lib().dmrgx_parse_args(len(argv), a, len(names), nm, C.byref(L), C.byref(m), C.byref(n), C.byref(s),
                                 vals.ctypes.data_as(C.POINTER(C.c_double)), buf, 4096), "dmrgx_parse_args")
    return L.value, m.value, n.value, s.value, list(vals[:len(names)]), buf.value.decode()


def splitmix64_stream(seed: int, n: int) -> np.ndarray:
    """n successive splitmix64 outputs of the generator state `seed`
    (state += 0x9E3779B97F4A7C15, then the xor-shift-multiply finaliser)."""
    M = (1 << 64) - 1
    out = np.zeros(n, dtype=np.uint64)
    x = seed & M
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out[i] = z ^ (z >> 31)
    return out


def batch_couplings(n_chains: int, seed: int) -> np.ndarray:
    """(Jx, Jy, Jz) of n_chains XYZ chains, i.i.d. U[0.5, 1.5): 0.5 + (x >> 11) 2^-53
    over one splitmix64 stream, chain-major."""
    x = splitmix64_stream(seed, 3 * n_chains)
    return (0.5 + (x >> np.uint64(11)).astype(np.float64) * 2.0 ** -53).reshape(n_chains, 3)


@dataclass
class BatchResult:
    energies: np.ndarray  # last E0 of each chain
    trunc: np.ndarray     # last truncation error of each chain
    stats: dict
    steps: np.ndarray     # (n_steps, BATCH_LOG_WIDTH)
    chain_steps: np.ndarray  # (n_chains, n_chain_steps, STEP_LOG_WIDTH)


def run_batch(model: str, L: int, m: int, sweeps: int, params, sector=0, prec="fp64", sign_rule="physical",
              ladder=True, seed=1234, tol=1e-10, tol_lo=1e-6, refine_k=8, verbose=False, nev_lo=0) -> BatchResult:
    """dmrgx_run_batch: chains with couplings params[i] swept in lockstep."""
    from . import check, init, lib
    init()
    p = np.zeros((len(params), 4))
    for i, q in enumerate(params):
        p[i, :len(q)] = q
    cfg = _config(model, L, m, sweeps, (), sector, "dmrg2", prec, sign_rule, ladder, seed, tol, tol_lo, refine_k,
                  verbose, nev_lo)
    n = len(params)
    e, tr = np.zeros(n), np.zeros(n)
    st = RunStats()
    cap = 2 * L * (sweeps + m // 100 + 3)  # above the schedule's step count
    log = np.zeros((cap, BATCH_LOG_WIDTH))
    clog = np.zeros((n, cap, STEP_LOG_WIDTH))
    ns = C.c_int()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    check(lib().dmrgx_run_batch(C.byref(cfg), n, dp(p), dp(e), dp(tr), C.byref(st), dp(log), cap, C.byref(ns),
                                dp(clog)), "dmrgx_run_batch")
    stats = {k: getattr(st, k) for k, _ in RunStats._fields_}
    k = min(ns.value, cap)
    return BatchResult(e, tr, stats, log[:k].copy(), clog[:, :k].copy())


def encode_desc(dim, sym, dirs=None):
    v = [len(dim)] + [len(d) for d in dim]
    for d in dim:
        v += list(d)
    v.append(len(sym))
    for s in sym:
        v += list(s)
    if dirs is not None:
        v += list(dirs)
    return np.asarray(v, dtype=np.int32)


def decode_desc(a):
    a = list(a)
    p = 0
    nleg = a[p]
    p += 1
    ns = a[p:p + nleg]
    p += nleg
    dim = []
    for k in ns:
        dim.append(a[p:p + k])
        p += k
    nb = a[p]
    p += 1
    sym = [a[p + b * nleg: p + (b + 1) * nleg] for b in range(nb)]
    return dim, sym


def get_map(symmetry, tdim, tsym, tdir, dim, sym, ta="N", tb="T", num=1, beta=0, back=False, sym_ret=None,
            fermion=True):
    """Structure-only qtensor::get_map through the library (integer bookkeeping)."""
    from . import check, lib
    td = encode_desc(tdim, tsym, tdir)
    ad = encode_desc(dim, sym)
    nleg = len(tdim) + len(dim) - 2 * num
    sr = np.zeros(1, dtype=np.int32)
    if back:
        sr = np.asarray([len(sym_ret)] + [x for s in sym_ret for x in s], dtype=np.int32)
    cap = max(16, 4 * len(tsym) * len(sym) + 16)
    out = np.zeros(cap * 7, dtype=np.int32)
    rcap = 1 << 16
    ret = np.zeros(rcap, dtype=np.int32)
    n = C.c_int()
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
    check(lib().dmrgx_get_map(symmetry, P(td), P(ad), ta.encode(), tb.encode(), num, int(beta), 1 if back else 0,
                              1 if fermion else 0, P(sr), P(out), cap, C.byref(n), P(ret), rcap), "dmrgx_get_map")
    maps = out[:n.value * 7].reshape(-1, 7).tolist()
    d, s = decode_desc(ret)
    assert len(d) == nleg
    return maps, d, s


def generate_dim_sym(symmetry, ldim, lsym, ldir, lnum, rdim, rsym, rdir, rnum, sector):
    from . import check, lib
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
    ld = encode_desc(ldim, lsym, ldir)
    rd = encode_desc(rdim, rsym, rdir)
    out = np.zeros(1 << 16, dtype=np.int32)
    check(lib().dmrgx_generate_dim_sym(symmetry, P(ld), lnum, P(rd), rnum, sector, P(out), len(out)),
          "dmrgx_generate_dim_sym")
    return decode_desc(out)


class SynthHeff:
    """A config-shaped synthetic two-site operator built by the library."""

    @classmethod
    def from_handle(cls, h, n_psi, dtype, psi_ptr=None):
        """Wrap an existing plan handle (e.g. State.centre_heff)."""
        o = cls.__new__(cls)
        o.h, o.n_psi, o.dtype, o.psi_ptr = h, n_psi, dtype, psi_ptr
        return o

    @property
    def flops_exec(self):
        from . import lib
        return lib().dmrgx_heff_flops_exec(self.h)

    def __init__(self, S, D, bond, complex_=True, prec="fp64", seed=20261203, fermion=False):
        from . import C64, PREC_NAMES, R64, check, init, lib
        init()
        b = (C.c_int * len(bond))(*bond)
        h = C.c_void_p()
        n = C.c_int64()
        check(lib().dmrgx_synth_heff_create(S, D, b, len(bond), C64 if complex_ else R64, PREC_NAMES[prec], seed,
                                            1 if fermion else 0, C.byref(h), C.byref(n)), "dmrgx_synth_heff_create")
        self.h = h
        self.n_psi = n.value
        self.dtype = C64 if complex_ else R64

    @property
    def flops(self):
        from . import lib
        return lib().dmrgx_heff_flops(self.h)

    def time(self, x, y, iters):
        """(ms per apply, ms R-pass kernel, ms L-pass kernel, R flops, L flops,
        R executed flops, L executed flops)"""
        from . import check, lib
        ms = np.zeros(7)
        check(lib().dmrgx_heff_time(self.h, x.ptr, y.ptr, int(iters), ms.ctypes.data_as(C.POINTER(C.c_double))),
              "dmrgx_heff_time")
        return ms

    def time_kernels(self, x, y, iters):
        """(ms per launch of the R-pass GEMM kernel, of the L-pass kernel),
        each launched `iters` times back to back (dmrgx_heff_time_kernels)"""
        from . import check, lib
        ms = np.zeros(2)
        check(lib().dmrgx_heff_time_kernels(self.h, x.ptr, y.ptr, int(iters),
                                            ms.ctypes.data_as(C.POINTER(C.c_double))), "dmrgx_heff_time_kernels")
        return ms

    def apply(self, x, y):
        from . import check, lib
        check(lib().dmrgx_heff_apply(self.h, x.ptr, y.ptr), "dmrgx_heff_apply")

    def set_split(self, r_target, l_target):
        """Pin the work decomposition of both passes (dmrgx_heff_set_split)."""
        from . import check, lib
        check(lib().dmrgx_heff_set_split(self.h, int(r_target), int(l_target)), "dmrgx_heff_set_split")

    def split(self):
        """(R target, L target, R work items, L work items)"""
        from . import check, lib
        out = (C.c_int * 4)()
        check(lib().dmrgx_heff_get_split(self.h, out), "dmrgx_heff_get_split")
        return tuple(out)

    def sizes(self):
        """(n_psi, n_store, renv length, lenv length)"""
        from . import check, lib
        out = np.zeros(4, np.int64)
        check(lib().dmrgx_heff_sizes(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "dmrgx_heff_sizes")
        return tuple(int(v) for v in out)

    def adjoint_synth(self, psi, g_out, g_psi):
        """Cotangents of psi and of both (plan-owned) environments; the
        environment cotangent buffers are kept between calls."""
        from . import DeviceArray, check, lib
        npd = np.complex128 if self.dtype != 0 else np.float64
        if getattr(self, "_genv", None) is None:
            _, _, rl, ll = self.sizes()
            self._genv = (DeviceArray(rl, npd), DeviceArray(ll, npd))
        gr, gl = self._genv
        check(lib().dmrgx_heff_adjoint(self.h, psi.ptr, g_out.ptr, g_psi.ptr, gr.ptr, gl.ptr), "dmrgx_heff_adjoint")

    def time_adjoint(self, psi, g_out, g_psi, iters):
        """HIP-event milliseconds per adjoint_synth call (dmrgx_heff_time_adjoint)."""
        from . import check, lib
        self.adjoint_synth(psi, g_out, g_psi)  # allocates the environment cotangents
        gr, gl = self._genv
        ms = np.zeros(1)
        check(lib().dmrgx_heff_time_adjoint(self.h, psi.ptr, g_out.ptr, g_psi.ptr, gr.ptr, gl.ptr, int(iters),
                                            ms.ctypes.data_as(C.POINTER(C.c_double))), "dmrgx_heff_time_adjoint")
        return float(ms[0])

    def close(self):
        from . import lib
        if getattr(self, "_genv", None) is not None:
            for b in self._genv:
                b.free()
            self._genv = None
        if self.h:
            lib().dmrgx_heff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
