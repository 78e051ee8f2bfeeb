"""GPU parity of the task-list GEMM and the H_eff matvec against the oracle's
av() (src/core.cpp:22-31 restated in oracle/dmrg.py)."""
import numpy as np
import pytest

from tests.cases import dmrg_case, synthetic_case

pytestmark = pytest.mark.gpu

# Tolerances (max |y - y_oracle| / max |y_oracle|):
TOL = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 3e-5}


def _run(gpu, case, prec):
    c = gpu.C64 if np.dtype(case.dtype).kind == "c" else gpu.R64
    R = gpu.DeviceArray.from_numpy(case.rflat)
    Lb = gpu.DeviceArray.from_numpy(case.lflat)
    h = gpu.Heff(c, gpu.PREC_NAMES[prec], R, case.roff, case.rsize, case.rmap, Lb, case.loff, case.lsize,
                 case.lmap, case.n_psi, case.n_store)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(case.n_psi)
    if c == gpu.C64:
        x = x + 1j * rng.standard_normal(case.n_psi)
    xd = gpu.DeviceArray.from_numpy(x.astype(case.dtype))
    yd = gpu.DeviceArray(case.n_psi, case.dtype)
    h.apply(xd, yd)
    y = yd.numpy()
    ref = case.oracle_apply(x)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert abs(h.flops - case.flops()) < 1e-6 * case.flops()
    return err, h


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("shape", [
    ("z2_small_real", 2, 4, 16, np.float64, None),
    ("z2_m64_complex", 2, 5, 64, np.complex128, None),
    ("z3_uneven_complex", 3, 4, 60, np.complex128, [27, 20, 13]),
    ("z2_ragged_real", 2, 5, 70, np.float64, [37, 33]),
    ("z3_m200_uneven_complex", 3, 4, 200, np.complex128, [81, 67, 52]),
])
def test_heff_synthetic(gpu, prec, shape):
    name, S, D, m, dt, uneven = shape
    print(f"case {name} prec {prec}", flush=True)
    case = synthetic_case(S, D, m, dtype=dt, uneven=uneven)
    err, _ = _run(gpu, case, prec)
    assert err < TOL[prec], f"{name} {prec}: rel err {err:.3e}"


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
@pytest.mark.parametrize("sign_rule", ["physical", "reference"])
def test_heff_real_envs_xyz(gpu, sign_rule, prec):
    print(f"case xyz L=8 m=16 sign={sign_rule} prec={prec}", flush=True)
    case, _ = dmrg_case("xyz", 8, 16, 1, sign_rule=sign_rule, jx=1.0, jy=0.8, jz=0.5)
    err, _ = _run(gpu, case, prec)
    assert err < TOL[prec]


def test_heff_real_envs_kitaev(gpu):
    print("case kitaev L=10 m=16", flush=True)
    case, _ = dmrg_case("kitaev", 10, 16, 1, t=-1.0, p=-1.0, u=0.25)
    assert any(t[6] == -1 for t in case.rmap), "fermion signs must appear in the Kitaev rmap"
    err, _ = _run(gpu, case, "fp64")
    assert err < TOL["fp64"]


def test_heff_xyz_m512_shape(gpu):
    """XYZ m=512 two-site shapes: 32+32 tasks of (1280,256,256)/(256,256,1280)."""
    print("case z2 m=512 D=5 complex", flush=True)
    case = synthetic_case(2, 5, 512, dtype=np.complex128)
    assert len(case.rmap) == 32 and len(case.lmap) == 32
    err, h = _run(gpu, case, "fp64")
    assert err < TOL["fp64"]
    assert abs(h.flops - 4.294967296e10) < 1e3


def test_task_gemm_direct(gpu):
    """dmrgx_task_gemm on one pass equals the oracle's contract_map."""
    import ctypes as C
    case = synthetic_case(2, 4, 32, dtype=np.complex128, seed=3)
    rng = np.random.default_rng(11)
    x = (rng.standard_normal(case.n_psi) + 1j * rng.standard_normal(case.n_psi))
    ref = np.zeros(case.n_store, dtype=np.complex128)
    case.renv.contract_map(ref, x, case.rmap, "N", "T", 2)
    A = gpu.DeviceArray.from_numpy(case.rflat)
    xd = gpu.DeviceArray.from_numpy(x)
    yd = gpu.DeviceArray(case.n_store, np.complex128)
    yd.zero()
    m = np.ascontiguousarray(np.asarray(case.rmap, dtype=np.int32))
    lib = gpu.lib()
    gpu.check(lib.dmrgx_task_gemm(gpu.C64, gpu.FP64, A.ptr, case.roff.ctypes.data_as(C.POINTER(C.c_int64)),
                                  case.rsize.ctypes.data_as(C.POINTER(C.c_int64)), len(case.roff),
                                  m.ctypes.data_as(C.POINTER(C.c_int32)), len(m), xd.ptr, yd.ptr, case.n_store))
    y = yd.numpy()
    assert np.max(np.abs(y - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("model,kw,L,m,ratio", [
    ("xyz", dict(jx=1.0, jy=0.8, jz=0.5), 12, 32, 0.30),
    ("potts3", dict(f=1.0, j=1.0, p=0.0, t=3.14159265359 / 8), 8, 54, 1 / 6),
    ("kitaev", dict(t=-1.0, p=-1.0, u=0.25), 12, 32, 0.25),
])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_heff_skips_zero_slices(gpu, model, kw, L, m, ratio, prec):
    """Real environments carry structurally zero MPO-bond slices (the sectors
    of a block allow only some operators on w).  The plan skips them: it
    executes about `ratio` of the algorithmic flops (the reference's count,
    which dmrgx_heff_flops keeps reporting), and the matvec still equals the
    oracle's full av(), with the skipping on and off."""
    print(f"case skip zeros {model} L={L} m={m} {prec}", flush=True)
    case, _ = dmrg_case(model, L, m, 1, **kw)
    lib = gpu.lib()
    err_on, h_on = _run(gpu, case, prec)
    assert err_on < TOL[prec], err_on
    assert abs(h_on.flops - case.flops()) < 1e-6 * case.flops()
    frac = h_on.flops_exec / h_on.flops
    print(f"  executed / algorithmic flops {frac:.3f}, rel err {err_on:.2e}", flush=True)
    # the bf16 plane kernels contract 8-aligned K ranges, so a few zero
    # columns next to each range are contracted too
    assert ratio - 0.02 < frac < ratio + (0.15 if prec == "bf16x3" else 0.02), frac
    try:
        lib.dmrgx_set_skip_zero_slices(0)
        err_off, h_off = _run(gpu, case, prec)
    finally:
        lib.dmrgx_set_skip_zero_slices(1)
    assert err_off < TOL[prec]
    assert h_off.flops_exec == h_off.flops
    # split-K decompositions: partial tiles of row-mapped output groups (their
    # A rows through the row map, the tile compact, the reduction mapping rows)
    import ctypes as C
    n = C.c_int(0)
    gpu.check(lib.dmrgx_device_info(C.byref(n), None, 0), "dmrgx_device_info")
    rng = np.random.default_rng(9)
    x = rng.standard_normal(case.n_psi) + (1j * rng.standard_normal(case.n_psi) if h_on.dtype == gpu.C64 else 0)
    x = x.astype(case.dtype)
    ref = case.oracle_apply(x)
    xd = gpu.DeviceArray.from_numpy(x)
    yd = gpu.DeviceArray(case.n_psi, case.dtype)
    items = set()
    for rt, lt in [(8 * n.value, 8 * n.value), (32 * n.value, 16 * n.value), (-n.value, -n.value)]:
        gpu.check(lib.dmrgx_heff_set_split(h_on.h, rt, lt), "dmrgx_heff_set_split")
        out = (C.c_int * 4)()
        gpu.check(lib.dmrgx_heff_get_split(h_on.h, out), "dmrgx_heff_get_split")
        items.add((out[2], out[3]))
        h_on.apply(xd, yd)
        err = np.max(np.abs(yd.numpy() - ref)) / np.max(np.abs(ref))
        assert err < TOL[prec], (rt, lt, err)
    assert len(items) > 1


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_time_kernels_leaves_the_apply_result(gpu, prec):
    """dmrgx_heff_time_kernels launches each pass kernel alone, back to back
    (split-K partials left unreduced in between), and returns with `out`
    holding H_eff in again: the oracle's av() to the precision's tolerance,
    and positive per-launch times."""
    case = synthetic_case(3, 4, 200, dtype=np.complex128, uneven=[81, 67, 52])
    err, h = _run(gpu, case, prec)
    assert err < TOL[prec]
    rng = np.random.default_rng(11)
    x = rng.standard_normal(case.n_psi) + 1j * rng.standard_normal(case.n_psi)
    xd = gpu.DeviceArray.from_numpy(x)
    yd = gpu.DeviceArray(case.n_psi, np.complex128)
    ms = h.time_kernels(xd, yd, 5)
    assert ms[0] > 0 and ms[1] > 0
    ref = case.oracle_apply(x)
    err = np.max(np.abs(yd.numpy() - ref)) / np.max(np.abs(ref))
    assert err < TOL[prec], err
    h.close()
