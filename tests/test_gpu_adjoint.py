"""Adjoint of the two-site H_eff matvec (dmrgx_heff_adjoint) with respect to
psi and the two environments.

The forward map out = H_eff(psi; R, L) is linear in each argument separately,
so the adjoint is pinned exactly, without finite differences:
  Re<g, H(dpsi)>        = Re<g_psi, dpsi>   (H applied by the oracle's av())
  Re<g, H_{R:=dR}(psi)> = Re<g_R, dR>       (the FP64 forward with dR as R)
  Re<g, H_{L:=dL}(psi)> = Re<g_L, dL>
Errors are scaled by |g| |forward(d)|, the size of the terms summed."""
import numpy as np
import pytest

from tests.cases import synthetic_case

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 3e-6, "bf16x3": 3e-6}


def _rand(rng, n, dt):
    x = rng.standard_normal(n)
    if np.dtype(dt).kind == "c":
        x = x + 1j * rng.standard_normal(n)
    return x.astype(dt)


def _plan(gpu, case, prec, R, L):
    c = gpu.C64 if np.dtype(case.dtype).kind == "c" else gpu.R64
    return gpu.Heff(c, gpu.PREC_NAMES[prec], R, case.roff, case.rsize, case.rmap, L, case.loff, case.lsize,
                    case.lmap, case.n_psi, case.n_store)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("shape", [
    ("z2_m32_complex", 2, 4, 32, np.complex128, None),
    ("z3_uneven_complex", 3, 4, 60, np.complex128, [27, 20, 13]),
    ("z2_ragged_real", 2, 5, 24, np.float64, [13, 11]),
])
def test_heff_adjoint(gpu, prec, shape):
    name, S, D, m, dt, uneven = shape
    print(f"case adjoint {name} {prec}", flush=True)
    case = synthetic_case(S, D, m, dtype=dt, uneven=uneven)
    rng = np.random.default_rng(13)
    R = gpu.DeviceArray.from_numpy(case.rflat)
    L = gpu.DeviceArray.from_numpy(case.lflat)
    h = _plan(gpu, case, prec, R, L)
    psi = _rand(rng, case.n_psi, dt)
    g = _rand(rng, case.n_psi, dt)
    psi_d, g_d = gpu.DeviceArray.from_numpy(psi), gpu.DeviceArray.from_numpy(g)
    gp = gpu.DeviceArray(case.n_psi, dt)
    gR = gpu.DeviceArray(case.rflat.size, dt)
    gL = gpu.DeviceArray(case.lflat.size, dt)
    h.adjoint(psi_d, g_d, gp, gR, gL)
    g_psi, g_R, g_L = gp.numpy(), gR.numpy(), gL.numpy()

    def check(lhs_vec, grad, d, what):
        lhs = np.real(np.vdot(g, lhs_vec))
        rhs = np.real(np.vdot(grad, d))
        scale = np.linalg.norm(g) * np.linalg.norm(lhs_vec)
        assert abs(lhs - rhs) <= TOL[prec] * scale, f"{name} {prec} {what}: {lhs} vs {rhs}"

    dpsi = _rand(rng, case.n_psi, dt)
    check(case.oracle_apply(dpsi), g_psi, dpsi, "psi")
    y = gpu.DeviceArray(case.n_psi, dt)
    dR = _rand(rng, case.rflat.size, dt)
    hR = _plan(gpu, case, "fp64", gpu.DeviceArray.from_numpy(dR), L)
    hR.apply(psi_d, y)
    check(y.numpy(), g_R, dR, "renv")
    dL = _rand(rng, case.lflat.size, dt)
    hL = _plan(gpu, case, "fp64", R, gpu.DeviceArray.from_numpy(dL))
    hL.apply(psi_d, y)
    check(y.numpy(), g_L, dL, "lenv")
