"""The work-decomposition autotuner of the H_eff plans (dmrgx_set_autotune):
it times each new sector-shape histogram once, reuses the choice for later
plans of that histogram, leaves small operators alone, and changes nothing
but the summation order."""
import ctypes as C

import numpy as np
import pytest

from dmrg_amd import driver

pytestmark = pytest.mark.gpu


def _apply(gpu, h, x):
    X = gpu.DeviceArray.from_numpy(x)
    Y = gpu.DeviceArray(h.n_psi, np.complex128)
    h.apply(X, Y)
    return Y.numpy()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("bf16x3", 1e-5)])
def test_autotune_cache_and_results(gpu, prec, tol):
    lib = gpu.lib()
    print(f"case autotune {prec}", flush=True)
    try:
        lib.dmrgx_set_autotune(0)
        ref = driver.SynthHeff(2, 5, [160, 96], prec=prec)
        rng = np.random.default_rng(1)
        x = rng.standard_normal(ref.n_psi) + 1j * rng.standard_normal(ref.n_psi)
        y0 = _apply(gpu, ref, x)
        lib.dmrgx_set_autotune(1)
        k0 = lib.dmrgx_autotune_keys()
        a = driver.SynthHeff(2, 5, [160, 96], prec=prec)
        k1 = lib.dmrgx_autotune_keys()
        # same histogram (sizes round to the same powers of two): no new timing
        b = driver.SynthHeff(2, 5, [150, 100], prec=prec, seed=7)
        k2 = lib.dmrgx_autotune_keys()
        # a small operator is not timed
        c = driver.SynthHeff(2, 5, [16, 16], prec=prec)
        k3 = lib.dmrgx_autotune_keys()
    finally:
        lib.dmrgx_set_autotune(1)
    assert a.flops >= 1e9 and c.flops < 1e9
    assert k1 in (k0, k0 + 1)  # +1 unless an earlier test in this process timed the same histogram
    assert k2 == k1 and k3 == k2
    y1 = _apply(gpu, a, x)
    assert np.max(np.abs(y1 - y0)) <= tol * np.max(np.abs(y0))
    for h in (ref, a, b, c):
        h.close()


@pytest.mark.parametrize("shape", [(2, 5, [160, 96]), (3, 4, [120, 98, 76])])
def test_pinned_decompositions_agree(gpu, shape):
    """Every work decomposition (the default, K cut for 1 and 8 items per CU,
    and the balanced one over 1 or 2 resident work-groups per CU, which cuts
    whole blocks into two K halves and orders items per XCD) gives the BF16x3
    matvec to its precision against the FP64 plan; the FP64 plan itself is
    unchanged by its decompositions to rounding."""
    S, D, bond = shape
    lib = gpu.lib()
    n = C.c_int(0)
    gpu.check(lib.dmrgx_device_info(C.byref(n), None, 0), "dmrgx_device_info")
    ncu = n.value
    print(f"case pinned decompositions S={S} bond={bond}", flush=True)
    lib.dmrgx_set_autotune(0)
    try:
        f64 = driver.SynthHeff(S, D, bond, prec="fp64")
        b3 = driver.SynthHeff(S, D, bond, prec="bf16x3")
    finally:
        lib.dmrgx_set_autotune(1)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(f64.n_psi) + 1j * rng.standard_normal(f64.n_psi)
    y64 = _apply(gpu, f64, x)
    scale = np.max(np.abs(y64))
    seen = set()
    for rt, lt in [(0, 0), (ncu, 8 * ncu), (-ncu, 0), (0, -ncu), (-2 * ncu, -2 * ncu), (8 * ncu, -ncu)]:
        b3.set_split(rt, lt)
        r_t, l_t, r_items, l_items = b3.split()
        assert (r_t, l_t) == (rt, lt) and r_items > 0 and l_items > 0
        seen.add((r_items, l_items))
        y = _apply(gpu, b3, x)
        assert np.max(np.abs(y - y64)) <= 2e-5 * scale, (rt, lt)
        f64.set_split(rt, lt)
        y2 = _apply(gpu, f64, x)
        assert np.max(np.abs(y2 - y64)) <= 1e-13 * scale, (rt, lt)
    assert len(seen) > 1  # the targets do change the decomposition
    f64.close()
    b3.close()
