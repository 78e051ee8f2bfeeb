"""The work-decomposition autotuner of the H_eff plans (dmrgx_set_autotune):
it times each new sector-shape histogram once, reuses the choice for later
plans of that histogram, leaves small operators alone, and changes nothing
but the summation order."""
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
