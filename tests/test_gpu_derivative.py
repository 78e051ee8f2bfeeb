"""dE0/dJ from the adjoint H_eff kernels with recomputed environment tangents
(dmrgx_state_derivative) against the Hellmann-Feynman expectation value
<psi|dH/dJ|psi> (calc(mps, mpo)) and a central finite difference (delta =
1e-4) of exact-diagonalisation energies."""
import pytest

from dmrg_amd import driver
from oracle import exact

pytestmark = pytest.mark.gpu

XYZ = (1.0, 0.8, 0.5)
KITAEV = (-1.0, -1.0, 0.25)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_xyz_dE_dJz(gpu, prec, tol):
    L, m = 10, 64
    print(f"case xyz L={L} m={m} dE/dJz {prec}", flush=True)
    res = driver.run_dmrg("xyz", L, m, 2, XYZ, keep_state=True)
    st = res.state
    der, e = st.derivative(XYZ, (0.0, 0.0, 1.0), prec=prec)
    assert abs(e - res.energy) < (1e-9 if prec == "fp64" else 1e-6), (e, res.energy)
    hf = st.expect((0.0, 0.0, 1.0)).real
    assert abs(der - hf) < tol * max(1.0, abs(hf)), (der, hf)
    d = 1e-4
    fd = (exact.xyz_e0(L, *XYZ[:2], XYZ[2] + d)[0] - exact.xyz_e0(L, *XYZ[:2], XYZ[2] - d)[0]) / (2 * d)
    assert abs(der - fd) < max(1e-6, tol), (der, fd)
    # a direction mixing all three couplings
    der2, _ = st.derivative(XYZ, (0.3, -0.2, 0.1), prec=prec)
    hf2 = 0.3 * st.expect((1.0, 0, 0)).real - 0.2 * st.expect((0, 1.0, 0)).real + 0.1 * hf
    assert abs(der2 - hf2) < tol * max(1.0, abs(hf2)), (der2, hf2)
    st.close()


def test_kitaev_dE_dmu(gpu):
    """Fermionic chain: the derivative through the environments the DMRG steps
    themselves use (graded signs) matches ED finite differences."""
    L, m = 12, 32
    print(f"case kitaev L={L} m={m} dE/dU", flush=True)
    res = driver.run_dmrg("kitaev", L, m, 2, KITAEV, keep_state=True)
    st = res.state
    der, e = st.derivative(KITAEV, (0.0, 0.0, 1.0))
    assert abs(e - res.energy) < 1e-9, (e, res.energy)
    d = 1e-4
    fd = (exact.kitaev_ed(L, *KITAEV[:2], KITAEV[2] + d)[0] - exact.kitaev_ed(L, *KITAEV[:2], KITAEV[2] - d)[0]) / (
        2 * d)
    assert abs(der - fd) < 1e-6, (der, fd)
    st.close()
