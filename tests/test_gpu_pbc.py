"""Periodic chains (the reference's PBC build): the edge environments the
mains start with add_edge (xyz.cpp:39-43, kitaev.cpp:43-46, potts.cpp:45-48)
are carried through the chain by identity operators (qtensor::id(bulk),
src/qtensor.cpp:157-182) and every matvec adds L_edge W_id W_id R_edge to the
bulk operator (src/core.cpp:27-30).  Pinned by exact diagonalisation of the
periodic Hamiltonians: the XYZ ring and the Kitaev ring written in fermion
operators (a physical check: the reference flips the pairing sign of the
Kitaev edge so that the boundary bond is the ring bond), and the Potts chain
plus the boundary term of its edge tensors (oracle.exact.mpo_edge_dense,
which the XYZ ring pins in tests/test_oracle.py)."""
import numpy as np
import pytest

from dmrg_amd import driver
from oracle import exact, models

pytestmark = pytest.mark.gpu

KITAEV = (-1.0, -1.0, 0.25)
XYZ = (1.0, 0.8, 0.5)
POTTS = (1.0, 1.0, 0.0, 3.14159265359 / 8)


@pytest.mark.parametrize("schedule", ["dmrg2", "dmrg"])
def test_xyz_ring(gpu, schedule):
    L = 10
    print(f"case xyz pbc L={L} {schedule}", flush=True)
    res = driver.run_dmrg("xyz", L, 64, 4, XYZ, schedule=schedule, pbc=True)
    e_ring = exact.xyz_e0(L, *XYZ, pbc=True)[0]
    e_open = exact.xyz_e0(L, *XYZ)[0]
    assert abs(e_ring - e_open) > 0.5  # the edge term matters
    assert abs(res.energy - e_ring) < 1e-8, (res.energy, e_ring)


@pytest.mark.parametrize("sector", [0, 1])
def test_kitaev_ring(gpu, sector):
    L = 10
    print(f"case kitaev pbc L={L} sector={sector}", flush=True)
    res = driver.run_dmrg("kitaev", L, 64, 4, KITAEV, sector=sector, schedule="dmrg2", pbc=True)
    e_ring = exact.kitaev_ed(L, *KITAEV, sector=sector, pbc=True)[0]
    assert abs(res.energy - e_ring) < 1e-8, (res.energy, e_ring, exact.kitaev_ed(L, *KITAEV, sector=sector)[0])


def test_potts3_periodic(gpu):
    L = 6
    print(f"case potts3 pbc L={L}", flush=True)
    res = driver.run_dmrg("potts3", L, 81, 4, POTTS, schedule="dmrg2", pbc=True)
    H = exact.potts3_hamiltonian(L, *POTTS, pbc=True)
    sel = exact.z_n_sector_indices(L, 3, 3, 0)
    A = H[sel][:, sel].toarray()
    e = np.linalg.eigvalsh(0.5 * (A + A.conj().T))[0]
    assert abs(res.energy - e) < 1e-8, (res.energy, e)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_xyz_ring_mixed_precision(gpu, prec):
    """The edge operator runs in the same arithmetic as the bulk one inside the
    reduced-precision Lanczos; the FP64 refinement brings E0 back to FP64."""
    L = 12
    print(f"case xyz pbc L={L} {prec}", flush=True)
    a = driver.run_dmrg("xyz", L, 64, 3, XYZ, pbc=True)
    b = driver.run_dmrg("xyz", L, 64, 3, XYZ, pbc=True, prec=prec, tol_lo=1e-5)
    assert abs(a.energy - b.energy) < 1e-8, (a.energy, b.energy)
    assert abs(a.energy - exact.xyz_e0(L, *XYZ, pbc=True)[0]) < 1e-7


def test_periodic_state_refuses_open_chain_operators(gpu):
    """A periodic run's state keeps its observables (calc(mps) is 1) but the
    centre-operator and derivative calls, which rebuild open-chain
    environments, refuse it."""
    res = driver.run_dmrg("xyz", 8, 16, 1, XYZ, pbc=True, keep_state=True)
    st = res.state
    assert abs(st.norm2() - 1) < 1e-10
    with pytest.raises(RuntimeError, match="periodic"):
        st.derivative(XYZ, (0.0, 0.0, 1.0))
    with pytest.raises(RuntimeError, match="periodic"):
        st.centre_heff()
    st.close()
