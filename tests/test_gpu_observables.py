"""Observables of finished GPU states: the reference's calc(mps), calc(mps, mps)
and calc(mps, mpo) (src/functions.cpp:394-511) through dmrgx_state_*.

Pinned by physics, not by retyping the contraction: the norm of a converged
state is 1, <psi|H|psi> equals the eigensolver's E0 (Rayleigh quotient), and
<psi|dH/dJ|psi> (Hellmann-Feynman, the MPO builders are linear in the
couplings) equals a central finite difference of exact-diagonalisation
energies."""
import numpy as np
import pytest

from dmrg_amd import driver
from oracle import exact, models

pytestmark = pytest.mark.gpu

KITAEV = (-1.0, -1.0, 0.25)
XYZ = (1.0, 0.8, 0.5)
POTTS = (1.0, 1.0, 0.0, 3.14159265359 / 8)


def test_xyz_norm_energy_and_hellmann_feynman(gpu):
    L, m = 10, 64
    print(f"case xyz L={L} m={m} observables", flush=True)
    res = driver.run_dmrg("xyz", L, m, 2, XYZ, keep_state=True)
    st = res.state
    n2 = st.norm2()
    assert abs(n2 - 1) < 1e-10, n2
    e = st.expect(XYZ)
    assert abs(e.imag) < 1e-10 and abs(e.real - res.energy) < 1e-9, (e, res.energy)
    # dE0/dJz = <sum sz sz>: the MPO of H_XYZ(0, 0, 1)
    hf = st.expect((0.0, 0.0, 1.0)).real
    d = 1e-4
    ep = exact.xyz_e0(L, XYZ[0], XYZ[1], XYZ[2] + d)[0]
    em = exact.xyz_e0(L, XYZ[0], XYZ[1], XYZ[2] - d)[0]
    fd = (ep - em) / (2 * d)
    assert abs(hf - fd) < 1e-6, (hf, fd)
    # the three couplings' derivatives add up to E0 (H is homogeneous in J)
    hx = st.expect((1.0, 0.0, 0.0)).real
    hy = st.expect((0.0, 1.0, 0.0)).real
    assert abs(XYZ[0] * hx + XYZ[1] * hy + XYZ[2] * hf - res.energy) < 1e-9
    st.close()


def test_kitaev_norm_and_fidelity(gpu):
    """Fermionic states: calc(mps) with the graded trace is 1 and two runs
    (FP64, BF16x3) overlap with fidelity 1.  calc(mps, mpo) is not checked
    against E0 here: the reference's left-to-right transfer (change = true,
    src/functions.cpp:482-490) does not reproduce the DMRG operator of the
    Kitaev MPO for any choice of the graded signs (the oracle restatement gives
    the same values)."""
    L, m = 12, 32
    print(f"case kitaev L={L} m={m} observables", flush=True)
    a = driver.run_dmrg("kitaev", L, m, 2, KITAEV, keep_state=True)
    b = driver.run_dmrg("kitaev", L, m, 2, KITAEV, keep_state=True, prec="bf16x3", tol_lo=1e-5, refine_k=10)
    assert abs(a.state.norm2() - 1) < 1e-10
    assert abs(a.state.fidelity(b.state) - 1) < 1e-8
    assert abs(a.energy - b.energy) < 1e-8
    a.state.close()
    b.state.close()


def test_potts3_expectation(gpu):
    L, m = 7, 81
    print(f"case potts3 L={L} observables", flush=True)
    res = driver.run_dmrg("potts3", L, m, 2, POTTS, keep_state=True)
    st = res.state
    assert abs(st.norm2() - 1) < 1e-10
    e = st.expect(POTTS)
    assert abs(e.imag) < 1e-10 and abs(e.real - res.energy) < 1e-9, (e, res.energy)
    st.close()


def test_overlap_fidelity_between_runs(gpu):
    """Two FP64 runs from different Lanczos seeds and a BF16x3 run reach the
    same (non-degenerate) ground state: fidelity 1 up to the energy error."""
    L, m = 12, 64
    print(f"case xyz L={L} m={m} fidelity", flush=True)
    a = driver.run_dmrg("xyz", L, m, 2, XYZ, keep_state=True, seed=1)
    b = driver.run_dmrg("xyz", L, m, 2, XYZ, keep_state=True, seed=99)
    c = driver.run_dmrg("xyz", L, m, 2, XYZ, keep_state=True, prec="bf16x3", tol_lo=1e-5, refine_k=10)
    f_ab = a.state.fidelity(b.state)
    f_ac = a.state.fidelity(c.state)
    assert abs(f_ab - 1) < 1e-9, f_ab
    assert abs(f_ac - 1) < 1e-8, f_ac
    assert abs(c.energy - a.energy) < 1e-8
    # <a|b> = conj(<b|a>)
    ab, ba = a.state.overlap(b.state), b.state.overlap(a.state)
    assert abs(ab - np.conj(ba)) < 1e-10
    for r in (a, b, c):
        r.state.close()


def test_move_cut_and_inverse(gpu):
    """mps::move_cut_left / move_cut_right and mps::inverse
    (src/mps.cpp:56-128).  Moving the centre splits two sites again with every
    singular value kept, so the state is unchanged: norm 1, <psi|H|psi> = E0,
    overlap 1 with an unmoved copy, and the derivative through the
    environments at the new centre still equals Hellmann-Feynman.  The
    reversed XYZ chain has the same energy (H is reflection symmetric), and
    reversing twice gives the state back."""
    L, m = 12, 32
    print(f"case xyz L={L} m={m} move_cut / inverse", flush=True)
    res = driver.run_dmrg("xyz", L, m, 2, XYZ, keep_state=True)
    st = res.state
    c0 = st.position()
    assert 1 <= c0 <= L - 1
    copy = st.inverse().inverse()
    assert copy.position() == c0
    assert abs(copy.fidelity(st) - 1) < 1e-12
    hf = st.expect((0.0, 0.0, 1.0)).real
    for steps in (-3, 5, -2):
        st.move_cut(steps)
        c0 += steps
        assert st.position() == c0
        assert abs(st.norm2() - 1) < 1e-10
        e = st.expect(XYZ)
        assert abs(e.real - res.energy) < 1e-9 and abs(e.imag) < 1e-10, (steps, e, res.energy)
        assert abs(abs(st.overlap(copy)) - 1) < 1e-10
        der, ec = st.derivative(XYZ, (0.0, 0.0, 1.0))
        assert abs(ec - res.energy) < 1e-9 and abs(der - hf) < 1e-9, (steps, der, hf)
    # the centre stays on bonds 1..L-1
    with pytest.raises(RuntimeError):
        st.move_cut(L)
    with pytest.raises(RuntimeError):
        st.move_cut(-c0)
    assert st.position() == c0
    inv = st.inverse()
    assert inv.position() == L - c0
    assert abs(inv.norm2() - 1) < 1e-10
    assert abs(inv.expect(XYZ).real - res.energy) < 1e-9
    back = inv.inverse()
    assert abs(back.fidelity(st) - 1) < 1e-12
    for s in (st, copy, inv, back):
        s.close()


def test_potts3_inverse_is_the_reflected_chain(gpu):
    """Reflection maps the chiral Potts chain H(theta) to H(-theta)
    (sigma+_i sigma_{i+1} e^{i theta} becomes sigma+_{i+1} sigma_i e^{i theta}),
    so the reversed ground state has energy E0 under H(-theta) and, being no
    eigenstate of H(theta) in general, a higher one under H(theta)."""
    L, m = 8, 81
    print(f"case potts3 L={L} m={m} inverse", flush=True)
    res = driver.run_dmrg("potts3", L, m, 2, POTTS, keep_state=True)
    inv = res.state.inverse()
    assert abs(inv.norm2() - 1) < 1e-10
    mirror = POTTS[:3] + (-POTTS[3],)
    assert abs(inv.expect(mirror).real - res.energy) < 1e-9
    assert inv.expect(POTTS).real > res.energy + 1e-6
    inv.close()
    res.state.close()
