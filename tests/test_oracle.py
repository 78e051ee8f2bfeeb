"""Pins the oracle itself against what the mathematics fixes: BdG closed form
vs brute-force Jordan-Wigner ED, DMRG energies vs ED, H_eff from the task
lists vs an independent dense Kronecker build, and the SVD truncation rule as
a pure function of the singular values.  CPU only."""
import itertools

import numpy as np
import pytest

from oracle import exact, models
from oracle import qtensor as qt
from oracle.dmrg import Runner, dense_heff
from oracle.qtensor import QTensor, tensor_svd
from tests.cases import dmrg_case

KITAEV = dict(t=-1.0, p=-1.0, u=0.25)
XYZ = dict(jx=1.0, jy=0.8, jz=0.5)


@pytest.mark.parametrize("L", [3, 4, 5, 6])
def test_kitaev_bdg_matches_jordan_wigner_ed(L):
    e0, gap = exact.kitaev_bdg_e0(L, -1.0, -1.0, 0.25)
    ev = exact.kitaev_ed(L, -1.0, -1.0, 0.25, 0)[0]
    od = exact.kitaev_ed(L, -1.0, -1.0, 0.25, 1)[0]
    assert abs(min(ev, od) - e0) < 1e-12
    assert abs(abs(ev - od) - gap) < 1e-12


@pytest.mark.parametrize("sector", [0, 1])
def test_kitaev_ring_hopping_is_the_periodic_free_fermion_ring(sector):
    """With the pairing off, the periodic ED of kitaev_ed is free fermions on a
    ring with periodic boundary conditions in both parity sectors:
    e_k = 2T cos(2 pi n / L) - 2U, E0 = the lowest N levels with N of the
    sector's parity (the ring bond is written in fermion operators, so no
    parity-dependent twist appears)."""
    L, T, U = 7, -1.0, 0.3
    e = np.sort(2 * T * np.cos(2 * np.pi * np.arange(L) / L) - 2 * U)
    best = min(np.sum(e[:n]) for n in range(L + 1) if n % 2 == sector)
    assert abs(exact.kitaev_ed(L, T, 0.0, U, sector, pbc=True)[0] - best) < 1e-12
    # the open chain differs
    assert abs(exact.kitaev_ed(L, T, 0.0, U, sector)[0] - best) > 1e-3


def test_edge_term_of_the_xyz_mpo_is_the_ring_bond():
    """mpo_edge_dense(W) for the XYZ MPO equals the bond (L-1, 0) of the ring:
    the XYZ bond is symmetric under exchanging its two sites."""
    L = 6
    W = models.H_XYZ(1.0, 0.8, 0.5).remove_symmetry().val[0]
    ring = exact.xyz_hamiltonian(L, 1.0, 0.8, 0.5, pbc=True)
    assert abs(ring - (exact.mpo_dense(W, L, 2) + exact.mpo_edge_dense(W, L, 2))).max() < 1e-14


def test_xy_free_fermions_match_ed():
    L = 8
    e, nf = exact.xy_free_fermion_e0(L, 1.0)
    sec = nf % 2
    assert abs(exact.xyz_e0(L, 1.0, 1.0, 0.0, sec)[0] - e) < 1e-12


@pytest.mark.parametrize("schedule", ["dmrg2", "dmrg"])
def test_oracle_kitaev_dmrg_closed_form(schedule):
    mpo, n, ferm = models.make_model("kitaev", 8, **KITAEV)
    R = Runner(mpo, n, ferm)
    e = R.dmrg2(16, 2) if schedule == "dmrg2" else R.dmrg(16, 2)
    assert abs(e - exact.kitaev_ed(8, -1.0, -1.0, 0.25)[0]) < 1e-10


def test_oracle_xyz_physical_rule_matches_ed():
    mpo, n, ferm = models.make_model("xyz", 8, **XYZ)
    R = Runner(mpo, n, ferm, sign_rule="physical")
    assert abs(R.dmrg2(32, 2) - exact.xyz_e0(8, 1.0, 0.8, 0.5)[0]) < 1e-10


def test_oracle_reference_rule_is_not_the_xyz_ground_state():
    """The reference's unconditional get_map sign applied to the bosonic XYZ
    model converges to a different (non-physical) energy."""
    mpo, n, ferm = models.make_model("xyz", 8, **XYZ)
    R = Runner(mpo, n, ferm, sign_rule="reference")
    e = R.dmrg2(32, 2)
    assert abs(e - exact.xyz_e0(8, 1.0, 0.8, 0.5)[0]) > 1e-3


def test_oracle_potts3_matches_ed():
    P = dict(f=1.0, j=1.0, p=0.0, t=3.14159265359 / 8)
    mpo, n, ferm = models.make_model("potts3", 5, **P)
    R = Runner(mpo, n, ferm)
    W = models.H_Potts3(1.0, 1.0, 0.0, 3.14159265359 / 8).remove_symmetry().val[0]
    assert abs(R.dmrg2(27, 2) - exact.mpo_e0(W, 5, 3, 3, 0)[0]) < 1e-10


def _psi_full_index(dim, sym):
    """flat psi position -> multi-index in the dense (symmetry-free) space."""
    idx = []
    for s in sym:
        shape = [dim[k][s[k]] for k in range(len(dim))]
        base = [sum(dim[k][:s[k]]) for k in range(len(dim))]
        for loc in itertools.product(*[range(x) for x in shape]):
            idx.append(tuple(b + l for b, l in zip(base, loc)))
    return idx


def test_heff_task_lists_match_dense_kronecker_build():
    case, _ = dmrg_case("xyz", 8, 12, 1, **XYZ)
    H = dense_heff(case.ctx, case.n_psi, case.dtype)
    Ld = case.lenv.remove_symmetry().val[0]   # [a, s1, w, a', s1']
    Rd = case.renv.remove_symmetry().val[0]   # [s2, b, w, s2', b']
    # H[(a s1 s2 b), (a' s1' s2' b')] = sum_w L[a,s1,w,a',s1'] R[s2,b,w,s2',b']
    full = np.einsum("aswcd,tbwxy->astbcdxy", Ld, Rd)
    idx = _psi_full_index(case.dim, case.sym)
    ref = np.array([[full[o + i] for i in idx] for o in idx])
    assert np.max(np.abs(H - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(H - H.conj().T)) < 1e-12 * np.max(np.abs(H))


def test_svd_reconstruction_and_unitarity():
    rng = np.random.default_rng(1)
    for r, c in [(5, 3), (3, 5), (4, 4), (1, 1)]:
        a = rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))
        u, s, v = tensor_svd(a)
        k = len(s)
        assert np.allclose((u[:, :k] * s) @ v[:k, :], a)
        assert np.allclose(u.conj().T @ u, np.eye(r))
        assert np.allclose(v @ v.conj().T, np.eye(c))


def test_svd_selection_rule_pure_function_of_s():
    """Global top-cutoff selection with ties resolved to the lowest sector:
    compare the oracle's max-scan (src/qtensor.cpp:1183-1195) with a stable
    sort of (-s, sector, rank)."""
    qt.set_symmetry(2)
    rng = np.random.default_rng(3)
    for trial in range(20):
        # two sector blocks with degenerate singular values
        sv = np.sort(rng.choice([0.5, 0.3, 0.1], size=4))[::-1]
        blocks = []
        for sec in range(2):
            q1, _ = np.linalg.qr(rng.standard_normal((4, 4)))
            q2, _ = np.linalg.qr(rng.standard_normal((4, 4)))
            blocks.append(q1 @ np.diag(sv) @ q2)
        t = QTensor(2, 2)
        t.dim = [[4, 4], [4, 4]]
        t.sym = [[0, 0], [1, 1]]
        t.val = blocks
        cutoff = int(rng.integers(1, 8))
        U, S, V, s = t.svd(1, cutoff)
        # the rule is applied to the computed values (degenerate ones differ in the last bits)
        comp = [tensor_svd(b)[1] for b in blocks]
        keyed = sorted([(-x, sec, k) for sec in range(2) for k, x in enumerate(comp[sec])])
        want = [0, 0]
        for _, sec, _ in keyed[:cutoff]:
            want[sec] += 1
        got = [U.dim[1][0], U.dim[1][1]]
        assert got == want, (sv, cutoff, got, want)
        assert np.allclose(s, sorted(np.concatenate([sv, sv]), reverse=True))


def test_oracle_calc_physical_xyz_observables():
    """calc(mps) = 1, calc(mps, mpo) = E0 and the Hellmann-Feynman derivative
    <sum sz sz> = dE0/dJz by central differences of ED (XYZ, physical rule:
    no graded signs)."""
    from oracle.dmrg import calc_expect, calc_overlap, prepared
    L = 8
    mpo, n, ferm = models.make_model("xyz", L, **XYZ)
    R = Runner(mpo, n, ferm, sign_rule="physical")
    e = R.dmrg2(32, 2)
    st = prepared(R.mps)
    assert abs(calc_overlap(st, st, False) - 1) < 1e-10
    assert abs(calc_expect(st, mpo, False) - e) < 1e-9
    dz, _, _ = models.make_model("xyz", L, jx=0.0, jy=0.0, jz=1.0)
    hf = np.real(calc_expect(st, dz, False))
    d = 1e-4
    fd = (exact.xyz_e0(L, 1.0, 0.8, 0.5 + d)[0] - exact.xyz_e0(L, 1.0, 0.8, 0.5 - d)[0]) / (2 * d)
    assert abs(hf - fd) < 1e-6


def test_oracle_calc_fermionic_norm():
    """The graded calc(mps) of a Kitaev ground state is 1."""
    from oracle.dmrg import calc_overlap, prepared
    mpo, n, ferm = models.make_model("kitaev", 8, **KITAEV)
    R = Runner(mpo, n, ferm)
    R.dmrg2(16, 2)
    st = prepared(R.mps)
    assert abs(calc_overlap(st, st, True) - 1) < 1e-10


def test_reference_rule_depends_on_svd_complement():
    """Under the reference's sign rule the energies of some two-site steps of
    an open XYZ chain depend on which orthonormal complement the SVD pads U / V
    with (src/qtensor.cpp:1221-1252): another valid complement moves them,
    while the converged final energy stays.  Under the physical rule the same
    change moves no step by more than rounding."""
    from oracle import qtensor as qt
    from oracle.dmrg import Runner

    def run(rule, seed):
        old = qt.COMPLEMENT_ROTATION_SEED
        qt.COMPLEMENT_ROTATION_SEED = seed
        try:
            mpo, n, ferm = models.make_model("xyz", 10, jx=1.0, jy=0.8, jz=0.5)
            R = Runner(mpo, n, ferm, sign_rule=rule)
            e = R.dmrg2(16, 1)
        finally:
            qt.COMPLEMENT_ROTATION_SEED = old
        return e, np.array([o.energies[0] for o in R.log])

    e0, a = run("reference", None)
    e1, b = run("reference", 5)
    assert np.max(np.abs(a - b)) > 1e-6
    assert abs(e0 - e1) < 1e-10
    p0, c = run("physical", None)
    p1, d = run("physical", 5)
    assert np.max(np.abs(c - d)) < 1e-11
    assert abs(p0 - p1) < 1e-11
