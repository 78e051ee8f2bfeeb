"""Independent physics references: closed forms and exact diagonalisation.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Nothing here restates the
reference; these fix the energies the reference's algorithm must reproduce.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .qtensor import PI


def kitaev_bdg_e0(L, T, P, U):
    """Ground-state energy of the open Kitaev chain in the reference's MPO
    convention (src/operators.cpp:396-431):

        H = sum_i [ T c+_i c_{i+1} - P c_i c_{i+1} + h.c. ] - 2U sum_i n_i

    from the Bogoliubov-de Gennes spectrum:  H = 1/2 Psi+ HB Psi + 1/2 tr h with
    Psi = (c, c+), HB = [[h, D], [-D*, -h*]] (D antisymmetric), so
    E0 = 1/2 tr h - 1/2 sum_k e_k over the L non-negative BdG eigenvalues.
    Returns (E0, smallest e_k): the ground state of the opposite fermion
    parity lies e_min above E0."""
    h = np.zeros((L, L))
    D = np.zeros((L, L))
    for i in range(L - 1):
        h[i, i + 1] = h[i + 1, i] = T
        # -P c_i c_{i+1} - P* c+_{i+1} c+_i  ==  1/2 sum_ij (D_ij c+_i c+_j + h.c.)
        D[i, i + 1] = P
        D[i + 1, i] = -P
    for i in range(L):
        h[i, i] = -2 * U
    HB = np.block([[h, D], [-D.conj(), -h.conj()]])
    w = np.sort(np.linalg.eigvalsh(HB))
    pos = w[L:]
    e0 = 0.5 * np.trace(h) - 0.5 * np.sum(pos)
    return float(e0), float(pos[0])


def kitaev_ed(L, T, P, U, sector=0, k=1, pbc=False):
    """Brute-force Jordan-Wigner diagonalisation of the same Hamiltonian in a
    fixed fermion-parity sector (local state 1 = occupied).  pbc adds the bond
    (L-1, 0) written in the fermion operators themselves (the Jordan-Wigner
    string of c_0 c_{L-1} comes out of the matrices): the periodic ring with
    no twist."""
    Z = sp.csr_matrix(np.diag([1.0, -1.0]))
    cm = sp.csr_matrix(np.array([[0.0, 1.0], [0.0, 0.0]]))  # <0|c|1> = 1
    I2 = sp.identity(2, format="csr")

    def c(i):
        op = None
        for j in range(L):
            o = Z if j < i else (cm if j == i else I2)
            op = o if op is None else sp.kron(op, o, format="csr")
        return op

    cs = [c(i) for i in range(L)]
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L if pbc else L - 1):
        j = (i + 1) % L
        hop = T * (cs[i].T @ cs[j])
        pair = -P * (cs[i] @ cs[j])
        H = H + hop + hop.T + pair + pair.T
    for i in range(L):
        H = H - 2 * U * (cs[i].T @ cs[i])
    sel = z_n_sector_indices(L, 2, 2, sector)
    A = H[sel][:, sel].toarray()
    return np.sort(np.linalg.eigvalsh(A))[:k]


def _pauli():
    sx = sp.csr_matrix(np.array([[0, 1], [1, 0]], dtype=complex))
    sy = sp.csr_matrix(np.array([[0, -1j], [1j, 0]], dtype=complex))
    sz = sp.csr_matrix(np.array([[1, 0], [0, -1]], dtype=complex))
    return sx, sy, sz


def _site_op(op, i, L, d):
    left = sp.identity(d ** i, format="csr")
    right = sp.identity(d ** (L - i - 1), format="csr")
    return sp.kron(sp.kron(left, op, format="csr"), right, format="csr")


def xyz_hamiltonian(L, Jx, Jy, Jz, pbc=False):
    """Open XYZ chain in Pauli units (H_XYZ, src/operators.cpp:355-391):
    H = sum_i Jx sx sx + Jy sy sy + Jz sz sz.  Basis: site 0 most significant,
    local state 0 = spin up (sz=+1).  pbc adds the bond (L-1, 0)."""
    sx, sy, sz = _pauli()
    H = sp.csr_matrix((2 ** L, 2 ** L), dtype=complex)
    for i in range(L if pbc else L - 1):
        for J, o in ((Jx, sx), (Jy, sy), (Jz, sz)):
            if J != 0:
                H = H + J * (_site_op(o, i, L, 2) @ _site_op(o, (i + 1) % L, L, 2))
    return H


def sector_projector_z2(L):
    """Indices of basis states with even number of down spins (Z2 charge 0)."""
    idx = np.arange(2 ** L)
    par = np.array([bin(x).count("1") % 2 for x in idx])
    return idx[par == 0], idx[par == 1]


def xyz_e0(L, Jx, Jy, Jz, sector=0, k=1, pbc=False):
    H = xyz_hamiltonian(L, Jx, Jy, Jz, pbc).tocsr()
    ev, od = sector_projector_z2(L)
    sel = ev if sector == 0 else od
    Hs = H[sel][:, sel]
    if Hs.shape[0] <= 256:
        w = np.linalg.eigvalsh(Hs.toarray())
        return np.sort(w)[:k]
    w = spla.eigsh(Hs, k=k, which="SA", tol=1e-12)[0]
    return np.sort(w)


def xy_free_fermion_e0(L, J):
    """XX chain (Jx=Jy=J, Jz=0) in Pauli units: H = 2J sum (s+s- + h.c.) maps by
    Jordan-Wigner to free fermions with hopping 2J; E0 = sum of negative modes
    (open chain, both parities; the even-parity value is returned with the
    parity of the filled Fermi sea)."""
    h = np.zeros((L, L))
    for i in range(L - 1):
        h[i, i + 1] = h[i + 1, i] = 2 * J
    e = np.sort(np.linalg.eigvalsh(h))
    return float(np.sum(e[e < 0])), int(np.sum(e < 0))


def potts3_hamiltonian(L, f, j, p, t, pbc=False):
    """Chiral 3-state Potts chain in the reference convention
    (H_Potts3, src/operators.cpp:293-350):
       H = -f sum_i (e^{ip} T_i + h.c.) - J sum_i (e^{it} S_i S+_{i+1} + h.c.)
    with T = diag(1, w, w^2), w = e^{2 pi i/3} (truncated PI), S the cyclic shift
    S|q> = |q-1>.  The MPO entries fix the exact operator ordering; this
    function builds the same operator from the MPO tensors directly, so it is
    independent of the DMRG machinery but not of the MPO.  pbc adds the edge
    term of H_Potts3(0, j, p, t) (see mpo_edge_dense)."""
    from .models import H_Potts3
    W = H_Potts3(f, j, p, t).remove_symmetry().val[0]  # (4,3,4,3)
    H = mpo_dense(W, L, 3)
    if pbc:
        H = H + mpo_edge_dense(H_Potts3(0.0, j, p, t).remove_symmetry().val[0], L, 3)
    return H.tocsr()


def mpo_dense(W, L, d):
    """Contract a uniform MPO W[w, s, w', s'] (reference index convention,
    left edge = last row, right edge = first column, include/mpo.h:45-53,
    src/qtensor.cpp:448-461) into a sparse matrix on L sites."""
    D = W.shape[0]
    # ops[w][w'] = sparse dxd operator <s|.|s'>
    ops = [[sp.csr_matrix(W[a, :, b, :]) for b in range(D)] for a in range(D)]
    # row vector over w' for the partial product of sites 0..i
    cur = [ops[D - 1][b] for b in range(D)]
    for _ in range(1, L):
        nxt = []
        for b in range(D):
            acc = None
            for a in range(D):
                if ops[a][b].nnz == 0 or cur[a].nnz == 0:
                    continue
                term = sp.kron(cur[a], ops[a][b], format="csr")
                acc = term if acc is None else acc + term
            if acc is None:
                acc = sp.csr_matrix((cur[0].shape[0] * d, cur[0].shape[1] * d), dtype=W.dtype)
            nxt.append(acc)
        cur = nxt
    return cur[0].tocsr()


def mpo_edge_dense(W, L, d):
    """The boundary term that a periodic chain's two edge tensors W.left()
    (on site 0) and W.right() (on site L-1) carry between them through
    identity operators: sum_w W[D-1, :, w, :] (site 0) x 1 x ... x 1 x
    W[w, :, 0, :] (site L-1).  Each term of W starts on site 0 and ends on
    site L-1: the bond (0, L-1) with the orientation of an open bond (i, i+1),
    not of the ring bond (L-1, 0); the two agree for bonds symmetric under
    exchanging the sites (XYZ) and differ for the chiral Potts bond."""
    D = W.shape[0]
    mid = sp.identity(d ** (L - 2), format="csr")
    H = None
    for w in range(D):
        a, b = W[D - 1, :, w, :], W[w, :, 0, :]
        if not a.any() or not b.any():
            continue
        t = sp.kron(sp.kron(sp.csr_matrix(a), mid, format="csr"), sp.csr_matrix(b), format="csr")
        H = t if H is None else H + t
    return H if H is not None else sp.csr_matrix((d ** L, d ** L), dtype=W.dtype)


def z_n_sector_indices(L, d, n, sector):
    """Basis indices (site 0 most significant) with sum of local charges % n == sector."""
    idx = np.arange(d ** L)
    charge = np.zeros_like(idx)
    x = idx.copy()
    for _ in range(L):
        charge += x % d
        x //= d
    return idx[(charge % n) == sector]


def mpo_e0(W, L, d, n, sector, k=1):
    H = mpo_dense(W, L, d)
    sel = z_n_sector_indices(L, d, n, sector)
    Hs = H[sel][:, sel]
    if Hs.shape[0] <= 400:
        A = Hs.toarray()
        w = np.linalg.eigvalsh(0.5 * (A + A.conj().T))
        return np.sort(w)[:k]
    w = spla.eigsh(Hs, k=k, which="SA", tol=1e-12)[0]
    return np.sort(w)
