"""numpy/scipy restatement of the reference's DMRG steps and sweep schedules.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The eigensolver is scipy's ARPACK (the same dsaupd/znaupd the reference links,
src/dsaupd.cpp:65-155, src/znaupd.cpp:78-197) with the reference's
parameters; tiny problems are diagonalised densely.  The SVD is LAPACK
?gesvd through scipy (oracle.qtensor.tensor_svd).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg
import scipy.sparse.linalg as spla

from . import qtensor as qt
from .qtensor import NEV, NSI, TOL, QTensor, generate_dim_sym, get_dimension


class MPS:
    """mps<T>, include/mps.h:9-75: state_ (site tensors), store_ (environments)
    and the centre singular values."""

    def __init__(self, L):
        self.state = [None] * L
        self.store = [None] * L
        self.singular = None
        self.cut = 0

    def size(self):
        return len(self.state)


@dataclass
class StepLog:
    l: int
    r: int
    energies: list
    singular: list
    entropy: float
    trunc_err: float
    n_matvec: int
    dim: int


@dataclass
class Lanczos:
    """lanczos<T>, include/val_for_lanczos.h:7-24."""
    lenv: QTensor = None
    renv: QTensor = None
    lmap: list = field(default_factory=list)
    rmap: list = field(default_factory=list)
    r_num: int = 0
    l_num: int = 0
    d_store: int = 0
    n_matvec: int = 0


def av(inp, ctx: Lanczos):
    """src/core.cpp:22-31: store = R * in (rmap), out = L * store (lmap)."""
    store = np.zeros(ctx.d_store, dtype=inp.dtype)
    out = np.zeros(inp.shape[0], dtype=inp.dtype)
    ctx.renv.contract_map(store, inp, ctx.rmap, "N", "T", ctx.r_num)
    ctx.lenv.contract_map(out, store, ctx.lmap, "N", "T", ctx.l_num)
    ctx.n_matvec += 1
    return out


def dense_heff(ctx: Lanczos, n, dtype):
    """Materialise H_eff column by column (as check_hermitian does, src/core.cpp:40-83)."""
    H = np.zeros((n, n), dtype=dtype)
    e = np.zeros(n, dtype=dtype)
    for j in range(n):
        e[:] = 0
        e[j] = 1
        H[:, j] = av(e, ctx)
    return H


def znaupd(n, ctx: Lanczos, dtype, seed=1234, nev=NEV, hermitian=True):
    """Lowest eigenpairs with the reference's ARPACK settings:
    real: dsaupd which=SA, ncv=4*nev, tol=1e-10, maxiter=3n (src/dsaupd.cpp:65-91);
    complex: znaupd which=SR, ncv=6*nev (src/znaupd.cpp:78-111), sorted by Re (:170-183).
    Returns (nev eigenvalues ascending, lowest eigenvector)."""
    complex_ = np.dtype(dtype).kind == "c"
    ncv = (6 if complex_ else 4) * nev
    ncv = min(ncv, n)
    if n <= max(2 * ncv, nev + 2):
        H = dense_heff(ctx, n, dtype)
        if hermitian:
            w, v = scipy.linalg.eigh(0.5 * (H + H.conj().T))
        else:
            w, v = scipy.linalg.eig(H)
            order = np.argsort(w.real)
            w, v = w[order].real, v[:, order]
        vals = np.zeros(nev)
        k = min(nev, n)
        vals[:k] = w[:k]
        vec = v[:, 0] / np.linalg.norm(v[:, 0])
        return vals, vec.astype(dtype)
    rng = np.random.default_rng(seed)
    v0 = rng.standard_normal(n)
    if complex_:
        v0 = v0 + 1j * rng.standard_normal(n)
    op = spla.LinearOperator((n, n), matvec=lambda x: av(np.asarray(x, dtype=dtype).ravel(), ctx),
                             dtype=dtype)
    if not complex_:
        w, v = spla.eigsh(op, k=nev, which="SA", ncv=ncv, tol=TOL, maxiter=3 * n, v0=v0)
        order = np.argsort(w)
    else:
        w, v = spla.eigs(op, k=nev, which="SR", ncv=ncv, tol=TOL, maxiter=3 * n, v0=v0)
        order = np.argsort(w.real)
        w = w.real
    w, v = w[order], v[:, order]
    return np.array(w), v[:, 0].astype(dtype)


def _log_singular(s, cutoff):
    """print_singular, src/useful.cpp:75-102: entropy and truncation error."""
    ent = 0.0
    for x in s:
        ent += -x * x * math.log(x * x + TOL)
    err = 1.0
    c = min(len(s), cutoff)
    for i in range(c):
        err -= s[i] * s[i]
    return ent, err


class Runner:
    """Step functions of src/core.cpp over an MPS/MPO pair.

    ``sign_rule``: "reference" applies get_map's alpha unconditionally (the
    reference's behaviour); "physical" applies it only when the model is
    fermionic (alpha=1 for bosonic models).
    """

    def __init__(self, mpo, sym_n, fermion, sector=0, sign_rule="physical", seed=1234):
        self.mpo = mpo
        self.sym_n = sym_n
        self.fermion = fermion
        self.sector = sector
        self.sign_rule = sign_rule
        self.seed = seed
        self.log: list[StepLog] = []
        qt.set_symmetry(sym_n)
        self.mps = MPS(mpo.size())

    @property
    def alpha_on(self):
        return self.sign_rule == "reference" or self.fermion

    def _solve(self, ctx, lnum, rnum, allow_ed):
        dim, sym = generate_dim_sym(ctx.lenv.dimension_all(), lnum, ctx.renv.dimension_all(), rnum,
                                    self.sector)
        d = get_dimension(dim, sym)
        dtype = np.result_type(ctx.lenv.dtype, ctx.renv.dtype)
        return dim, sym, d, dtype

    def _lanczos(self, ctx, dim, sym, d, dtype, r_num, l_num):
        ctx.r_num, ctx.l_num = r_num, l_num
        ctx.rmap, dim_ret, sym_ret = ctx.renv.get_map(dim, sym, "N", "T", r_num, 0, False,
                                                      fermion=self.alpha_on)
        ctx.d_store = get_dimension(dim_ret, sym_ret)
        ctx.lmap, _, _ = ctx.lenv.get_map(dim_ret, sym_ret, "N", "T", l_num, 0, True, sym_ret=sym,
                                          fermion=self.alpha_on)
        herm = not (self.sign_rule == "reference" and not self.fermion)
        vals, vec = znaupd(d, ctx, dtype, seed=self.seed + len(self.log), hermitian=herm)
        return vals, vec

    def _ed(self, whole, d):
        whole = whole.exchange(2, 3).combine(2, 3).combine(0, 1)
        w, v = whole.eig(self.sector)
        vals = np.zeros(NEV)
        k = min(NEV, len(w))
        vals[:k] = w[:k]
        return vals, v

    def _finish(self, l, r, vals, vec, dim, sym, num, cutoff, d, nmv):
        vecq = QTensor.from_flat(vec, dim, sym)
        U, S, V, s = vecq.svd(num, cutoff)
        ent, err = _log_singular(s, cutoff)
        self.log.append(StepLog(l, r, list(vals), list(s[:NSI]), ent, err, nmv, d))
        return U, S, V, s

    def two_sites(self, l, r, cutoff):
        """src/core.cpp:88-172."""
        m, mps = self.mpo, self.mps
        ctx = Lanczos(lenv=m[l].copy(), renv=m[r].copy())
        dim, sym, d, dtype = self._solve(ctx, 1, 1, True)
        if d < NEV:
            whole = QTensor(dtype=dtype).contract_index(ctx.lenv, 1, ctx.renv, 1).simplify()
            vals, vec = self._ed(whole, d)
        else:
            vals, vec = self._lanczos(ctx, dim, sym, d, dtype, 1, 2)
        U, S, V, s = self._finish(l, r, vals, vec, dim, sym, 1, cutoff, d, ctx.n_matvec)
        mps.state[l] = U
        mps.singular = S
        mps.cut = l + 1
        mps.state[r] = V
        vecq = QTensor().contract_gemm(ctx.lenv, U, "N", "N")
        U.conjugate()
        mps.store[l] = QTensor().contract_gemm(U, vecq, "T", "N")
        vecq = QTensor().contract_gemm(ctx.renv, V, "N", "T")
        V.conjugate()
        mps.store[r] = QTensor().contract_gemm(V, vecq, "N", "N")
        return vals[0]

    def _lenv(self, l):
        le = QTensor().contract_index(self.mps.store[l - 1], 1, self.mpo[l], 0)
        return le.exchange(3, 4)

    def _renv(self, r):
        ope = self.mpo[r].exchange(0, 2)
        re = QTensor().contract_index(self.mps.store[r + 1], 1, ope, 0, True)
        return re.exchange(0, 1)

    def update_two(self, l, r, cutoff):
        """src/core.cpp:182-283 (two-site L.W.W.R step)."""
        mps = self.mps
        ctx = Lanczos(lenv=self._lenv(l), renv=self._renv(r))
        dim, sym, d, dtype = self._solve(ctx, 2, 2, False)
        vals, vec = self._lanczos(ctx, dim, sym, d, dtype, 2, 3)
        U, S, V, s = self._finish(l, r, vals, vec, dim, sym, 2, cutoff, d, ctx.n_matvec)
        mps.state[l] = U
        mps.singular = S
        mps.cut = l + 1
        mps.state[r] = V
        vecq = QTensor().contract_gemm(ctx.lenv, U, "N", "N", 2)
        U.conjugate()
        mps.store[l] = QTensor().contract_gemm(U, vecq, "T", "N", 2)
        vecq = QTensor().contract_gemm(ctx.renv, V, "N", "T", 2)
        V.conjugate()
        mps.store[r] = QTensor().contract_gemm(V, vecq, "N", "N", 2)
        return vals[0]

    def move2right(self, l, r, cutoff):
        """src/core.cpp:293-378 (single explicit site l, renv = store(r))."""
        mps = self.mps
        ctx = Lanczos(lenv=self._lenv(l), renv=mps.store[r].copy())
        dim, sym, d, dtype = self._solve(ctx, 2, 1, True)
        if d < NEV:
            tl = ctx.lenv.combine(3, 4).combine(0, 1)
            whole = QTensor(dtype=dtype).contract_index(tl, 1, ctx.renv, 1).simplify()
            vals, vec = self._ed(whole, d)
        else:
            vals, vec = self._lanczos(ctx, dim, sym, d, dtype, 1, 3)
        U, S, V, s = self._finish(l, r, vals, vec, dim, sym, 2, cutoff, d, ctx.n_matvec)
        mps.state[l] = U
        mps.singular = S
        mps.cut = l + 1
        vecq = QTensor().contract_gemm(ctx.lenv, U, "N", "N", 2)
        U.conjugate()
        mps.store[l] = QTensor().contract_gemm(U, vecq, "T", "N", 2)
        return vals[0]

    def move2left(self, l, r, cutoff):
        """src/core.cpp:388-474 (single explicit site r, lenv = store(l))."""
        mps = self.mps
        ctx = Lanczos(lenv=mps.store[l].copy(), renv=self._renv(r))
        dim, sym, d, dtype = self._solve(ctx, 1, 2, True)
        if d < NEV:
            tr = ctx.renv.combine(3, 4).combine(0, 1)
            whole = QTensor(dtype=dtype).contract_index(ctx.lenv, 1, tr, 1).simplify()
            vals, vec = self._ed(whole, d)
        else:
            vals, vec = self._lanczos(ctx, dim, sym, d, dtype, 2, 2)
        U, S, V, s = self._finish(l, r, vals, vec, dim, sym, 1, cutoff, d, ctx.n_matvec)
        mps.singular = S
        mps.cut = l + 1
        mps.state[r] = V
        vecq = QTensor().contract_gemm(ctx.renv, V, "N", "T", 2)
        V.conjugate()
        mps.store[r] = QTensor().contract_gemm(V, vecq, "N", "N", 2)
        return vals[0]

    # ---------------------------------------------------------------- sweeps
    def dmrg(self, cutoff, sweep):
        """Single-site schedule, src/functions.cpp:157-227."""
        L = self.mps.size()
        pre_cutoff = cutoff if sweep == -1 else 10
        pre_sweep = cutoff // 100
        e = self.two_sites(0, L - 1, cutoff * 10)
        for i in range(1, L // 2):
            e = self.update_two(i, L - i - 1, pre_cutoff)
        if sweep != -1:
            for i in range(L // 2, L - 1):
                e = self.move2right(i, i + 1, pre_cutoff)
            for s in range(pre_sweep - 1):
                for i in range(L - 2, 0, -1):
                    e = self.move2left(i - 1, i, 100 * (s + 1))
                for i in range(1, L - 1):
                    e = self.move2right(i, i + 1, 100 * (s + 1))
        for s in range(max(sweep, 0)):
            for i in range(L - 2, 0, -1):
                e = self.move2left(i - 1, i, cutoff)
            for i in range(1, L - 1):
                e = self.move2right(i, i + 1, cutoff)
        if sweep != -1:
            for i in range(L - 2, L // 2 - 1, -1):
                e = self.move2left(i - 1, i, cutoff)
        return e

    def dmrg2(self, cutoff, sweep, ladder=True):
        """Two-site schedule, src/functions.cpp:234-299.  ``ladder=False``
        skips the pre-sweep ladder (cutoff 100, 200, ...)."""
        L = self.mps.size()
        pre_cutoff = cutoff if sweep == -1 else 10
        pre_sweep = cutoff // 100 if ladder else 0
        e = self.two_sites(0, L - 1, 0)
        for i in range(1, L // 2):
            e = self.update_two(i, L - i - 1, pre_cutoff)
        if sweep != -1:
            for i in range(L // 2, L - 2):
                e = self.update_two(i, i + 1, pre_cutoff)
            for s in range(pre_sweep - 1):
                for i in range(L - 3, 1, -1):
                    e = self.update_two(i - 1, i, 100 * (s + 1))
                for i in range(2, L - 2):
                    e = self.update_two(i, i + 1, 100 * (s + 1))
        for s in range(max(sweep, 0)):
            for i in range(L - 3, 1, -1):
                e = self.update_two(i - 1, i, cutoff)
            for i in range(2, L - 2):
                e = self.update_two(i, i + 1, cutoff)
        if sweep != -1:
            for i in range(L - 3, L // 2 - 1, -1):
                e = self.update_two(i - 1, i, cutoff)
        return e


# ---------------------------------------------------------------------------
# Observables: mps::prep_calc and the three calc functions.


def prepared(mps: MPS):
    """mps::prep_calc (src/mps.cpp:33-54) on a copy: the centre singular
    values absorbed into the site at the cut; returns the site tensors."""
    st = list(mps.state)
    S = mps.singular.copy()
    site = st[mps.cut]
    st[mps.cut] = QTensor().contract_gemm(S if site.dtype == np.float64 else S.comp(), site)
    return st


def calc_overlap(l_st, r_st, fermion=True):
    """calc(mps, mps) = <r|l> (src/functions.cpp:428-460)."""
    n = len(l_st)
    bra = r_st[0].copy()
    bra.conjugate()
    left = QTensor().contract_gemm(bra, l_st[0], "T", "N", 1)
    for i in range(1, n - 1):
        end = QTensor().contract_gemm(left, l_st[i], "N", "N", 1)
        bra = r_st[i].copy()
        bra.conjugate()
        left = QTensor().contract_gemm(bra, end, "T", "N", 2)
    bra = r_st[n - 1].copy()
    bra.conjugate()
    end = QTensor().contract_gemm(bra, l_st[n - 1], "N", "T", 1)
    return left.trace(end, fermion)


def calc_expect(st, mpo, fermion=True):
    """calc(mps, mpo) = <m|O|m> (src/functions.cpp:467-511): the transfer
    matrix with the MPO leg, contracted with change = true at every interior
    site, simplified, and closed by the graded trace."""
    n = len(st)
    end = QTensor().contract_gemm(mpo[0], st[0], "N", "N", 1)
    tmp = st[0].copy()
    tmp.conjugate()
    left = QTensor().contract_gemm(tmp, end, "T", "N", 1)
    for i in range(1, n - 1):
        tmp = QTensor().contract_index(left, 1, mpo[i], 0, True, fermion)
        left = tmp.exchange(3, 4)
        end = QTensor().contract_gemm(left, st[i], "N", "N", 2)
        tmp = st[i].copy()
        tmp.conjugate()
        left = QTensor().contract_gemm(tmp, end, "T", "N", 2).simplify()
    store = QTensor().contract_gemm(mpo[n - 1], st[n - 1], "N", "T", 1)
    tmp = st[n - 1].copy()
    tmp.conjugate()
    end = QTensor().contract_gemm(tmp, store, "N", "N", 1).simplify()
    return left.trace(end, fermion)


def exact_chain_spectrum(mpo, sector, k=1, edges=None, alpha=True, tol=1e-12):
    """The k lowest eigenvalues of a four-tensor chain from its exact two-site
    operator on tensors (1, 2): the stores of tensors 0 and 3 are the MPO edge
    tensors themselves (no basis rotation, nothing truncated), so H_eff is the
    whole Hamiltonian (update_two's environments, src/core.cpp:188-192, and
    av, src/core.cpp:22-31).  `edges` = (edge of tensor 0, edge of tensor 3)
    adds the periodic edge term as update_two does under PBC
    (src/core.cpp:193-199, 241-247): both edge environments through the
    identity operator built from the left edge, the edge R pass over its own
    store layout and the edge L pass accumulated onto the result."""
    assert mpo.size() == 4
    lenv = QTensor().contract_index(mpo[0], 1, mpo[1], 0).exchange(3, 4)
    renv = QTensor().contract_index(mpo[3], 1, mpo[2].exchange(0, 2), 0, True).exchange(0, 1)
    dim, sym = generate_dim_sym(lenv.dimension_all(), 2, renv.dimension_all(), 2, sector)
    n = get_dimension(dim, sym)
    ctx = Lanczos(lenv=lenv, renv=renv, r_num=2, l_num=3)
    ctx.rmap, dim_ret, sym_ret = renv.get_map(dim, sym, "N", "T", 2, 0, False, fermion=alpha)
    ctx.d_store = get_dimension(dim_ret, sym_ret)
    ctx.lmap, _, _ = lenv.get_map(dim_ret, sym_ret, "N", "T", 3, 0, True, sym_ret=sym, fermion=alpha)
    ops = [ctx]
    if edges is not None:
        e0, e3 = edges
        ope = e0.id(mpo[1])
        ledge = QTensor().contract_index(e0, 1, ope, 0).exchange(3, 4)
        redge = QTensor().contract_index(e3, 1, ope, 0, True).exchange(0, 1)
        ectx = Lanczos(lenv=ledge, renv=redge, r_num=2, l_num=3)
        ectx.rmap, edim_ret, esym_ret = redge.get_map(dim, sym, "N", "T", 2, 0, False, fermion=alpha)
        ectx.d_store = get_dimension(edim_ret, esym_ret)
        ectx.lmap, _, _ = ledge.get_map(edim_ret, esym_ret, "N", "T", 3, 0, True, sym_ret=sym, fermion=alpha)
        ops.append(ectx)
    dtype = np.result_type(lenv.dtype, renv.dtype)
    op = spla.LinearOperator((n, n), matvec=lambda x: sum(av(np.asarray(x, dtype=dtype).ravel(), c) for c in ops),
                             dtype=dtype)
    w = spla.eigsh(op, k=k, which="SA", tol=tol, v0=np.ones(n, dtype=dtype))[0]
    return np.sort(w), n
