"""Synthetic two-site environments with the block structure of the
reference's update_two step and the oracle's av() over them.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): used by the tests and by
the cpu_baseline leg of bench.py.
"""
from __future__ import annotations

import numpy as np

from . import qtensor as qt
from .dmrg import Lanczos, av
from .qtensor import QTensor, generate_dim_sym, get_dimension


def synthetic_env(S, D, bond, dtype, rng, legs="R"):
    """A dense-in-w two-site environment (R: [s, b, w, s', b']; L: [a, s, w, a', s'])
    with every sector combination present, physical sectors of dimension 1."""
    phys = [1] * S
    if legs == "R":
        dims = [phys, list(bond), [D], phys, list(bond)]
    else:
        dims = [list(bond), phys, [D], list(bond), phys]
    t = QTensor(*[len(d) for d in dims], dtype=dtype)
    t.dim = [list(d) for d in dims]
    for a in range(S):
        for b in range(S):
            for c in range(S):
                for e in range(S):
                    sym = [a, b, 0, c, e]
                    shape = [dims[i][sym[i]] for i in range(5)]
                    v = rng.standard_normal(shape)
                    if np.dtype(dtype).kind == "c":
                        v = v + 1j * rng.standard_normal(shape)
                    t.sym.append(sym)
                    t.val.append(v.astype(dtype))
    t.dir = [0, -1, 0, 0, 1] if legs == "R" else [-1, 0, 0, 1, 0]
    return t


def flat_blocks(t: QTensor):
    sizes = np.array([v.size for v in t.val], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    return t.flat(), offs, sizes


class HeffCase:
    """Everything dmrgx_heff_create needs plus the oracle's av()."""

    def __init__(self, lenv, renv, sector=0, r_num=2, l_num=3, lnum=2, rnum=2, fermion=True):
        self.lenv, self.renv = lenv, renv
        self.dim, self.sym = generate_dim_sym(lenv.dimension_all(), lnum, renv.dimension_all(), rnum, sector)
        self.n_psi = get_dimension(self.dim, self.sym)
        self.rmap, dim_ret, sym_ret = renv.get_map(self.dim, self.sym, "N", "T", r_num, 0, False,
                                                   fermion=fermion)
        self.n_store = get_dimension(dim_ret, sym_ret)
        self.lmap, _, _ = lenv.get_map(dim_ret, sym_ret, "N", "T", l_num, 0, True, sym_ret=self.sym,
                                       fermion=fermion)
        self.ctx = Lanczos(lenv=lenv, renv=renv, lmap=self.lmap, rmap=self.rmap, r_num=r_num, l_num=l_num,
                           d_store=self.n_store)
        self.dtype = np.result_type(lenv.dtype, renv.dtype)
        self.rflat, self.roff, self.rsize = flat_blocks(renv)
        self.lflat, self.loff, self.lsize = flat_blocks(lenv)

    def oracle_apply(self, x):
        return av(np.asarray(x, dtype=self.dtype), self.ctx)

    def flops(self):
        c = 8 if np.dtype(self.dtype).kind == "c" else 2
        f = 0
        for maps, t in ((self.rmap, self.renv), (self.lmap, self.lenv)):
            for (i, ib, ob, row, col, beta, alpha) in maps:
                f += c * (t.val[i].size // col) * row * col
        return f


def synthetic_case(S, D, m, dtype=np.complex128, seed=20261203, uneven=None, fermion=True):
    rng = np.random.default_rng(seed)
    bond = uneven if uneven is not None else [m // S] * S
    R = synthetic_env(S, D, bond, dtype, rng, "R")
    Lt = synthetic_env(S, D, bond, dtype, rng, "L")
    qt.set_symmetry(S)
    return HeffCase(Lt, R, fermion=fermion)


