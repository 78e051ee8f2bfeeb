"""Bit-exact parity of the library's integer bookkeeping (get_map tuples,
dim_ret/sym_ret, generate_dim_sym) with the oracle's restatement of
src/qtensor.cpp:997-1119 and src/useful.cpp:151-194.  Structure only: no GPU."""
import itertools

import numpy as np
import pytest

from dmrg_amd import driver
from oracle import qtensor as qt
from oracle.qtensor import QTensor, generate_dim_sym


def _random_structure(rng, S, nleg, max_dim=4, fill=0.7, dirs=None):
    dim = [[int(rng.integers(0, max_dim + 1)) for _ in range(S)] for _ in range(nleg)]
    for d in dim:
        if sum(d) == 0:
            d[0] = 1
    syms = [list(s) for s in itertools.product(range(S), repeat=nleg) if rng.random() < fill]
    t = QTensor(*([S] * nleg))
    t.dim = dim
    t.sym = syms
    t.dir = list(dirs) if dirs is not None else [int(x) for x in rng.integers(-1, 2, size=nleg)]
    return t


@pytest.mark.parametrize("S", [2, 3])
@pytest.mark.parametrize("seed", range(6))
def test_get_map_two_site_passes_bit_exact(S, seed):
    rng = np.random.default_rng(100 * S + seed)
    qt.set_symmetry(S)
    # renv [s, b, w, s', b'], lenv [a, s, w, a', s'] with random sector dims / sparsity
    R = _random_structure(rng, S, 5, fill=0.8)
    R.dim[2] = [int(rng.integers(1, 6))]
    R.sym = [s for s in R.sym if s[2] == 0] or [[0] * 5]
    Lt = _random_structure(rng, S, 5, fill=0.8)
    Lt.dim[2] = list(R.dim[2])
    Lt.sym = [s for s in Lt.sym if s[2] == 0] or [[0] * 5]
    Lt.dim[3], Lt.dim[4] = list(Lt.dim[0]), list(Lt.dim[1])
    R.dim[3], R.dim[4] = list(R.dim[0]), list(R.dim[1])
    for sector in range(S):
        dim, sym = generate_dim_sym(Lt.dim, 2, R.dim, 2, sector)
        d2, s2 = driver.generate_dim_sym(S, Lt.dim, Lt.sym, Lt.dir, 2, R.dim, R.sym, R.dir, 2, sector)
        assert d2 == dim and s2 == sym
        if not sym:
            continue
        for fermion in (True, False):
            rmap, dim_ret, sym_ret = R.get_map(dim, sym, "N", "T", 2, 0, False, fermion=fermion)
            m2, dr2, sr2 = driver.get_map(S, R.dim, R.sym, R.dir, dim, sym, "N", "T", 2, 0, False,
                                          fermion=fermion)
            assert m2 == rmap and dr2 == dim_ret and sr2 == sym_ret
            lmap, dim_b, _ = Lt.get_map(dim_ret, sym_ret, "N", "T", 3, 0, True, sym_ret=sym, fermion=fermion)
            m3, dr3, sr3 = driver.get_map(S, Lt.dim, Lt.sym, Lt.dir, dim_ret, sym_ret, "N", "T", 3, 0, True,
                                          sym_ret=sym, fermion=fermion)
            assert m3 == lmap and dr3 == dim_b and sr3 == sym


def test_get_map_invariants_xyz_m512_shape():
    """Every output block gets exactly one beta=0 task and it comes first;
    offsets are prefix sums; the flop count matches 4*D*S^2*m^3*(c/2)."""
    qt.set_symmetry(2)
    S, D, m = 2, 5, 512
    bond = [m // 2] * 2
    dims_R = [[1, 1], bond, [D], [1, 1], bond]
    R = QTensor(2, 2, 1, 2, 2)
    R.dim = dims_R
    R.sym = [[a, b, 0, c, e] for a in range(2) for b in range(2) for c in range(2) for e in range(2)]
    R.dir = [0, -1, 0, 0, 1]
    Lt = QTensor(2, 2, 1, 2, 2)
    Lt.dim = [bond, [1, 1], [D], bond, [1, 1]]
    Lt.sym = list(R.sym)
    Lt.dir = [-1, 0, 0, 1, 0]
    dim, sym = driver.generate_dim_sym(2, Lt.dim, Lt.sym, Lt.dir, 2, R.dim, R.sym, R.dir, 2, 0)
    rmap, dim_ret, sym_ret = driver.get_map(2, R.dim, R.sym, R.dir, dim, sym, "N", "T", 2, 0, False)
    lmap, _, _ = driver.get_map(2, Lt.dim, Lt.sym, Lt.dir, dim_ret, sym_ret, "N", "T", 3, 0, True, sym_ret=sym)
    assert len(rmap) == 32 and len(lmap) == 32
    for maps in (rmap, lmap):
        seen = set()
        for t in maps:
            if t[2] not in seen:
                assert t[5] == 0
                seen.add(t[2])
            else:
                assert t[5] == 1
    # M = |A_i| / col
    sizes_R = [int(np.prod([R.dim[k][s[k]] for k in range(5)])) for s in R.sym]
    sizes_L = [int(np.prod([Lt.dim[k][s[k]] for k in range(5)])) for s in Lt.sym]
    flops = sum(8 * (sizes_R[t[0]] // t[4]) * t[3] * t[4] for t in rmap)
    flops += sum(8 * (sizes_L[t[0]] // t[4]) * t[3] * t[4] for t in lmap)
    assert flops == 4 * D * S ** 2 * m ** 3 * 4
