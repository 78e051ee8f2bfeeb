"""numpy restatement of the reference's symmetry algebra and block-sparse tensor.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Each function cites the
reference file:line it restates.  Blocks are numpy arrays in row-major order
with shape == the per-leg sector dimensions, exactly as ``tensor<T>::val_``
with ``index_`` (include/tensor.h:24-25).
"""
from __future__ import annotations

import copy
import itertools

import numpy as np
import scipy.linalg

TOL = 1e-10          # global.h:19
NEV = 6              # global.h:21
NSI = 20             # global.h:22
PI = 3.14159265359   # global.h:25 (truncated pi, used by H_Potts3)

# Z_N group order; the reference fixes it at compile time (global.h:16).
# S4E = -8 selects the Z2 x Z2 x Z2 group of the reference's S4E build.
S4E = -8
_SYMMETRY = 2
_S4E = False
# src/useful.cpp:9-16 and 24: the fusion table and the fermion parities of
# the eight Z2^3 sectors (bits 000 .. 111)
_S4E_TABLE = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 3, 2, 5, 4, 7, 6], [2, 3, 0, 1, 6, 7, 4, 5], [3, 2, 1, 0, 7, 6, 5, 4],
              [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 7, 6, 1, 0, 3, 2], [6, 7, 4, 5, 2, 3, 0, 1], [7, 6, 5, 4, 3, 2, 1, 0]]
_S4E_PARITY = [0, 0, 1, 1, 1, 1, 0, 0]


def set_symmetry(n: int) -> None:
    global _SYMMETRY, _S4E
    _S4E = int(n) == S4E
    _SYMMETRY = 8 if _S4E else int(n)


def symmetry() -> int:
    return _SYMMETRY


def ss(a: int, b: int) -> int:
    """Charge fusion, src/useful.cpp:30: (a+b+N)%N; the S4E table lookup
    p[a][|b|] (src/useful.cpp:6-20) for Z2^3."""
    if _S4E:
        return _S4E_TABLE[a][-b if b < 0 else b]
    return (a + b + _SYMMETRY) % _SYMMETRY


def ff(a: int) -> int:
    """Fermion parity of a charge, src/useful.cpp:32: a%2; the S4E table
    (src/useful.cpp:22-26) for Z2^3."""
    if _S4E:
        return _S4E_PARITY[a]
    return a % 2


def generate_map(index):
    """All multi-indices, last index fastest (src/useful.cpp:119-134)."""
    return [list(t) for t in itertools.product(*[range(i) for i in index])]


def get_dimension(dim, sym) -> int:
    """src/useful.cpp:137-148."""
    d = 0
    for s in sym:
        t = 1
        for j in range(len(dim)):
            t *= dim[j][s[j]]
        d += t
    return d


def generate_dim_sym(ldim_all, lnum, rdim_all, rnum, sector):
    """psi sector list (src/useful.cpp:151-194).

    ``ldim_all``/``rdim_all`` are ``dimension_all()`` of lenv/renv.
    """
    dim = [list(ldim_all[len(ldim_all) - lnum + i]) for i in range(lnum)]
    dim += [list(rdim_all[len(rdim_all) - rnum + i]) for i in range(rnum)]
    lmap = generate_map([len(dim[i]) for i in range(lnum)])
    rmap = generate_map([len(dim[lnum + i]) for i in range(rnum)])
    sym = []
    if sector != -1:
        for s in range(_SYMMETRY):
            t = ss(sector, -s)
            for lt in lmap:
                lsym = 0
                for v in lt:
                    lsym = ss(lsym, v)
                for rt in rmap:
                    rsym = 0
                    for v in rt:
                        rsym = ss(rsym, v)
                    if lsym == s and rsym == t:
                        sym.append(list(lt) + list(rt))
    else:
        sym = [[0] * len(dim)]
    return dim, sym


def _prod(xs):
    p = 1
    for x in xs:
        p *= x
    return p


class QTensor:
    """Restatement of ``qtensor<T>`` (include/qtensor.h:21-31)."""

    def __init__(self, *index, dtype=np.float64):
        # qtensor(i0..i7) ctor: src/qtensor.cpp:4-21
        self.dim = [[0] * i for i in index]
        self.sym: list[list[int]] = []
        self.val: list[np.ndarray] = []
        self.dir = [0] * len(index)
        self.dtype = np.dtype(dtype)

    # ------------------------------------------------------------ basics
    def copy(self) -> "QTensor":
        r = QTensor(dtype=self.dtype)
        r.dim = copy.deepcopy(self.dim)
        r.sym = copy.deepcopy(self.sym)
        r.dir = list(self.dir)
        r.val = [v.copy() for v in self.val]
        return r

    @staticmethod
    def from_flat(flat, dim, sym, dtype=None):
        """qtensor(T* val, dim, sym), src/qtensor.cpp:24-48 (dir_ all 0)."""
        flat = np.asarray(flat)
        r = QTensor(dtype=dtype or flat.dtype)
        r.dim = copy.deepcopy(dim)
        r.sym = copy.deepcopy(sym)
        r.dir = [0] * len(dim)
        count = 0
        for s in sym:
            shape = [dim[j][s[j]] for j in range(len(dim))]
            d = _prod(shape)
            r.val.append(np.array(flat[count:count + d], dtype=r.dtype).reshape(shape))
            count += d
        return r

    def _check_dim(self, n):
        """check_dim_, include/qtensor.h:34-45."""
        for i in range(len(self.dim)):
            s = self.sym[n][i]
            if self.dim[i][s] == 0:
                self.dim[i][s] = self.val[n].shape[i]
            elif self.dim[i][s] != self.val[n].shape[i]:
                raise ValueError("check_dim mismatch")

    def update(self, block, *loc):
        """src/qtensor.cpp:51-85 (empty blocks are skipped by the reference:
        ``tensor::empty`` is true when the sum of |values| is below TOL,
        src/tensor.cpp:176-182)."""
        block = np.asarray(block, dtype=self.dtype)
        if np.sum(np.abs(block)) < TOL:
            return
        assert len(loc) == len(self.dim)
        self.sym.append(list(loc))
        self.val.append(block.copy())
        self._check_dim(len(self.sym) - 1)

    def add_sign(self, *dirs):
        """src/qtensor.cpp:184-202."""
        self.dir = list(dirs)
        assert len(self.dir) == len(self.dim)

    def u_dim(self, *dims):
        """src/qtensor.cpp:90-124: every sector of leg i gets dimension dims[i]."""
        assert len(dims) == len(self.dim)
        for i, d in enumerate(dims):
            for j in range(len(self.dim[i])):
                assert self.dim[i][j] in (0, d)
                self.dim[i][j] = d

    def _identity(self):
        for s0 in range(len(self.dim[0])):
            for s1 in range(len(self.dim[1])):
                d0, d1 = self.dim[0][s0], self.dim[1][s1]
                self.sym.append([s0, s1, s0, s1])
                self.val.append(np.einsum("ac,bd->abcd", np.eye(d0), np.eye(d1)).astype(self.dtype))
        return self

    def id(self, bulk=None):
        """qtensor::id() (src/qtensor.cpp:128-154) and id(bulk_mpo)
        (src/qtensor.cpp:157-182): identity blocks with legs (0, 2) and (1, 3)
        paired; with a bulk MPO, its legs with bond legs 0 and 2 given this
        tensor's leg-1 sectors."""
        r = QTensor(dtype=self.dtype)
        if bulk is None:
            assert self.dim[0] == self.dim[2] and self.dim[1] == self.dim[3]
            r.dim, r.dir = copy.deepcopy(self.dim), list(self.dir)
        else:
            r.dim = copy.deepcopy(bulk.dim)
            r.dim[0], r.dim[2] = list(self.dim[1]), list(self.dim[1])
            r.dir = list(bulk.dir)
            r.dir[0] = r.dir[2] = self.dir[1]
        return r._identity()

    def dimension_all(self):
        return copy.deepcopy(self.dim)

    def dimension(self, n=None):
        if n is None:
            return _prod([sum(d) for d in self.dim])
        return sum(self.dim[n])

    def flat(self):
        """``val(out,dim,sym)`` (include/qtensor.h:99-106)."""
        if not self.val:
            return np.zeros(0, dtype=self.dtype)
        return np.concatenate([v.ravel() for v in self.val])

    def conjugate(self):
        """src/qtensor.cpp:297-303: flips dir_ and conjugates in place."""
        self.dir = [-d for d in self.dir]
        self.val = [np.conj(v) for v in self.val]
        return self.copy()

    def comp(self):
        r = self.copy()
        r.dtype = np.dtype(np.complex128)
        r.val = [v.astype(np.complex128) for v in r.val]
        return r

    def simplify(self):
        """src/qtensor.cpp:327-346."""
        r = QTensor(dtype=self.dtype)
        r.dir = list(self.dir)
        r.dim = copy.deepcopy(self.dim)
        for s, v in zip(self.sym, self.val):
            if np.any(np.abs(v) > TOL):
                r.sym.append(list(s))
                r.val.append(v.copy())
        return r

    def exchange(self, a=0, b=1):
        """src/qtensor.cpp:349-373 (+ tensor::exchange src/tensor.cpp:335-366)."""
        if not self.dim:
            return self.copy()
        r = self.copy()
        r.dim[a], r.dim[b] = copy.deepcopy(self.dim[b]), copy.deepcopy(self.dim[a])
        r.dir[a], r.dir[b] = self.dir[b], self.dir[a]
        for i in range(len(self.val)):
            r.sym[i][a], r.sym[i][b] = self.sym[i][b], self.sym[i][a]
            r.val[i] = np.ascontiguousarray(np.swapaxes(self.val[i], a, b))
        return r

    def take(self, n, k):
        """src/qtensor.cpp:401-446 (+ tensor::take src/tensor.cpp:402-429)."""
        assert len(self.dim[n]) == 1
        r = QTensor(dtype=self.dtype)
        r.dim = [copy.deepcopy(d) for i, d in enumerate(self.dim) if i != n]
        r.dir = [d for i, d in enumerate(self.dir) if i != n]
        for s, v in zip(self.sym, self.val):
            r.sym.append([x for i, x in enumerate(s) if i != n])
            r.val.append(np.ascontiguousarray(np.take(v, k, axis=n)))
        return r.simplify()

    def left(self):
        """src/qtensor.cpp:448-452: last row of the MPO's left bond."""
        return self.take(0, self.dim[0][0] - 1)

    def right(self):
        """src/qtensor.cpp:455-461: first column of the right bond, then swap."""
        return self.take(2, 0).exchange(0, 1)

    # ------------------------------------------------------- combine/split
    def combine(self, mn=0, mx=1):
        """src/qtensor.cpp:466-585."""
        if mn == mx:
            return self.copy()
        for i in range(mn, mx):
            if self.dir[i] != self.dir[mx]:
                raise ValueError("combine: legs in different directions")
        r = QTensor(dtype=self.dtype)
        r.dir = self.dir[:mn] + self.dir[mx:]
        cmap = generate_map([len(self.dim[s]) for s in range(mn, mx + 1)])
        sym_dim = [_prod([self.dim[mn + s][m[s]] for s in range(mx - mn + 1)]) for m in cmap]
        sym_sum = []
        for m in cmap:
            t = 0
            for s in range(mx - mn + 1):
                t = ss(t, m[s])
            sym_sum.append(t)
        nsec = max([0] + sym_sum) + 1
        r.dim = [copy.deepcopy(d) for d in self.dim[:mn]] + [[0] * nsec] + \
            [copy.deepcopy(d) for d in self.dim[mx + 1:]]
        for i in range(len(sym_dim)):
            r.dim[mn][sym_sum[i]] += sym_dim[i]
        index = [len(d) for d in r.dim]
        rmap = generate_map(index)
        check = [False] * len(rmap)
        for s in self.sym:
            loc = list(s[:mn])
            c = 0
            for i in range(mn, mx + 1):
                c = ss(c, s[i])
            loc.append(c)
            loc += list(s[mx + 1:])
            k = 0
            for i in range(len(loc)):
                k = k * index[i] + loc[i]
            check[k] = True
        for t in range(len(check)):
            if check[t]:
                r.sym.append(list(rmap[t]))
                shape = [r.dim[i][rmap[t][i]] for i in range(len(index))]
                r.val.append(np.zeros(shape, dtype=self.dtype))
        # beginning of each sub-sector inside the combined sector
        store = [0] * index[mn]
        beg_of = []
        for i in range(len(sym_dim)):
            beg_of.append(store[sym_sum[i]])
            store[sym_sum[i]] += sym_dim[i]
        for t, (s, v) in enumerate(zip(self.sym, self.val)):
            c = 0
            for i in range(mn, mx + 1):
                c = ss(c, s[i])
            loc = -1
            for q, rs in enumerate(r.sym):
                if c == rs[mn] and list(s[:mn]) == rs[:mn] and list(s[mx + 1:]) == rs[mn + 1:]:
                    loc = q
            if loc == -1:
                raise ValueError("combine: cannot find sector")
            b = 0
            for i in range(mn, mx + 1):
                b = b * len(self.dim[i]) + s[i]
            beg = beg_of[b]
            l_ = _prod(v.shape[:mn])
            m_ = _prod(v.shape[mn:mx + 1])
            r_ = _prod(v.shape[mx + 1:])
            dst = r.val[loc].reshape(l_, -1, r_)
            dst[:, beg:beg + m_, :] = v.reshape(l_, m_, r_)
        return r

    def split(self, n, dim):
        """src/qtensor.cpp:588-656 (drops blocks with every |x| <= TOL)."""
        r = QTensor(dtype=self.dtype)
        r.dir = self.dir[:n] + [self.dir[n]] * len(dim) + self.dir[n + 1:]
        r.dim = [copy.deepcopy(d) for d in self.dim[:n]] + [list(d) for d in dim] + \
            [copy.deepcopy(d) for d in self.dim[n + 1:]]
        smap = generate_map([len(d) for d in dim])
        begs = [0] * len(self.dim[n])
        sym_dim = []
        for m in smap:
            c = 0
            for s in range(len(dim)):
                c = ss(c, m[s])
            d = _prod([dim[s][m[s]] for s in range(len(dim))])
            sym_dim.append((c, d, begs[c]))
            begs[c] += d
        for s, v in zip(self.sym, self.val):
            for q, (c, d, beg) in enumerate(sym_dim):
                if c != s[n]:
                    continue
                l_ = _prod(v.shape[:n])
                m_ = v.shape[n]
                r_ = _prod(v.shape[n + 1:])
                piece = v.reshape(l_, m_, r_)[:, beg:beg + d, :]
                if np.any(np.abs(piece) > TOL):
                    nsym = list(s[:n]) + list(smap[q]) + list(s[n + 1:])
                    shape = [r.dim[i][nsym[i]] for i in range(len(nsym))]
                    r.sym.append(nsym)
                    r.val.append(np.ascontiguousarray(piece).reshape(shape))
        return r

    def remove_symmetry(self):
        """src/qtensor.cpp:659-689: one dense block."""
        r = QTensor(dtype=self.dtype)
        r.dir = list(self.dir)
        r.dim = [[self.dimension(i)] for i in range(len(self.dim))]
        r.sym = [[0] * len(self.dim)]
        full = np.zeros([d[0] for d in r.dim], dtype=self.dtype)
        for s, v in zip(self.sym, self.val):
            sl = []
            for i in range(len(self.dim)):
                o = sum(self.dim[i][:s[i]])
                sl.append(slice(o, o + v.shape[i]))
            full[tuple(sl)] = v
        r.val = [full]
        return r

    # ------------------------------------------------------------ contractions
    def contract_gemm(self, A: "QTensor", B: "QTensor", ta="N", tb="N", num=1):
        """qtensor::contract(A,B,ta,tb,num), src/qtensor.cpp:824-899 over
        tensor::contract(A,B,...) src/tensor.cpp:603-649.  Writes into self."""
        ia = len(A.dim) - num if ta in "Nn" else 0
        ib = 0 if tb in "Nn" else len(B.dim) - num
        for i in range(num):
            if A.dim[ia + i] != B.dim[ib + i]:
                raise ValueError("contract: dimensions differ")
        keepa = [i for i in range(len(A.dim)) if i < ia or i >= ia + num]
        keepb = [i for i in range(len(B.dim)) if i < ib or i >= ib + num]
        self.dim = [copy.deepcopy(A.dim[i]) for i in keepa] + [copy.deepcopy(B.dim[i]) for i in keepb]
        self.dir = [A.dir[i] for i in keepa] + [B.dir[i] for i in keepb]
        self.dtype = np.result_type(A.dtype, B.dtype)
        self.sym, self.val = [], []
        locs = []
        for i, (sa, va) in enumerate(zip(A.sym, A.val)):
            for j, (sb, vb) in enumerate(zip(B.sym, B.val)):
                if any(sa[ia + k] != sb[ib + k] for k in range(num)):
                    continue
                sym = [sa[k] for k in keepa] + [sb[k] for k in keepb]
                loc = 0
                for s in range(len(self.dim)):
                    loc = loc * len(self.dim[s]) + sym[s]
                k_ = _prod(va.shape[ia:ia + num])
                ma = va.size // k_
                nb = vb.size // k_
                amat = va.reshape(ma, k_) if ta in "Nn" else va.reshape(k_, ma).T
                bmat = vb.reshape(k_, nb) if tb in "Nn" else vb.reshape(nb, k_).T
                prod = amat @ bmat
                shape = [va.shape[k] for k in keepa] + [vb.shape[k] for k in keepb]
                if loc in locs:
                    t = locs.index(loc)
                    self.val[t] = self.val[t] + prod.reshape(shape)
                else:
                    locs.append(loc)
                    self.sym.append(sym)
                    self.val.append(prod.reshape(shape).astype(self.dtype))
        for i in range(len(self.val)):
            self._check_dim(i)
        return self

    def contract_index(self, A: "QTensor", a: int, B: "QTensor", b: int,
                       change=False, fermion=True):
        """qtensor::contract(A,a,B,b,change,fermion), src/qtensor.cpp:902-983 over
        tensor::contract(A,a,B,b,cover,sign) src/tensor.cpp:652-695.

        Result legs: A<a, B<b, B>b, A>a.  Writes into self.
        """
        if A.dim[a] != B.dim[b]:
            raise ValueError("contract: dimension mismatch")
        nb = len(B.dim)
        self.dim = [copy.deepcopy(x) for x in A.dim[:a]] + [copy.deepcopy(x) for x in B.dim[:b]] + \
            [copy.deepcopy(x) for x in B.dim[b + 1:]] + [copy.deepcopy(x) for x in A.dim[a + 1:]]
        self.dir = A.dir[:a] + B.dir[:b] + B.dir[b + 1:] + A.dir[a + 1:]
        self.dtype = np.result_type(A.dtype, B.dtype)
        self.sym, self.val = [], []
        locs = []
        for sa, va in zip(A.sym, A.val):
            for sb, vb in zip(B.sym, B.val):
                if sa[a] != sb[b]:
                    continue
                sym = sa[:a] + sb[:b] + sb[b + 1:] + sa[a + 1:]
                sign = 1
                if fermion:
                    if change:
                        d_a = 0
                        for i in range(len(A.dir)):
                            d_a = ss(d_a, sa[i] * A.dir[i])
                        b_p = 0
                        for i in range(len(B.dir)):
                            if B.dir[i] == 1:
                                b_p = ss(b_p, sb[i])
                        sign = 1 - 2 * ((ff(d_a) * ff(b_p)) % 2)
                    else:
                        d_b = 0
                        for i in range(len(B.dir)):
                            d_b = ss(d_b, sb[i] * B.dir[i])
                        a_p = 0
                        for i in range(len(A.dir)):
                            if A.dir[i] == 1:
                                a_p = ss(a_p, sa[i])
                        sign = 1 - 2 * ((ff(d_b) * ff(a_p)) % 2)
                # out[i1,j1,j2,i2] = sum_k A[i1,k,i2] B[j1,k,j2] * sign
                al = _prod(va.shape[:a])
                ar = _prod(va.shape[a + 1:])
                bl = _prod(vb.shape[:b])
                br = _prod(vb.shape[b + 1:])
                mid = va.shape[a]
                A3 = va.reshape(al, mid, ar)
                B3 = vb.reshape(bl, mid, br)
                out = np.einsum("ikr,jks->ijsr", A3, B3) * sign
                shape = list(va.shape[:a]) + list(vb.shape[:b]) + list(vb.shape[b + 1:]) + list(va.shape[a + 1:])
                out = out.reshape(shape)
                loc = 0
                for s in range(len(sym)):
                    loc = loc * len(self.dim[s]) + sym[s]
                if loc in locs:
                    t = locs.index(loc)
                    self.val[t] = self.val[t] + out
                else:
                    locs.append(loc)
                    self.sym.append(sym)
                    self.val.append(out.astype(self.dtype))
        for i in range(len(self.val)):
            self._check_dim(i)
        return self

    # ------------------------------------------------------------ hot path
    def get_map(self, dim, sym, ta="N", tb="T", num=1, beta=0, back=False,
                sym_ret=None, fermion=True):
        """qtensor::get_map, src/qtensor.cpp:997-1119.

        Returns (map, dim_ret, sym_ret).  ``sym_ret`` is read when back=True
        and built (first-appearance order) when back=False.  ``fermion=False``
        forces alpha=1 for every task (the physical rule for bosonic models);
        ``fermion=True`` is the reference's unconditional rule.
        """
        ia = len(self.dim) - num if ta in "Nn" else 0
        ib = 0 if tb in "Nn" else len(dim) - num
        for i in range(num):
            if self.dim[ia + i] != dim[ib + i]:
                raise ValueError("get_map: dimensions differ")
        dim_ret = [copy.deepcopy(self.dim[i]) for i in range(len(self.dim)) if i < ia or i >= ia + num]
        dim_ret += [copy.deepcopy(dim[i]) for i in range(len(dim)) if i < ib or i >= ib + num]
        sym_dim = []
        for s in sym:
            tot = _prod([dim[q][s[q]] for q in range(len(dim))])
            if ib:
                row = _prod([dim[q][s[q]] for q in range(ib)])
            else:
                row = _prod([dim[q][s[q]] for q in range(num)])
            sym_dim.append([tot, row, 0])
        for i in range(1, len(sym)):
            sym_dim[i][2] = sym_dim[i - 1][2] + sym_dim[i - 1][0]
        maps = []
        if not back:
            sym_ret = []
            locs = []
            map_dim = [0]
        else:
            sym_ret = [list(s) for s in sym_ret]
            locs = []
            for s in sym_ret:
                v = 0
                for j in range(len(dim_ret)):
                    v = v * len(dim_ret[j]) + s[j]
                locs.append(v)
            map_dim = [0] * len(sym_ret)
            for i in range(len(sym_ret) - 1):
                map_dim[i + 1] = map_dim[i] + _prod([dim_ret[j][sym_ret[i][j]] for j in range(len(dim_ret))])
        first = [True] * len(locs)
        for i, si in enumerate(self.sym):
            for j, sj in enumerate(sym):
                if any(si[ia + k] != sj[ib + k] for k in range(num)) or sym_dim[j][0] == 0:
                    continue
                ms = [i, sym_dim[j][2], map_dim[-1], sym_dim[j][1],
                      sym_dim[j][0] // sym_dim[j][1], int(beta), 1]
                if not back:
                    d_b = 0
                    for k in range(len(self.dir)):
                        d_b = ss(d_b, si[k] * self.dir[k])
                    a_p = 0
                    for k in range(len(dim) - num):
                        a_p = ss(a_p, sj[k])
                    ms[6] = 1 - 2 * ((ff(d_b) * ff(a_p)) % 2) if fermion else 1
                st = [si[k] for k in range(len(self.dim)) if k < ia or k >= ia + num]
                st += [sj[k] for k in range(len(dim)) if k < ib or k >= ib + num]
                loc = 0
                for s in range(len(dim_ret)):
                    loc = loc * len(dim_ret[s]) + st[s]
                same = False
                for t in range(len(locs)):
                    if loc == locs[t]:
                        ms[2] = map_dim[t]
                        if back and first[t]:
                            first[t] = False
                        else:
                            ms[5] = 1
                        maps.append(list(ms))
                        same = True
                if not same and not back:
                    d = _prod([dim_ret[s][st[s]] for s in range(len(dim_ret))])
                    map_dim.append(map_dim[-1] + d)
                    locs.append(loc)
                    first.append(True)
                    sym_ret.append(st)
                    maps.append(list(ms))
        return maps, dim_ret, sym_ret

    def contract_map(self, out: np.ndarray, inp: np.ndarray, maps, ta="N", tb="T", num=1):
        """qtensor::contract(T* out, T* in, map, ...) src/qtensor.cpp:986-994 over
        tensor::contract(out,in,row,col,...) src/tensor.cpp:698-718 and the row-major
        ?gemm wrappers (src/dgemm.cpp:29-47, src/zgemm.cpp:32-51):
        out_blk(m x n) = alpha * A_i(m x k) . in_blk(n x k)^T + beta * out_blk."""
        assert ta in "Nn" and tb in "Tt"
        for (i, ib, ob, row, col, beta, alpha) in maps:
            A = self.val[i]
            k = _prod(A.shape[len(A.shape) - num:])
            m = A.size // k
            n, kk = row, col
            assert kk == k
            if m <= 0 or n <= 0 or k <= 0:
                continue
            Am = A.reshape(m, k)
            Bm = inp[ib:ib + n * k].reshape(n, k)
            res = alpha * (Am @ Bm.T)
            if beta:
                out[ob:ob + m * n] = res.ravel() + beta * out[ob:ob + m * n]
            else:
                out[ob:ob + m * n] = res.ravel()

    # ------------------------------------------------------------ SVD / eig
    def svd(self, num=1, cutoff=0):
        """qtensor::svd, src/qtensor.cpp:1122-1262.  Returns (U, S, V, s)."""
        comb = self.combine(num, len(self.dim) - 1).combine(0, num - 1)
        sym0 = ss(comb.sym[0][0], comb.sym[0][1])
        if any(ss(s[0], s[1]) != sym0 for s in comb.sym):
            raise ValueError("svd: symmetry sectors do not have constraint")
        U = QTensor(dtype=self.dtype)
        V = QTensor(dtype=self.dtype)
        S = QTensor(dtype=np.float64)
        U.dir = [1, 1]
        V.dir = [1, 1]
        S.dir = [1, 1]
        U.dim = copy.deepcopy(comb.dim)
        U.dim[1] = copy.deepcopy(U.dim[0])
        V.dim = copy.deepcopy(comb.dim)
        V.dim[0] = copy.deepcopy(V.dim[1])
        S.dim = copy.deepcopy(comb.dim)
        nsym = len(comb.val)
        Ss = []
        total = 0
        for i in range(nsym):
            u, s, v = tensor_svd(comb.val[i])
            U.val.append(u)
            V.val.append(v)
            Ss.append(s)
            total += len(s)
            su = list(comb.sym[i])
            su[1] = su[0]
            sv = list(comb.sym[i])
            sv[0] = sv[1]
            U.sym.append(su)
            V.sym.append(sv)
            S.sym.append(list(comb.sym[i]))
        s_sorted = [0.0] * total
        keep = total
        if cutoff != 0 and cutoff < total:
            keep = cutoff
        u_num = [0] * nsym
        for t in range(total):
            mx, loc = -1.0, -1
            for i in range(nsym):
                if u_num[i] < len(Ss[i]) and mx < Ss[i][u_num[i]]:
                    mx = Ss[i][u_num[i]]
                    loc = i
            s_sorted[t] = mx
            u_num[loc] += 1
            if t == keep - 1:
                v_num = list(u_num)
        if keep == 0:
            v_num = [0] * nsym
        u_num = list(v_num)
        for i in range(nsym):
            r_, c_ = comb.val[i].shape
            sd = np.zeros((len(Ss[i]), len(Ss[i])))
            np.fill_diagonal(sd, Ss[i])
            # tensor::diag(2) then resize to the full rectangular shape
            if r_ < c_:
                sd = _resize(sd, 1, c_)
            else:
                sd = _resize(sd, 0, r_)
            S.val.append(sd)
        if cutoff != 0:
            diff = min(comb.dimension(0), cutoff) - keep
            while diff > 0:
                for i in range(nsym):
                    if u_num[i] < comb.dim[0][i] and diff > 0:
                        u_num[i] += 1
                        diff -= 1
            diff = min(comb.dimension(1), cutoff) - keep
            while diff > 0:
                for i in range(nsym):
                    if v_num[i] < comb.dim[1][i] and diff > 0:
                        v_num[i] += 1
                        diff -= 1
            for i in range(nsym):
                U.val[i] = _resize(U.val[i], 1, u_num[i])
                U.dim[1][U.sym[i][1]] = u_num[i]
                V.val[i] = _resize(V.val[i], 0, v_num[i])
                V.dim[0][V.sym[i][0]] = v_num[i]
                S.val[i] = _resize(S.val[i], 0, u_num[i])
                S.dim[0][S.sym[i][0]] = u_num[i]
                S.val[i] = _resize(S.val[i], 1, v_num[i])
                S.dim[1][S.sym[i][1]] = v_num[i]
        U = U.split(0, [copy.deepcopy(self.dim[i]) for i in range(num)])
        V = V.split(1, [copy.deepcopy(self.dim[i]) for i in range(num, len(self.dim))])
        return U, S, V, s_sorted

    def eig(self, sector=-1):
        """qtensor::eig, src/qtensor.cpp:1265-1313 -> tensor::eig -> ?syev/?heev.
        Returns (all eigenvalues ascending, lowest eigenvector)."""
        assert len(self.dim) == 2
        if any(s[0] != s[1] for s in self.sym) or (sector == -1 and len(self.sym) != 1):
            return self.remove_symmetry().eig(-1)
        if sector == -1:
            return dense_eig(self.val[0])
        loc = -1
        for i, s in enumerate(self.sym):
            if s[0] == sector:
                loc = i
        if loc == -1:
            d = self.dim[0][sector]
            v = np.zeros(d, dtype=self.dtype)
            v[0] = 1
            return np.zeros(d), v
        return dense_eig(self.val[loc])

    def trace(self, A: "QTensor", fermion=True):
        """qtensor::trace, src/qtensor.cpp:789-821."""
        ret = 0.0
        for sl, vl in zip(self.sym, self.val):
            for sr, vr in zip(A.sym, A.val):
                if sl != sr:
                    continue
                sign = 1.0
                if fermion:
                    d_b = 0
                    for i in range(len(A.dir)):
                        d_b = ss(d_b, sr[i] * A.dir[i])
                    a_p = 0
                    for i in range(len(self.dir)):
                        if self.dir[i] == 1:
                            a_p = ss(a_p, sl[i])
                    sign = 1 - 2 * ((ff(d_b) * ff(a_p)) % 2)
                ret = ret + np.sum(vl * vr) * sign
        return ret


def _resize(a: np.ndarray, k: int, n: int) -> np.ndarray:
    """tensor::resize(k,n), src/tensor.cpp:482-513: cut or zero-pad leg k."""
    old = a.shape[k]
    if n == old:
        return a
    shape = list(a.shape)
    shape[k] = n
    out = np.zeros(shape, dtype=a.dtype)
    keep = min(n, old)
    sl = [slice(None)] * a.ndim
    sl[k] = slice(0, keep)
    out[tuple(sl)] = a[tuple(sl)]
    return out


# Test knob (None in every restatement run): an integer seed makes tensor_svd
# replace LAPACK's orthonormal complement -- the columns of U and rows of V
# beyond min(rows, cols), which carry no singular value -- by another
# orthonormal basis of the same complement (a seeded random rotation).  The
# reference pads truncated U / V with these vectors (src/qtensor.cpp:1221-1252),
# so a run's sensitivity to this LAPACK-specific choice can be measured.
COMPLEMENT_ROTATION_SEED = None


def _rotate_complement(U, V, mn, seed):
    rng = np.random.default_rng(seed)

    def unitary(n, dtype):
        z = rng.standard_normal((n, n))
        if np.dtype(dtype).kind == "c":
            z = z + 1j * rng.standard_normal((n, n))
        q, r = np.linalg.qr(z)
        return (q * (np.diag(r) / np.abs(np.diag(r)))).astype(dtype)

    if U.shape[1] > mn:
        U = U.copy()
        U[:, mn:] = U[:, mn:] @ unitary(U.shape[1] - mn, U.dtype)
    if V.shape[0] > mn:
        V = V.copy()
        V[mn:, :] = unitary(V.shape[0] - mn, V.dtype) @ V[mn:, :]
    return U, V


def tensor_svd(a: np.ndarray):
    """tensor::svd(U,S,V,1) on a matrix block, src/tensor.cpp:736-785, with the
    row-major ?gesvd wrappers (src/dgesvd.cpp:32-97, src/zgesvd.cpp:36-104):
    full U (r x r), s (min), full V (c x c) with a = U diag(s) V."""
    r_, c_ = a.shape
    mn = min(r_, c_)
    if not np.any(np.abs(a) > TOL):
        return np.eye(r_, dtype=a.dtype), np.zeros(mn), np.eye(c_, dtype=a.dtype)
    if r_ * c_ == 1:
        x = a[0, 0]
        return np.ones((1, 1), dtype=a.dtype), np.array([abs(x)]), np.array([[x / abs(x)]], dtype=a.dtype)
    # LAPACK sees the row-major matrix as its column-major transpose
    Up, s, VTp = scipy.linalg.svd(a.T, full_matrices=True, lapack_driver="gesvd")
    U, V = np.ascontiguousarray(VTp.T), np.ascontiguousarray(Up.T)
    if COMPLEMENT_ROTATION_SEED is not None:
        U, V = _rotate_complement(U, V, mn, COMPLEMENT_ROTATION_SEED + r_ * 7919 + c_)
    return U, s, V


def dense_eig(h: np.ndarray):
    """?syev / ?heev wrappers (src/dsyev.cpp:66-112, src/zheev.cpp:72-121): the
    upper triangle is used; returns ascending eigenvalues and the lowest vector."""
    w, v = scipy.linalg.eigh(h, lower=False)
    return w, v[:, 0].copy()
