"""MPO builders and the MPO container, restated from the reference.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import math

import numpy as np

from .qtensor import PI, QTensor


def _blk(D, entries, dtype):
    t = np.zeros((D, 1, D, 1), dtype=dtype)
    for val, w0, w1 in entries:
        t[w0, 0, w1, 0] = val
    return t


def H_XYZ(Jx, Jy, Jz):
    """src/operators.cpp:361-391 (Pauli matrices, complex, Z2 = Sz parity)."""
    c = np.complex128
    i = 1j
    r = QTensor(1, 2, 1, 2, dtype=c)
    r.update(_blk(5, [(1, 0, 0), (1, 3, 0), (Jz, 4, 3), (1, 4, 4)], c), 0, 0, 0, 0)
    r.update(_blk(5, [(1, 1, 0), (-i, 2, 0), (Jx, 4, 1), (-i * Jy, 4, 2)], c), 0, 0, 0, 1)
    r.update(_blk(5, [(1, 1, 0), (i, 2, 0), (Jx, 4, 1), (i * Jy, 4, 2)], c), 0, 1, 0, 0)
    r.update(_blk(5, [(1, 0, 0), (-1, 3, 0), (-Jz, 4, 3), (1, 4, 4)], c), 0, 1, 0, 1)
    return r


def H_Kitaev(t, p, u):
    """Real Kitaev chain MPO, src/operators.cpp:406-431 (fermionic: add_sign)."""
    d = np.float64
    r = QTensor(1, 2, 1, 2, dtype=d)
    r.update(_blk(4, [(1, 0, 0), (1, 3, 3)], d), 0, 0, 0, 0)
    r.update(_blk(4, [(1, 1, 0), (-p, 3, 1), (-t, 3, 2)], d), 0, 0, 0, 1)
    r.update(_blk(4, [(1, 2, 0), (t, 3, 1), (p, 3, 2)], d), 0, 1, 0, 0)
    r.update(_blk(4, [(1, 0, 0), (-2 * u, 3, 0), (1, 3, 3)], d), 0, 1, 0, 1)
    r.add_sign(0, -1, 0, 1)
    return r


def H_Potts3(f, j, p, t):
    """Chiral 3-state Potts MPO, src/operators.cpp:301-350 (Z3, truncated PI)."""
    c = np.complex128
    r = QTensor(1, 3, 1, 3, dtype=c)
    diag = [-2 * f * math.cos(p), -2 * f * math.cos(p + 2 * PI / 3), -2 * f * math.cos(p + 4 * PI / 3)]
    v1 = -j * complex(math.cos(t), math.sin(t))
    v2 = -j * complex(math.cos(t), -math.sin(t))
    blocks = {
        (0, 0): [(1, 0, 0), (diag[0], 3, 0), (1, 3, 3)],
        (0, 1): [(1, 1, 0), (v1, 3, 2)],
        (0, 2): [(1, 2, 0), (v2, 3, 1)],
        (1, 0): [(1, 2, 0), (v2, 3, 1)],
        (1, 1): [(1, 0, 0), (diag[1], 3, 0), (1, 3, 3)],
        (1, 2): [(1, 1, 0), (v1, 3, 2)],
        (2, 0): [(1, 1, 0), (v1, 3, 2)],
        (2, 1): [(1, 2, 0), (v2, 3, 1)],
        (2, 2): [(1, 0, 0), (diag[2], 3, 0), (1, 3, 3)],
    }
    for (a, b) in [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2), (2, 0), (2, 1), (2, 2)]:
        r.update(_blk(4, blocks[(a, b)], c), 0, a, 0, b)
    return r


def H_Ising(h, j):
    """Ising MPO without symmetry, src/operators.cpp:239-256."""
    x = np.zeros((3, 2, 3, 2))
    for val, i0, i1, i2, i3 in [(1, 0, 0, 0, 0), (1, 0, 1, 0, 1), (1, 1, 0, 0, 0), (-1, 1, 1, 0, 1),
                                (h, 2, 0, 0, 1), (h, 2, 1, 0, 0), (j, 2, 0, 1, 0), (-j, 2, 1, 1, 1),
                                (1, 2, 0, 2, 0), (1, 2, 1, 2, 1)]:
        x[i0, i1, i2, i3] = val
    r = QTensor(1, 1, 1, 1)
    r.update(x, 0, 0, 0, 0)
    return r


def H_Potts2(h, j):
    """Z2 Potts / transverse Ising MPO, src/operators.cpp:265-288."""
    d = np.float64
    r = QTensor(1, 2, 1, 2, dtype=d)
    r.update(_blk(3, [(1, 0, 0), (h, 2, 0), (1, 2, 2)], d), 0, 0, 0, 0)
    r.update(_blk(3, [(1, 1, 0), (j, 2, 1)], d), 0, 0, 0, 1)
    r.update(_blk(3, [(1, 1, 0), (j, 2, 1)], d), 0, 1, 0, 0)
    r.update(_blk(3, [(1, 0, 0), (-h, 2, 0), (1, 2, 2)], d), 0, 1, 0, 1)
    return r


class MPO:
    """mpo<T>, include/mpo.h:9-61 (uniform bulk operator + edges)."""

    def __init__(self, sites, ham: QTensor):
        self.sites = sites
        self.ops = [ham]
        self.l_edge = None
        self.r_edge = None

    def edge(self, left=None, right=None):
        """include/mpo.h:45-53."""
        if left is None:
            self.l_edge = self.ops[0].left()
            self.r_edge = self.ops[(self.sites - 1) % len(self.ops)].right()
        else:
            self.l_edge, self.r_edge = left, right

    def __getitem__(self, n):
        """include/mpo.h:55-59."""
        if n == 0:
            return self.l_edge
        if n == self.sites - 1:
            return self.r_edge
        return self.ops[n % len(self.ops)]

    def size(self):
        return self.sites


def make_model(name, L, **p):
    """MPO + symmetry order + fermion flag for a named model, set up as the
    mains do (xyz.cpp:36-38, kitaev.cpp:37-40, potts.cpp:43-44)."""
    if name == "xyz":
        H = H_XYZ(p.get("jx", 1.0), p.get("jy", 1.0), p.get("jz", 1.0))
        m = MPO(L, H)
        m.edge()
        return m, 2, False
    if name == "kitaev":
        H = H_Kitaev(p.get("t", 1.0), p.get("p", 0.0), p.get("u", 0.0))
        m = MPO(L, H)
        m.edge(H.left(), H.right())
        return m, 2, True
    if name == "potts3":
        H = H_Potts3(p.get("f", 1.0), p.get("j", 0.0), p.get("p", 0.0), p.get("t", 0.0))
        m = MPO(L, H)
        m.edge()
        return m, 3, False
    if name == "potts2":
        H = H_Potts2(p.get("h", 1.0), p.get("j", 1.0))
        m = MPO(L, H)
        m.edge()
        return m, 2, False
    if name == "4e":
        # L physical sites, two per MPO tensor (4e.cpp:46-47); Z2^3, graded
        m = mpo_4e(L // 2, p.get("t1", 1.0), p.get("t3", 0.0), p.get("f0", 0.0), p.get("f1", 0.0),
                   p.get("frac", 0.0))
        from .qtensor import S4E
        return m, S4E, True
    raise ValueError(name)


# ---------------------------------------------------------------- 4e model
# The reference's matrix-element tables of 4e.cpp:203-391, as data: each row
# (sector in, sector out, state in, state out, sign) of an operator on the
# two-site space of one 4e MPO tensor (eight Z2^3 sectors of dimension 2).
_LCU = """0 4 0 0 1|0 4 1 1 1|0 5 0 0 1|0 5 1 1 -1|1 4 0 0 1|1 4 1 1 -1|1 5 0 0 1|1 5 1 1 1|2 6 0 0 1|2 6 1 1 -1|
2 7 0 0 1|2 7 1 1 1|3 6 0 0 1|3 6 1 1 -1|3 7 0 0 1|3 7 1 1 1|4 0 0 0 1|4 0 1 1 1|4 1 0 0 -1|4 1 1 1 1|5 0 0 0 -1|
5 0 1 1 1|5 1 0 0 1|5 1 1 1 1|6 2 0 0 1|6 2 1 1 -1|6 3 0 0 -1|6 3 1 1 1|7 2 0 0 -1|7 2 1 1 -1|7 3 0 0 1|7 3 1 1 1"""
_LCD = """0 2 0 1 1|0 2 1 0 1|0 3 0 1 1|0 3 1 0 -1|1 2 0 1 1|1 2 1 0 1|1 3 0 1 1|1 3 1 0 -1|2 0 0 1 1|2 0 1 0 1|
2 1 0 1 -1|2 1 1 0 -1|3 0 0 1 1|3 0 1 0 -1|3 1 0 1 -1|3 1 1 0 1|4 6 0 1 1|4 6 1 0 -1|4 7 0 1 -1|4 7 1 0 1|5 6 0 1 1|
5 6 1 0 -1|5 7 0 1 -1|5 7 1 0 1|6 4 0 1 -1|6 4 1 0 1|6 5 0 1 1|6 5 1 0 -1|7 4 0 1 -1|7 4 1 0 1|7 5 0 1 1|7 5 1 0 -1"""
_RCU = """0 4 1 0 -1|0 4 0 1 1|0 5 1 0 1|0 5 0 1 1|1 4 1 0 1|1 4 0 1 1|1 5 1 0 -1|1 5 0 1 1|2 6 1 0 1|2 6 0 1 1|
2 7 1 0 -1|2 7 0 1 1|3 6 1 0 1|3 6 0 1 1|3 7 1 0 -1|3 7 0 1 1|4 0 1 0 1|4 0 0 1 -1|4 1 1 0 -1|4 1 0 1 -1|5 0 1 0 -1|
5 0 0 1 -1|5 1 1 0 1|5 1 0 1 -1|6 2 1 0 1|6 2 0 1 1|6 3 1 0 -1|6 3 0 1 -1|7 2 1 0 -1|7 2 0 1 1|7 3 1 0 1|7 3 0 1 -1"""
_RCD = """0 2 0 0 1|0 2 1 1 -1|0 3 0 0 1|0 3 1 1 1|1 2 0 0 1|1 2 1 1 -1|1 3 0 0 1|1 3 1 1 1|2 0 0 0 1|2 0 1 1 -1|
2 1 0 0 -1|2 1 1 1 1|3 0 0 0 -1|3 0 1 1 -1|3 1 0 0 1|3 1 1 1 1|4 6 0 0 -1|4 6 1 1 -1|4 7 0 0 -1|4 7 1 1 -1|
5 6 0 0 -1|5 6 1 1 -1|5 7 0 0 -1|5 7 1 1 -1|6 4 0 0 -1|6 4 1 1 -1|6 5 0 0 1|6 5 1 1 1|7 4 0 0 1|7 4 1 1 1|
7 5 0 0 -1|7 5 1 1 -1"""
_LP = """0 6 0 1 1|0 7 0 1 -1|1 6 0 1 1|1 7 0 1 -1|2 4 0 1 -1|2 5 0 1 1|3 4 0 1 -1|3 5 0 1 1|4 2 1 0 1|4 3 1 0 -1|
5 2 1 0 1|5 3 1 0 -1|6 0 1 0 -1|6 1 1 0 1|7 0 1 0 -1|7 1 1 0 1"""
_RP = """0 6 0 1 -1|0 7 0 1 -1|1 6 0 1 -1|1 7 0 1 -1|2 4 1 0 -1|2 5 1 0 1|3 4 1 0 -1|3 5 1 0 1|4 2 0 1 1|4 3 0 1 -1|
5 2 0 1 1|5 3 0 1 -1|6 0 1 0 1|6 1 1 0 -1|7 0 1 0 -1|7 1 1 0 1"""
_PP = """0 0 0 0 1|0 1 0 0 -1|1 0 0 0 1|1 1 0 0 -1"""
_LN = """0 1 0 0 -1|1 0 0 0 -1|2 3 0 0 -1|3 2 0 0 -1|4 5 1 1 -1|5 4 1 1 -1|6 7 1 1 -1|7 6 1 1 -1"""
_RN = """0 1 0 0 -1|1 0 0 0 -1|2 3 1 1 -1|3 2 1 1 -1|4 5 0 0 -1|5 4 0 0 -1|6 7 1 1 1|7 6 1 1 1"""


def _rows(text):
    return [tuple(int(x) for x in r.split()) for r in text.replace("\n", "").split("|")]


T4E = {k: _rows(v) for k, v in dict(lcu=_LCU, lcd=_LCD, rcu=_RCU, rcd=_RCD, lp=_LP, rp=_RP, pp=_PP, ln=_LN,
                                   rn=_RN).items()}


def _assemble_4e(t, D):
    """4e.cpp:478-485: blocks (0, j, 0, k), k outer; dirs (0, -1, 0, 1); simplify."""
    r = QTensor(1, 8, 1, 8, dtype=np.complex128)
    r.u_dim(D, 2, D, 2)
    for k in range(8):
        for j in range(8):
            r.update(t[j][k], 0, j, 0, k)
    r.add_sign(0, -1, 0, 1)
    return r.simplify()


def _add(t, j, k, v, *idx):
    """tensor::update adds and skips |v| < TOL (src/tensor.cpp:68-102)."""
    if abs(v) >= 1e-10:
        t[j][k][idx] += v


def H_4e(second, t1, t3, f0, f1, p1, frac=0.0):
    """H_4e1 (second=False, 4e.cpp:397-487) / H_4e2 (second=True,
    4e.cpp:490-580): bond 9, complex, graded."""
    i = 1j
    t = [[np.zeros((9, 2, 9, 2), dtype=np.complex128) for _ in range(8)] for _ in range(8)]
    for k in range(8):
        for idx in ((0, 0, 0, 0), (0, 1, 0, 1), (8, 0, 8, 0), (8, 1, 8, 1)):
            _add(t, k, k, 1.0, *idx)
    for k in (0, 1):
        _add(t, k, k, -2 * f0, 8, 0, 0, 0)
        _add(t, k, k, 2 * f0, 8, 1, 0, 1)
    for k in (6, 7):
        _add(t, k, k, 2 * f0, 8, 0, 0, 0)
        _add(t, k, k, -2 * f0, 8, 1, 0, 1)
    for k in (2, 3):
        _add(t, k, k, -2 * i * t1, 8, 0, 0, 1)
        _add(t, k, k, 2 * i * t1, 8, 1, 0, 0)
    for k in (4, 5):
        _add(t, k, k, 2 * i * t1, 8, 0, 0, 1)
        _add(t, k, k, -2 * i * t1, 8, 1, 0, 0)
    _add(t, 6, 6, 4 * i * t1, 8, 0, 0, 1)
    _add(t, 6, 6, -4 * i * t1, 8, 1, 0, 0)
    ou, cu = (T4E["rcu"], T4E["lcu"]) if second else (T4E["lcu"], T4E["rcu"])
    od, cd = (T4E["rcd"], T4E["lcd"]) if second else (T4E["lcd"], T4E["rcd"])
    for k in range(32):
        a, b = ou[k], cu[k]
        _add(t, a[0], a[1], 0.5 * a[4], 1, a[2], 0, a[3])
        _add(t, b[0], b[1], i * (t3 * b[4]), 8, b[2], 2, b[3])
        _add(t, a[1], a[0], 0.5 * a[4], 2, a[3], 0, a[2])
        _add(t, b[1], b[0], i * (t3 * b[4]), 8, b[3], 1, b[2])
        a, b = od[k], cd[k]
        _add(t, a[0], a[1], 0.5 * a[4], 3, a[2], 0, a[3])
        _add(t, b[0], b[1], i * (t3 * b[4]), 8, b[2], 4, b[3])
        _add(t, a[1], a[0], 0.5 * a[4], 4, a[3], 0, a[2])
        _add(t, b[1], b[0], i * (t3 * b[4]), 8, b[3], 3, b[2])
    pc = p1 * complex(math.cos(2 * PI * frac), math.sin(2 * PI * frac))
    pd = p1 * complex(math.cos(2 * PI * frac), -math.sin(2 * PI * frac))
    for a, b in zip(T4E["lp"], T4E["rp"]):
        _add(t, a[0], a[1], 0.5 * a[4], 5, a[2], 0, a[3])
        _add(t, b[0], b[1], pc * (4.0 * b[4]), 8, b[2], 5, b[3])
        _add(t, a[1], a[0], 0.5 * a[4], 6, a[3], 0, a[2])
        _add(t, b[1], b[0], pd * (4.0 * b[4]), 8, b[3], 6, b[2])
    for a, b in zip(T4E["ln"], T4E["rn"]):
        _add(t, a[0], a[1], a[4], 7, a[2], 0, a[3])
        _add(t, b[0], b[1], -4 * f1 * b[4], 8, b[2], 7, b[3])
    return _assemble_4e(t, 9)


def op_4e(table, coef=0.5, flip=False):
    """c_up1 .. c_pair_dagger2 (4e.cpp:587-779): a bond-1 operator from one table;
    flip swaps in and out (the daggered pair operators, coef -0.5)."""
    t = [[np.zeros((1, 2, 1, 2), dtype=np.complex128) for _ in range(8)] for _ in range(8)]
    for a in T4E[table]:
        if flip:
            _add(t, a[1], a[0], coef * a[4], 0, a[3], 0, a[2])
        else:
            _add(t, a[0], a[1], coef * a[4], 0, a[2], 0, a[3])
    return _assemble_4e(t, 1)


def mpo_4e(sites, t1, t3, f0, f1, frac):
    """4e.cpp:49-59: H_4e1 on even, H_4e2 on odd tensors, pairing phase frac*i/sites."""
    m = MPO(sites, None)
    m.ops = [None] * sites
    for n in range(1, sites - 1):
        m.ops[n] = H_4e(n % 2 == 1, t1, t3, f0, f1, f1, frac * n / sites)
    m.edge(H_4e(False, t1, t3, f0, f1, f1, 0).left(), H_4e(True, t1, t3, f0, f1, f1, frac * (sites - 1) / sites).right())
    return m
