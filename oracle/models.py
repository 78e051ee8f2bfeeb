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
    raise ValueError(name)
