"""The host symmetric / Hermitian eigensolvers (dmrg_amd/csrc/host_eig.cpp)
built with the plain C++ compiler and checked against LAPACK (numpy) on the
matrices they serve: Lanczos projected matrices (tridiagonal, and arrowhead
plus tridiagonal after a thick restart), degenerate spectra, and small
Hermitian operators."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

SRC = Path(__file__).resolve().parents[1] / "dmrg_amd" / "csrc"

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "host_eig.hpp"
int main(int argc, char** argv) {
  FILE* f = std::fopen(argv[1], "rb");
  int kind, n;
  std::fread(&kind, 4, 1, f);
  std::fread(&n, 4, 1, f);
  FILE* o = std::fopen(argv[2], "wb");
  std::vector<double> w;
  if (kind == 0) {
    std::vector<double> a((size_t)n * n), v;
    std::fread(a.data(), 8, a.size(), f);
    dmrgx::host_symeig(n, a, w, v);
    std::fwrite(w.data(), 8, n, o);
    std::fwrite(v.data(), 8, v.size(), o);
  } else {
    std::vector<dmrgx::cplx> a((size_t)n * n), v;
    std::fread(a.data(), 16, a.size(), f);
    if (kind == 1) dmrgx::host_heeig(n, a, w, v);
    else dmrgx::host_hermitian_eig(n, a, true, w, v);
    std::fwrite(w.data(), 8, n, o);
    std::fwrite(v.data(), 16, v.size(), o);
  }
  std::fclose(o);
  return 0;
}
"""


@pytest.fixture(scope="module")
def solver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("host_eig")
    (d / "drv.cpp").write_text(DRIVER)
    exe = d / "drv"
    subprocess.run([gxx, "-O2", "-std=c++17", f"-I{SRC}", str(d / "drv.cpp"), str(SRC / "host_eig.cpp"), "-o",
                    str(exe)], check=True)

    def run(a, real_aware=False):
        n = a.shape[0]
        cplx = np.iscomplexobj(a) or real_aware
        inp, out = d / "in.bin", d / "out.bin"
        with open(inp, "wb") as f:
            np.array([2 if real_aware else (1 if cplx else 0), n], dtype=np.int32).tofile(f)
            a.astype(np.complex128 if cplx else np.float64).tofile(f)
        subprocess.run([str(exe), str(inp), str(out)], check=True)
        raw = np.fromfile(out, dtype=np.uint8)
        w = raw[:8 * n].view(np.float64)
        v = raw[8 * n:].view(np.complex128 if cplx else np.float64).reshape(n, n)
        return w, v

    return run


def _check(a, w, v, tol=1e-12):
    ref = np.linalg.eigvalsh(a)
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - ref)) < tol * scale
    assert np.max(np.abs(v.conj().T @ v - np.eye(len(w)))) < tol * 10
    assert np.max(np.abs(a @ v - v * w)) < tol * scale * 10


def test_random_symmetric(solver):
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 7, 36, 60):
        a = rng.standard_normal((n, n))
        a = a + a.T
        _check(a, *solver(a))


def test_lanczos_projected_matrices(solver):
    """Tridiagonal T of a Lanczos cycle (diagonal far from zero, as for DMRG
    operators), and the arrowhead + tridiagonal form after a thick restart."""
    rng = np.random.default_rng(2)
    m, keep = 36, 21
    d = -25 + rng.standard_normal(m)
    e = 5 + rng.standard_normal(m - 1)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    _check(T, *solver(T))
    A = np.zeros((m, m))
    A[np.arange(keep), np.arange(keep)] = np.sort(-37 + 3 * rng.standard_normal(keep))
    spike = 1e-6 * rng.standard_normal(keep)
    A[:keep, keep] = A[keep, :keep] = spike
    A[keep:, keep:] = T[keep:, keep:]
    _check(A, *solver(A))


def test_degenerate_and_diagonal(solver):
    a = np.diag([3.0, 1.0, 1.0, 1.0, -2.0, 0.0])
    w, v = solver(a)
    assert np.array_equal(w, [-2.0, 0.0, 1.0, 1.0, 1.0, 3.0])
    _check(a, w, v)
    rng = np.random.default_rng(3)
    q = np.linalg.qr(rng.standard_normal((20, 20)))[0]
    a = q @ np.diag(np.repeat([-1.0, 0.5, 2.0, 7.0], 5)) @ q.T
    _check(a, *solver(a))


def test_hermitian(solver):
    rng = np.random.default_rng(4)
    for n in (1, 5, 32, 64):
        x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        a = x + x.conj().T
        w, v = solver(a)
        ref = np.linalg.eigvalsh(a)
        assert np.max(np.abs(w - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
        # the lowest eigenvector (the one the exact-diagonalisation branches use)
        assert np.linalg.norm(a @ v[:, 0] - w[0] * v[:, 0]) < 1e-11 * max(1.0, np.max(np.abs(ref)))


def test_real_input_gives_real_vectors(solver):
    """The exact-diagonalisation branches of the eigensolvers read the real part
    of the eigenvectors of real operators: they must not come back as i v."""
    rng = np.random.default_rng(5)
    for n in (3, 16, 40):
        a = rng.standard_normal((n, n))
        a = a + a.T
        w, v = solver(a.astype(np.complex128), real_aware=True)
        assert np.max(np.abs(v.imag)) == 0
        _check(a, w, v.real)
