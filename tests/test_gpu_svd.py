"""GPU batched one-sided Jacobi SVD (dmrgx_block_svd) against numpy/LAPACK and the
special cases of tensor<T>::svd (src/tensor.cpp:736-785)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _svd(gpu, blocks, cplx):
    dt = np.complex128 if cplx else np.float64
    rows = np.array([b.shape[0] for b in blocks], dtype=np.int32)
    cols = np.array([b.shape[1] for b in blocks], dtype=np.int32)
    a_off = np.concatenate([[0], np.cumsum([b.size for b in blocks])[:-1]]).astype(np.int64)
    u_off = np.concatenate([[0], np.cumsum([r * r for r in rows])[:-1]]).astype(np.int64)
    v_off = np.concatenate([[0], np.cumsum([c * c for c in cols])[:-1]]).astype(np.int64)
    s_off = np.concatenate([[0], np.cumsum([min(r, c) for r, c in zip(rows, cols)])[:-1]]).astype(np.int64)
    A = gpu.DeviceArray.from_numpy(np.concatenate([b.ravel() for b in blocks]).astype(dt))
    U = gpu.DeviceArray(int(sum(r * r for r in rows)), dt)
    V = gpu.DeviceArray(int(sum(c * c for c in cols)), dt)
    S = gpu.DeviceArray(int(sum(min(r, c) for r, c in zip(rows, cols))), np.float64)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    gpu.check(gpu.lib().dmrgx_block_svd(gpu.C64 if cplx else gpu.R64, len(blocks), P(rows, C.c_int32),
                                        P(cols, C.c_int32), A.ptr, P(a_off, C.c_int64), U.ptr,
                                        P(u_off, C.c_int64), S.ptr, P(s_off, C.c_int64), V.ptr,
                                        P(v_off, C.c_int64)), "block_svd")
    u, s, v = U.numpy(), S.numpy(), V.numpy()
    out = []
    for b, blk in enumerate(blocks):
        r, c = blk.shape
        out.append((u[u_off[b]:u_off[b] + r * r].reshape(r, r), s[s_off[b]:s_off[b] + min(r, c)],
                    v[v_off[b]:v_off[b] + c * c].reshape(c, c)))
    return out


@pytest.mark.parametrize("cplx", [False, True])
def test_block_svd_random(gpu, cplx):
    rng = np.random.default_rng(5)
    shapes = [(40, 40), (64, 24), (24, 64), (1, 7), (9, 1), (100, 100), (3, 5)]
    blocks = []
    for r, c in shapes:
        a = rng.standard_normal((r, c))
        if cplx:
            a = a + 1j * rng.standard_normal((r, c))
        blocks.append(a)
    # graded spectrum (DMRG-like decay) to exercise tiny singular values
    q1, _ = np.linalg.qr(rng.standard_normal((60, 60)))
    q2, _ = np.linalg.qr(rng.standard_normal((60, 60)))
    blocks.append(q1 @ np.diag(np.exp(-0.5 * np.arange(60))) @ q2)
    res = _svd(gpu, blocks, cplx)
    for a, (u, s, v) in zip(blocks, res):
        ref = np.linalg.svd(a, compute_uv=False)
        assert np.allclose(s, ref, rtol=1e-12, atol=1e-14 * ref[0])
        assert np.all(np.diff(s) <= 1e-15 * s[0])
        k = len(s)
        recon = (u[:, :k] * s) @ v[:k, :]
        assert np.max(np.abs(recon - a)) < 1e-12 * max(1, np.max(np.abs(a)))
        assert np.max(np.abs(u.conj().T @ u - np.eye(u.shape[0]))) < 1e-12
        assert np.max(np.abs(v @ v.conj().T - np.eye(v.shape[0]))) < 1e-12


def test_block_svd_special_cases(gpu):
    z = np.zeros((3, 4))
    one = np.array([[-2.5]])
    tiny = np.full((2, 2), 1e-12)
    res = _svd(gpu, [z, one, tiny], False)
    u, s, v = res[0]
    assert np.all(s == 0) and np.allclose(u, np.eye(3)) and np.allclose(v, np.eye(4))
    u, s, v = res[1]
    assert u[0, 0] == 1 and s[0] == 2.5 and v[0, 0] == -1
    u, s, v = res[2]  # every |x| <= 1e-10 counts as zero
    assert np.all(s == 0) and np.allclose(u, np.eye(2))


def test_block_svd_rank_deficient_completion(gpu):
    rng = np.random.default_rng(9)
    a = rng.standard_normal((30, 3)) @ rng.standard_normal((3, 20))  # rank 3, tall
    b = rng.standard_normal((4, 2)) @ rng.standard_normal((2, 12))   # rank 2, wide
    res = _svd(gpu, [a, b], False)
    for x, (u, s, v) in zip([a, b], res):
        assert np.max(np.abs(u.T @ u - np.eye(u.shape[0]))) < 1e-12
        assert np.max(np.abs(v @ v.T - np.eye(v.shape[0]))) < 1e-12
        k = len(s)
        assert np.max(np.abs((u[:, :k] * s) @ v[:k, :] - x)) < 1e-11 * np.max(np.abs(x))


@pytest.mark.parametrize("cplx", [False, True])
def test_block_svd_large_graded_sectors(gpu, cplx):
    """Sector blocks at DMRG sizes with an exponentially decaying spectrum: a
    square 512 block, a tall block with more than 1024 rows (the wide panel
    path) and a wide one whose dimensions are not multiples of the panel or
    column-block widths."""
    rng = np.random.default_rng(21)

    def graded(r, c, decay):
        k = min(r, c)
        q1 = np.linalg.qr(rng.standard_normal((r, k)) + (1j * rng.standard_normal((r, k)) if cplx else 0))[0]
        q2 = np.linalg.qr(rng.standard_normal((c, k)) + (1j * rng.standard_normal((c, k)) if cplx else 0))[0]
        return (q1 * np.exp(-decay * np.arange(k) / k)) @ q2.conj().T

    blocks = [graded(512, 512, 30.0), graded(1100, 1030, 25.0), graded(203, 517, 20.0)]
    res = _svd(gpu, blocks, cplx)
    for a, (u, s, v) in zip(blocks, res):
        ref = np.linalg.svd(a, compute_uv=False)
        # backward-stable bound: both this SVD (Jacobi stopping tolerance
        # max(rows, cols) eps) and LAPACK's are accurate to ~dim * eps * s_max
        assert np.max(np.abs(s - ref)) < 4 * max(a.shape) * 2.2e-16 * ref[0]
        assert np.all(np.diff(s) <= 1e-15 * s[0])
        k = len(s)
        assert np.max(np.abs((u[:, :k] * s) @ v[:k, :] - a)) < 1e-12 * np.max(np.abs(a))
        assert np.max(np.abs(u.conj().T @ u - np.eye(u.shape[0]))) < 1e-12
        assert np.max(np.abs(v @ v.conj().T - np.eye(v.shape[0]))) < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
def test_block_svd_many_blocks(gpu, cplx):
    """128 sector blocks in one call (the blocks of 64 chains' two-site states):
    every block is factorised as if it were alone."""
    rng = np.random.default_rng(5)

    def graded(r, c, decay):
        k = min(r, c)
        q1 = np.linalg.qr(rng.standard_normal((r, k)) + (1j * rng.standard_normal((r, k)) if cplx else 0))[0]
        q2 = np.linalg.qr(rng.standard_normal((c, k)) + (1j * rng.standard_normal((c, k)) if cplx else 0))[0]
        return (q1 * np.exp(-decay * np.arange(k) / k)) @ q2.conj().T

    shapes = [(128, 128), (128, 128), (96, 160), (160, 96)]
    blocks = [graded(*shapes[i % 4], 10.0 + (i % 7)) for i in range(128)]
    print(f"case block_svd 128 blocks cplx={cplx}", flush=True)
    res = _svd(gpu, blocks, cplx)
    for i, (a, (u, s, v)) in enumerate(zip(blocks, res)):
        ref = np.linalg.svd(a, compute_uv=False)
        assert np.max(np.abs(s - ref)) < 4 * max(a.shape) * 2.2e-16 * ref[0], i
        k = len(s)
        assert np.max(np.abs((u[:, :k] * s) @ v[:k, :] - a)) < 1e-12 * np.max(np.abs(a)), i
        assert np.max(np.abs(u.conj().T @ u - np.eye(u.shape[0]))) < 1e-12, i
        assert np.max(np.abs(v @ v.conj().T - np.eye(v.shape[0]))) < 1e-12, i
