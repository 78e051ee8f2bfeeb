"""GPU FP64 Lanczos (dmrgx_eigsh) against dense diagonalisation of the oracle's
H_eff and against the oracle's ARPACK call (src/dsaupd.cpp / src/znaupd.cpp)."""
import numpy as np
import pytest
import scipy.linalg

from oracle.dmrg import dense_heff
from tests.cases import dmrg_case, dmrg_step_case

pytestmark = pytest.mark.gpu


def _plan(gpu, case, prec):
    c = gpu.C64 if np.dtype(case.dtype).kind == "c" else gpu.R64
    R = gpu.DeviceArray.from_numpy(case.rflat)
    Lb = gpu.DeviceArray.from_numpy(case.lflat)
    return gpu.Heff(c, gpu.PREC_NAMES[prec], R, case.roff, case.rsize, case.rmap, Lb, case.loff, case.lsize,
                    case.lmap, case.n_psi, case.n_store)


@pytest.mark.parametrize("model,kw", [
    ("xyz", dict(jx=1.0, jy=0.8, jz=0.5)),
    ("kitaev", dict(t=-1.0, p=-1.0, u=0.25)),
])
def test_eigsh_matches_dense(gpu, model, kw):
    print(f"case eigsh {model}", flush=True)
    case, _ = dmrg_case(model, 10, 24, 1, **kw)
    H = dense_heff(case.ctx, case.n_psi, case.dtype)
    assert np.max(np.abs(H - H.conj().T)) < 1e-12 * np.max(np.abs(H))
    w, v = scipy.linalg.eigh(H)
    h = _plan(gpu, case, "fp64")
    ev, vec, info, res = h.eigsh(nev=6, tol=1e-10)
    assert info[2] == 1 and info[4] == 1
    assert np.max(np.abs(ev - w[:6])) < 1e-9 * max(1.0, abs(w[0]))
    x = vec.numpy()
    ov = abs(np.vdot(v[:, 0], x))
    assert abs(ov - 1) < 1e-8
    # residual of the returned pair, and the one the entry point reports
    r = H @ x - ev[0] * x
    assert np.linalg.norm(r) < 1e-7 * max(1.0, abs(ev[0]))
    assert abs(res - np.linalg.norm(r)) < 1e-9 * max(1.0, abs(ev[0]))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_eigsh_mixed_with_fp64_polish(gpu, prec):
    print(f"case eigsh mixed {prec}", flush=True)
    case, _ = dmrg_case("xyz", 10, 24, 1, jx=1.0, jy=0.8, jz=0.5)
    H = dense_heff(case.ctx, case.n_psi, case.dtype)
    w = scipy.linalg.eigvalsh(H)
    lo = _plan(gpu, case, prec)
    hi = _plan(gpu, case, "fp64")
    ev, vec, info, res = lo.eigsh(nev=6, tol=1e-6 if prec == "fp32" else 1e-4, refine=hi, refine_tol=1e-10)
    assert abs(ev[0] - w[0]) < 1e-8 * max(1.0, abs(w[0])), (ev[0], w[0])
    assert info[3] > 0 and info[4] == 1
    x = vec.numpy()
    r = np.linalg.norm(H @ x - ev[0] * x)
    assert res <= 1e-10 * abs(ev[0]) and r < 1e-9 * abs(ev[0]), (res, r)


def test_eigsh_spectrum_far_from_zero(gpu):
    """A DMRG operator whose spectrum lies far from zero (all eigenvalues
    within [-38, -10], so |alpha| is about 5 beta at every Lanczos step): the
    warm-up step l=8 of an open XYZ chain L=32 at bond dimension 10.  One
    classical Gram-Schmidt pass per step loses orthogonality here within a
    dozen steps and the convergence test then passes on wrong Ritz pairs."""
    print("case eigsh far-from-zero spectrum", flush=True)
    case = dmrg_step_case("xyz", 32, 128, 8, jx=1.1268656294706476, jy=1.4919361875572013, jz=1.0733433358684978)
    H = dense_heff(case.ctx, case.n_psi, case.dtype)
    w = scipy.linalg.eigvalsh(H)
    assert w[-1] < 0 and abs(w[0]) > 30
    h = _plan(gpu, case, "fp64")
    ev, vec, info, _ = h.eigsh(nev=6, tol=1e-10, max_restart=5000)
    assert info[2] == 1
    assert np.max(np.abs(ev - w[:6])) < 1e-9 * abs(w[0]), (ev, w[:6])
    x = vec.numpy()
    assert np.linalg.norm(H @ x - ev[0] * x) < 1e-7 * abs(ev[0])
    # the same operator three times in one batched solve
    from tests.test_gpu_batch import _merged_plan
    hb = _merged_plan(gpu, [case] * 3, "fp64")
    evb, vb, infob, _ = hb.eigsh_batched([case.n_psi] * 3, nev=6, tol=1e-10, seeds=[1, 2, 3])
    xb = vb.numpy().reshape(3, -1)
    for k in range(3):
        assert infob[k, 2] == 1
        assert np.max(np.abs(evb[k] - w[:6])) < 1e-9 * abs(w[0]), (k, evb[k], w[:6])
        assert np.linalg.norm(H @ xb[k] - evb[k, 0] * xb[k]) < 1e-7 * abs(w[0])


@pytest.mark.parametrize("nev,ncv", [(6, None), (1, 24)])
def test_eigsh_mixed_refine_near_degenerate_pair(gpu, nev, ncv):
    """The mixed-precision C ABI path on a step whose two lowest states lie
    2.2e-5 apart: the warm-up step 14 of an open XYZ chain L=32 with Jz
    dominant (chain 63 of the batched workload's couplings).  The BF16x3
    Lanczos stops at 1e-4, below that gap (on six pairs, or on the lowest
    pair only with the sweep engine's 24-vector basis); the FP64 refinement
    starts from the four lowest Ritz vectors either way, so it must return the
    true ground state (energy to 1e-10, overlap with the exact vector to
    1e-8), not its partner."""
    print(f"case eigsh bf16x3 refine near-degenerate nev={nev}", flush=True)
    J = (0.54790153, 0.57295701, 1.46374402)
    case = dmrg_step_case("xyz", 32, 16, 14, jx=J[0], jy=J[1], jz=J[2])
    H = dense_heff(case.ctx, case.n_psi, case.dtype)
    w, v = scipy.linalg.eigh(H)
    assert 1e-6 < w[1] - w[0] < 1e-4, w[:2]
    lo = _plan(gpu, case, "bf16x3")
    hi = _plan(gpu, case, "fp64")
    ev, vec, info, res = lo.eigsh(nev=nev, ncv=ncv, tol=1e-4, refine=hi, refine_tol=1e-10)
    assert info[4] == 1 and info[3] >= 4
    assert abs(ev[0] - w[0]) < 1e-10 * abs(w[0]), (ev[0], w[0], w[1])
    x = vec.numpy()
    assert abs(abs(np.vdot(v[:, 0], x)) - 1) < 1e-8
    assert res <= 1e-10 * abs(w[0])
    assert abs(res - np.linalg.norm(H @ x - ev[0] * x)) < 1e-10 * abs(w[0])


def test_eigsh_refine_plan_must_match(gpu):
    """A refine plan of another psi length or element type, or of reduced
    precision, is an error, not an out-of-bounds read."""
    print("case eigsh refine mismatch", flush=True)
    from oracle.synth import synthetic_case
    a = synthetic_case(2, 5, 16, dtype=np.complex128)
    b = synthetic_case(2, 5, 20, dtype=np.complex128)
    r = synthetic_case(2, 5, 16, dtype=np.float64)
    assert a.n_psi == r.n_psi != b.n_psi
    lo = _plan(gpu, a, "bf16x3")
    for bad in (_plan(gpu, b, "fp64"), _plan(gpu, r, "fp64"), _plan(gpu, a, "fp32")):
        with pytest.raises(gpu.DmrgxError, match="refine plan"):
            lo.eigsh(nev=4, tol=1e-4, refine=bad)
        with pytest.raises(gpu.DmrgxError, match="refine plan"):
            lo.eigsh_batched([a.n_psi], nev=4, tol=1e-4, refine=bad)
