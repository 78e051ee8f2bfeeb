"""Full DMRG runs on the GPU through the C++ host mirror (dmrgx_run_dmrg)
against the oracle's restatement of the same schedule and against exact
energies (BdG closed form, exact diagonalisation)."""
import numpy as np
import pytest

from dmrg_amd import driver
from oracle import exact, models
from oracle.dmrg import Runner

pytestmark = pytest.mark.gpu

KITAEV = (-1.0, -1.0, 0.25)          # t=Delta=1, mu=0.5 in the reference's (T, P, U)
XYZ = (1.0, 0.8, 0.5)
POTTS = (1.0, 1.0, 0.0, 3.14159265359 / 8)


def _oracle(model, L, m, sweeps, schedule, sign_rule, params):
    names = {"kitaev": ("t", "p", "u"), "xyz": ("jx", "jy", "jz"), "potts3": ("f", "j", "p", "t")}[model]
    kw = dict(zip(names, params))
    mpo, n, ferm = models.make_model(model, L, **kw)
    R = Runner(mpo, n, ferm, sign_rule=sign_rule)
    e = R.dmrg2(m, sweeps) if schedule == "dmrg2" else R.dmrg(m, sweeps)
    return e, R


@pytest.mark.parametrize("schedule", ["dmrg2", "dmrg"])
def test_kitaev_closed_form(gpu, schedule):
    L = 12
    print(f"case kitaev L={L} {schedule}", flush=True)
    res = driver.run_dmrg("kitaev", L, 32, 2, KITAEV, schedule=schedule)
    e_bdg, _ = exact.kitaev_bdg_e0(L, *KITAEV)
    e_ed = exact.kitaev_ed(L, *KITAEV)[0]
    assert abs(res.energy - e_ed) < 1e-9, (res.energy, e_ed, e_bdg)


@pytest.mark.parametrize("schedule", ["dmrg2", "dmrg"])
def test_xyz_exact_diagonalisation(gpu, schedule):
    L = 10
    print(f"case xyz L={L} {schedule}", flush=True)
    res = driver.run_dmrg("xyz", L, 64, 2, XYZ, schedule=schedule)
    e_ed = exact.xyz_e0(L, *XYZ)[0]
    assert abs(res.energy - e_ed) < 1e-9, (res.energy, e_ed)


def test_potts3_exact_diagonalisation(gpu):
    L = 7
    print(f"case potts3 L={L}", flush=True)
    res = driver.run_dmrg("potts3", L, 81, 2, POTTS)
    W = models.H_Potts3(*POTTS).remove_symmetry().val[0]
    e_ed = exact.mpo_e0(W, L, 3, 3, 0)[0]
    assert abs(res.energy - e_ed) < 1e-9, (res.energy, e_ed)


def test_xyz_matches_oracle_step_by_step(gpu):
    """Energies, dimensions and truncation of every step agree with the
    oracle's schedule (same schedule, same truncation)."""
    L, m = 10, 16
    print("case xyz oracle-vs-gpu physical", flush=True)
    res = driver.run_dmrg("xyz", L, m, 1, XYZ, schedule="dmrg2")
    e, R = _oracle("xyz", L, m, 1, "dmrg2", "physical", XYZ)
    assert len(R.log) == len(res.steps)
    for g, o in zip(res.steps, R.log):
        assert (int(g[0]), int(g[1])) == (o.l, o.r)
        assert int(g[16]) == o.dim
        assert abs(g[2] - o.energies[0]) < 1e-8 * max(1, abs(o.energies[0])), (g[:3], o.energies[0])
    assert abs(res.energy - e) < 1e-8


def _oracle_steps(model, L, m, sweeps, sign_rule, params, rotation_seed=None):
    from oracle import qtensor as qt
    old = qt.COMPLEMENT_ROTATION_SEED
    qt.COMPLEMENT_ROTATION_SEED = rotation_seed
    try:
        e, R = _oracle(model, L, m, sweeps, "dmrg2", sign_rule, params)
    finally:
        qt.COMPLEMENT_ROTATION_SEED = old
    return e, np.array([o.energies[0] for o in R.log]), R


def test_xyz_reference_sign_rule_step_by_step(gpu):
    """The reference's unconditional sign rule on the bosonic chain, step by
    step.  When fewer singular values than the cutoff exist, the schedule pads
    U / V with orthonormal complement vectors (src/qtensor.cpp:1221-1252) that
    carry no weight and depend on the SVD routine (LAPACK's here, a Householder
    complement on the GPU).  The oracle itself, rerun with two other orthonormal
    complements, marks the steps whose energy depends on that choice; every
    other step and the final energy must agree with the oracle to 1e-8."""
    L, m = 10, 16
    print("case xyz oracle-vs-gpu reference rule", flush=True)
    res = driver.run_dmrg("xyz", L, m, 1, XYZ, schedule="dmrg2", sign_rule="reference")
    e, ref, R = _oracle_steps("xyz", L, m, 1, "reference", XYZ)
    _, alt1, _ = _oracle_steps("xyz", L, m, 1, "reference", XYZ, rotation_seed=5)
    _, alt2, _ = _oracle_steps("xyz", L, m, 1, "reference", XYZ, rotation_seed=6)
    sensitive = (np.abs(alt1 - ref) > 1e-9) | (np.abs(alt2 - ref) > 1e-9)
    assert sensitive.any()  # the reference rule makes these energies LAPACK-specific
    assert len(R.log) == len(res.steps)
    gpu_e = res.steps[:, 2]
    for k in range(len(ref)):
        assert (int(res.steps[k, 0]), int(res.steps[k, 1])) == (R.log[k].l, R.log[k].r)
        if not sensitive[k]:
            assert abs(gpu_e[k] - ref[k]) < 1e-8 * max(1, abs(ref[k])), (k, gpu_e[k], ref[k])
    assert abs(res.energy - e) < 1e-8, (res.energy, e)
    assert abs(e - exact.xyz_e0(L, *XYZ)[0]) > 1e-3  # not the physical ground state


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_energy(gpu, prec):
    L, m = 16, 32
    print(f"case xyz mixed {prec}", flush=True)
    ref = driver.run_dmrg("xyz", L, m, 2, XYZ, ladder=False)
    mix = driver.run_dmrg("xyz", L, m, 2, XYZ, ladder=False, prec=prec,
                          tol_lo=1e-7 if prec == "fp32" else 1e-5, refine_k=10)
    assert abs(mix.energy - ref.energy) < 1e-8, (mix.energy, ref.energy)


@pytest.mark.parametrize("model,params", [("xyz", XYZ), ("potts3", POTTS)])
def test_zero_slice_skipping_sweep(gpu, model, params):
    """A sweep whose operators are large enough for split-K partial tiles in
    the row-mapped R-pass groups: every step's energy with the zero MPO-bond
    slices skipped equals the one with every slice contracted."""
    L, m = (20, 96) if model == "xyz" else (12, 108)
    print(f"case skip zeros sweep {model} L={L} m={m}", flush=True)
    lib = gpu.lib()
    on = driver.run_dmrg(model, L, m, 1, params, schedule="dmrg2", ladder=False)
    try:
        lib.dmrgx_set_skip_zero_slices(0)
        off = driver.run_dmrg(model, L, m, 1, params, schedule="dmrg2", ladder=False)
    finally:
        lib.dmrgx_set_skip_zero_slices(1)
    assert len(on.steps) == len(off.steps)
    d = np.abs(on.steps[:, 2] - off.steps[:, 2])
    print(f"  max step |dE| {d.max():.2e}, E0 {on.energy:.12f}", flush=True)
    assert d.max() < 1e-8 * abs(on.energy), d.max()


@pytest.mark.parametrize("model,params,prec,L,m", [("xyz", XYZ, "fp64", 16, 32), ("xyz", XYZ, "bf16x3", 16, 32),
                                                   ("potts3", POTTS, "bf16x3", 10, 54),
                                                   ("kitaev", KITAEV, "fp64", 16, 32)])
def test_warm_start(gpu, model, params, prec, L, m):
    """Sweep steps started from White's prediction of the previous step's
    state reach the energies of random starts (FP64, six pairs) with fewer
    eigensolver matvecs; the warm steps converge and report the lowest pair
    only."""
    print(f"case warm start {model} {prec} L={L} m={m}", flush=True)
    plain = driver.run_dmrg(model, L, m, 2, params, ladder=False)
    warm = driver.run_dmrg(model, L, m, 2, params, ladder=False, prec=prec, warm_start=True, tol_lo=1e-4,
                           refine_k=8)
    assert len(warm.steps) == len(plain.steps)
    assert (warm.steps[:, :2] == plain.steps[:, :2]).all()
    grow = L // 2  # two_sites plus the growth steps: random starts, six pairs
    assert (warm.steps[:grow, 3] != 0).any()
    assert (warm.steps[grow:, 3:8] == 0).all()  # E1..E5 of the warm steps
    tail = 2 * (L - 4)  # the last full sweep and the final half sweep
    d = np.abs(warm.steps[-tail:, 2] - plain.steps[-tail:, 2])
    mv_w, mv_p = warm.steps[grow:, 10].sum(), plain.steps[grow:, 10].sum()
    print(f"  E0 {warm.energy:.12f} vs {plain.energy:.12f}, max step |dE| {d.max():.2e}, "
          f"sweep matvecs {mv_w:.0f} vs {mv_p:.0f}", flush=True)
    assert abs(warm.energy - plain.energy) < 1e-8 * max(1, abs(plain.energy))
    assert d.max() < 1e-7 * max(1, abs(plain.energy))
    assert mv_w < mv_p
