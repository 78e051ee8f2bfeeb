"""Many independent chains at once: the segmented batched Lanczos
(dmrgx_eigsh_batched) on a block-diagonal H_eff plan built from several
chains' get_map tasks, and lockstep batched sweeps (dmrgx_run_batch).

Pinned per chain: eigenvalues and vectors against dense diagonalisation of the
oracle's H_eff of that chain alone, and whole batched runs against individual
runs of the same chains (dmrgx_run_dmrg, itself pinned against exact
diagonalisation) and, for short chains, against exact diagonalisation."""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg

from dmrg_amd import driver
from oracle import exact
from oracle.dmrg import dense_heff
from tests.cases import dmrg_case

pytestmark = pytest.mark.gpu

COUPLINGS = json.loads((Path(__file__).parent / "golden" / "batch_couplings_64.json").read_text())["couplings"]


def _merged_plan(gpu, cases, prec):
    """One plan over the chains' concatenated psi / store / environments."""
    rflat, lflat, roff, rsize, loff, lsize, rmap, lmap = [], [], [], [], [], [], [], []
    P = S = RT = LT = 0
    for c in cases:
        rb, lb = len(roff), len(loff)
        rflat.append(c.rflat)
        lflat.append(c.lflat)
        roff += [o + RT for o in c.roff]
        loff += [o + LT for o in c.loff]
        rsize += list(c.rsize)
        lsize += list(c.lsize)
        rmap += [(i + rb, ib + P, ob + S, row, col, beta, alpha) for (i, ib, ob, row, col, beta, alpha) in c.rmap]
        lmap += [(i + lb, ib + S, ob + P, row, col, beta, alpha) for (i, ib, ob, row, col, beta, alpha) in c.lmap]
        P += c.n_psi
        S += c.n_store
        RT += c.rflat.size
        LT += c.lflat.size
    dt = gpu.C64 if np.dtype(cases[0].dtype).kind == "c" else gpu.R64
    R = gpu.DeviceArray.from_numpy(np.concatenate(rflat))
    L = gpu.DeviceArray.from_numpy(np.concatenate(lflat))
    h = gpu.Heff(dt, gpu.PREC_NAMES[prec], R, np.array(roff), np.array(rsize), rmap, L, np.array(loff),
                 np.array(lsize), lmap, P, S)
    return h


def _cases():
    """Three XYZ chains of the batched workload's couplings and one chain small
    enough (n <= 64) for the dense branch."""
    cases = [dmrg_case("xyz", 10, 24, 1, jx=j[0], jy=j[1], jz=j[2])[0] for j in COUPLINGS[:3]]
    j = COUPLINGS[3]
    cases.insert(1, dmrg_case("xyz", 6, 4, 1, jx=j[0], jy=j[1], jz=j[2])[0])
    assert cases[1].n_psi <= 64 < min(c.n_psi for c in cases[:1] + cases[2:])
    return cases


def test_eigsh_batched_matches_dense(gpu):
    print("case eigsh_batched fp64", flush=True)
    cases = _cases()
    h = _merged_plan(gpu, cases, "fp64")
    lens = [c.n_psi for c in cases]
    ev, vec, info, res = h.eigsh_batched(lens, nev=6, tol=1e-10, seeds=[11, 12, 13, 14])
    x = vec.numpy()
    off = np.cumsum([0] + lens)
    for k, c in enumerate(cases):
        H = dense_heff(c.ctx, c.n_psi, c.dtype)
        w, v = scipy.linalg.eigh(H)
        ne = min(6, c.n_psi)
        assert info[k, 2] == 1
        assert np.max(np.abs(ev[k, :ne] - w[:ne])) < 1e-9 * max(1.0, abs(w[0])), (k, ev[k], w[:6])
        xk = x[off[k]:off[k + 1]]
        assert abs(np.linalg.norm(xk) - 1) < 1e-12
        r = H @ xk - ev[k, 0] * xk
        assert np.linalg.norm(r) < 1e-7 * max(1.0, abs(ev[k, 0])), k
        assert abs(res[k] - np.linalg.norm(r)) < 1e-9 * max(1.0, abs(ev[k, 0])), k


def test_eigsh_batched_equals_single(gpu):
    """Each chain of the batch reaches the eigenvalues of its own dmrgx_eigsh."""
    print("case eigsh_batched vs single", flush=True)
    cases = _cases()
    h = _merged_plan(gpu, cases, "fp64")
    ev, _, _, _ = h.eigsh_batched([c.n_psi for c in cases], nev=6, tol=1e-10)
    for k, c in enumerate(cases):
        hk = _merged_plan(gpu, [c], "fp64")
        e1, _, _, _ = hk.eigsh(nev=6, tol=1e-10)
        ne = min(6, c.n_psi)
        assert np.max(np.abs(ev[k, :ne] - e1[:ne])) < 1e-9 * max(1.0, abs(e1[0]))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_eigsh_batched_mixed_polish(gpu, prec):
    print(f"case eigsh_batched {prec}", flush=True)
    cases = _cases()
    lo = _merged_plan(gpu, cases, prec)
    hi = _merged_plan(gpu, cases, "fp64")
    ev, _, info, res = lo.eigsh_batched([c.n_psi for c in cases], nev=6, tol=1e-6 if prec == "fp32" else 1e-4,
                                        refine=hi, refine_tol=1e-10)
    assert info[0, 3] > 0 and np.all(info[:, 4] == 1)
    for k, c in enumerate(cases):
        w = scipy.linalg.eigvalsh(dense_heff(c.ctx, c.n_psi, c.dtype))
        assert abs(ev[k, 0] - w[0]) < 1e-8 * max(1.0, abs(w[0])), (k, ev[k, 0], w[0])
        assert res[k] <= 1e-10 * max(1.0, abs(w[0])), (k, res[k])


def test_run_batch_matches_individual_runs(gpu):
    L, m, sweeps = 12, 32, 2
    J = COUPLINGS[:4]
    print(f"case run_batch xyz L={L} m={m} chains={len(J)}", flush=True)
    b = driver.run_batch("xyz", L, m, sweeps, J)
    assert b.stats["steps"] == len(b.steps) > 0
    for k, j in enumerate(J):
        one = driver.run_dmrg("xyz", L, m, sweeps, j)
        assert abs(b.energies[k] - one.energy) < 1e-9 * abs(one.energy), (k, b.energies[k], one.energy)
        assert abs(b.trunc[k] - one.stats["trunc_err"]) < 1e-9
    # L = 10: the bond dimension 32 holds the exact ground state
    b10 = driver.run_batch("xyz", 10, m, 1, J)
    for k, j in enumerate(J):
        e_ed = exact.xyz_e0(10, *j)[0]
        assert abs(b10.energies[k] - e_ed) < 1e-9 * abs(e_ed), (k, b10.energies[k], e_ed)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_run_batch_mixed_precision(gpu, prec):
    L, m = 12, 32
    J = COUPLINGS[4:8]
    print(f"case run_batch {prec} L={L} m={m}", flush=True)
    ref = driver.run_batch("xyz", L, m, 2, J)
    mix = driver.run_batch("xyz", L, m, 2, J, prec=prec, tol_lo=1e-5, refine_k=10)
    assert np.max(np.abs(mix.energies - ref.energies)) < 1e-8, (mix.energies, ref.energies)


def test_merged_plan_applies_each_chain(gpu):
    """The merged plan maps every chain's segment as that chain's own plan does."""
    print("case merged apply", flush=True)
    cases = _cases()
    h = _merged_plan(gpu, cases, "fp64")
    rng = np.random.default_rng(3)
    x = [rng.standard_normal(c.n_psi) + 1j * rng.standard_normal(c.n_psi) for c in cases]
    X = gpu.DeviceArray.from_numpy(np.concatenate(x))
    Y = gpu.DeviceArray(h.n_psi, np.complex128)
    h.apply(X, Y)
    y = Y.numpy()
    off = np.cumsum([0] + [c.n_psi for c in cases])
    for k, c in enumerate(cases):
        ref = c.oracle_apply(x[k])
        assert np.max(np.abs(y[off[k]:off[k + 1]] - ref)) < 1e-12 * np.max(np.abs(ref)), k


def test_eigsh_batched_graphs_equal_eager(gpu):
    """The hipGraph replays of the Lanczos steps give bit-identical results to
    the eager launches (the kernels are deterministic), across the restart
    cycles in which every graph is replayed."""
    print("case eigsh_batched graphs vs eager", flush=True)
    cases = _cases()
    h = _merged_plan(gpu, cases, "fp64")
    lens = [c.n_psi for c in cases]
    lib = gpu.lib()
    try:
        lib.dmrgx_set_lanczos_graphs(0)
        e0, v0, i0, _ = h.eigsh_batched(lens, nev=6, tol=1e-10, seeds=[5, 6, 7, 8])
        lib.dmrgx_set_lanczos_graphs(1)
        e1, v1, i1, _ = h.eigsh_batched(lens, nev=6, tol=1e-10, seeds=[5, 6, 7, 8])
        e2, v2, i2, _ = h.eigsh_batched(lens, nev=6, tol=1e-10, seeds=[5, 6, 7, 8])
    finally:
        lib.dmrgx_set_lanczos_graphs(1)
    assert i0[:, 1].max() > 0  # restart cycles happened: graphs were replayed
    assert np.array_equal(e0, e1) and np.array_equal(e1, e2)
    assert np.array_equal(v0.numpy(), v1.numpy()) and np.array_equal(v1.numpy(), v2.numpy())
    assert np.array_equal(i0, i1)
