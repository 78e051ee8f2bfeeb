#!/usr/bin/env python3
"""Benchmark of the DMRG two-site H_eff matvec and of full DMRG sweeps on one
MI355X.

Workload (the JSON line's value): the XYZ spin-1/2 chain at L=100, m=512 (Z2
sectors of 256, MPO bond D=5, complex): one step = one H_eff matvec over the
exact block structure of that step (32 R-pass + 32 L-pass tasks of
(1280,256,256) / (256,256,1280), 4.29e10 algorithmic flop), synthetic
uniform(-1,1) environments built by the library on the device.  The
mixed-precision path (BF16x3 products, FP32 accumulate) is the headline value;
FP64 and FP32 run on the same shapes.

Beside it, in the same JSON line:
* ``sweep``: full two-site DMRG of that chain (XYZ L=100, m=512, the
  reference's dmrg2 schedule: growth at cutoff 10, then sweeps at m; the
  cutoff ladder is off unless --ladder), FP64 against BF16x3 with the FP64
  refinement: seconds of the last full sweep (steady state at m) and of every
  sweep, E0 of both and |dE0|, truncation error, fidelity of the final states.
  --sweep-model potts3 --sweep-L 64 --sweep-m 1024 runs the chiral Z3 Potts
  chain instead.
* ``batch`` (--batch N): the 64-chain batched workload (open XYZ chains L=32, m=128,
  couplings U[0.5, 1.5) from a seeded splitmix64 stream) swept in lockstep,
  FP64 and BF16x3, beside the same chains run one at a time.
* ``cpu_baseline``: the oracle's av() (numpy, one OpenBLAS GEMM per get_map
  task) at the same m=512 shapes on the host, all threads and one thread,
  with the host CPU model; ``cpu_sweep`` (--cpu-sweep): the Kitaev chain
  L=40, m=64, 5 two-site sweeps (Z2, real FP64) run by the oracle on the host
  and by the FP64 GPU path.
``--part svd`` / ``--part copy`` time the step SVD alone / the 256 MiB copy
used to calibrate the HBM byte counters.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "DMRG sweep s + H_eff matvec TFLOP/s, mixed vs FP64 GPU at |ΔE0|<1e-8"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}  # MI355X dense matrix peaks, TFLOP/s
HBM_PEAK = 8000.0  # GB/s

# HBM bytes per launch of k_gemm_bf3p<complex>, per workload and work
# decomposition the autotuner can pick: (workload, pass, work items) ->
# (FETCH_SIZE KiB, WRITE_SIZE KiB per dispatch), from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of `bench.py --only bf16x3` (each candidate is
# dispatched while the plan is tuned); each counter is scaled by what it reads
# for a plain 256 MiB copy kernel (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  The
# roofline's traffic is the mean over the R and L launches of the chosen plan,
# None if the chosen decomposition was not measured.
TRAFFIC_KIB = {
    ("synthetic", "R", 320): (47233.0, 47852.5), ("synthetic", "R", 400): (52067.8, 56371.7),
    ("synthetic", "R", 640): (61582.8, 81920.0), ("synthetic", "L", 64): (82023.6, 8192.0),
    ("synthetic", "L", 256): (82220.4, 32768.0), ("synthetic", "L", 512): (82258.0, 65536.0),
    ("synthetic", "L", 640): (82051.6, 81920.0),
    # the centre operator of a converged XYZ m=512 chain (zero MPO slices
    # skipped), measured on the L=20 chain's centre (the same block structure)
    ("real_xyz_m512", "R", 192): (15532.0, 27819.2), ("real_xyz_m512", "L", 192): (24694.1, 24576.0),
}
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0


def workload_key(real, a):
    """The traffic table's workload: the synthetic shapes, or the centre
    operator of a converged chain, whose block structure (sectors, MPO bond,
    the zero slices the plan skips) depends on the model and m, not on L."""
    if not real:
        return "synthetic"
    return f"real_{a.sweep_model}_m{a.sweep_m}"


def traffic_bytes(workload, r_items, l_items):
    r, l = TRAFFIC_KIB.get((workload, "R", r_items)), TRAFFIC_KIB.get((workload, "L", l_items))
    if r is None or l is None:
        return None
    return round(0.5 * sum(f * FETCH_SCALE + w * WRITE_SCALE for f, w in (r, l)) * 1024)


# algorithmic bytes per launch: env blocks (5 x 1024^2 complex) and the
# input operand as four bf16 planes (8 B per complex), output FP32 complex
ALG_BYTES = (5 * 1024 * 1024 + 524288) * 8 + 5 * 1024 * 1024 * 8

S, D, M = 2, 5, 512
L_CHAIN = 100
JXYZ = (1.0, 0.8, 0.5)
POTTS = (1.0, 1.0, 0.0, 3.14159265359 / 8)  # f, j, p, theta of H_Potts3 (the reference's truncated pi)
# tolerance of the BF16x3 Lanczos solve ahead of the FP64 refinement (its
# lowest pair; measured fastest of 1e-3 .. 3e-5 on the L=100 m=512 XYZ and
# L=64 m=1024 Potts sweeps)
TOL_LO = 3e-5
# FP64 residual tolerance of the refinement that follows it: half a digit
# past the FP64 solves' 1e-10, which leaves the refined vector's error in the
# high-energy directions small enough that the block SVD needs no more
# Jacobi sweeps than on the FP64 run's states (measured fastest of 1e-10 ..
# 1e-12 on the two sweeps above)
TOL_REFINE = 5e-11


def time_plan(h, prec, x, steps, warmup):
    """HIP-event timing of `steps` applies of plan h (after `warmup`)."""
    import dmrg_amd
    xd = dmrg_amd.DeviceArray.from_numpy(x)
    yd = dmrg_amd.DeviceArray(h.n_psi, np.complex128)
    if warmup:
        h.time(xd, yd, warmup)
    ms = h.time(xd, yd, steps)
    # each pass's GEMM kernel alone, launched back to back: the per-launch
    # duration of the roofline (the events around single launches inside an
    # apply add the event packets' own latency)
    kr, kl = h.time_kernels(xd, yd, steps)
    flops = h.flops
    _, _, r_items, l_items = h.split()
    out = dict(ms=ms[0], ms_r=kr, ms_l=kl, ms_r_bracketed=ms[1], ms_l_bracketed=ms[2], flops=flops,
               flops_r=ms[3], flops_l=ms[4],
               flops_exec=h.flops_exec, exec_r=ms[5], exec_l=ms[6],
               tflops=flops / (ms[0] * 1e-3) / 1e12, exec_tflops=h.flops_exec / (ms[0] * 1e-3) / 1e12,
               n_psi=h.n_psi, r_items=r_items, l_items=l_items)
    print(f"[bench] {prec}: {ms[0]:.3f} ms/matvec, {out['tflops']:.1f} TFLOP/s algorithmic, "
          f"{out['exec_tflops']:.1f} executed (kernels back to back: R {kr:.4f} ms, L {kl:.4f} ms; inside the "
          f"apply: R {ms[1]:.4f} ms, L {ms[2]:.4f} ms)", file=sys.stderr, flush=True)
    xd.free()
    yd.free()
    return out


def gpu_leg(steps, warmup, only=None):
    """The matvec on dense synthetic environments of the m=512 shapes (every
    MPO-bond slice nonzero, so nothing is skipped)."""
    import dmrg_amd
    from dmrg_amd.driver import SynthHeff

    dmrg_amd.init(0)
    bond = [M // S] * S
    out = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        if only and prec not in only:
            continue
        h = SynthHeff(S, D, bond, complex_=True, prec=prec, seed=20261203)
        rng = np.random.default_rng(1)
        x = rng.uniform(-1, 1, h.n_psi) + 1j * rng.uniform(-1, 1, h.n_psi)
        out[prec] = time_plan(h, prec, x, steps, warmup)
        h.close()
    return out


def state_leg(state, steps, warmup, only=None):
    """The matvec of the real operator: the two-site H_eff at the centre bond
    of a converged state (environments rebuilt from its MPS), applied to that
    state's two-site vector."""
    import ctypes as C
    import dmrg_amd

    out = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        if only and prec not in only:
            continue
        h = state.centre_heff(prec)
        x = np.zeros(h.n_psi, np.complex128)
        dmrg_amd.check(dmrg_amd.lib().dmrgx_memcpy_d2h(x.ctypes.data_as(C.c_void_p), C.c_void_p(h.psi_ptr),
                                                       x.nbytes), "d2h")
        out[prec] = time_plan(h, prec, x, steps, warmup)
        h.close()
    return out


def quick_state(model, L, m):
    """A converged-enough state for profiling the real operator without the
    full sweep comparison: BF16x3 growth and one sweep at m."""
    from dmrg_amd import driver
    params = {"xyz": JXYZ, "potts3": POTTS}[model]
    r = driver.run_dmrg(model, L, m, 1, params, schedule="dmrg2", prec="bf16x3", tol_lo=TOL_LO, refine_k=8,
                        ladder=False, keep_state=True)
    return r.state


def configs_leg(reps=5, warmup=2):
    """Matvec throughput at the other large configurations' shapes:
    Z3 chiral Potts at m=1024 with uneven sectors (0.40, 0.33, 0.27 m; d=3,
    MPO bond 4; 243 + 243 tasks) and the XYZ chain at m=2048 (Z2, MPO bond 5;
    FP32 is that configuration's forward precision)."""
    import dmrg_amd
    from dmrg_amd.driver import SynthHeff

    dmrg_amd.init(0)
    cases = {
        "potts3_m1024_uneven": (3, 4, [410, 338, 276], ("fp64", "fp32", "bf16x3")),
        "xyz_m2048": (2, 5, [1024, 1024], ("fp64", "fp32", "bf16x3")),
    }
    out = {}
    for name, (S_, D_, bond, precs) in cases.items():
        res = {"sectors": bond, "S": S_, "D": D_}
        for prec in precs:
            h = SynthHeff(S_, D_, bond, complex_=True, prec=prec, seed=20261203)
            rng = np.random.default_rng(3)
            x = rng.uniform(-1, 1, h.n_psi) + 1j * rng.uniform(-1, 1, h.n_psi)
            xd = dmrg_amd.DeviceArray.from_numpy(x)
            yd = dmrg_amd.DeviceArray(h.n_psi, np.complex128)
            h.time(xd, yd, warmup)
            ms = h.time(xd, yd, reps)
            res[prec] = {"ms": round(ms[0], 3), "tflops": round(h.flops / (ms[0] * 1e-3) / 1e12, 2)}
            if name == "xyz_m2048" and prec == "fp32":
                # the adjoint (cotangents of psi and both environments): one R pass
                # plus four gradient GEMM passes = 2.5 matvecs of algorithmic flop
                gp = dmrg_amd.DeviceArray(h.n_psi, np.complex128)
                h.time_adjoint(xd, yd, gp, warmup)
                ams = h.time_adjoint(xd, yd, gp, reps)
                res["fp32_adjoint"] = {"ms": round(ams, 3), "tflops": round(2.5 * h.flops / (ams * 1e-3) / 1e12, 2),
                                       "note": "cotangents of psi, renv, lenv; HIP events around the whole call, "
                                               "including the operand conjugation and FP32 conversions"}
                print(f"[bench] {name} fp32 adjoint: {ams:.3f} ms, {res['fp32_adjoint']['tflops']:.1f} TFLOP/s",
                      file=sys.stderr, flush=True)
                gp.free()
            res["flop"] = h.flops
            res["n_psi"] = h.n_psi
            print(f"[bench] {name} {prec}: {ms[0]:.3f} ms/matvec, {res[prec]['tflops']:.1f} TFLOP/s", file=sys.stderr,
                  flush=True)
            xd.free()
            yd.free()
            h.close()
        out[name] = res
    return out


def svd_leg(reps=5):
    """dmrgx_block_svd on the two sector blocks of the L=100 m=512 two-site
    state (512 x 512 complex each) with a DMRG-like spectrum exp(-30 k / 512)."""
    import ctypes as C
    import dmrg_amd

    dmrg_amd.init(0)
    rng = np.random.default_rng(7)
    n = M * 2 // S  # (m d / S) rows and columns per sector block (d = 2)
    blocks = []
    for _ in range(S):
        q1 = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
        q2 = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
        blocks.append((q1 * np.exp(-30.0 * np.arange(n) / n)) @ q2.conj().T)
    nb = len(blocks)
    rows = np.full(nb, n, np.int32)
    cols = np.full(nb, n, np.int32)
    off = np.arange(nb, dtype=np.int64) * n * n
    soff = np.arange(nb, dtype=np.int64) * n
    A = dmrg_amd.DeviceArray.from_numpy(np.concatenate([b.ravel() for b in blocks]))
    U = dmrg_amd.DeviceArray(nb * n * n, np.complex128)
    V = dmrg_amd.DeviceArray(nb * n * n, np.complex128)
    Sv = dmrg_amd.DeviceArray(nb * n, np.float64)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    lib = dmrg_amd.lib()

    def run():
        dmrg_amd.check(lib.dmrgx_block_svd(dmrg_amd.C64, nb, P(rows, C.c_int32), P(cols, C.c_int32), A.ptr,
                                           P(off, C.c_int64), U.ptr, P(off, C.c_int64), Sv.ptr,
                                           P(soff, C.c_int64), V.ptr, P(off, C.c_int64)), "block_svd")

    run()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    ms = (time.perf_counter() - t0) / reps * 1e3
    s = Sv.numpy()[:n]
    ref = np.exp(-30.0 * np.arange(n) / n)
    return {"svd_ms": round(ms, 3), "sweeps": int(lib.dmrgx_block_svd_sweeps()), "blocks": f"{nb} x {n}x{n} complex",
            "max_sv_err": float(np.max(np.abs(s - ref)))}


def cpu_info():
    """Host CPU model and logical CPU count (from /proc/cpuinfo)."""
    model, n = "unknown", os.cpu_count() or 1
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return model, n


def cpu_leg(seconds):
    """The oracle's av() (numpy, one BLAS GEMM per get_map task) on the same
    shapes, timed on this host's cores for about `seconds`, with all BLAS
    threads and with one."""
    from oracle.synth import synthetic_case
    from threadpoolctl import threadpool_info, threadpool_limits

    threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [os.cpu_count() or 1])
    model, ncpu = cpu_info()
    case = synthetic_case(S, D, M, dtype=np.complex128)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, case.n_psi) + 1j * rng.uniform(-1, 1, case.n_psi)
    flops = case.flops()

    def rate(limit):
        case.oracle_apply(x)  # warm
        n, t0 = 0, time.perf_counter()
        while True:
            case.oracle_apply(x)
            n += 1
            el = time.perf_counter() - t0
            if el > limit or n >= 200:
                return flops * n / el / 1e12, n, el

    all_r, all_n, all_el = rate(seconds)
    with threadpool_limits(limits=1):
        one_r, one_n, one_el = rate(seconds / 2)
    print(f"[bench] cpu av(): {all_r:.3f} TFLOP/s on {threads} threads, {one_r:.4f} on 1 ({model})", file=sys.stderr,
          flush=True)
    return {"value": all_r, "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
            "sample": f"{all_n} H_eff matvecs of the oracle's av() (numpy/OpenBLAS per-task zgemm) at the "
                      f"same m=512 Z2 D=5 shapes, {all_el:.1f} s",
            "one_core": {"value": one_r, "matvecs": one_n, "seconds": round(one_el, 2)},
            "cpu_model": model, "host_cpus": ncpu}


def cpu_sweep_leg(L=40, m=64, sweeps=5):
    """The Kitaev chain (t = Delta = 1, mu = 0.5; the reference's t=-1 p=-1
    u=0.25), open, L=40, Z2 parity, real FP64, two-site DMRG at m=64 with 5
    sweeps: the oracle on the host (numpy/scipy: OpenBLAS GEMMs, ARPACK,
    LAPACK gesvd) and the FP64 GPU path, both checked against the BdG closed
    form.  The reference's own timing sites are kitaev.cpp:19-20,56-57."""
    from dmrg_amd import driver
    from oracle import exact
    from oracle.dmrg import Runner
    from oracle.models import make_model
    from threadpoolctl import threadpool_info

    p = (-1.0, -1.0, 0.25)
    e_bdg = exact.kitaev_bdg_e0(L, *p)[0]
    t0 = time.perf_counter()
    g = driver.run_dmrg("kitaev", L, m, sweeps, p, schedule="dmrg2")
    gpu_s = time.perf_counter() - t0
    from threadpoolctl import threadpool_limits

    threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    model, ncpu = cpu_info()

    def host_run(limit):
        mpo, n, ferm = make_model("kitaev", L, t=p[0], p=p[1], u=p[2])
        R = Runner(mpo, n, ferm)
        t0 = time.perf_counter()
        if limit:
            with threadpool_limits(limits=limit):
                e = R.dmrg2(m, sweeps)
        else:
            e = R.dmrg2(m, sweeps)
        return e, time.perf_counter() - t0, len(R.log)

    e_cpu, cpu_s, nsteps = host_run(None)
    e_one, one_s, _ = host_run(1)
    print(f"[bench] kitaev L={L} m={m} {sweeps} sweeps: GPU {gpu_s:.2f} s, CPU oracle {cpu_s:.1f} s on {threads} "
          f"threads, {one_s:.1f} s on 1 ({model})", file=sys.stderr, flush=True)
    return {"workload": f"Kitaev L={L} m={m}, {sweeps} two-site sweeps (Z2, real FP64)", "steps": nsteps,
            "gpu_fp64_s": round(gpu_s, 3), "cpu_s": round(cpu_s, 3), "cpu_threads": int(threads),
            "cpu_one_thread_s": round(one_s, 3), "cpu_model": model, "host_cpus": ncpu, "kind": "port",
            "gpu_E0": g.energy, "cpu_E0": e_cpu, "cpu_one_thread_E0": e_one, "bdg_E0": e_bdg,
            "gpu_err": abs(g.energy - e_bdg), "cpu_err": abs(e_cpu - e_bdg)}


DERIV_J = (1.0, 0.9, 0.7)


def deriv_leg(L=128, m=1024, sweeps=3, delta=1e-4, reps=3, fd=True):
    """dE0/dJz of the open XYZ chain (Jx, Jy, Jz) = (1, 0.9, 0.7): the adjoint
    H_eff kernels with environment tangents recomputed from the MPS, at the
    centre of the FP64 DMRG state (FP64 and FP32 adjoints), against
    Hellmann-Feynman <psi|H_XYZ(0,0,1)|psi> (calc(mps, mpo), the reference's
    src/functions.cpp:469-511) and against the central difference
    (E0(Jz + delta) - E0(Jz - delta)) / (2 delta) of two more runs.  The runs
    are warm-started FP64 two-site DMRG.  The FP32 adjoint of the centre
    operator is timed with HIP events."""
    import dmrg_amd
    from dmrg_amd import driver

    def run(jz):
        t0 = time.perf_counter()
        r = driver.run_dmrg("xyz", L, m, sweeps, (DERIV_J[0], DERIV_J[1], jz), schedule="dmrg2", ladder=False,
                            keep_state=True, warm_start=True, verbose="stderr")
        print(f"[bench] deriv run Jz={jz}: E0={r.energy:.13f} in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
              flush=True)
        return r

    base = run(DERIV_J[2])
    st = base.state
    res = {"model": "xyz", "L": L, "m": m, "sweeps": sweeps, "J": list(DERIV_J), "energy": base.energy,
           "trunc_err": base.stats["trunc_err"], "delta": delta}
    hf = st.expect((0.0, 0.0, 1.0)).real / st.norm2().real
    res["hellmann_feynman"] = hf
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        der, e = st.derivative(DERIV_J, (0.0, 0.0, 1.0), prec=prec)
        res[f"adjoint_{prec}"] = {"dE_dJz": der, "centre_energy": e, "vs_hf": abs(der - hf),
                                  "vs_hf_rel": abs(der - hf) / abs(hf), "host_s": round(time.perf_counter() - t0, 3)}
        print(f"[bench] deriv {prec}: {der:.12f} (HF {hf:.12f}, |diff| {abs(der - hf):.2e})", file=sys.stderr,
              flush=True)
    h = st.centre_heff("fp32")
    rng = np.random.default_rng(5)
    x = dmrg_amd.DeviceArray.from_numpy(rng.uniform(-1, 1, h.n_psi) + 1j * rng.uniform(-1, 1, h.n_psi))
    g = dmrg_amd.DeviceArray.from_numpy(rng.uniform(-1, 1, h.n_psi) + 1j * rng.uniform(-1, 1, h.n_psi))
    gp = dmrg_amd.DeviceArray(h.n_psi, np.complex128)
    h.time_adjoint(x, g, gp, 1)
    ams = h.time_adjoint(x, g, gp, reps)
    res["fp32_adjoint_kernel"] = {"ms": round(ams, 3), "n_psi": h.n_psi, "flop": 2.5 * h.flops,
                                  "tflops": round(2.5 * h.flops / (ams * 1e-3) / 1e12, 2),
                                  "timing": "HIP events around dmrgx_heff_adjoint on the library stream"}
    for a in (x, g, gp):
        a.free()
    h.close()
    st.close()
    if fd:
        ep = run(DERIV_J[2] + delta)
        em = run(DERIV_J[2] - delta)
        ep.state.close()
        em.state.close()
        fdv = (ep.energy - em.energy) / (2 * delta)
        res["central_difference"] = fdv
        res["adjoint_fp64"]["vs_fd"] = abs(res["adjoint_fp64"]["dE_dJz"] - fdv)
        res["hf_vs_fd"] = abs(hf - fdv)
    return res


def sweep_leg(model, L, m, sweeps, ladder, keep=None, precs=("fp64", "bf16x3")):
    """Full two-site DMRG (the reference's dmrg2 schedule: growth at cutoff 10,
    the optional pre-sweep cutoff ladder, then full sweeps at m), FP64 against
    BF16x3 with the FP64 Jacobi-Davidson refinement; sweep_s is the wall time
    of the last full sweep (steady state at m), sweeps_s every full sweep;
    fidelity is |<psi|psi_fp64>| of the final states.  A run named
    "<prec>+warm" starts every sweep step from the previous step's state
    (White's prediction) and converges the lowest pair only; the plain runs
    start from random vectors, and the FP64 ones converge six pairs, as the
    reference's ARPACK calls do.  "+nevK" makes a reduced-precision solve test
    the residuals of its K lowest pairs (default 1), "+tolX" sets its tolerance
    (default TOL_LO), "+rkN" gives the refinement's inner solves 2N
    conjugate-gradient steps (default 8), "+rtolX" sets the FP64 tolerance
    (the refinement's, for a reduced-precision run; default 1e-10 for FP64,
    TOL_REFINE for the others)."""
    from dmrg_amd import driver
    params = {"xyz": JXYZ, "potts3": POTTS}[model]
    res = {"model": model, "L": L, "m": m, "sweeps": sweeps, "ladder": bool(ladder), "params": list(params),
           "timed": "last full sweep"}
    states = {}
    for spec in precs:
        prec, *mods = spec.split("+")
        warm = "warm" in mods
        nev_lo = next((int(x[3:]) for x in mods if x.startswith("nev")), 0)
        tol_lo = next((float(x[3:]) for x in mods if x.startswith("tol")), TOL_LO)
        refine_k = next((int(x[2:]) for x in mods if x.startswith("rk")), 8)
        tol = next((float(x[4:]) for x in mods if x.startswith("rtol")), 1e-10 if prec == "fp64" else TOL_REFINE)
        t0 = time.perf_counter()
        r = driver.run_dmrg(model, L, m, sweeps, params, schedule="dmrg2", prec=prec, tol=tol, tol_lo=tol_lo,
                            refine_k=refine_k, ladder=ladder, keep_state=True, verbose="stderr", warm_start=warm,
                            nev_lo=nev_lo)
        states[spec] = r.state
        st = r.stats
        per = sweep_seconds(r.steps, L, sweeps)
        steps = np.asarray(r.steps)
        full = steps[steps[:, 16] == steps[:, 16].max()]  # the steps at full two-site dimension
        res[spec] = dict(energy=r.energy, sweep_s=round(st["last_sweep_seconds"], 4),
                         sweeps_s=[round(x, 3) for x in per], total_s=round(time.perf_counter() - t0, 3),
                         matvecs=st["matvecs"], polish=st["polish_matvecs"], trunc_err=st["trunc_err"],
                         t_eig=round(st["t_eig"], 3), t_svd=round(st["t_svd"], 3), t_env=round(st["t_env"], 3),
                         full_dim_steps=int(len(full)), svd_jacobi_sweeps_full=round(float(full[:, 17].mean()), 2))
        print(f"[bench] sweep {spec} {model} L={L} m={m}: E0={r.energy:.12f} last sweep {st['last_sweep_seconds']:.2f} s "
              f"(eig {st['t_eig']:.2f} svd {st['t_svd']:.2f} env {st['t_env']:.2f} s in all steps)", file=sys.stderr,
              flush=True)
    if "fp64" in states:
        for spec in states:
            if spec == "fp64":
                continue
            res[spec]["dE0"] = abs(res[spec]["energy"] - res["fp64"]["energy"])
            # |<psi|psi_fp64>| of the two final MPSs (the reference's calc(mps, mps))
            res[spec]["fidelity"] = states[spec].fidelity(states["fp64"])
            res[spec]["speedup"] = round(res["fp64"]["sweep_s"] / max(res[spec]["sweep_s"], 1e-12), 3)
        if "bf16x3" in states:
            for k in ("dE0", "fidelity", "speedup"):
                res[k] = res["bf16x3"][k]
    for p, st in states.items():
        if p != keep:
            st.close()
    if keep:
        return res, states[keep]
    return res


def sweep_seconds(steps, L, sweeps):
    """Wall seconds of each full sweep of the dmrg2 schedule: the last
    sweeps * 2 (L - 4) steps before the final half sweep (L - 3 - L/2 + 1
    steps); column 12 of the step log is the step's wall time."""
    per, fin = 2 * (L - 4), (L - 3) - (L // 2) + 1
    end = len(steps) - max(fin, 0)
    out = []
    for k in range(sweeps):
        a = end - per * (sweeps - k)
        if a >= 0:
            out.append(float(np.sum(steps[a:a + per, 12])))
    return out


def batch_leg(n_chains=64, L=32, m=128, sweeps=1):
    """The batched workload: n_chains independent open XYZ chains with
    couplings i.i.d. U[0.5, 1.5) from splitmix64 (seed 20261203), swept in
    lockstep (one H_eff plan and one block SVD over all chains per step,
    segmented batched Lanczos), FP64 and BF16x3 with the FP64 polish; two
    chains also run alone (FP64) for the sequential comparison."""
    from dmrg_amd import driver
    J = driver.batch_couplings(n_chains, 20261203)
    res = {"chains": n_chains, "L": L, "m": m, "sweeps": sweeps, "coupling_seed": 20261203}
    runs = {}
    for prec in ("fp64", "bf16x3"):
        t0 = time.perf_counter()
        b = driver.run_batch("xyz", L, m, sweeps, J, prec=prec, tol_lo=TOL_LO, refine_k=8)
        el = time.perf_counter() - t0
        st = b.stats
        runs[prec] = b
        res[prec] = dict(total_s=round(el, 3), sweep_s=round(st["sweep_seconds"], 4), matvec_calls=st["matvecs"],
                         polish_calls=st["polish_matvecs"], matvec_tflop=round(st["matvec_flops"] / 1e12, 3),
                         t_eig=round(st["t_eig"], 3), t_svd=round(st["t_svd"], 3), t_env=round(st["t_env"], 3),
                         chain_sweeps_per_s=round(n_chains * sweeps / max(st["sweep_seconds"] * sweeps, 1e-12), 3))
        print(f"[bench] batch {prec}: {n_chains} chains L={L} m={m} in {el:.1f} s, sweep {st['sweep_seconds']:.2f} s",
              file=sys.stderr, flush=True)
    res["max_dE0_bf16x3_vs_fp64"] = float(np.max(np.abs(runs["bf16x3"].energies - runs["fp64"].energies)))
    res["max_trunc_err"] = float(np.max(runs["fp64"].trunc))
    single = []
    for k in range(2):
        t0 = time.perf_counter()
        one = driver.run_dmrg("xyz", L, m, sweeps, J[k])
        single.append(time.perf_counter() - t0)
        res.setdefault("single_dE0", []).append(abs(one.energy - runs["fp64"].energies[k]))
    res["single_fp64_s"] = round(float(np.mean(single)), 3)
    res["sequential_fp64_s_extrapolated"] = round(float(np.mean(single)) * n_chains, 2)
    res["speedup_vs_sequential"] = {p: round(res["sequential_fp64_s_extrapolated"] / res[p]["total_s"], 2)
                                    for p in ("fp64", "bf16x3")}
    return res


def copy_leg(reps):
    """A plain copy of 256 MiB through dmrgx_copy_boxes (one box): the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    import dmrg_amd

    dmrg_amd.init(0)
    n = 32 * 1024 * 1024  # float64 elements
    src = dmrg_amd.DeviceArray(n, np.float64)
    dst = dmrg_amd.DeviceArray(n, np.float64)
    src.zero()
    task = np.array([0, 0, 0, 0, 0, 0, 1, 1, n, 0], dtype=np.int64)
    lib = dmrg_amd.lib()
    t0 = time.perf_counter()
    for _ in range(reps):
        dmrg_amd.check(lib.dmrgx_copy_boxes(dmrg_amd.R64, src.ptr, dst.ptr,
                                            task.ctypes.data_as(C.POINTER(C.c_int64)), 1), "copy_boxes")
    lib.dmrgx_synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    return {"copy_bytes": 8 * n, "ms": round(ms, 4), "GB/s (read+write)": round(16 * n / ms / 1e6, 1)}


MFMA_KINDS = {0: ("v_mfma_f32_32x32x16_bf16", "bf16x3"), 1: ("v_mfma_f32_16x16x32_bf16", "bf16x3"),
              2: ("v_mfma_f64_16x16x4f64", "fp64"), 3: ("v_mfma_f32_16x16x4f32", "fp32")}


def mfma_leg(rounds=20000):
    """The matrix-core ceilings: each kind of MFMA instruction the kernels use,
    issued back to back on four independent register accumulators per
    wavefront (no memory traffic), at 1 and 2 wavefronts per SIMD of every CU
    (dmrgx_calibrate_mfma), against the published dense peak of its type."""
    import ctypes as C
    import dmrg_amd

    dmrg_amd.init(0)
    lib = dmrg_amd.lib()
    out = {}
    for kind, (name, prec) in MFMA_KINDS.items():
        row = {"peak": PEAK[prec]}
        for w in (1, 2):
            v = C.c_double(0)
            n = rounds if kind != 2 else rounds // 4
            dmrg_amd.check(lib.dmrgx_calibrate_mfma(kind, w, n, C.byref(v)), "calibrate_mfma")
            row[f"tflops_{w}wave"] = round(v.value, 1)
        row["frac_of_peak"] = round(max(row["tflops_1wave"], row["tflops_2wave"]) / PEAK[prec], 3)
        out[name] = row
        print(f"[bench] {name}: {row['tflops_1wave']:.1f} TFLOP/s (1 wave/SIMD), {row['tflops_2wave']:.1f} (2), "
              f"published peak {PEAK[prec]}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-configs", action="store_true", help="skip the Potts m=1024 / XYZ m=2048 matvec shapes")
    ap.add_argument("--no-cpu", action="store_true", help="skip the host matvec baseline")
    ap.add_argument("--sweep", type=int, default=2, help="full sweeps in the sweep comparison (0 skips it)")
    ap.add_argument("--sweep-model", default="xyz", choices=["xyz", "potts3"])
    ap.add_argument("--sweep-L", type=int, default=L_CHAIN, help="chain length of the sweep comparison")
    ap.add_argument("--sweep-m", type=int, default=M, help="bond dimension of the sweep comparison")
    ap.add_argument("--ladder", action="store_true", help="run the reference's pre-sweep cutoff ladder 100, 200, ..")
    ap.add_argument("--sweep-precs", default="fp64,bf16x3",
                    help="--part sweep: runs to compare (fp64, fp32, bf16x3, each optionally +warm)")
    ap.add_argument("--batch", type=int, default=0, help="sweeps of the 64-chain batched workload (0 skips it)")
    ap.add_argument("--cpu-sweep", action="store_true",
                    help="also run the host Kitaev L=40 m=64 5-sweep baseline beside the GPU (about 1-2 minutes)")
    ap.add_argument("--only", default="", help="profiling aid: comma-separated precisions to run (no JSON line)")
    ap.add_argument("--matvec-from", default="sweep", choices=["sweep", "quick", "synthetic"],
                    help="operator of the matvec timing: the centre of the sweep comparison's FP64 state "
                         "(default; synthetic when --sweep 0), of a quick BF16x3 run (growth + one sweep), or dense "
                         "synthetic environments")
    ap.add_argument("--deriv-L", type=int, default=128, help="--part deriv: chain length")
    ap.add_argument("--deriv-m", type=int, default=1024, help="--part deriv: bond dimension")
    ap.add_argument("--deriv-sweeps", type=int, default=3, help="--part deriv: full sweeps per run")
    ap.add_argument("--no-fd", action="store_true", help="--part deriv: skip the two finite-difference runs")
    ap.add_argument("--part", default="", choices=["", "svd", "copy", "sweep", "cpusweep", "deriv", "mfma"],
                    help="run one part alone: the step SVD, the copy kernel used to calibrate PMC counters, the "
                         "sweep comparison, the host Kitaev sweep, the derivative at scale or the matrix-core "
                         "ceilings (JSON of that part only)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the multi-GPU layout is design-only.",
              file=sys.stderr)
        sys.exit(2)
    if a.part == "svd":
        print(json.dumps(svd_leg(a.steps)), flush=True)
        return
    if a.part == "copy":
        print(json.dumps(copy_leg(a.steps)), flush=True)
        return
    if a.part == "mfma":
        print(json.dumps(mfma_leg()), flush=True)
        return
    if a.part == "sweep":
        print(json.dumps(sweep_leg(a.sweep_model, a.sweep_L, a.sweep_m, a.sweep, a.ladder,
                                   precs=tuple(a.sweep_precs.split(",")))), flush=True)
        return
    if a.part == "deriv":
        print(json.dumps(deriv_leg(a.deriv_L, a.deriv_m, a.deriv_sweeps, fd=not a.no_fd)), flush=True)
        return
    if a.part == "cpusweep":
        print(json.dumps(cpu_sweep_leg()), flush=True)
        return
    if a.only:
        if a.matvec_from == "quick":
            st = quick_state(a.sweep_model, a.sweep_L, a.sweep_m)
            state_leg(st, a.steps, a.warmup, a.only.split(","))
            st.close()
        else:
            gpu_leg(a.steps, a.warmup, a.only.split(","))
        return
    # the sweep comparison first: the matvec is then timed on the real operator
    # at the centre of its converged FP64 state
    sweep, state = None, None
    if a.sweep:
        sweep, state = sweep_leg(a.sweep_model, a.sweep_L, a.sweep_m, a.sweep, a.ladder, keep="fp64")
        if a.matvec_from != "sweep":
            state.close()
            state = None
    if a.matvec_from == "quick":
        state = quick_state(a.sweep_model, a.sweep_L, a.sweep_m)
    if state is not None:
        g = state_leg(state, a.steps, a.warmup)
        state.close()
        name = {"xyz": "XYZ", "potts3": "chiral Z3 Potts"}[a.sweep_model]
        workload = (f"two-site H_eff matvec at the centre bond of the converged open {name} chain "
                    f"L={a.sweep_L} m={a.sweep_m} ({'FP64 sweep state' if a.sweep else 'BF16x3 state after one sweep'}"
                    f"; its real environments, whose structurally zero MPO-bond slices the plan skips)")
        data = "real operator of a DMRG run on the model (no external data)"
    else:
        g = gpu_leg(a.steps, a.warmup)
        workload = ("two-site H_eff matvec with the block structure of the open XYZ chain L=100 m=512 (Z2 sectors "
                    "256+256, MPO bond 5, complex; 32+32 get_map tasks, 4.295e10 flop) on dense synthetic "
                    "environments")
        data = "synthetic"
    mixed, fp64 = g["bf16x3"], g["fp64"]
    # dominant kernel: the BF16x3 grouped GEMM (R and L pass launches)
    per_launch_ms = 0.5 * (mixed["ms_r"] + mixed["ms_l"])
    per_launch_flop = 0.5 * (mixed["flops_r"] + mixed["flops_l"])
    per_launch_exec = 0.5 * (mixed["exec_r"] + mixed["exec_l"])
    achieved = per_launch_flop / (per_launch_ms * 1e-3) / 1e12
    achieved_exec = per_launch_exec / (per_launch_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(mixed["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(mixed["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128 vectors; bf16x3 split products, fp32 accumulate",
        "data": data,
        "config": {"workload": workload, "model": a.sweep_model, "L": a.sweep_L, "m": a.sweep_m,
                   "params": list(JXYZ if a.sweep_model == "xyz" else POTTS), "prec": "bf16x3",
                   "flop_algorithmic": mixed["flops"], "flop_executed": mixed["flops_exec"]},
        "exec_tflops": round(mixed["exec_tflops"], 3),
        "fp64_tflops": round(fp64["tflops"], 3),
        "fp64_exec_tflops": round(fp64["exec_tflops"], 3),
        "fp32_tflops": round(g["fp32"]["tflops"], 3),
        "fp64_ms_per_matvec": round(fp64["ms"], 4),
        "mixed_vs_fp64_matvec": round(fp64["ms"] / mixed["ms"], 3),
        "fp64_roofline_frac_executed": round(fp64["exec_tflops"] / PEAK["fp64"], 4),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK["bf16x3"], "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK["bf16x3"], 5),
                     "achieved_executed": round(achieved_exec, 3),
                     "frac_executed": round(achieved_exec / PEAK["bf16x3"], 5),
                     "traffic": traffic_bytes(workload_key(state is not None, a), mixed["r_items"], mixed["l_items"]),
                     "work_items": {"R": mixed["r_items"], "L": mixed["l_items"]},
                     "algorithmic_bytes": ALG_BYTES,
                     "kernel": "k_gemm_bf3p<complex> (R and L pass launches)",
                     "issued_tflops_executed": round(3 * achieved_exec, 3)},
        "cpu_baseline": None,
        "configs": None,
        "sweep": sweep,
        "batch": None,
        "cpu_sweep": None,
    }
    if not a.no_configs:
        line["configs"] = configs_leg()
    if a.batch:
        line["batch"] = batch_leg(sweeps=a.batch)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_leg(a.cpu_seconds)
    if a.cpu_sweep:
        line["cpu_sweep"] = cpu_sweep_leg()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
