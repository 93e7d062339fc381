"""Benchmark: advection apply GDoF/s, 3D+3D, k=3, FP64 and FP32-mixed.

Workload: periodic [0,1]^6, 8^3 x 8^3 = 262,144 cells, k=3 (4096 DoF per
cell, 1,073,741,824 DoF, 8.6 GB per FP64 vector), velocity a_i ~ U(-1,1)
(counter-based SplitMix64, seed 20261202), f0 = prod sin(2 pi x_i) +
1e-3 U(-1,1) noise (seed 20261203).  Inputs are generated on the GPU and are
resident in HBM before timing.

A step is one operator apply (src -> dst) over the whole vector, the
reference's operator benchmark unit (performance/operators_advection_01.
likwid.cc:341-348: GDoF/s = n_dofs * n_applies / time).  `value` is the FP64
apply rate; the FP32-mixed apply and the fused LSRK stage rates are reported
next to it.  Kernel times come from HIP events recorded on the stream the
kernels are launched on.

cpu_baseline: the CPU oracle (a restatement of hyper.deal's ECL quadrature
algorithm, OpenMP over macro cells) timed on this host on a bounded sample of
the same workload: 3+3, k=3, 4^3 x 8^3 cells.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X dense vector peaks (TFLOP/s)
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# the collapsed apply's arithmetic per DoF at 3+3, k=3: D (N + 1) FMA
FLOP_PER_DOF = 2 * 6 * (3 + 2)
SEED_VELOCITY = 20261202
SEED_NOISE = 20261203
# HBM bytes per FP64 apply launch, measured with rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE and scaled by the same counters on a plain copy (the command and
# figures are in this file)
TRAFFIC_FILE = os.path.join(ROOT, "profiles",
                            "collapsed_apply_3d3d_k3_8p6_traffic.txt")


def measured_traffic():
    """bytes per launch of the FP64 apply from the committed counter summary,
    or None"""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def splitmix_uniform(seed, i):
    m = (1 << 64) - 1
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z = z ^ (z >> 31)
    return (z >> 11) * (2.0 / 2 ** 53) - 1.0


def timed(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1000.0 / steps  # seconds per call


def cpu_baseline(velocity):
    """Time the oracle's ECL apply on 3+3, k=3, 4^3 x 8^3 cells."""
    import numpy as np

    import oracle as O
    lib_path = os.path.join("/tmp", "hd_oracle_native_%d" % os.getpid(),
                            "libhd_oracle.so")
    try:
        O.build(native_out=lib_path)
        lib = O.load(lib_path)
        arch = "native"
    except Exception:  # fall back to the portable build in the tree
        lib = O.load()
        arch = "x86-64-v3"
    n_cells = [4, 4, 4, 8, 8, 8]
    P = O.Problem(3, 3, 3, n_cells, [0.0] * 6, [1.0] * 6, velocity, lib=lib,
                  func=O.FUNC_PRODUCT_SINES, func_params=[0.0] * 6)
    u = P.interpolate()
    P.add_noise(u, SEED_NOISE, 1e-3)
    dst = P.vector()
    P.apply(u, dst=dst)  # warm-up
    n_apply = 0
    t0 = time.perf_counter()
    while True:
        P.apply(u, dst=dst)
        n_apply += 1
        el = time.perf_counter() - t0
        if el > 10.0 or n_apply >= 20:
            break
    gdofs = P.n_dofs * n_apply / el / 1e9
    cores = lib.hdo_num_threads()
    return {"value": round(gdofs, 4), "unit": "GDoF/s", "cores": cores,
            "kind": "port",
            "sample": "oracle ECL apply (restated hyper.deal quadrature "
                      "algorithm), 3+3 k=3, 4^3 x 8^3 cells = %d DoF, %d "
                      "applies after 1 warm-up, %s build, %d OpenMP threads; "
                      "the reference's algorithm (about 420 flop/DoF), not "
                      "the GPU's collapsed form (60 flop/DoF), so the "
                      "GPU/CPU ratio compares algorithms as well as hardware"
                      % (P.n_dofs, n_apply, arch, cores)}


def batch_workload(H, torch, stream, steps):
    """64 independent 2+2 advection instances (instance i: seed 20261204 + i
    draws cells per direction from {4, 6, 8, 12, 16}, degree from
    {2, 3, 4, 5} and the velocity from U(-1, 1)^4, SplitMix64), FP32-mixed
    rk45 steps: one launch per stage for all instances, captured in a
    hipGraph, after the autotuner picked the work split per degree; against
    64 sequential single-instance steps."""
    insts, n_total = [], 0
    for i in range(64):
        seed = 20261204 + i
        nc = [4, 6, 8, 12, 16][int((splitmix_uniform(seed, 0) + 1) / 2 * 5) % 5]
        k = [2, 3, 4, 5][int((splitmix_uniform(seed, 1) + 1) / 2 * 4) % 4]
        a = [splitmix_uniform(seed, 2 + d) for d in range(4)]
        mf = H.MatrixFree(2, 2, k, [nc] * 4, [0.0] * 4, [1.0] * 4)
        op = H.AdvectionOperation(mf, a)
        sol = torch.empty(mf.n_dofs(), dtype=torch.float64, device="cuda")
        H.VectorTools.interpolate(mf, sol, H.FUNC_PRODUCT_SINES, [0.0] * 4)
        ta = torch.empty(mf.n_dofs(), dtype=torch.float32, device="cuda")
        tb = torch.empty_like(ta)
        insts.append((mf, op, sol, ta, tb))
        n_total += mf.n_dofs()
    batch = H.Batch([x[1] for x in insts])
    sols = [x[2] for x in insts]
    tas = [x[3] for x in insts]
    tbs = [x[4] for x in insts]
    for t, s_ in zip(tas, sols):
        t.copy_(s_)
    batch.autotune(tas, tbs, precision=H.FP32)
    dt = 1e-4
    graph = batch.graph("rk45", sols, tas, tbs, dt, precision=H.FP32)
    t_graph = timed(lambda: graph.launch(stream), steps, 1, stream)
    t_eager = timed(lambda: batch.lsrk_step("rk45", sols, tas, tbs, dt,
                                            precision=H.FP32, stream=stream),
                    steps, 1, stream)
    integ = H.LowStorageRungeKuttaIntegrator("rk45")

    def sequential():
        for mf, op, sol, ta, tb in insts:
            integ.perform_time_step(op, sol, ta, tb, 0.0, dt, precision=H.FP32,
                                    stream=stream)
    t_seq = timed(sequential, steps, 1, stream)
    stages = integ.n_stages()
    tuned = {k: batch.cells_per_task(k, precision=H.FP32) for k in (2, 3, 4, 5)}
    del graph
    return {"instances": 64, "dof": n_total, "precision": "fp32-mixed",
            "rk": "rk45", "unit": "GDoF/s (stage applies)",
            "value": round(stages * n_total / t_graph / 1e9, 2),
            "ms_per_step": round(t_graph * 1e3, 3),
            "eager_batch_GDoFs": round(stages * n_total / t_eager / 1e9, 2),
            "sequential_GDoFs": round(stages * n_total / t_seq / 1e9, 2),
            "sequential_ms_per_step": round(t_seq * 1e3, 3),
            "speedup_vs_sequential": round(t_seq / t_graph, 3),
            "cells_per_task": tuned}


def landau_workload(H, torch, stream, steps):
    """2+2 Vlasov-Poisson Landau damping, x in [0, 4 pi]^2, v in [-6, 6]^2,
    32^4 cells, k=3 (268M DoF), rk45 at CFL 0.3: the time per step, the CG
    iterations per Poisson solve (relative residual 1e-10) and the share of
    the step spent on the field update (one minus frozen-field step time over
    step time), for the Poisson start from the previous phi (FP64 and
    FP32-mixed), the learned start (FP32-mixed; trained online from the
    run's first solves) and the exact start (FP32-mixed).  The damping rate
    needs the run to t = 10: hyperdeal_amd/examples/landau.py."""
    import numpy as np
    L, n, k = 4 * np.pi, 32, 3
    dt = 0.3 * (L / n / 6.0) / k ** 1.5
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    out = {"workload": "Landau 2D+2D, 32^4 cells, k=3 (268435456 DoF), "
                       "rk45, dt %.5f, Poisson CG to 1e-10" % dt,
           "runs": {}}
    for name in ("fp64", "fp32", "fp32-learned", "fp32-exact"):
        mf = H.MatrixFree(2, 2, k, [n] * 4, [0, 0, -6, -6], [L, L, 6, 6])
        op = H.VlasovPoissonOperation(mf, poisson_tol=1e-10)
        if name.endswith("-exact"):
            op.exact_start()
        if name.endswith("-learned"):
            op.enable_learned_guess(seed=20261205, capacity=64,
                                    train_every=1, train_batch=4,
                                    train_steps=1, learning_rate=1e-3,
                                    min_pairs=8)
        prec = H.FP64 if name == "fp64" else H.FP32
        dtype = torch.float64 if prec == H.FP64 else torch.float32
        f = mf.initialize_dof_vector()
        H.VectorTools.interpolate(mf, f, H.FUNC_LANDAU, [0.01, 0.5, 2.0])
        ta = torch.empty(f.numel(), dtype=dtype, device="cuda")
        tb = torch.empty_like(ta)

        def step():
            integ.perform_time_step(op, f, ta, tb, 0.0, dt, precision=prec,
                                    stream=stream)
        for _ in range(3):
            step()
        torch.cuda.synchronize()
        it0, s0, _ = op.solver_stats()
        t = timed(step, steps, 0, stream)
        it1, s1, _ = op.solver_stats()
        op.set_auto_update(False)
        t_frozen = timed(step, steps, 0, stream)
        out["runs"][name] = {
            "ms_per_step": round(t * 1e3, 3),
            "ms_per_step_frozen_field": round(t_frozen * 1e3, 3),
            "field_update_share": round(1.0 - t_frozen / t, 4),
            "cg_iterations_per_solve": round((it1 - it0) / max(s1 - s0, 1),
                                             2)}
        del op, f, ta, tb
        torch.cuda.empty_cache()
    # the learned start against the reference's previous-phi start
    # (vlasov_poisson/include/application.h:575-576), both FP32-mixed
    pl, ln = out["runs"]["fp32"], out["runs"]["fp32-learned"]
    out["learned_vs_previous_phi"] = {
        "cg_iterations_ratio": round(pl["cg_iterations_per_solve"] and
                                     ln["cg_iterations_per_solve"] /
                                     pl["cg_iterations_per_solve"], 4),
        "step_time_ratio": round(ln["ms_per_step"] / pl["ms_per_step"], 4),
        "note": "below 1: the learned start wins; the network's training "
                "and inference cost more per solve than the CG iterations "
                "they save when the step time ratio is above 1"}
    return out


def landau_damping(H, torch, t_end=10.0):
    """The Landau damping rate of the 32^4, k=3 workload to t = 10 (rk45 at
    CFL 0.3, Poisson CG to 1e-10 from the exact start), FP64 and
    FP32-mixed: the rate fitted to the maxima of the field energy (written
    every step; examples/vlasov_poisson/include/application.h:630-660),
    against the FP64 run and the linear dispersion relation, the relative
    mass drift and the FP32-mixed run's relative L2 difference to the FP64
    run at t_end."""
    import numpy as np

    from hyperdeal_amd.diagnostics import fit_damping_rate, landau_theory_rate
    L, n, k = 4 * np.pi, 32, 3
    dt = 0.3 * (L / n / 6.0) / k ** 1.5
    n_steps = int(np.ceil(t_end / dt))
    dt = t_end / n_steps
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    omega, gamma_th = landau_theory_rate(np.sqrt(0.5))
    out = {"workload": "Landau 2D+2D, 32^4 cells, k=3, rk45, %d steps of "
                       "dt %.6f to t = %g, exact Poisson start, CG to 1e-10"
                       % (n_steps, dt, t_end),
           "theory_rate": round(gamma_th, 5), "runs": {}}
    ref = None
    for name, prec, dtype in (("fp64", H.FP64, torch.float64),
                              ("fp32", H.FP32, torch.float32)):
        mf = H.MatrixFree(2, 2, k, [n] * 4, [0, 0, -6, -6], [L, L, 6, 6])
        op = H.VlasovPoissonOperation(mf, poisson_tol=1e-10)
        op.exact_start()
        f = mf.initialize_dof_vector()
        H.VectorTools.interpolate(mf, f, H.FUNC_LANDAU, [0.01, 0.5, 2.0])
        m0 = H.vp_moments(mf, f)[0]
        ta = torch.empty(f.numel(), dtype=dtype, device="cuda")
        tb = torch.empty_like(ta)
        ts, en = [], []
        t0 = time.perf_counter()
        last = t0
        for s in range(n_steps):
            integ.perform_time_step(op, f, ta, tb, s * dt, dt, precision=prec)
            ts.append((s + 1) * dt)
            en.append(sum(op.field_energy()))
            if time.perf_counter() - last > 20:
                last = time.perf_counter()
                print("landau damping %s: step %d / %d" % (name, s + 1,
                                                           n_steps),
                      file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gamma, peaks = fit_damping_rate(ts, en)
        run = {"damping_rate": round(gamma, 6), "fit_maxima": peaks,
               "mass_drift_rel": (H.vp_moments(mf, f)[0] - m0) / m0,
               "ms_per_step_incl_diagnostics": round(1e3 * wall / n_steps, 3)}
        if ref is None:
            ref = f
        else:
            run["rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(f - ref) /
                                          torch.linalg.vector_norm(ref))
            run["rate_minus_fp64_rate"] = gamma - out["runs"]["fp64"][
                "damping_rate"]
            del f
        run["rate_minus_theory"] = gamma - gamma_th
        out["runs"][name] = run
        del op, ta, tb
        torch.cuda.empty_cache()
    del ref
    torch.cuda.empty_cache()
    return out


def gl_weights(N):
    """Gauss-Lobatto quadrature weights of N points on [0, 1]."""
    import numpy as np
    from numpy.polynomial import legendre as L
    c = np.zeros(N)
    c[-1] = 1.0
    x = np.concatenate([[-1.0], np.sort(L.legroots(L.legder(c))), [1.0]])
    return 1.0 / ((N - 1) * N * L.legval(x, c) ** 2)


def lsrk_precision_run(H, torch, stream, mf, op, u, vel, k, h):
    """20 rk45 steps in FP64 and in FP32-mixed from the same f0 = u on a
    uniform 3+3 mesh of cell size h, dt = 0.1 * dt_crit / k^1.5: the
    relative L2 difference FP32-mixed vs FP64 after 20 steps and the mass
    drift of each run (both with the Gauss-Lobatto rule on the nodes), and
    the time per step."""
    import numpy as np
    N = k + 1
    n = mf.n_dofs()
    cell = N ** 6
    w1 = gl_weights(N) * h
    W = w1
    for _ in range(5):
        W = np.multiply.outer(W, w1)
    W = torch.from_numpy(W.reshape(-1)).cuda()   # node index x0 fastest

    def mass(f):
        return float((f.view(-1, cell) @ W).sum())

    dt = 0.1 * H.compute_critical_time_step(mf, vel) / k ** 1.5
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    runs = {}
    for name, prec, dtype in (("fp64", H.FP64, torch.float64),
                              ("fp32_mixed", H.FP32, torch.float32)):
        sol = u.clone()
        ta = torch.empty(n, dtype=dtype, device="cuda")
        tb = torch.empty_like(ta)
        m0 = mass(sol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(20):
            integ.perform_time_step(op, sol, ta, tb, s * dt, dt,
                                    precision=prec, stream=stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        del ta, tb
        torch.cuda.empty_cache()
        runs[name] = (sol, m0, mass(sol), wall)
    s64, s32 = runs["fp64"][0], runs["fp32_mixed"][0]
    num = den = 0.0
    chunk = 4096 * cell
    for o in range(0, n, chunk):
        a = s64[o:o + chunk].view(-1, cell)
        b = s32[o:o + chunk].view(-1, cell)
        num += float((((b - a) ** 2) @ W).sum())
        den += float(((a ** 2) @ W).sum())
    l1 = float((u.abs().view(-1, cell) @ W).sum())
    out = {
        "dt": dt,
        "rel_l2_fp32_mixed_vs_fp64": (num / den) ** 0.5,
        "mass_drift_fp64": runs["fp64"][2] - runs["fp64"][1],
        "mass_drift_fp32_mixed":
            runs["fp32_mixed"][2] - runs["fp32_mixed"][1],
        "mass_scale": l1,
        "note": "mass drifts are absolute; mass_scale = int |f0|",
        "s_per_step_fp64": round(runs["fp64"][3] / 20, 4),
        "s_per_step_fp32_mixed": round(runs["fp32_mixed"][3] / 20, 4)}
    del runs, s64, s32
    torch.cuda.empty_cache()
    return out


def c3_workload(H, torch, stream, steps):
    """3+3, k=5, 6^3 x 6^3 cells (46,656 DoF per cell, 2,176,782,336 DoF):
    the split collapsed apply in FP64 and FP32-mixed, then 20 rk45 steps in
    each precision from the same f0 (product of sines + 1e-3 noise, the C2
    seeds), dt = 0.1 * dt_crit / k^1.5.  Reported: apply GDoF/s and their
    ratio, the relative L2 difference FP32-mixed vs FP64 after 20 steps and
    the mass drift of each run (both with the Gauss-Lobatto rule on the
    nodes)."""
    k, nc = 5, [6] * 6
    vel = [splitmix_uniform(SEED_VELOCITY, i) for i in range(6)]
    mf = H.MatrixFree(3, 3, k, nc, [0.0] * 6, [1.0] * 6)
    op = H.AdvectionOperation(mf, vel)
    n = mf.n_dofs()
    u = torch.empty(n, dtype=torch.float64, device="cuda")
    H.VectorTools.interpolate(mf, u, H.FUNC_PRODUCT_SINES, [0.0] * 6)
    H.VectorTools.add_noise(mf, u, SEED_NOISE, 1e-3)
    dst = torch.empty_like(u)
    t64 = timed(lambda: op.apply(dst, u, 0.0, H.FP64, stream), steps, 2,
                stream)
    del dst
    u32 = u.float()
    d32 = torch.empty_like(u32)
    t32 = timed(lambda: op.apply(d32, u32, 0.0, H.FP32, stream), steps, 2,
                stream)
    del u32, d32
    torch.cuda.empty_cache()
    lsrk = lsrk_precision_run(H, torch, stream, mf, op, u, vel, k, 1.0 / 6)
    out = {
        "workload": "3D+3D apply, periodic [0,1]^6, 6^3 x 6^3 cells, k=5 "
                    "(46656 DoF/cell, %d DoF), split collapsed form" % n,
        "fp64": {"value": round(n / t64 / 1e9, 3), "unit": "GDoF/s",
                 "ms_per_apply": round(t64 * 1e3, 3),
                 "achieved_GBs": round(16.0 * n / t64 / 1e9, 1)},
        "fp32_mixed": {"value": round(n / t32 / 1e9, 3), "unit": "GDoF/s",
                       "ms_per_apply": round(t32 * 1e3, 3),
                       "achieved_GBs": round(8.0 * n / t32 / 1e9, 1)},
        "ratio_fp32_vs_fp64": round(t64 / t32, 3),
        "lsrk_20_steps": lsrk,
    }
    del u
    torch.cuda.empty_cache()
    return out


def quadrature_forms(H, torch, stream, steps):
    """The kernels of the reference's algorithm itself, on a 2+2, k=3 mesh
    of 24^4 cells (84,934,656 DoF) on [-1, 1]^4 with inhomogeneous Dirichlet
    boundaries (the hyperrectangle case, velocity (1, 0.15, -0.05, 0)): the
    element-centric quadrature kernel (FP64 and FP32-mixed), the
    face-centric cell / face / inverse-mass kernels (FP64) and the
    quadrature kernel on the deformed mesh (DeformedCubeManifold 0.1, 2;
    FP64); and the automatic choice for the Cartesian operator, the
    collapsed kernel with its Dirichlet boundary correction (FP64, FP32).
    Rates in GDoF/s and the algorithmic 16 (FP64) / 8 (FP32) bytes per DoF
    per apply; about 420 flop per DoF for the quadrature steps."""
    k, n = 3, [24] * 4
    vel = [1.0, 0.15, -0.05, 0.0]
    out = {"workload": "2D+2D apply, [-1,1]^4 Dirichlet, 24^4 cells, k=3 "
                       "(84934656 DoF), Gauss k+1", "runs": {}}
    for name, deform, algo, prec in (
            ("ecl_fp64", False, H.ALGO_QUADRATURE, H.FP64),
            ("ecl_fp32", False, H.ALGO_QUADRATURE, H.FP32),
            ("fcl_fp64", False, H.ALGO_FCL, H.FP64),
            ("ecl_deformed_fp64", True, H.ALGO_QUADRATURE, H.FP64),
            ("fcl_deformed_fp64", True, H.ALGO_FCL, H.FP64),
            ("collapsed_fp64", False, H.ALGO_COLLAPSED, H.FP64),
            ("collapsed_fp32", False, H.ALGO_COLLAPSED, H.FP32)):
        mf = H.MatrixFree(2, 2, k, n, [-1.0] * 4, [1.0] * 4,
                          periodic_x=False, periodic_v=False)
        if deform:
            mf.set_deformation(0.1, 2)
        op = H.AdvectionOperation(mf, vel, boundary=H.BC_INHOMOGENEOUS,
                                  boundary_function=H.FUNC_HYPERRECTANGLE,
                                  boundary_params=vel)
        op.set_algorithm(algo)
        nd = mf.n_dofs()
        dtype = torch.float64 if prec == H.FP64 else torch.float32
        u = mf.initialize_dof_vector()
        H.VectorTools.interpolate(mf, u, H.FUNC_HYPERRECTANGLE, vel)
        u = u.to(dtype)
        d = torch.empty_like(u)
        t = timed(lambda: op.apply(d, u, 0.1, prec, stream), steps, 2, stream)
        bpd = 16.0 if prec == H.FP64 else 8.0
        out["runs"][name] = {"GDoFs": round(nd / t / 1e9, 2),
                             "ms_per_apply": round(t * 1e3, 3),
                             "achieved_GBs": round(bpd * nd / t / 1e9, 1)}
        del op, mf, u, d
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-landau-rate", action="store_true",
                    help="skip the Landau damping-rate runs to t = 10 "
                    "(about two minutes)")
    ap.add_argument("--only", choices=["all", "fp64", "fp32", "stage",
                                       "batch", "c3", "landau", "quadrature"],
                    default="all", help="restrict the timed kernels "
                    "(profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run the plain device copy of one FP64 vector (8.6 GB "
                    "read + 8.6 GB written) before the timed kernels too, the "
                    "reference point for scaling the FETCH_SIZE / WRITE_SIZE "
                    "counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the "
              "operator would shard into x/v phase-space blocks exchanging "
              "inflow face halos). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import hyperdeal_amd as H

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    n_cells = [8] * 6
    velocity = [splitmix_uniform(SEED_VELOCITY, i) for i in range(6)]
    mf = H.MatrixFree(3, 3, 3, n_cells, [0.0] * 6, [1.0] * 6)
    op = H.AdvectionOperation(mf, velocity)
    n = mf.n_dofs()
    results = {}

    u = torch.empty(n, dtype=torch.float64, device="cuda")
    H.VectorTools.interpolate(mf, u, H.FUNC_PRODUCT_SINES, [0.0] * 6)
    H.VectorTools.add_noise(mf, u, SEED_NOISE, 1e-3)

    # a plain device copy of the vector: what HBM delivers to a streaming
    # kernel on this box (read + write bytes / time)
    cp = torch.empty_like(u)
    t = timed(lambda: cp.copy_(u), 3 if args.copy_calibration else 10, 2,
              stream)
    results["copy"] = t
    del cp
    if args.only in ("all", "fp64"):
        dst = torch.empty_like(u)
        t = timed(lambda: op.apply(dst, u, 0.0, H.FP64, stream), args.steps,
                  args.warmup, stream)
        results["fp64"] = t
        del dst
    if args.only in ("all", "fp32"):
        u32 = u.float()
        d32 = torch.empty_like(u32)
        t = timed(lambda: op.apply(d32, u32, 0.0, H.FP32, stream), args.steps,
                  args.warmup, stream)
        results["fp32"] = t
        del u32, d32
    c2_lsrk = None
    if args.only in ("all", "fp32"):
        # the precision run on the same mesh: 20 rk45 steps per precision
        c2_lsrk = lsrk_precision_run(H, torch, stream, mf, op, u, velocity,
                                     3, 1.0 / 8)
    if args.only in ("all", "stage"):
        # fused LSRK stage, FP64 (Ti, sol in FP64) and FP32-mixed (Ti FP32)
        for prec, dtype in ((H.FP64, torch.float64), (H.FP32, torch.float32)):
            sol = u.clone()
            ta = u.to(dtype)
            tb = torch.empty_like(ta)
            t = timed(lambda: op.stage(ta, tb, sol, 0.0, 1e-4, 1e-4, prec,
                                       stream), args.steps, args.warmup,
                      stream)
            results["stage_fp64" if prec == H.FP64 else "stage_fp32"] = t
            del sol, ta, tb
    torch.cuda.synchronize()
    del u
    torch.cuda.empty_cache()
    batch_result = c3_result = landau_result = None
    if args.only in ("all", "c3"):
        c3_result = c3_workload(H, torch, stream, max(3, args.steps // 10))
    if args.only in ("all", "batch"):
        batch_result = batch_workload(H, torch, stream, max(3, args.steps // 5))
        torch.cuda.empty_cache()
    if args.only in ("all", "landau"):
        landau_result = landau_workload(H, torch, stream,
                                        max(5, args.steps // 5))
    damping_result = None
    if args.only == "all" and not args.no_landau_rate:
        damping_result = landau_damping(H, torch)
    quad_result = None
    if args.only in ("all", "quadrature"):
        quad_result = quadrature_forms(H, torch, stream,
                                       max(3, args.steps // 5))

    def gdofs(t):
        return n / t / 1e9

    line = {
        "metric": "advection apply GDoF/s 3D+3D k=3, FP64 vs FP32-mixed; "
                  "% of HBM roofline",
        "value": None, "unit": "GDoF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded product of sines + 1e-3 SplitMix64 noise, "
                "generated on the GPU)",
        "config": {"workload": "3D+3D linear advection apply, periodic "
                               "[0,1]^6, 8^3 x 8^3 cells, k=3 (4096 DoF/cell,"
                               " 1073741824 DoF), Gauss k+1, upwind, skew 0",
                   "algorithm": ["auto", "quadrature", "collapsed"][
                       op.algorithm(H.FP64)],
                   "global_batch": n, "parallelism": "single GPU"},
    }
    if "fp64" in results:
        t = results["fp64"]
        line["value"] = round(gdofs(t), 3)
        line["ms_per_step"] = round(t * 1e3, 4)
        achieved = 16.0 * n / t / 1e9
        traffic = measured_traffic()
        line["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "traffic": traffic}
    if "copy" in results:
        line["device_copy"] = {
            "GBs": round(16.0 * n / results["copy"] / 1e9, 1),
            "note": "torch copy_ of one FP64 vector (8.6 GB read + 8.6 GB "
                    "written), same stream and events"}
    flop = {}
    for key, peak in (("fp64", FP64_PEAK_TFLOPS), ("fp32", FP32_PEAK_TFLOPS)):
        if key in results:
            tf = FLOP_PER_DOF * n / results[key] / 1e12
            flop[key] = {"TFLOPs": round(tf, 2), "peak": peak,
                         "frac": round(tf / peak, 4)}
    if flop:
        line["flop_rate"] = dict(flop, flop_per_dof=FLOP_PER_DOF,
                                 note="the collapsed apply's FMAs, D (N+1) "
                                      "per DoF; vector peaks")
    if "fp32" in results:
        t = results["fp32"]
        line["fp32_mixed"] = {
            "value": round(gdofs(t), 3), "unit": "GDoF/s",
            "ms_per_apply": round(t * 1e3, 4),
            "achieved_GBs": round(8.0 * n / t / 1e9, 1),
            "frac_hbm": round(8.0 * n / t / 1e9 / HBM_PEAK_GBS, 4)}
        if "fp64" in results:
            line["fp32_mixed"]["ratio_vs_fp64"] = round(
                results["fp64"] / t, 3)
        if c2_lsrk is not None:
            line["fp32_mixed"]["lsrk_20_steps"] = c2_lsrk
    for key, bpd in (("stage_fp64", 32.0), ("stage_fp32", 24.0)):
        if key in results:
            t = results[key]
            line[key] = {"value": round(gdofs(t), 3), "unit": "GDoF/s",
                         "ms_per_stage": round(t * 1e3, 4),
                         "bytes_per_dof": bpd,
                         "achieved_GBs": round(bpd * n / t / 1e9, 1)}
    if batch_result is not None:
        line["batch_2d2d"] = batch_result
    if c3_result is not None:
        line["c3_3d3d_k5"] = c3_result
    if landau_result is not None:
        line["landau_2d2d"] = landau_result
    if damping_result is not None:
        line["landau_damping"] = damping_result
    if quad_result is not None:
        line["quadrature_forms"] = quad_result
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(velocity)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
