"""2D+2D Vlasov-Poisson Landau damping on the GPU, with and without the
learned Poisson initial guess.

    python hyperdeal_amd/examples/landau.py [--cells 32] [--degree 3]
        [--t-end 10] [--cfl 0.3] [--runs fp64,fp32,fp32-learned]
        [--json out.json]

Case: x in [0, 4 pi]^2 periodic, v in [-6, 6]^2, f0 = (1 + 0.01 cos(x1/2)
cos(x2/2)) exp(-|v|^2/2) / (2 pi), cells^4 cells of degree k, LSRK rk45 with
dt = cfl * (h_x / 6) / k^1.5 (the fastest transport, |v| = 6 along x), the
field recomputed from the stage input in every stage (the reference's RK
right-hand side, examples/vlasov_poisson/include/application.h:516-602),
Poisson CG to 1e-10 relative residual.

Runs (any subset, in this order):
  fp64          the plain method in FP64 (the reference point)
  fp32          FP32-mixed apply, CG from the previous phi (the method
                without the learned part)
  fp32-learned  FP32-mixed apply, CG from the combination of the previous
                phi and the network's correction for its residual; the
                network trains online on the run's own solves
  fp64-learned  the same with the FP64 apply
  fp64-exact    FP64, CG from the exact solution of the periodic SIPG system
                (fast diagonalisation, hd_vp_exact_start)
  fp32-exact    the same with the FP32-mixed apply

Reported per run: wall time per step, CG iterations per solve, the share
of the step spent on the field update (density, network, CG, training:
one minus the ratio of the frozen-field step time to the full step time),
the damping rate from a fit of log(field energy) at its maxima, mass drift,
and for the FP32 runs the relative L2 (nodal) difference to the FP64 run.
The damping rate of the linear dispersion relation
1 + (1 + z Z(z)) / k^2 = 0, z = omega / (sqrt(2) k), |k| = 1/sqrt(2) is
printed beside it.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))))


from hyperdeal_amd.diagnostics import fit_damping_rate as fit_rate  # noqa
from hyperdeal_amd.diagnostics import landau_theory_rate as theory_rate  # noqa


def run(H, torch, name, args, ref_f=None):
    L = 4 * np.pi
    n, k = args.cells, args.degree
    mf = H.MatrixFree(2, 2, k, [n] * 4, [0, 0, -6, -6], [L, L, 6, 6])
    op = H.VlasovPoissonOperation(mf, poisson_tol=1e-10)
    prec = H.FP64 if name.startswith("fp64") else H.FP32
    if name.endswith("-exact"):
        op.exact_start()
    if name.endswith("-learned"):
        op.enable_learned_guess(seed=20261205, capacity=64, train_every=1,
                                train_batch=4, train_steps=1,
                                learning_rate=1e-3, min_pairs=8)
    f = mf.initialize_dof_vector()
    H.VectorTools.interpolate(mf, f, H.FUNC_LANDAU, [0.01, 0.5, 2.0])
    m0 = H.vp_moments(mf, f)[0]
    dtype = torch.float64 if prec == H.FP64 else torch.float32
    ta = torch.empty(f.numel(), dtype=dtype, device="cuda")
    tb = torch.empty_like(ta)
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    hx = L / n
    dt = args.cfl * (hx / 6.0) / k ** 1.5
    n_steps = int(np.ceil(args.t_end / dt))
    dt = args.t_end / n_steps
    ts, en = [], []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    last_print = t_start
    for s in range(n_steps):
        integ.perform_time_step(op, f, ta, tb, s * dt, dt, precision=prec)
        # the field of the last stage's input, at about t + dt
        ts.append((s + 1) * dt)
        en.append(sum(op.field_energy()))
        now = time.perf_counter()
        if now - last_print > 20:
            print("  %s: step %d / %d" % (name, s + 1, n_steps), flush=True)
            last_print = now
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    it, solves, res = op.solver_stats()
    # the step time with the field frozen (no density / solve / training)
    op.set_auto_update(False)
    nf = min(20, n_steps)
    g = f.clone()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s in range(nf):
        integ.perform_time_step(op, g, ta, tb, s * dt, dt, precision=prec)
    torch.cuda.synchronize()
    frozen = (time.perf_counter() - t1) / nf
    del g
    gamma, n_peaks = fit_rate(ts, en)
    out = {
        "run": name, "steps": n_steps, "dt": dt,
        "ms_per_step": 1e3 * wall / n_steps,
        "ms_per_step_frozen_field": 1e3 * frozen,
        "field_update_share": 1.0 - frozen / (wall / n_steps),
        "cg_iterations": it, "solves": solves,
        "cg_iterations_per_solve": it / max(solves, 1),
        "last_rel_residual": res,
        "damping_rate": gamma, "fit_peaks": n_peaks,
        "mass_drift": (H.vp_moments(mf, f)[0] - m0) / m0,
        "gdof_per_s_stage": 5 * n_steps * mf.n_dofs() / wall / 1e9,
    }
    if name.endswith("-learned"):
        st, guessed, pairs, loss = op.learned_guess_stats()
        r0, rp, w, _ = op.solver_start()
        out.update({"train_steps": st, "guessed_solves": guessed,
                    "last_train_loss": loss,
                    "last_start_rel_residual": r0,
                    "last_previous_phi_rel_residual": rp,
                    "last_start_weights": list(w)})
    if ref_f is not None and name != "fp64":
        out["rel_l2_diff_vs_fp64"] = float(
            torch.linalg.vector_norm(f - ref_f) /
            torch.linalg.vector_norm(ref_f))
    out["energy_trace"] = [[round(t, 5), e] for t, e in
                           zip(ts[::max(1, len(ts) // 200)],
                               en[::max(1, len(en) // 200)])]
    return out, f


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=32)
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--t-end", type=float, default=10.0)
    ap.add_argument("--cfl", type=float, default=0.3)
    ap.add_argument("--runs", default="fp64,fp32,fp32-learned")
    ap.add_argument("--json", default=None)
    args = ap.parse_args()
    import torch
    import hyperdeal_amd as H
    H.load_library()
    omega, gamma_th = theory_rate(np.sqrt(0.5) * 1.0)
    results = {"case": "landau 2d2d", "cells": args.cells,
               "degree": args.degree, "t_end": args.t_end, "cfl": args.cfl,
               "theory": {"omega": omega, "gamma": gamma_th}, "runs": []}
    ref = None
    for name in args.runs.split(","):
        print("run", name, flush=True)
        out, f = run(H, torch, name, args, ref)
        if name == "fp64" and ref is None:
            ref = f
        else:
            del f
        slim = {k: v for k, v in out.items() if k != "energy_trace"}
        print(json.dumps(slim), flush=True)
        results["runs"].append(out)
        torch.cuda.empty_cache()
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(results, fh)


if __name__ == "__main__":
    main()
