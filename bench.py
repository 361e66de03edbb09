"""Benchmark of the hot path: forward finite-volume steps + the discrete
adjoint sweep (checkpointed, recompute in reverse) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--n 2048]

Workload: uniform 2048 x 2048 grid (2.5 m) on a seeded 1025^2 diamond-square
DEM (5 m, 800 m relief, H = 0.8), three seeded paraboloid piles, phi_int 32 /
phi_bed 22 deg; K forward steps, then the adjoint of "h integrated over a
seeded 20 x 20-cell target region" back to step 0 with FP64 checkpoints every
50 steps.  A "step" is one forward step plus its reverse step over every
cell; value = cells * K / (forward + adjoint kernel time).

Prints one JSON line.  Diagnostics go to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+adjoint cell-steps/s & %HBM roofline, mixed/batched vs FP64 GPU at fixed err"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per cell per launch (unique bytes, perfect neighbour
# reuse): forward update = U in + U out + 7 static coefficients; adjoint
# gather = U^s + lambda^{s+1} in, lambda^s out, 7 static coefficients
ALGO_BYTES = {"coef": 24 + 8, "update": 24 + 24 + 56, "adjoint": 24 + 24 + 24 + 56}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(pb, nsteps):
    """The oracle (FP64 C restatement, one core) on a bounded sample: the
    first `nsteps` forward steps of the same workload and their adjoint."""
    import numpy as np

    import oracle
    orc = oracle.Oracle.from_problem(pb)
    lam = np.zeros((3, pb.n))
    lam[0] = 1.0
    # initialisation (not timed): fault in the oracle's scratch buffers
    W = [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]
    dt0, _ = orc.step(W, 0)
    orc.adjoint_step([pb.h, pb.hu, pb.hv], 0, dt0, lam)
    orc.time = 0.0
    U = [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]
    traj = [[a.copy() for a in U]]
    dts = []
    tf = 0.0
    for s in range(nsteps):
        t0 = time.perf_counter()
        dt, _ = orc.step(U, s)
        tf += time.perf_counter() - t0
        dts.append(dt)
        traj.append([a.copy() for a in U])   # trajectory store: bookkeeping, not timed
    t0 = time.perf_counter()
    for s in range(nsteps - 1, -1, -1):
        lam, _ = orc.adjoint_step(traj[s], s, dts[s], lam)
    ta = time.perf_counter() - t0
    return tf, ta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--checkpoint-every", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this hot path has no multi-GPU mode "
              "(ensembles shard by instance, one mesh by SFC range; neither is built)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from discrete_adjoint_amd import problems
    from discrete_adjoint_amd.solver import Simulation

    n, K, W = args.n, args.steps, args.warmup
    dx = 5120.0 / n
    t_setup = time.perf_counter()
    pb = problems.fractal_problem(n=n, dx=dx, nsteps=K, npiles=3, seed=11, dem_seed=1,
                                  checkpoint_every=args.checkpoint_every)
    log(f"setup {time.perf_counter() - t_setup:.1f}s: {n}x{n} cells, dx={dx}, "
        f"volume {pb.h.sum() * dx * dx:.4g} m^3")

    sim = Simulation(pb, checkpoint_every=args.checkpoint_every, max_steps=max(K, W, 1))
    # warmup: W forward steps + their adjoint (untimed)
    if W > 0:
        sim.run(W)
        sim.dual_solver(np.concatenate([np.full(pb.n, 1.0), np.zeros(2 * pb.n)]))
        sim.set_state(pb.h, pb.hu, pb.hv)
        log(f"warmup {W} steps done")

    wall0 = time.perf_counter()
    st = sim.run(K)
    wall_f = time.perf_counter() - wall0
    fwd_ms, _, _ = sim.timing()
    h, _, _ = sim.get_state()
    box = problems.target_box(pb, h, cells=20, seed=5)
    w = pb.box_weights(box)
    J = float(np.sum(w[: pb.n] * h))
    wall0 = time.perf_counter()
    lam0, dphi = sim.dual_solver(w, want_dphi=True)
    wall_a = time.perf_counter() - wall0
    _, adj_ms, rec_ms = sim.timing()
    total_s = (fwd_ms + adj_ms) / 1e3
    value = pb.n * K / total_s
    log(f"forward {K} steps: {fwd_ms:.1f} ms kernels ({wall_f * 1e3:.1f} ms wall), "
        f"t_end={st.time:.3f}s, mass={st.mass:.6g}")
    log(f"adjoint: {adj_ms:.1f} ms kernels incl. recompute {rec_ms:.1f} ms "
        f"({wall_a * 1e3:.1f} ms wall); J={J:.6g} dJ/dphi_bed={dphi:.6g}")
    if not (np.isfinite(lam0).all() and math.isfinite(dphi)):
        raise RuntimeError("non-finite adjoint")

    # dominant kernel, timed live with HIP events on the library's stream
    kt = {name: sim.time_kernel(i, args.kernel_reps)
          for i, name in enumerate(("coef", "update", "adjoint"))}
    k = args.checkpoint_every
    per_step = {"coef": kt["coef"] * (2 - 1.0 / k), "update": kt["update"] * (2 - 1.0 / k),
                "adjoint": kt["adjoint"]}
    dom = max(per_step, key=per_step.get)
    achieved = ALGO_BYTES[dom] * pb.n / (kt[dom] * 1e-3) / 1e9
    log("kernel avg ms: " + ", ".join(f"{a}={b:.3f}" for a, b in kt.items()) +
        f"; dominant {dom}: {achieved:.0f} GB/s algorithmic")
    for name in kt:
        log(f"  {name}: {ALGO_BYTES[name] * pb.n / (kt[name] * 1e-3) / 1e9:.0f} GB/s")
    sim.close()

    cpu = None
    if not args.no_cpu:
        cs = args.cpu_steps
        tf, ta = cpu_baseline(pb, cs)
        cpu = {"value": pb.n * cs / (tf + ta), "unit": "cell-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"first {cs} forward steps of the same {n}x{n} workload and their "
                         f"{cs} adjoint steps (oracle, FP64 C, -O3, one thread); "
                         f"forward {pb.n * cs / tf:.3g}, adjoint {pb.n * cs / ta:.3g} cell-steps/s"}
        log(f"cpu baseline: fwd {tf:.2f}s adj {ta:.2f}s")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": (fwd_ms + adj_ms) / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"uniform {n}x{n} ({dx} m) fractal DEM, 3 piles, {K} forward steps + "
                        f"discrete adjoint, FP64 checkpoints every {k}",
            "cells": pb.n,
            "checkpoint_every": k,
            "precision": "fp64",
            "forward_cell_steps_per_s": pb.n * K / (fwd_ms / 1e3),
            "adjoint_cell_steps_per_s": pb.n * K / (adj_ms / 1e3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": dom,
            "algorithmic_bytes_per_cell": ALGO_BYTES[dom],
            "avg_launch_ms": kt[dom],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
