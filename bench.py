"""Benchmark of the hot path: forward finite-volume steps + the discrete
adjoint sweep (checkpointed, recomputed in reverse) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--members 256]
                    [--workload ensemble|grid|adaptive] [--precision fp64|adj32|mixed|fp32]

Default workload ("ensemble"): 256 independent members advanced together in
every launch.  Member b is a uniform side x side block (side drawn from
{64, 128, 256, 512}, 5 m cells) on its own seeded diamond-square DEM, with
phi_bed ~ U(15, 30) deg, phi_int ~ U(28, 38) deg and one paraboloid pile of
volume lognormal(ln 3e5 m^3, 0.7); K = 500 forward steps (the default; any
--steps), then the adjoint of "h integrated over the middle half of the
block at the last step" back to step 0, with dJ/dphi_bed per member.

The measured path (the headline value) is the FP64 batched path (bit-exact
with the CPU restatement; --precision selects another mode) with checkpoints
every 25 steps and nothing else kept: the reverse sweep restores each
checkpoint and recomputes its segment, as the memory-bounded schedule does.
Sweeps are captured once and replayed (hipGraph), tiles come from the
project's autotuner.  The same ensemble is also run (a) with the last states
of the forward run kept whole in HBM (stored trajectory: less recompute),
(b) member by member in FP64 (the plain FP64 GPU path, warmed, graph
replayed, with the same tiles), and (c) batched in the other precisions
(errors against FP64 per member).

A "step" is one forward step plus its reverse step over every cell of every
member; value = total cell-steps / wall time of the timed forward run plus
reverse sweep (synchronised on both sides, inputs resident on the GPU).  The
kernels skip blocks known to hold +0.0 and take exact short paths on dry or
void tiles, so the line also gives the cell-steps that took the full path.

"grid" workload: one 2048 x 2048 block (2.5 m) with three piles, 1500 steps,
a reduced-precision mode against FP64 (L2 / Linf errors of h, hu, hv and the
functional).

"adaptive" workload: the FP64 leaf-table path on a fixed adaptive mesh of
0.88 M leaves (40 m base cells refined to 5 m around two piles on the
fractal DEM), 200 steps + adjoint.  The ensemble run also reports this leg
(config.adaptive_leaves_fp64) unless --no-adaptive.

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
KERNELS = ("coef", "update", "adjoint")


def algo_bytes(prec):
    """Algorithmic HBM bytes per cell for one launch of each kernel (unique
    bytes, perfect neighbour reuse).  es = state/adjoint element size,
    ss = static-coefficient / kact element size.
      coef:    read U (3), write kact (1)
      update:  read U (3), write U (3), read 7 static coefficients
      adjoint: read U^s (3), read lambda^{s+1} (3), write lambda^s (3), 7 statics
    The forward step (coef + update) is charged the update's bytes: kact is
    an intermediate."""
    from discrete_adjoint_amd import _lib
    es = 4 if prec == _lib.TG_FP32 else 8
    ss = 8 if prec in (_lib.TG_FP64, _lib.TG_FWD64_ADJ32) else 4
    sa = 8 if prec == _lib.TG_FP64 else 4  # statics read by the adjoint
    return {"coef": 3 * es + ss, "update": 6 * es + 7 * ss, "adjoint": 9 * es + 7 * sa,
            "forward_step": 6 * es + 7 * ss}


PRECISIONS = {"adj32": "TG_FWD64_ADJ32", "mixed": "TG_MIXED", "fp32": "TG_FP32",
              "fp64": "TG_FP64"}
DTYPES = {"adj32": "f64 forward + f32 adjoint", "mixed": "f32 increments on f64 state",
          "fp32": "f32", "fp64": "f64"}


def prec_of(name):
    from discrete_adjoint_amd import _lib
    return getattr(_lib, PRECISIONS[name])


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stats(e):
    import numpy as np
    e = np.asarray(e, dtype=float)
    return {"max": float(e.max()), "median": float(np.median(e)),
            "p90": float(np.percentile(e, 90))}


def cpu_info():
    """Host CPU model and the cores this process may run on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return model, usable, os.cpu_count()


def cpu_baseline(problems_, U_mid, s_mid, nsteps, omp):
    """The oracle (FP64 C restatement of the reference) on a bounded sample
    of the same workload: `nsteps` forward steps of every member from its
    state U^{s_mid} of the GPU run (a representative, moving state) and the
    adjoint steps back over them.  omp: the OpenMP build on all usable
    cores, else one thread.  Returns (cell-steps, forward s, adjoint s)."""
    import numpy as np

    import oracle
    tf = ta = 0.0
    cs = 0
    o = 0
    for pb in problems_:
        orc = oracle.Oracle.from_problem(pb, omp=omp)
        U = [np.array(a[o:o + pb.n], copy=True) for a in U_mid]
        o += pb.n
        lam = pb.box_weights().reshape(3, -1).copy()
        traj = [[a.copy() for a in U]]
        dts = []
        for s in range(s_mid, s_mid + nsteps):
            t0 = time.perf_counter()
            dt, _ = orc.step(U, s)
            tf += time.perf_counter() - t0
            dts.append(dt)
            traj.append([a.copy() for a in U])   # trajectory store: bookkeeping, not timed
        t0 = time.perf_counter()
        for i in range(nsteps - 1, -1, -1):
            lam, _ = orc.adjoint_step(traj[i], s_mid + i, dts[i], lam)
        ta += time.perf_counter() - t0
        cs += pb.n * nsteps
    return cs, tf, ta


def warm_oracle(pb, omp=False):
    """Faults in the oracle's reusable scratch buffers (untimed)."""
    import numpy as np

    import oracle
    orc = oracle.Oracle.from_problem(pb, omp=omp)
    U = [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]
    dt, _ = orc.step(U, 0)
    lam = np.zeros((3, pb.n))
    orc.adjoint_step([pb.h, pb.hu, pb.hv], 0, dt, lam)


def timed_run(sim, K, weights, h0, want_lambda=False):
    """Forward K steps + reverse sweep, each timed by wall clock around a
    synchronous library call, inputs resident on the device before the
    clock starts (initial state, terminal weights).  Returns a dict."""
    import numpy as np
    sim.set_functional(weights)
    sim.set_state(*h0)
    t0 = time.perf_counter()
    st = sim.run(K)
    wall_f = time.perf_counter() - t0
    fwd_ms, _, _ = sim.timing()
    H = sim.get_state()
    t0 = time.perf_counter()
    lam, dphi = sim.dual_solver(None, want_dphi=True, want_lambda=want_lambda)
    wall_a = time.perf_counter() - t0
    _, adj_ms, rec_ms = sim.timing()
    stored_ms, stored_steps = sim.stored_sweep_timing()
    return dict(st=st, H=H, lam=lam, dphi=np.atleast_1d(np.asarray(dphi, dtype=float)),
                wall_f=wall_f, wall_a=wall_a, fwd_ms=fwd_ms, adj_ms=adj_ms, rec_ms=rec_ms,
                stored_ms=stored_ms, stored_steps=stored_steps)


def forward_step_json(rl, prec, ab=None, sb=None):
    ab = ab or algo_bytes(prec)
    sb = sb or short_bytes(prec)
    f = rl["forward_step"]
    return {"achieved": f["gbs"], "frac": f["gbs"] / HBM_PEAK_GBS,
            "avg_step_ms": f["avg_launch_ms"],
            "algorithmic_bytes_per_cell": ab["forward_step"],
            "short_path_bytes_per_cell": sb["update"]}


def roofline_json(rl, dom, kin, prec, workload, ab=None, sb=None, run_config=None):
    """The roofline object of the JSON line for the dominant role `dom`."""
    ab = ab or algo_bytes(prec)
    sb = sb or short_bytes(prec)
    k = kin["update" if dom == "forward_step" else dom]
    traffic, tinfo = (traffic_from_profile(dom, workload, run_config) if run_config
                      else (None, {"source": None}))
    out = {"bound": "hbm", "achieved": rl[dom]["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": rl[dom]["gbs"] / HBM_PEAK_GBS,
           "traffic": traffic, "kernel": dom,
           "algorithmic_bytes_per_cell": ab["forward_step" if dom == "forward_step" else dom],
           "short_path_bytes_per_cell": sb["update" if dom == "forward_step" else dom],
           "algorithmic_bytes_per_launch": rl[dom]["bytes_per_launch"],
           "full_path_share": k["full_path_share"], "short_path_share": k["short_path_share"],
           "avg_launch_ms": rl[dom]["avg_launch_ms"],
           "launches_timed": rl[dom]["launches_timed"],
           "traffic_info": tinfo}
    if traffic:
        out["traffic_over_algorithmic"] = traffic / rl[dom]["bytes_per_launch"]
    return out


def recomputed_steps(K, k, stored):
    """Forward steps the reverse sweep recomputes: below the stored part
    (K - stored states), each checkpoint segment [a, b) recomputes
    U^{a+1} .. U^{b-1}."""
    s0 = K - stored
    return sum(min(a + k, s0) - a - 1 for a in range(0, s0, k))


def replay_roofline(r, kin, K, k):
    """Launch durations from the timed (graph-replayed) run itself, HIP
    events on the library's stream around the replayed sweeps: the forward
    step (coefficients + dt + update, checkpoint copies included) = the
    forward sweep / K; the reverse step = the reverse sweep minus its
    checkpoint-segment recompute (timed apart by events between the
    segment graphs) / K; the recompute step = that recompute / its steps.
    Bytes per launch from the instrumented run's counters (the same inputs
    give the same counts).  Returns per-role figures and the dominant role
    (largest time in the run)."""
    out = {}
    rec = max(r["rec_ms"], 0.0)
    ms = (r["adj_ms"] - rec) / K
    b = kin["adjoint"]["bytes_per_launch"]
    out["adjoint"] = {"avg_launch_ms": ms, "launches_timed": K, "bytes_per_launch": b,
                      "gbs": b / (ms * 1e-3) / 1e9}
    ms = r["fwd_ms"] / K
    b = kin["update"]["bytes_per_launch"]
    out["forward_step"] = {"avg_launch_ms": ms, "launches_timed": K, "bytes_per_launch": b,
                           "gbs": b / (ms * 1e-3) / 1e9}
    nrec = recomputed_steps(K, k, r["stored_steps"])
    if nrec > 0 and rec > 0:
        ms = rec / nrec
        out["recompute_step"] = {"avg_launch_ms": ms, "launches_timed": nrec,
                                 "bytes_per_launch": b, "gbs": b / (ms * 1e-3) / 1e9,
                                 "note": "checkpoint restores included"}
    dom = max(("adjoint", "forward_step"), key=lambda n: out[n]["avg_launch_ms"])
    return dom, out


def path_rates(kin, K, wall, ncells):
    """Cell-steps per second that took each kernel's full path, and that
    were visited at all (full or short path; the rest were skipped by the
    zero marks), averaged over the update and the adjoint launches."""
    full = sum(kin[n]["full_path_share"] for n in ("update", "adjoint")) / 2
    short = sum(kin[n]["short_path_share"] for n in ("update", "adjoint")) / 2
    cs = ncells * K
    return {"full_path_cell_steps_per_s": full * cs / wall,
            "visited_cell_steps_per_s": (full + short) * cs / wall,
            "full_path_share": full, "short_path_share": short}


def short_bytes(prec):
    """Bytes per cell of each kernel's short path: a cell below tiny in the
    coefficient pass reads h only (its kact is never read); a cell of an
    entirely dry update tile reads U and writes it unchanged; a cell of a
    void adjoint tile reads h and lambda^{s+1} and writes lambda^s = 0."""
    from discrete_adjoint_amd import _lib
    es = 4 if prec == _lib.TG_FP32 else 8
    return {"coef": es, "update": 6 * es, "adjoint": 7 * es}


def path_cells(cells, full, short):
    """(full-path cells, short-path cells) out of `cells` visited
    cell-launches, from the kernels' counters: an uncounted full figure (-1)
    means every cell took the full path, an uncounted short figure every
    other cell the short path.  Cells in neither were skipped (known +0.0
    from the zero marks: nothing read or written)."""
    fc = cells if full < 0 else full
    sc = cells - fc if short < 0 else short
    return fc, sc


def instrumented_roofline(sim, prec, ncells, K, weights, h0, ab=None, sb=None):
    """One more forward run + reverse sweep, launched kernel by kernel with a
    HIP event pair around every launch (on the library's stream) and the
    kernels counting their full-path and short-path cells.  Per kernel:
    summed launch time, launches, and algorithmic bytes = full-path cells x
    the full per-cell bytes + short-path cells x the short-path bytes
    (skipped cells move none).  Returns the dominant kernel (largest summed
    time) and per-kernel figures."""
    ab = ab or algo_bytes(prec)
    sb = sb or short_bytes(prec)
    sim.instrument(True)
    timed_run(sim, K, weights, h0)
    ms, launches, full, short = sim.instrumentation()
    sim.instrument(False)
    out = {}
    for i, name in enumerate(KERNELS):
        cells = launches[i] * ncells
        fc, sc = path_cells(cells, full[i], short[i])
        nbytes = fc * ab[name] + sc * sb[name]
        out[name] = {"ms_total": ms[i], "launches": launches[i],
                     "avg_launch_ms": ms[i] / max(launches[i], 1),
                     "full_path_share": fc / max(cells, 1),
                     "short_path_share": sc / max(cells, 1),
                     "bytes_per_launch": nbytes / max(launches[i], 1),
                     "gbs": nbytes / (ms[i] * 1e-3) / 1e9 if ms[i] > 0 else None}
        log(f"  {name}: {launches[i]} launches, avg {out[name]['avg_launch_ms']:.4f} ms, "
            f"full path {out[name]['full_path_share']:.3f}, short path "
            f"{out[name]['short_path_share']:.3f}, {out[name]['gbs']:.0f} GB/s")
    upd = out["update"]
    fwd_ms = ms[0] + ms[1]
    fwd_gbs = upd["bytes_per_launch"] * upd["launches"] / (fwd_ms * 1e-3) / 1e9
    dom = max(KERNELS, key=lambda n: out[n]["ms_total"])
    return dom, out, fwd_gbs


def tile_shape(v):
    import ctypes as C

    from discrete_adjoint_amd import _lib
    tx, ty, band = C.c_int(), C.c_int(), C.c_int()
    _lib.load().tg_tile_variant(v, C.byref(tx), C.byref(ty), C.byref(band))
    return f"{tx.value}x{ty.value}" + (" xcd-banded" if band.value else " row-major")


def traffic_from_profile(kernel_name, workload, run_config):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE runs of --probe, scaled by a copy kernel's counters), from the
    committed summary for this kernel role and workload -- used only when the
    summary's run configuration (members, steps, checkpoint interval,
    precision, stored trajectory, tiles) is this run's, so that traffic and
    algorithmic bytes describe the same launches.  Returns (traffic, info)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel_name}_{workload}.json")
    if not os.path.exists(path):
        return None, {"source": None, "reason": f"no {os.path.relpath(path, ROOT)}"}
    with open(path) as f:
        t = json.load(f)
    info = {"source": os.path.relpath(path, ROOT), "profile_config": t.get("config"),
            "profile_algorithmic_bytes_per_launch": t.get("algorithmic_bytes_per_launch")}
    if t.get("config") != run_config:
        info["reason"] = "profile taken on another configuration: not used"
        return None, info
    return t.get("hbm_bytes_per_launch"), info


def store_budget(args):
    """Bytes of device memory a context may use to keep the last forward
    states whole (the reverse sweep recomputes only below them): a fraction
    of the free HBM."""
    import ctypes as C

    from discrete_adjoint_amd import _lib
    free, total = C.c_int64(), C.c_int64()
    rc = _lib.load().tg_device_memory(0, C.byref(free), C.byref(total))
    if rc:
        raise RuntimeError(f"tg_device_memory returned {rc}")
    return int(args.store_fraction * free.value)


def ensemble_run_config(args, store, tiles):
    """What a traffic profile must share with this run to describe the same
    launches."""
    return {"members": args.members, "seed": args.seed, "steps": args.steps,
            "checkpoint_every": args.checkpoint_every, "precision": args.precision,
            "stored_trajectory": bool(store), "tiles": list(tiles)}


def plain_fp64(E, weights, K, W, k, graph, tiles):
    """The plain FP64 GPU path: one member at a time, each warmed untimed
    (first launches, checkpoint store, graph capture of its K-step sweeps),
    its functional handed over before the clock, the batched run's tiles,
    checkpoints only like the measured path.  Returns wall seconds."""
    from discrete_adjoint_amd.solver import Simulation
    total = 0.0
    for b, p in enumerate(E):
        with Simulation(p, checkpoint_every=k, max_steps=max(K, W, 1), use_graph=graph,
                        store_bytes=0) as sim:
            sim.select_tiles(tiles)
            sim.run(K)
            sim.set_functional(weights[b])  # the timed call's graphs captured here
            sim.dual_solver(None, want_dphi=True, want_lambda=False)
            sim.set_state(p.h, p.hu, p.hv)
            t0 = time.perf_counter()
            sim.run(K)
            sim.dual_solver(None, want_dphi=True, want_lambda=False)
            total += time.perf_counter() - t0
    return total


def bench_ensemble(args):
    import numpy as np

    from discrete_adjoint_amd import _lib, problems
    from discrete_adjoint_amd.solver import Ensemble

    K, W, k = args.steps, args.warmup, args.checkpoint_every
    t0 = time.perf_counter()
    E = problems.ensemble_problems(args.members, seed=args.seed, nsteps=K, checkpoint_every=k)
    ncells = sum(p.n for p in E)
    weights = [p.box_weights().reshape(3, -1) for p in E]
    h0 = tuple(np.concatenate([getattr(p, f) for p in E]) for f in ("h", "hu", "hv"))
    log(f"setup {time.perf_counter() - t0:.1f}s: {len(E)} members, {ncells} cells")
    prec = prec_of(args.precision)
    cell_steps = ncells * K

    def member_J(H):
        out, o = [], 0
        for b, p in enumerate(E):
            out.append(float(np.sum(weights[b][0] * H[0][o:o + p.n])))
            o += p.n
        return np.array(out)

    graph = 0 if args.no_graph else 1
    store_head = store_budget(args) if args.store else 0

    def measure(mode, store, roofline=False, mid_state=False):
        with Ensemble(E, precision=mode, checkpoint_every=k, max_steps=max(K, W, 1),
                      use_graph=graph, store_bytes=store) as ens:
            # untimed: first launches, allocation of the checkpoint store
            ens.run(max(W, 1))
            ens.dual_solver(weights, want_lambda=False)
            if not args.no_autotune:  # tile variant per kernel (untimed)
                var, _ = ens.autotune(reps=5)
                log(f"precision {mode}: autotuned tile variants {var}")
            if graph:  # capture the K-step sweeps (untimed), as the timed run calls them
                ens.set_state(*h0)
                ens.run(K)
                ens.set_functional(weights)
                ens.dual_solver(None, want_dphi=True, want_lambda=False)
            r = timed_run(ens, K, weights, h0)
            r["tiles"] = ens.tile_selection()
            r["J"] = member_J(r["H"])
            del r["H"]
            log(f"precision {mode} store {store > 0}: forward {r['wall_f'] * 1e3:.1f} ms wall "
                f"({r['fwd_ms']:.1f} ms events), reverse {r['wall_a'] * 1e3:.1f} ms wall "
                f"({r['adj_ms']:.1f} ms events, recompute {r['rec_ms']:.1f}), "
                f"{cell_steps / (r['wall_f'] + r['wall_a']):.3g} cell-steps/s")
            if roofline:
                r["roof"] = instrumented_roofline(ens, mode, ncells, K, weights, h0)
            if mid_state:
                r["mid"] = ens.trajectory_state(K // 2)
            return r

    res = {}
    names = [args.precision]
    if not args.product_only:
        names += ["fp64"] if args.precision != "fp64" else []
        names += [c for c in args.compare.split(",") if c and c not in names]
    for name in names:
        mode = prec_of(name)
        res[mode] = measure(mode, store_head, roofline=(mode == prec),
                            mid_state=(mode == _lib.TG_FP64 and not args.no_cpu))
    m, d = res[prec], res.get(_lib.TG_FP64)
    wall = m["wall_f"] + m["wall_a"]
    value = cell_steps / wall

    stored = None
    if not args.product_only and not args.store:
        r = measure(prec, store_budget(args))
        stored = {"cell_steps_per_s": cell_steps / (r["wall_f"] + r["wall_a"]),
                  "stored_states": r["stored_steps"], "bytes": store_budget(args),
                  "kernel_ms": {"forward": r["fwd_ms"], "reverse": r["adj_ms"],
                                "recompute": r["rec_ms"]}}

    plain = None
    if not args.product_only:
        plain_s = plain_fp64(E, weights, K, W, k, graph, m["tiles"])
        plain = cell_steps / plain_s
        log(f"plain FP64 per-member: {plain_s:.2f} s wall, {plain:.3g} cell-steps/s")

    def errors(r):
        Jd, Jm = d["J"], r["J"]
        okJ = np.abs(Jd) > 0
        eJ = np.abs(Jm[okJ] - Jd[okJ]) / np.abs(Jd[okJ])
        okg = np.abs(d["dphi"]) > 0
        eg = np.abs(r["dphi"][okg] - d["dphi"][okg]) / np.abs(d["dphi"][okg])
        gmax = np.abs(d["dphi"]).max()
        return {"functional_rel": stats(eJ),
                "functional_members_le_1e-6": int((eJ <= 1e-6).sum()),
                "dJ_dphibed_rel": stats(eg),
                "dJ_dphibed_members_le_1e-6": int((eg <= 1e-6).sum()),
                "dJ_dphibed_abs_err_over_max": float(np.abs(r["dphi"] - d["dphi"]).max() / gmax),
                "members": len(E)}

    errs = errors(m) if (d is not None and prec != _lib.TG_FP64) else None
    others = {}
    for name in names:
        mode = prec_of(name)
        if mode in (prec, _lib.TG_FP64):
            continue
        r = res[mode]
        others[name] = {"cell_steps_per_s": cell_steps / (r["wall_f"] + r["wall_a"]),
                        "errors_vs_fp64": errors(r)}
        log(f"{name}: {others[name]}")

    _, kin, _ = m["roof"]
    dom, rl = replay_roofline(m, kin, K, k)
    workload = f"ensemble{args.members}_{args.precision}"
    run_config = ensemble_run_config(args, store_head, m["tiles"])
    rates = path_rates(kin, K, wall, ncells)

    cpu = None
    if not (args.no_cpu or args.product_only):
        model, usable, total = cpu_info()
        s_mid = K // 2
        warm_oracle(E[0], omp=True)
        cs, tf, ta = cpu_baseline(E, d["mid"], s_mid, args.cpu_steps, omp=True)
        threads = int(os.environ.get("OMP_NUM_THREADS", usable))
        warm_oracle(E[0])
        cs1, tf1, ta1 = cpu_baseline(E, d["mid"], s_mid, args.cpu_steps, omp=False)
        cpu = {"value": cs / (tf + ta), "unit": "cell-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"{args.cpu_steps} forward steps of each of the {len(E)} members from "
                         f"its state at step {s_mid} of the GPU run, and their adjoint steps "
                         f"(oracle: FP64 C restatement of the reference, -O3, OpenMP over "
                         f"cells, {threads} threads); forward {cs / tf:.3g}, adjoint "
                         f"{cs / ta:.3g} cell-steps/s",
               "single_thread_value": cs1 / (tf1 + ta1),
               "cpu_model": model, "usable_cores": usable, "host_cpus": total}
        log(f"cpu baseline: {threads} threads {cpu['value']:.3g}, one thread "
            f"{cpu['single_thread_value']:.3g} cell-steps/s over {cs} cell-steps")

    adaptive = None
    if not args.product_only and not args.no_adaptive:
        adaptive = adaptive_leg(graph=bool(graph))[0]

    return {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"ensemble of {len(E)} members (uniform blocks, side in "
                        f"{{64,128,256,512}}, 5 m, seeded fractal DEMs, one pile each), "
                        f"{K} forward steps + discrete adjoint with dJ/dphi_bed, checkpoints "
                        f"every {k}" + (" and the last states kept whole" if store_head else
                                        " (nothing else kept: segments recomputed)") +
                        ", all members in every launch",
            "members": len(E),
            "cells": ncells,
            "checkpoint_every": k,
            "precision": args.precision,
            "forward_cell_steps_per_s": cell_steps / m["wall_f"],
            "adjoint_cell_steps_per_s": cell_steps / m["wall_a"],
            **rates,
            "kernel_ms": {"forward": m["fwd_ms"], "reverse": m["adj_ms"],
                          "recompute": m["rec_ms"]},
            "stored_trajectory_variant": stored,
            "fp64_batched_cell_steps_per_s":
                cell_steps / (d["wall_f"] + d["wall_a"]) if d is not None else None,
            "fp64_plain_cell_steps_per_s": plain,
            "speedup_vs_fp64_plain": value / plain if plain else None,
            "errors_vs_fp64": errs,
            "other_precisions": others,
            "forward_step_hbm": forward_step_json(rl, prec),
            "timed_run_launches": rl,
            "kernels_event_pair_per_launch": kin,
            "graph_replay": bool(graph),
            "stored_trajectory_bytes": store_head,
            "tile_variants": {"coef_update_adjoint": m["tiles"],
                              "shapes": [tile_shape(v) for v in m["tiles"]]},
            "adaptive_leaves_fp64": adaptive,
        },
        "roofline": roofline_json(rl, dom, kin, prec, workload, run_config=run_config),
        "cpu_baseline": cpu,
    }


# Leaf-table bytes per leaf a full-path kernel reads: nbr 8 x 4, the packed
# side-kind / back-slot word 4, cell size 16, centre 16.
LEAF_TABLE_BYTES = 68
# ... and what the short paths add to the uniform short-path bytes: the dry
# update reads the packed word, nbr and cell size; the void adjoint reads the
# packed word and nbr and writes kact and its phi_bed derivative.
LEAF_SHORT_EXTRA = {"coef": 0, "update": 52, "adjoint": 52}


# Algorithmic bytes per error element of one error-indicator evaluation:
# reconstruction reads the rule (4), the stencil's cells (16) and points
# (64) and a quarter of the father leaves' nine fields (18), writes nine
# fields (72); kact reads the reconstructed previous state (24) and the
# tables (68: nbr 32, packed word 4, size 16, centre 16), writes kact (8);
# the residual reads the three reconstructions (72), kact (8), seven statics
# (56), the tables (68) and a quarter of the father's adjoint (6), writes the
# residual, error and correction (40).
ERROR_BYTES = {"recon": 4 + 16 + 64 + 18 + 72, "kact": 24 + 68 + 8,
               "residual": 72 + 8 + 56 + 68 + 6 + 40}


def error_indicator_leg(sim, pb, K, w, reps=5):
    """The error indicator on the mesh's error grid (every leaf refined once),
    evaluated as the reference's first error_compute call: the last step's
    previous state U^{K-1} and state U^K with the functional's sensitivity
    as the adjoint.  Kernel times from HIP events, median of reps calls."""
    import numpy as np

    from discrete_adjoint_amd import errorgrid, problems
    t0 = time.perf_counter()
    geom = (pb.nx0, pb.ny0, pb.xmin, pb.ymin, pb.dx0, pb.dy0)
    eg = errorgrid.build(geom, errorgrid.cells_of_table(pb.table, pb.xmin, pb.ymin),
                         problems.leaf_dem_statics(pb))
    setup = time.perf_counter() - t0
    prev = np.array(sim.trajectory_state(K - 1))
    state = np.array(sim.trajectory_state(K))
    dt = float(sim.dt_record()[K - 1])
    lam = np.asarray(w, dtype=np.float64).reshape(3, -1)
    sim.set_error_grid(eg)
    ms = []
    for _ in range(reps):
        out = sim.error_indicator(prev, state, lam, K - 1, dt)
        ms.append(sim.error_kernel_ms())
    ms = np.median(np.array(ms), axis=0)
    names = ("recon", "kact", "residual")
    per = {nm: {"ms": float(m), "bytes_per_element": ERROR_BYTES[nm],
                "gbs": ERROR_BYTES[nm] * eg.n / (m * 1e-3) / 1e9} for nm, m in zip(names, ms)}
    tot = float(ms.sum())
    res = {"elements": eg.n, "setup_s": setup, "kernels": per, "kernel_ms": tot,
           "gbs": sum(ERROR_BYTES.values()) * eg.n / (tot * 1e-3) / 1e9,
           "rules": np.bincount(eg.kind, minlength=5).tolist(),
           "error_sum": float(out["error"].sum()), "correction_sum": float(out["correction"].sum())}
    log(f"error indicator: {eg.n} elements, kernels {tot:.3f} ms ({res['gbs']:.0f} GB/s), "
        f"host set-up {setup:.1f} s")
    return res


def adaptive_leg(K=200, k=25, graph=True, cpu_steps=0):
    """The adaptive-mesh (leaf table) FP64 path on the C1-size mesh: about
    0.88 M leaves (5 m around two piles, 40 m base cells, three levels of
    2:1-balanced refinement) on the seeded fractal DEM; K forward steps +
    adjoint with dJ/dphi_bed, checkpoints every k.  Returns the figures
    (wall-clock cell-steps/s, per-kernel launch times and algorithmic GB/s:
    the uniform FP64 bytes plus the leaf tables)."""
    import numpy as np

    from discrete_adjoint_amd import _lib, problems
    from discrete_adjoint_amd.solver import Simulation

    t0 = time.perf_counter()
    pb = problems.adaptive_fractal_problem(nsteps=K, checkpoint_every=k)
    log(f"adaptive setup {time.perf_counter() - t0:.1f}s: {pb.n} leaves, levels "
        f"{pb.meta['levels']}")
    w = pb.box_weights()
    h0 = (pb.h, pb.hu, pb.hv)
    # every state of the run kept whole (a few GB): the reverse sweep is
    # reverse steps back to back, timed as such
    with Simulation(pb, checkpoint_every=k, max_steps=K, use_graph=int(graph),
                    store_bytes=32 * pb.n * (K + 1)) as sim:
        sim.run(K)  # untimed: capture of the sweeps, first launches (as timed)
        sim.set_functional(w)
        sim.dual_solver(None, want_dphi=True, want_lambda=False)
        r = timed_run(sim, K, w, h0)
        ab = {n: v + LEAF_TABLE_BYTES for n, v in algo_bytes(_lib.TG_FP64).items()}
        sb = {n: v + LEAF_SHORT_EXTRA[n] for n, v in short_bytes(_lib.TG_FP64).items()}
        _, kin, _ = instrumented_roofline(sim, _lib.TG_FP64, pb.n, K, w, h0, ab=ab, sb=sb)
        err_ind = error_indicator_leg(sim, pb, K, w)
    dom, rl = replay_roofline(r, kin, K, k)
    cs = pb.n * K
    out = {
        "leaves": pb.n, "levels": pb.meta["levels"], "steps": K, "checkpoint_every": k,
        "cell_steps_per_s": cs / (r["wall_f"] + r["wall_a"]),
        "forward_cell_steps_per_s": cs / r["wall_f"],
        "adjoint_cell_steps_per_s": cs / r["wall_a"],
        "timed_run_launches": rl,
        "kernels_event_pair_per_launch": kin,
        "algorithmic_bytes_per_leaf": ab, "short_path_bytes_per_leaf": sb,
        "forward_step_hbm": forward_step_json(rl, _lib.TG_FP64, ab, sb),
        "dJ_dphibed": float(r["dphi"][0]),
        "error_indicator": err_ind,
    }
    log(f"adaptive: {out['cell_steps_per_s']:.3g} cell-steps/s fwd+adj, timed run "
        + ", ".join(f"{a} {v['avg_launch_ms']:.4f} ms/launch ({v['gbs']:.0f} GB/s)"
                    for a, v in rl.items()))
    if cpu_steps:
        import oracle
        orc = oracle.LeafOracle.from_problem(pb)
        U = [np.array(a, copy=True) for a in h0]
        traj, dts = [], []
        t0 = time.perf_counter()
        for s in range(cpu_steps):
            traj.append([a.copy() for a in U])
            dt, _ = orc.step(U, s)
            dts.append(dt)
        tf = time.perf_counter() - t0
        lam = w.reshape(3, -1)
        t0 = time.perf_counter()
        for s in range(cpu_steps - 1, -1, -1):
            lam, _ = orc.adjoint_step(traj[s], s, dts[s], lam, want_dphi=True)
        ta = time.perf_counter() - t0
        ccs = pb.n * cpu_steps
        out["cpu"] = {"value": ccs / (tf + ta), "unit": "cell-steps/s", "cores": 1, "kind": "port",
                      "sample": f"first {cpu_steps} forward steps of the {pb.n}-leaf mesh and "
                                f"their adjoint steps (oracle: FP64 C leaf-table restatement, "
                                f"-O3, one thread); forward {ccs / tf:.3g}, adjoint "
                                f"{ccs / ta:.3g} cell-steps/s"}
    return out, (rl, dom, kin, ab, sb), r, pb


def bench_adaptive(args):
    K, W, k = args.steps, args.warmup, args.checkpoint_every
    out, (rl, dom, kin, ab, sb), r, pb = adaptive_leg(K, k, not args.no_graph,
                                                      0 if (args.no_cpu or args.product_only)
                                                      else args.cpu_steps)
    wall = r["wall_f"] + r["wall_a"]
    cpu = out.pop("cpu", None)
    return {
        "metric": METRIC, "value": pb.n * K / wall, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": wall * 1e3 / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(workload=f"adaptive mesh of {pb.n} leaves (fractal DEM, two piles, "
                                f"levels {pb.meta['levels']}), {K} forward steps + discrete "
                                f"adjoint with dJ/dphi_bed, checkpoints every {k}",
                       **out),
        "roofline": roofline_json(rl, dom, kin, None, "adaptive882k", ab, sb),
        "cpu_baseline": cpu,
    }


def bench_grid(args):
    import numpy as np

    from discrete_adjoint_amd import _lib, problems
    from discrete_adjoint_amd.solver import Simulation

    n, K, W, k = args.n, args.steps, args.warmup, args.checkpoint_every
    dx = 5120.0 / n
    t0 = time.perf_counter()
    pb = problems.fractal_problem(n=n, dx=dx, nsteps=K, npiles=3, seed=11, dem_seed=1,
                                  checkpoint_every=k)
    log(f"setup {time.perf_counter() - t0:.1f}s: {n}x{n} cells, dx={dx}")
    prec = prec_of(args.precision)
    h0 = (pb.h, pb.hu, pb.hv)
    res = {}
    roof = None
    w = None
    for mode in ([_lib.TG_FP64, prec] if prec != _lib.TG_FP64 else [prec]):
        with Simulation(pb, precision=mode, checkpoint_every=k, max_steps=max(K, W, 1)) as sim:
            if W > 0:
                sim.run(W)
                sim.dual_solver(np.concatenate([np.ones(pb.n), np.zeros(2 * pb.n)]),
                                want_lambda=False)
            if w is None:  # target region placed on the FP64 deposit
                sim.set_state(*h0)
                sim.run(K)
                box = problems.target_box(pb, sim.get_state()[0], cells=20, seed=5)
                w = pb.box_weights(box)
            r = timed_run(sim, K, w, h0, want_lambda=True)
            r["J"] = float(np.sum(w[:pb.n] * r["H"][0]))
            res[mode] = r
            log(f"precision {mode}: forward {r['wall_f'] * 1e3:.1f} ms, reverse "
                f"{r['wall_a'] * 1e3:.1f} ms wall")
            if mode == prec:
                roof = instrumented_roofline(sim, prec, pb.n, K, w, h0)
    m, d = res[prec], res[_lib.TG_FP64]
    cell_steps = pb.n * K
    wall = m["wall_f"] + m["wall_a"]
    errs = None
    if prec != _lib.TG_FP64:
        errs = {}
        for i, name in enumerate(("h", "hu", "hv")):
            a, b = m["H"][i], d["H"][i]
            errs[name] = {"l2_rel": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                          "linf_rel": float(np.abs(a - b).max() / np.abs(b).max())}
        errs["functional_rel"] = abs(m["J"] - d["J"]) / abs(d["J"])
        errs["dJ_dphibed_rel"] = float(abs(m["dphi"][0] - d["dphi"][0]) / abs(d["dphi"][0]))
        errs["lambda0_l2_rel"] = float(np.linalg.norm(m["lam"] - d["lam"]) /
                                       np.linalg.norm(d["lam"]))
        log(f"errors vs fp64: {errs}")
    _, kin, _ = roof
    dom, rl = replay_roofline(m, kin, K, k)
    return {
        "metric": METRIC, "value": cell_steps / wall, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": wall * 1e3 / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"uniform {n}x{n} ({dx} m) fractal DEM, 3 piles, {K} forward steps + "
                        f"discrete adjoint, checkpoints every {k}",
            "cells": pb.n, "precision": args.precision,
            "fp64_cell_steps_per_s": cell_steps / (d["wall_f"] + d["wall_a"]),
            "errors_vs_fp64": errs,
            "forward_step_hbm": forward_step_json(rl, prec),
            "timed_run_launches": rl,
            "kernels_event_pair_per_launch": kin,
        },
        "roofline": roofline_json(rl, dom, kin, prec, f"grid{n}"),
        "cpu_baseline": None,
    }


def sweeps(args):
    """Run for kernel traces (rocprofv3 --kernel-trace): `--sweeps R`
    repetitions of the measured sweep (initial state, K forward steps,
    reverse sweep with dJ/dphi_bed; the bench's stored-trajectory budget and
    tile variants), nothing else.  Repetitions after the first see every
    stored slot's zero marks, as the bench's timed run does."""
    import numpy as np

    from discrete_adjoint_amd import problems
    from discrete_adjoint_amd.solver import Ensemble

    K, k = args.steps, args.checkpoint_every
    E = problems.ensemble_problems(args.members, seed=args.seed, nsteps=K, checkpoint_every=k)
    h0 = [np.concatenate([getattr(p, f) for p in E]) for f in ("h", "hu", "hv")]
    prec = prec_of(args.precision)
    with Ensemble(E, precision=prec, checkpoint_every=k, max_steps=max(K, 1),
                  use_graph=0 if args.no_graph else 1,
                  store_bytes=store_budget(args) if args.store else 0) as ens:
        ens.select_tiles([int(v) for v in (args.tiles or "3,3,3").split(",")])
        ens.set_functional([p.box_weights().reshape(3, -1) for p in E])
        for r in range(args.sweeps):
            ens.set_state(*h0)
            t0 = time.perf_counter()
            ens.run(K)
            t1 = time.perf_counter()
            ens.dual_solver(None, want_dphi=True, want_lambda=False)
            t2 = time.perf_counter()
            log(f"sweep {r}: forward {1e3 * (t1 - t0):.1f} ms, reverse {1e3 * (t2 - t1):.1f} ms")


# the two counter passes of the gather's profile (at most 8 SQ counters each)
COUNTER_PASSES = [
    ("SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_INSTS_LDS "
     "SQ_LDS_BANK_CONFLICT SQ_WAIT_INST_ANY", "sqa"),
    ("SQ_INSTS_VALU_FLOPS_FP64 SQ_INSTS_VALU_TRANS_F64 SQ_INSTS_VMEM_RD SQ_INSTS_SALU "
     "SQ_WAIT_ANY SQ_ACTIVE_INST_LDS SQ_ACTIVE_INST_VMEM SQ_LDS_IDX_ACTIVE", "sqb"),
]


def counter_summary(csvs, symbol, last, out_path, commands):
    """Per-dispatch averages of every counter in rocprofv3 --pmc CSVs over
    the last `last` dispatches of the GPU kernel `symbol` (the probe's
    instrumented launches on the final state), as a markdown table."""
    import csv

    vals = {}
    for path in csvs:
        rows = list(csv.DictReader(open(path)))
        name_key = "Kernel_Name" if "Kernel_Name" in rows[0] else "Kernel-Name"
        val_key = "Counter_Value" if "Counter_Value" in rows[0] else "Counter-Value"
        for r in rows:
            nm = r[name_key]
            if f"{symbol}<" in nm or f"{symbol}(" in nm:
                vals.setdefault(r.get("Counter_Name", r.get("Counter-Name")), []).append(
                    float(r[val_key]))
    lines = ["Commands:", ""] + [f"    {c}" for c in commands] + [
        "", f"Averages over the last {last} dispatches of `{symbol}` per counter.", "",
        "| counter | per dispatch |", "|---|---:|"]
    avg = {}
    for c in sorted(vals):
        v = vals[c][-last:]
        avg[c] = sum(v) / len(v)
        lines.append(f"| `{c}` | {avg[c]:.4g} |")
    open(out_path, "w").write("\n".join(lines) + "\n")
    return avg


def trace_summary(trace_csv, out_path, command):
    """Per-kernel dispatch counts and durations of the last forward run and
    the last reverse sweep of a `--sweeps` kernel trace (the forward run ends
    with its statistics kernel k_stats), as a markdown table."""
    import csv
    import re

    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))

    def name(r):
        m = re.search(r"(k_\w+(?:<\d+>)?)", r["Kernel_Name"])
        return m.group(1) if m else r["Kernel_Name"][:40]

    ks = [i for i, r in enumerate(rows) if name(r) == "k_stats"]
    # the last forward run starts after the previous reverse sweep's last
    # adjoint dispatch
    start = 1 + max([i for i in range(ks[-2] + 1, ks[-1]) if name(rows[i]).startswith(
        "k_adjoint")] or [ks[-2]])
    parts = [("last forward run", start, ks[-1] + 1), ("last reverse sweep", ks[-1] + 1,
                                                       len(rows))]
    lines = [f"Command: `{command}`", ""]
    for label, lo, hi in parts:
        per = {}
        for r in rows[lo:hi]:
            per.setdefault(name(r), []).append(
                (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        span = (int(rows[hi - 1]["End_Timestamp"]) - int(rows[lo]["Start_Timestamp"])) / 1e6
        busy = sum(sum(v) for v in per.values()) / 1e3
        lines += [f"### {label}: {hi - lo} dispatches, {span:.1f} ms span, {busy:.1f} ms in "
                  f"kernels", "", "| kernel | dispatches | average (us) | total (ms) |",
                  "|---|---:|---:|---:|"]
        for k, v in sorted(per.items(), key=lambda x: -sum(x[1])):
            lines.append(f"| `{k}` | {len(v)} | {sum(v) / len(v):.1f} | {sum(v) / 1e3:.2f} |")
        lines.append("")
    with open(out_path, "w") as f:
        f.write("\n".join(lines))


PROBE_WARM_SWEEPS = 2


def probe(args):
    """Run for hardware-counter profiles (rocprofv3 --pmc): the measured
    ensemble run -- same members, K, checkpoints, stored trajectory and
    (--tiles) tile variants as the bench's timed run -- twice untimed, as the
    bench's capture and timed sweeps precede its instrumented one (so every
    buffer's zero marks hold what that sweep finds), then once launched kernel by
    kernel with the kernels counting their full-path and short-path cells; per role (forward_step: coefficients + dt + update, forward and
    recompute launches; adjoint: the reverse steps) it logs the launches and
    the algorithmic bytes per launch of exactly these launches.  Then a 1 GiB
    device copy whose counters calibrate FETCH_SIZE / WRITE_SIZE against a
    known byte count."""
    import numpy as np
    import torch

    from discrete_adjoint_amd import problems
    from discrete_adjoint_amd.solver import Ensemble

    K, k = args.steps, args.checkpoint_every
    E = problems.ensemble_problems(args.members, seed=args.seed, nsteps=K, checkpoint_every=k)
    weights = [p.box_weights().reshape(3, -1) for p in E]
    h0 = tuple(np.concatenate([getattr(p, f) for p in E]) for f in ("h", "hu", "hv"))
    prec = prec_of(args.precision)
    tiles = [int(v) for v in args.tiles.split(",")] if args.tiles else None
    store = store_budget(args) if args.store else 0
    ab, sb = algo_bytes(prec), short_bytes(prec)
    with Ensemble(E, precision=prec, checkpoint_every=k, max_steps=max(K, 1), use_graph=0,
                  store_bytes=store) as ens:
        if tiles:  # reproduce a tile selection (the bench's autotuned one)
            ens.select_tiles(tiles)
        # sweeps first, as the bench's graph-capture and timed sweeps precede
        # its instrumented one: every buffer's zero marks (segment slots and
        # adjoint buffers included) hold what the bench's sweep finds there
        for _ in range(PROBE_WARM_SWEEPS):
            ens.set_functional(weights)
            ens.set_state(*h0)
            ens.run(K)
            ens.dual_solver(None, want_dphi=True, want_lambda=False)
        log(f"probe warm sweeps done: {PROBE_WARM_SWEEPS}")
        ens.set_functional(weights)
        ens.set_state(*h0)
        ens.instrument(True)
        ens.run(K)
        ens.dual_solver(None, want_dphi=True, want_lambda=False)
        ms, launches, full, short = ens.instrumentation()
        ens.instrument(False)
        cfg = ensemble_run_config(args, store, ens.tile_selection())
        log("probe config " + json.dumps(cfg))
        for role, i, name in (("forward_step", 1, "update"), ("adjoint", 2, "adjoint")):
            cells = launches[i] * ens.ncell
            fc, sc = path_cells(cells, full[i], short[i])
            per = (fc * ab[name] + sc * sb[name]) / max(launches[i], 1)
            log(f"probe {role}: {ms[i] / max(launches[i], 1):.4f} ms/launch over {launches[i]} "
                f"launches, full path {fc / max(cells, 1):.4f}, short path "
                f"{sc / max(cells, 1):.4f}, algorithmic bytes per launch {per:.0f}")
    n = 1 << 28  # 2^28 FP32 = 1 GiB read + 1 GiB written
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    log(f"probe copy: {4 * n} bytes read, {4 * n} bytes written")
    assert float(y[-1]) == 1.0


def traffic_summary(fetch_csv, write_csv, probe_log, kernel, out_path, commands):
    """HBM bytes per launch of one role from two rocprofv3 --pmc runs of the
    probe (FETCH_SIZE, WRITE_SIZE in KiB per dispatch) summed over every
    dispatch of the role's GPU kernels in the probe's run and divided by the
    role's launches, each counter scaled by what it reads for the probe's
    1 GiB copy; written to out_path with the probe's run configuration and
    the algorithmic bytes per launch of the same launches."""
    import csv

    role, symbols = kernel.split(":")
    symbols = symbols.split("+")  # the GPU kernels one launch of `role` consists of

    def per_dispatch(path, counter):
        rows = list(csv.DictReader(open(path)))
        name_key = "Kernel_Name" if "Kernel_Name" in rows[0] else "Kernel-Name"
        val_key = "Counter_Value" if "Counter_Value" in rows[0] else "Counter-Value"
        ker, copy = [], []
        for r in rows:
            if r.get("Counter_Name", r.get("Counter-Name")) != counter:
                continue
            nm = r[name_key]
            if any(f"{sym}<" in nm or f"{sym}(" in nm for sym in symbols):
                ker.append((nm, float(r[val_key])))
            elif "copy" in nm.lower() or "elementwise" in nm.lower():
                copy.append(float(r[val_key]))
        return ker, copy

    fk, fc = per_dispatch(fetch_csv, "FETCH_SIZE")
    wk, wc = per_dispatch(write_csv, "WRITE_SIZE")

    def last_share(ker, nsweeps):
        # the probe ran the sweep nsweeps times (untimed, then instrumented)
        # with the same dispatch sequence: keep each kernel's last 1/nsweeps
        by = {}
        for nm, v in ker:
            by.setdefault(nm, []).append((nm, v))
        return [x for lst in by.values() for x in lst[len(lst) - len(lst) // nsweeps:]]

    for line in open(probe_log):
        if line.startswith("probe warm sweeps done: "):
            n = int(line.split()[-1]) + 1
            fk, wk = last_share(fk, n), last_share(wk, n)
        elif line.startswith("probe warm sweep done"):  # one warm sweep
            fk, wk = last_share(fk, 2), last_share(wk, 2)
    launches = algo = cfg = None
    for line in open(probe_log):
        if line.startswith("probe config "):
            cfg = json.loads(line[len("probe config "):])
        if line.startswith(f"probe {role}:"):
            parts = line.split()
            launches = int(parts[parts.index("over") + 1])
            algo = float(parts[-1])
    copy_bytes = 2.0 ** 30
    fscale = copy_bytes / (max(fc) * 1024.0)
    wscale = copy_bytes / (max(wc) * 1024.0)
    f_kib = sum(v for _, v in fk) / launches
    w_kib = sum(v for _, v in wk) / launches
    hbm = f_kib * 1024.0 * fscale + w_kib * 1024.0 * wscale
    out = {"kernel": role, "dispatched_kernels": sorted({nm for nm, _ in fk}),
           "fetch_kib_per_launch": f_kib, "write_kib_per_launch": w_kib,
           "fetch_scale": fscale, "write_scale": wscale,
           "hbm_bytes_per_launch": hbm, "algorithmic_bytes_per_launch": algo,
           "traffic_over_algorithmic": hbm / algo if algo else None,
           "launches": launches, "dispatches": len(fk), "config": cfg,
           "command_fetch": commands[0], "command_write": commands[1]}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("ensemble", "grid", "adaptive"), default="ensemble")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--precision", choices=tuple(PRECISIONS), default="fp64")
    ap.add_argument("--compare", default="adj32,mixed,fp32",
                    help="further precisions run on the same ensemble (errors and speed)")
    ap.add_argument("--checkpoint-every", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="forward (+ adjoint) steps per member of the CPU baseline's sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--store", action="store_true",
                    help="measure with the last forward states kept whole in HBM (the "
                         "reverse sweep recomputes only below them) instead of checkpoints "
                         "only")
    ap.add_argument("--store-fraction", type=float, default=0.85,
                    help="share of free HBM for the stored trajectory")
    ap.add_argument("--product-only", action="store_true",
                    help="only the measured path (no FP64 / plain / CPU legs): for profiling")
    ap.add_argument("--probe", action="store_true",
                    help="short run for hardware-counter profiles (see probe())")
    ap.add_argument("--tiles", default="",
                    help="with --probe: tile variants coef,update,adjoint")
    ap.add_argument("--traffic-summary", default="",
                    help="fetch.csv,write.csv,probe.log,ROLE:SYM+SYM,out.json: HBM traffic "
                         "of one kernel role (coef, update, adjoint) made of the GPU kernels "
                         "SYM, from two --pmc runs of --probe")
    ap.add_argument("--no-adaptive", action="store_true",
                    help="skip the adaptive-mesh leg of the ensemble run")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernel by kernel instead of replaying captured sweeps")
    ap.add_argument("--sweeps", type=int, default=0,
                    help="R repetitions of the measured sweep only, for kernel traces")
    ap.add_argument("--counter-summary", default="",
                    help="a.csv+b.csv,SYMBOL,N,out.md: per-dispatch counter averages of the "
                         "last N dispatches of one GPU kernel from --pmc runs of --probe")
    ap.add_argument("--trace-summary", default="",
                    help="trace.csv,out.md: per-kernel table of the last forward run and "
                         "reverse sweep of a --sweeps kernel trace")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this hot path has no multi-GPU mode "
              "(ensembles would shard by member, one large block by SFC range; neither is "
              "built)", file=sys.stderr)
        sys.exit(2)
    if args.traffic_summary:
        fetch_csv, write_csv, probe_log, kernel, out_path = args.traffic_summary.split(",")
        probe_args = (f"--probe --steps {args.steps} --warmup {args.warmup} --tiles {args.tiles}"
                      + (" --store" if args.store else ""))
        traffic_summary(fetch_csv, write_csv, probe_log, kernel, out_path,
                        [f"cd /tmp && TMPDIR=/tmp timeout -s KILL 300 rocprofv3 --pmc {c} -d "
                         f"/tmp/pmc_{d} -o probe --output-format csv -- python3 "
                         f"$REPO/bench.py {probe_args}"
                         for c, d in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write"))])
        return
    if args.counter_summary:
        paths, symbol, last, out_path = args.counter_summary.split(",")
        counter_summary(paths.split("+"), symbol, int(last), out_path,
                        ["cd /tmp && TMPDIR=/tmp timeout -s KILL 300 rocprofv3 --pmc " + c +
                         " -d /tmp/pmc_" + d + " -o probe --output-format csv -- python3 "
                         "$REPO/bench.py --probe --tiles 3,3,3" for c, d in COUNTER_PASSES])
        return
    if args.trace_summary:
        trace_csv, out_path = args.trace_summary.split(",")
        trace_summary(trace_csv, out_path,
                      "cd /tmp && TMPDIR=/tmp rocprofv3 --kernel-trace -d /tmp/prof_sw -o sw "
                      "--output-format csv -- python3 $REPO/bench.py --sweeps 2 --no-graph"
                      + (f" --steps {args.steps}" if args.steps else ""))
        return
    if args.sweeps:
        args.steps = args.steps or 500
        args.checkpoint_every = args.checkpoint_every or 25
        sweeps(args)
        return
    if args.probe:
        args.steps = args.steps or 500
        args.checkpoint_every = args.checkpoint_every or 25
        probe(args)
        return
    if args.workload == "ensemble":
        args.steps = args.steps or 500
        args.checkpoint_every = args.checkpoint_every or 25
        out = bench_ensemble(args)
    elif args.workload == "adaptive":
        args.steps = args.steps or 200
        args.checkpoint_every = args.checkpoint_every or 25
        out = bench_adaptive(args)
    else:
        args.steps = args.steps or 1500
        args.checkpoint_every = args.checkpoint_every or 50
        out = bench_grid(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
