"""Benchmark of the PT-ICM hot path on one MI355X.

Default workload (the GPU configuration of the metric): plain PT-ICM on G022
(Erdos-Renyi n=2000, M=19,990, unit weights) with the 64-slot ladder from the
FP64 multihistogram preprocess (GPU pilot: 5000 burn-in + 2000 logged steps),
2 chains, ICM by the reference's slot rule.  A "step" is one PT step: one
Metropolis sweep of all 128 replicas, 2(Nt-1) exchange attempts, ICM.

Prints ONE JSON line (rank 0).  value = attempted spin updates per second over
the timed steps (inputs resident in HBM); roofline = the sweep kernel's
algorithmic LDS bytes (4 d_i + 6 per attempted update) per launch over its
average launch time, measured with HIP events on the engine stream;
cpu_baseline = the oracle in reference order (index-order sweep, maintained
dE, draws only for uphill moves), one run per host core at once, over a
bounded sample (about --cpu-seconds of work per core).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
GRAPHS = os.path.join(REPO, "tests", "golden", "graphs")
LDS_PEAK_GBS = 256 * 128 * 2.4  # gfx950: 256 CUs x 128 B/clk x 2.4 GHz (GB/s)
# HBM bytes per launch of the batched workload's dominant kernel, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of `microbench.py batched`,
# scaled by the k_copy16 calibration (FETCH_SIZE x 2, WRITE_SIZE x 1):
# k_sweep of the split kernels (one wave per replica; the multispin lattice
# sweep, after 2000 warm-up steps), and k_fused with 16 waves and 16 steps
# per launch.
BATCHED_HBM_BYTES = {"k_sweep": (2 * 40010.0 + 33392.0) * 1024,
                     "k_fused": (2 * 69305.5 + 475067.6) * 1024}
# HBM bytes of one k_coop launch on G022 (22 slots x 2 chains, 8 waves) as a
# function of the PT steps it runs: a per-launch part (staging the adjacency,
# thresholds and states into each workgroup's LDS) plus a per-step part (the
# tagged energy slots, written and polled at agent scope, which crosses the
# XCDs' L2s).  From rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# `microbench.py coop --launches 2 --short 20` (launches of 500 and of 20
# steps), scaled by the k_copy16 calibration (FETCH_SIZE x 2, WRITE_SIZE x 1):
# 500 steps 4,793 KiB, 20 steps 1,585 KiB.
COOP_HBM_KIB_PER_LAUNCH = 1451.0
COOP_HBM_KIB_PER_STEP = 6.68


def load_ptx():
    spec = importlib.util.spec_from_file_location("ptx", os.path.join(REPO, "pt-icm_amd", "ptx.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ptx"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def preprocess_ladder(ptx, problem, nt, T0, T1, seed, burn_in=5000, samples=2000):
    """--preprocess (src/initialization.cc:109-111 -> get_beta): GPU pilot run
    of one chain with per-slot energy histograms, then the FP64 DOS and
    ladder search on the host."""
    hb, bs, be = ptx.pilot_betas(T0, T1, nt)
    pilot = ptx.Engine(problem, hb, seed=seed, chains=1, flags=ptx.FLAG_PILOT | ptx.FLAG_HISTOGRAM,
                       hist_start=burn_in)
    pilot.step(burn_in + samples)
    counts, e_base = pilot.histograms()
    pilot.close()
    return ptx.ladder_from_histograms(counts, e_base, hb, bs, be, mode=ptx.LADDER_FAST)


def cpu_baseline(problem_path, betas, seconds, seed=0, cores=None):
    """The oracle in reference order (index-order sweep, maintained dE, draws
    only for uphill proposals) on the host's cores: one independent run (own
    seed) per core, each sized to about `seconds` of work; value = the summed
    attempted spin updates per second.  The oracle's C code runs outside the
    GIL and keeps no global state, so one thread per core is one run per core."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O
    from concurrent.futures import ThreadPoolExecutor
    if cores is None:
        cores = min(len(os.sched_getaffinity(0)), 16)
    g = O.Graph.read(problem_path)
    probe = O.Engine(g, betas, chains=2, seed=seed, flags=O.FLAG_ICM | O.FLAG_REFERENCE_ORDER)
    t0 = time.perf_counter()
    probe.step(2)
    steps = max(2, int(seconds / max((time.perf_counter() - t0) / 2, 1e-6)))

    def one(k):
        eng = O.Engine(g, betas, chains=2, seed=seed + k, flags=O.FLAG_ICM | O.FLAG_REFERENCE_ORDER)
        t = time.perf_counter()
        eng.step(steps)
        return time.perf_counter() - t

    t0 = time.perf_counter()
    single = one(0)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        times = list(ex.map(one, range(1, cores + 1)))
    wall = time.perf_counter() - t0
    per_run = 2 * len(betas) * g.n * steps
    return {"value": cores * per_run / wall, "unit": "spin updates/s", "cores": cores, "kind": "port",
            "value_1core": per_run / single,
            "sample": f"oracle reference-order PT-ICM on G022, same ladder ({len(betas)} slots x 2 chains): "
                      f"{cores} independent runs of {steps} PT steps, one per host core, in {wall:.1f} s "
                      f"(slowest run {max(times):.1f} s); one run alone {single:.1f} s"}


def batched_mixed_at_scale(ptx, steps, seed, n_inst=1024, warmup=100, ladders="surrogate", log=print):
    """Throughput of the batched engine at the scale of the batched + surrogate
    configuration: half 2D +-J tori L = 100 (10,000 spins), half ER +-1 graphs
    G(5000, 0.004), T 0.5..3.0, 2 chains, ICM.  ladders="surrogate": the
    configuration's ladders -- each instance's network ladder from its 8-slot
    pilot, judged by a 2,000-step validation run (every adjacent swap rate
    within 0.4 +- 0.05) and replaced by the 64-slot multihistogram ladder where
    it fails; "linear": 64 inverse temperatures linear in beta.  Ladders go to
    the engine ascending (the slot rule then gives the cold slots ICM).  The
    launch shape comes from the autotuner (8 steps per trial)."""
    sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
    import instances
    from concurrent.futures import ThreadPoolExecutor

    def make(i):
        if i % 2 == 0:
            c = instances.torus_2d(100, seed=seed * 100003 + i)
        else:
            c = instances.erdos_renyi(5000, 0.004, seed=seed * 100003 + i, pm=True)
        return c, ptx.Problem.from_coo(*c)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(8) as ex:
        made = list(ex.map(make, range(n_inst)))
    probs = [m[1] for m in made]
    t_build = time.perf_counter() - t0
    lad_info = {"ladders": ladders}
    if ladders == "surrogate":
        import surrogate as S
        stats = [S.problem_stats(m[0]) for m in made]
        sur = S.Surrogate(ptx)
        t0 = time.perf_counter()
        lads, ok, _, stages = S.choose_ladders(ptx, sur, probs, stats, 0.5, 3.0, seed=seed)
        sur.close()
        lads = [np.sort(l) for l in lads]
        lad_info.update({"ladder_s": round(time.perf_counter() - t0, 2), "surrogate_accepted": int(sum(ok)),
                         "fallback": int(len(ok) - sum(ok)),
                         "ladder_stages_s": {k: round(v, 2) for k, v in stages.items()},
                         "mean_slots": float(np.mean([len(l) for l in lads]))})
        log(f"at-scale ladders {lad_info}")
    else:
        lads = [np.sort(ptx.beta_linspace(0.5, 3.0, 64, False))] * n_inst
    del made
    t0 = time.perf_counter()
    kw, trials = ptx.autotune(probs, lads, seed=seed, steps=8, flags=ptx.FLAG_ICM)
    t_tune = time.perf_counter() - t0
    eng = ptx.Engine(probs, lads, seed=seed, flags=ptx.FLAG_ICM, use_graph=4, **kw)
    eng.step(warmup)  # past the first steps from random states, whose cluster moves span half the lattice
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    dt = time.perf_counter() - t0
    info = eng.info()
    icm = {"torus": float(np.mean([eng.instance_info(i).nc for i in range(0, n_inst, 2)])),
           "er": float(np.mean([eng.instance_info(i).nc for i in range(1, n_inst, 2)]))}
    eng.close()
    for q in probs:
        q.close()
    lad_text = ("surrogate ladders (8-slot pilot, network, 2,000-step judge, 64-slot multihistogram fallback)"
                if ladders == "surrogate" else "64 betas linear")
    return {"workload": f"{n_inst // 2} 2D +-J tori L=100 + {n_inst - n_inst // 2} ER +-1 G(5000, 0.004); "
                        f"{lad_text} over T 0.5..3.0, 2 chains, ICM",
            "value": info.spin_updates_per_step * steps / dt, "unit": "spin updates/s",
            "ms_per_step": dt * 1e3 / steps, "steps": steps, "warmup": warmup, "replicas": info.total_replicas,
            "mean_icm_slots": icm, **lad_info,
            "launch_shape": {"mode": info.mode, "waves_per_group": info.waves_per_group,
                             "icm_waves": info.icm_waves},
            "build_s": round(t_build, 2), "autotune_s": round(t_tune, 2),
            "autotune_trials": [{k: (round(v, 1) if isinstance(v, float) else v) for k, v in t.items()}
                                for t in trials]}


def batched_ea(ptx, steps, seed, n_inst=4096):
    """The batched workload: 4096 independent 3D Edwards-Anderson +-J instances,
    L uniform in {6, 8, 10, 12}, 32 inverse temperatures 0.2..3.0 x 2 chains,
    ICM by the slot rule.  The launch shape comes from the autotuner (timed
    separately).  The timed run launches eagerly (per-kernel HIP events for the
    roofline); the same steps are then replayed from hipGraphs and timed by wall
    clock."""
    sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
    import instances
    _, coos = instances.ea_batch(n_inst, seed)
    probs = [ptx.Problem.from_coo(*c) for c in coos]
    lad = ptx.beta_linspace(0.2, 3.0, 32, True)
    t0 = time.perf_counter()
    kw, trials = ptx.autotune(probs, [lad] * n_inst, seed=seed, steps=32, flags=ptx.FLAG_ICM)
    t_tune = time.perf_counter() - t0
    eng = ptx.Engine(probs, [lad] * n_inst, seed=seed, flags=ptx.FLAG_ICM, use_graph=4, **kw)
    eng.step(16)
    eng.sync()
    eng.timing(True)
    eng.timing()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    dt = time.perf_counter() - t0
    tm = eng.timing()
    # the same number of steps replayed from hipGraphs (per-kernel events off)
    eng.timing(False)
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    dt_graph = time.perf_counter() - t0
    info = eng.info()
    if info.mode == 2:  # the timed launches cover `steps` PT steps between them
        kern, avg_ms = "k_fused", tm.fused_ms / max(tm.fused_launches, 1)
        steps_per_launch = steps / max(tm.fused_launches, 1)
    else:
        kern, avg_ms = "k_sweep", tm.sweep_ms / max(tm.sweep_launches, 1)
        steps_per_launch = steps / max(tm.sweep_launches, 1)
    bpl = info.lds_bytes_per_step * steps_per_launch
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    eng.close()
    for q in probs:
        q.close()
    return {"workload": "4096 x 3D EA +-J, L in {6,8,10,12}, 32 betas 0.2..3.0, 2 chains, ICM",
            "value": info.spin_updates_per_step * steps / dt, "unit": "spin updates/s",
            "ms_per_step": dt * 1e3 / steps, "steps": steps, "instances_per_hour_at_10k_steps":
                n_inst * 3600.0 / (dt / steps * 10000),
            "launch_shape": {"mode": info.mode, "waves_per_group": info.waves_per_group,
                             "csr_in_lds": info.csr_in_lds, "steps_per_launch": info.steps_per_launch},
            "ms_per_step_graph_replay": dt_graph * 1e3 / steps,
            "graph": "4 launches (fused) or 4 PT steps (split) per hipGraph replay",
            "autotune_s": t_tune, "autotune_trials": [{k: (round(v, 1) if k == "us_per_step" else v)
                                                       for k, v in t.items()} for t in trials],
            "roofline": {"bound": "lds", "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / LDS_PEAK_GBS,
                         "traffic": BATCHED_HBM_BYTES["k_sweep"] if info.mode == 1 and info.waves_per_group == 1
                         else (BATCHED_HBM_BYTES["k_fused"] if info.mode == 2 and info.waves_per_group == 16
                               and info.steps_per_launch == 16 else None), "kernel": kern, "avg_launch_ms": avg_ms,
                         "bytes_per_launch": bpl, "steps_per_timed_launch": steps_per_launch}}


def ttt_summary(r):
    """The time-to-target result without per-stage detail."""
    keep = ("instances", "copies", "start_ladders", "hit_rate", "p50_ttt_s", "instances_per_hour", "surrogate_accepted",
            "done_in_validation", "fallback", "ladder_s", "run_s", "judge_at_s", "fallback_ready_s", "wall_s",
            "ms_per_step", "mean_slots", "mean_icm_slots", "mean_preprocess_s", "mode")
    out = {k: r[k] for k in ("workload", "max_steps", "calibration", "icm_semantics", "batched_copies") if k in r}
    for k, v in r.items():
        if isinstance(v, dict):
            out[k] = {kk: vv for kk, vv in v.items() if kk in keep}
        elif k.startswith("instances_per_hour_ratio"):
            out[k] = v
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: seconds of oracle work per host core (all cores run at once)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", type=int, default=0,
                    help="0 auto (cooperative per-replica kernel for one instance), 1 split, 2 fused, 3 cooperative")
    ap.add_argument("--scale-ladders", choices=["surrogate", "linear"], default="surrogate",
                    help="ladders of the at-scale line: the configuration's surrogate pipeline, or linear in beta")
    ap.add_argument("--scale-steps", type=int, default=16,
                    help="timed PT steps of the 1024-instance mixed batch (0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=64,
                    help="PT steps of the batched 3D EA workload reported beside the main line (0 = skip)")
    ap.add_argument("--ttt-c5-targets",
                    default=os.path.join(REPO, "pt-icm_amd", "data", "ttt_targets_c5.json"),
                    help="second time-to-target workload at the batched + surrogate configuration's instance "
                         "sizes ('' = skip)")
    ap.add_argument("--ttt-plain", type=int, default=32,
                    help="instances of the plain path timed for time-to-target (0 = skip the time-to-target part)")
    ap.add_argument("--ttt-targets", default=os.path.join(REPO, "pt-icm_amd", "data", "ttt_targets.json"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design "
                                   "(instances or temperature slots per GPU) and is not built", "n_gpus": args.gpus}))
        sys.exit(2)
    ptx = load_ptx()
    path = os.path.join(GRAPHS, "G022")
    p = ptx.Problem.read(path)
    t0 = time.perf_counter()
    betas = preprocess_ladder(ptx, p, 64, 0.6, 2.8, args.seed)
    t_pre = time.perf_counter() - t0
    nt = len(betas)
    log(f"preprocess: {nt} temperatures in {t_pre:.2f} s; beta {betas[0]:.4f} .. {betas[-1]:.4f}")
    tune = None
    if args.mode == 0:  # launch shape from the autotuner (the trajectories do not depend on it)
        kw, trials = ptx.autotune(p, betas, seed=args.seed, steps=500, flags=ptx.FLAG_ICM, steps_per_launch=500)
        tune = {"chosen": kw, "trials": [{k: (round(v, 2) if isinstance(v, float) else v) for k, v in t.items()}
                                         for t in trials]}
    else:
        kw = {"mode": args.mode}
    eng = ptx.Engine(p, betas, seed=args.seed, flags=ptx.FLAG_ICM, steps_per_launch=500, **kw)
    nc = eng.instance_info().nc
    eng.step(args.warmup)
    eng.sync()
    eng.timing(True)
    eng.timing()  # reset accumulated counts
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    dt = time.perf_counter() - t0
    tm = eng.timing()
    info = eng.info()
    updates = info.spin_updates_per_step * args.steps
    value = updates / dt
    # algorithmic bytes per launch: the timed launches cover args.steps PT
    # steps between them (a launch runs at most steps_per_launch of them)
    if info.mode in (2, 3):
        kern = "k_fused" if info.mode == 2 else "k_coop"
        launches = max(tm.fused_launches, 1)
        avg_ms = tm.fused_ms / launches
    else:
        kern, launches = "k_sweep", max(tm.sweep_launches, 1)
        avg_ms = tm.sweep_ms / launches
    steps_per_timed_launch = args.steps / launches
    bytes_per_launch = info.lds_bytes_per_step * steps_per_timed_launch
    # measured HBM bytes per launch (the counter runs' shape: G022, 22 slots, k_coop with 8 waves)
    traffic = None
    if kern == "k_coop" and nt == 22 and info.waves_per_group == 8:
        traffic = (COOP_HBM_KIB_PER_LAUNCH + COOP_HBM_KIB_PER_STEP * steps_per_timed_launch) * 1024
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    lds_peak = max(ptx.lds_bandwidth_gbs() for _ in range(3))  # pinned by the microbenchmark
    be, bs, _ = eng.best(0, with_state=False)
    best = int(be.min())
    out = {
        "metric": "spin updates/s",
        "value": value,
        "unit": "spin updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "G022 from the reference's example_graphs (ER n=2000, M=19990, w=+1); random initial states",
        "config": {"workload": "plain PT-ICM single instance: G022, multihistogram ladder (64 requested, "
                               f"{nt} chosen), 2 chains, ICM slots {nc}, 1 sweep/step",
                   "n": p.n, "replicas": 2 * nt, "icm_slots": nc, "preprocess_s": round(t_pre, 3),
                   "best_energy": best, "best_cut": (19990 - best) // 2,
                   "launch_shape": {"mode": info.mode, "waves_per_group": info.waves_per_group,
                                    "steps_per_launch": info.steps_per_launch}, "autotune": tune},
        "kernel_ms": {"sweep": tm.sweep_ms / max(tm.sweep_launches, 1),
                      "exchange": tm.exchange_ms / max(tm.exchange_launches, 1),
                      "icm": tm.icm_ms / max(tm.icm_launches, 1), "fused": tm.fused_ms / max(tm.fused_launches, 1)},
        "roofline": {"bound": "lds", "achieved": achieved, "peak": lds_peak, "unit": "GB/s",
                     "frac": achieved / lds_peak, "traffic": traffic, "peak_nominal": LDS_PEAK_GBS,
                     "peak_source": "k_lds_read microbenchmark (16-byte LDS reads, all CUs)",
                     "kernel": kern, "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms,
                     "timed_launches": launches, "steps_per_timed_launch": steps_per_timed_launch},
    }
    eng.close()
    if args.batched_steps > 0:
        out["batched"] = batched_ea(ptx, args.batched_steps, args.seed)
        out["batched"]["roofline"]["peak"] = lds_peak
        out["batched"]["roofline"]["frac"] = out["batched"]["roofline"]["achieved"] / lds_peak
    if args.scale_steps > 0:
        out["batched_at_scale"] = batched_mixed_at_scale(ptx, args.scale_steps, args.seed, ladders=args.scale_ladders,
                                                         log=log)
    if args.ttt_plain > 0:
        sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
        import ttt
        for key, tpath, full in (("time_to_target", args.ttt_targets, True),
                                 ("time_to_target_c5_sizes", args.ttt_c5_targets, False)):
            if not tpath or not os.path.exists(tpath):
                continue
            r = ttt.run(ptx, n_plain=args.ttt_plain, log=log, path=tpath, linear=full, judge_before=full,
                        copies=ttt.COPIES if full else 1)
            out[key] = ttt_summary(r)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(path, betas, args.cpu_seconds, args.seed)
    p.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
