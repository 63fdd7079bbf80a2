"""Time to target energy: the batched engine with surrogate ladders against the
plain single-instance path, on one GPU.

Target energy E* of an instance: the lowest energy that the plain path
(multihistogram --preprocess ladder, ICM, src/solver.cc:67-127) finds in a
calibration of several seeds at 4x the measured run's step budget; stored in
data/ttt_targets.json with the workload that produced it.

Time to target of an instance = wall time from the start of its ladder
construction until any of its replicas has energy <= E*.  The plain path
reads the step from the per-replica best-energy sweep (its time per step is
constant); the batched path takes the wall time of the poll that first saw
E* (an upper bound within `poll` steps: retiring instances changes its time
per step).  The two paths:

  plain     per instance, one after another: 64-slot GPU pilot + FP64
            multihistogram ladder (Newton solve), then the single-instance
            engine (cooperative kernel), polled every `poll` steps until E*
            or the step budget;
  batched   all instances at once: 8-slot pilot + network ladders +
            validation run + multihistogram fallback, then one batched
            engine (autotuned launch shape) polled every `poll` steps;
            instances that reached E* are retired from the batch, and the run
            ends when 99 % have or the budget is spent.

instances/hour = instances finished (hit or budget spent) per wall hour.

Both paths hand the engine their ladders in ascending beta order.  The
reference's Nc rule (src/initialization.cc:158-164) counts ICM slots from the
last slot down while beta >= the characteristic scale, so on the descending
preprocess ladder it gives all slots or none (here none: the hottest beta is
below 1); ascending, ICM runs on the cold slots, which is what "ICM on"
means for this workload.

    python pt-icm_amd/ttt.py calibrate     # writes data/ttt_targets.json (GPU)
    python pt-icm_amd/ttt.py run           # one JSON line (GPU)
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import surrogate as S  # noqa: E402

TARGETS = os.path.join(HERE, "data", "ttt_targets.json")
T0, T1 = 0.5, 3.0
COPIES = 3  # independent runs of every instance in the batched paths


def ttt_specs(n_inst, seed=0, L=32, n_er=1000, p_er=0.02):
    """Half 2D +-J tori L x L, half Erdos-Renyi +-1 graphs G(n_er, p_er)."""
    out = []
    for i in range(n_inst):
        if i % 2 == 0:
            out.append({"family": "torus", "L": L, "pm": True, "seed": int(seed * 100003 + i)})
        else:
            out.append({"family": "er", "n": n_er, "p": p_er, "pm": True, "seed": int(seed * 100003 + i)})
    return out


def build_problems(ptx, specs):
    coos = [S.build(s) for s in specs]
    return [ptx.Problem.from_coo(*c) for c in coos], [S.problem_stats(c) for c in coos]


def first_hits(eng, targets):
    """Per instance: the first sweep at which a replica reached <= E* (-1: not reached)."""
    out = np.full(len(targets), -1, np.int64)
    for i, t in enumerate(targets):
        be, bs, _ = eng.best(i, with_state=False)
        hit = be <= t
        if hit.any():
            out[i] = int(bs[hit].min())
    return out


def calibrate(ptx, specs, seeds, steps, log=print):
    probs, _ = build_problems(ptx, specs)
    t0 = time.perf_counter()
    lads, _, _ = S.multihistogram_ladders(ptx, probs, S.N_OUT, T0, T1, 5000, 2000, seed=7)
    lads = [np.sort(l) for l in lads]
    log(f"multihistogram ladders in {time.perf_counter() - t0:.1f} s")
    best = np.full(len(probs), np.iinfo(np.int64).max)
    for sd in seeds:
        t0 = time.perf_counter()
        eng = ptx.Engine(probs, lads, seed=sd, flags=ptx.FLAG_ICM, use_graph=50)
        for done in range(0, steps, 5000):
            eng.step(min(5000, steps - done))
            eng.sync()
            log(f"seed {sd}: {done + min(5000, steps - done)} steps, {time.perf_counter() - t0:.1f} s")
        m = eng.min_energies()
        log(f"seed {sd}: {steps} steps in {time.perf_counter() - t0:.1f} s, improved {int(np.sum(m < best))}")
        best = np.minimum(best, m)
        eng.close()
    return best, [len(l) for l in lads]


def run_batched(ptx, specs, targets, max_steps, poll, seed, log=print, surrogate=True, copies=COPIES):
    """surrogate=False: every ladder from the batched multihistogram
    preprocess.  copies: independent runs of every instance in the batch
    (batch entry e runs instance e % n on its own random stream, the engine's
    streams being keyed by batch position); an instance is done when any of
    its copies reaches E*, and each copy has the max_steps budget."""
    probs, stats = build_problems(ptx, specs)
    targets = np.asarray(targets)
    n = len(specs)
    t_start = time.perf_counter()
    if surrogate:
        sur = S.Surrogate(ptx)
        lads, ok, _, stages = S.choose_ladders(ptx, sur, probs, stats, T0, T1, seed=seed)
        sur.close()
    else:
        ft = {}
        lads, _, _ = S.multihistogram_ladders(ptx, probs, S.N_OUT, T0, T1, 5000, 2000, seed=seed + 2, timing=ft)
        ok, stages = [False] * len(probs), {"pilot_s": ft["pilot_s"], "host_s": ft["host_s"]}
    lads = [np.sort(l) for l in lads]
    t_lad = time.perf_counter() - t_start
    log(f"batched ladders in {t_lad:.2f} s ({sum(ok)} of {len(ok)} surrogate ladders accepted) {stages}")
    # launch shape for this batch (a one-off per workload shape, reported and not counted)
    t_tune = time.perf_counter()
    kw, _ = ptx.autotune(probs * copies, lads * copies, seed=seed + 10, steps=32, flags=ptx.FLAG_ICM)
    t_tune = time.perf_counter() - t_tune
    eng = ptx.Engine(probs * copies, lads * copies, seed=seed + 10, flags=ptx.FLAG_ICM,
                     use_graph=(50 if kw["mode"] == 1 else 4), small_batch_coop=1, **kw)
    steps = 0
    t0 = time.perf_counter()
    # instances that reached E* are retired from the batch at every poll; the
    # run ends when 99 % have or the budget is spent
    done = np.zeros(n, bool)
    t_hit = np.full(n, np.inf)
    while steps < max_steps:
        k = min(poll, max_steps - steps)
        eng.step(k)
        steps += k
        m = eng.min_energies().reshape(copies, n).min(axis=0)
        now = time.perf_counter() - t0
        newly = (m <= targets) & ~done
        if newly.any():
            t_hit[newly] = now
            done |= newly
            if done.mean() >= 0.99:
                break
            eng.set_active(np.tile(~done, copies))
    t_run = time.perf_counter() - t0
    per_step = t_run / steps
    hit = (first_hits(eng, np.tile(targets, copies)) >= 0).reshape(copies, n).any(axis=0)
    info = eng.info()
    mean_icm = float(np.mean([eng.instance_info(i).nc for i in range(n)]))
    eng.close()
    # time to target: the poll that first saw E* (an upper bound within `poll` steps)
    ttt = np.where(hit, t_lad + t_hit, np.inf)
    wall = t_lad + t_run
    return {"instances": n, "copies": copies, "hit_rate": float(np.mean(hit)),
            "p50_ttt_s": float(np.median(ttt)),
            "wall_s": wall, "ladder_s": t_lad, "run_s": t_run, "steps": steps, "ms_per_step": per_step * 1e3,
            "instances_per_hour": n * 3600.0 / wall, "surrogate_accepted": int(sum(ok)),
            "ladder_stages_s": {k: round(v, 3) for k, v in stages.items()},
            "spin_updates_per_s": info.spin_updates_per_step * steps / t_run, "mode": info.mode,
            "waves_per_group": info.waves_per_group, "autotune_s": t_tune,
            "mean_slots": float(np.mean([len(l) for l in lads])), "mean_icm_slots": mean_icm}


def _rates(eng, entries):
    """Adjacent-pair swap acceptance of the batch entries (copies of one
    instance on one ladder), both chains and every copy pooled."""
    att = suc = 0.0
    for e in entries:
        nt = eng.instance_info(e).nt
        a, s, _ = eng.swap_stats(e)
        att = att + a[: nt - 1].astype(np.float64) + a[nt: 2 * nt - 1]
        suc = suc + s[: nt - 1].astype(np.float64) + s[nt: 2 * nt - 1]
    return suc / np.maximum(att, 1)


def warm_start(src, si, lad_src, dst, di, lad_dst, chains=2):
    """Sets the replicas of instance di of engine dst (ladder lad_dst, as
    handed to it) to the current states of instance si of engine src: each
    slot of a chain takes the state at the slot of the same chain in src whose
    beta is nearest.  dst's permutation is read, so it may have stepped."""
    sp = src.spins(si)
    s2r_src = src.slot2rep(si)
    s2r_dst = dst.slot2rep(di)
    n1, n2 = len(lad_src), len(lad_dst)
    near = np.abs(np.asarray(lad_src)[None, :] - np.asarray(lad_dst)[:, None]).argmin(axis=1)
    out = np.empty((chains * n2, sp.shape[1]), np.uint8)
    for c in range(chains):
        for t in range(n2):
            out[s2r_dst[c * n2 + t]] = sp[s2r_src[c * n1 + near[t]]]
    dst.set_spins(out, di)


def run_batched_judged_in_run(ptx, specs, targets, max_steps, poll, seed, log=print, validate_steps=2000,
                              start="surrogate", copies=COPIES):
    """Batched + surrogate with the judge inside the run: every instance starts
    the main run (ICM on) on its network ladder right after the 8-slot pilot,
    and the first validate_steps steps of that run are the validation run.
    start="linear" runs the same protocol from ladders linear in beta with
    the network's slot count (the network's spacing replaced by even steps),
    which separates what the network adds from what the protocol adds.
    Then an instance whose measured adjacent swap rates (its copies pooled)
    are not all within TARGET +- TOL, and that has not reached E* yet, falls
    back to the multihistogram preprocess ladder, computed on a worker thread
    while the run goes on; until that ladder exists the instance keeps
    running on its network ladder (a hit there is a state of energy <= E*,
    whatever the ladder), then its copies leave the run and go on from their
    current states on the fallback ladder (each slot from the slot of nearest
    beta) with what is left of their max_steps budget.  Instances that
    reached E* during the validation window are done whatever their rates.
    The main engine and the fallback engine are both stepped before either is
    read.  copies: independent runs of every instance, as in run_batched."""
    probs, stats = build_problems(ptx, specs)
    targets = np.asarray(targets)
    n = len(specs)
    excluded = [0.0]  # the one-off autotune is reported, not counted
    t_start = time.perf_counter()

    def now():
        return time.perf_counter() - t_start - excluded[0]

    sur = S.Surrogate(ptx)
    X, bs, be = S.short_features(ptx, probs, stats, T0, T1, 500, 1000, seed)
    lads = sur.ladders(X, bs, be)
    sur.close()
    if start == "linear":
        lads = [np.linspace(bs, be, len(l) + 1)[:-1] for l in lads]
    lads = [np.sort(l) for l in lads]
    t_lad = now()
    t0 = time.perf_counter()
    kw, _ = ptx.autotune(probs * copies, lads * copies, seed=seed + 10, steps=32, flags=ptx.FLAG_ICM)
    excluded[0] += time.perf_counter() - t0
    graph = 50 if kw["mode"] == 1 else 4
    eng = ptx.Engine(probs * copies, lads * copies, seed=seed + 10, flags=ptx.FLAG_ICM, use_graph=graph,
                     small_batch_coop=1, **kw)
    done = np.zeros(n, bool)
    t_hit = np.full(n, np.inf)
    timeline = [("ladders", round(t_lad, 3), 0), ("engine", round(now(), 3), 0)]

    def collect(e, idx, tag):
        """Reads engine e, whose entries are copies x idx (instances)."""
        m = e.min_energies().reshape(copies, len(idx)).min(axis=0)
        t = now()
        newly = (m <= targets[idx]) & ~done[idx]
        if newly.any():
            t_hit[idx[newly]] = t
            done[idx[newly]] = True
        timeline.append((tag, round(t, 3), int(done.sum())))
        return newly.any()

    all_idx = np.arange(n)
    steps1 = 0
    while steps1 < validate_steps and done.mean() < 0.99:
        k = min(poll, validate_steps - steps1)
        eng.step(k)
        steps1 += k
        if collect(eng, all_idx, "validate"):
            eng.set_active(np.tile(~done, copies))
    # the judge: rates of the instances still running
    ok = np.zeros(n, bool)
    for i in np.nonzero(~done)[0]:
        ok[i] = S.in_band(_rates(eng, [i + c * n for c in range(copies)]))
    fb = np.nonzero(~done & ~ok)[0]
    in_run = ~done & ok
    waiting = np.zeros(n, bool)  # fallback instances still on their network ladders
    waiting[fb] = True
    t_judge = now()
    timeline.append(("judge", round(t_judge, 3), int(done.sum())))
    res = {}

    def fallback():
        ft = {}
        mh, _, _ = S.multihistogram_ladders(ptx, [probs[i] for i in fb], S.N_OUT, T0, T1, 5000, 2000, seed + 2,
                                            timing=ft)
        res["lads"] = [np.sort(l) for l in mh]
        res["timing"] = ft
        res["t_ready"] = now()

    th = threading.Thread(target=fallback) if len(fb) else None
    if th:
        th.start()
    eng2, steps2, budget2, fb_steps1 = None, 0, 0, 0
    while done.mean() < 0.99:
        if th is not None and eng2 is None and not th.is_alive():
            th.join()
            # the fallback instances leave the main run with fb_steps1 steps spent
            waiting[:] = False
            fb_steps1 = steps1
            budget2 = max_steps - fb_steps1
            if steps1 < max_steps and (in_run & ~done).any():
                eng.set_active(np.tile(in_run & ~done, copies))
            nf = len(fb)
            eng2 = ptx.Engine([probs[i] for i in fb] * copies, res["lads"] * copies, seed=seed + 20,
                              flags=ptx.FLAG_ICM, use_graph=graph, small_batch_coop=1, **kw)
            if budget2 > 0:
                for c in range(copies):  # the fallback ladder takes over the copies' current states
                    for j, i in enumerate(fb):
                        if not done[i]:
                            warm_start(eng, int(i) + c * n, lads[i], eng2, j + c * nf, res["lads"][j])
            if done[fb].any():
                eng2.set_active(np.tile(~done[fb], copies))
            timeline.append(("fallback engine", round(now(), 3), int(done.sum())))
        act1 = (in_run | waiting) & ~done
        k1 = min(poll, max_steps - steps1) if steps1 < max_steps and act1.any() else 0
        k2 = min(poll, budget2 - steps2) if eng2 is not None and steps2 < budget2 and (~done[fb]).any() else 0
        if k1:
            eng.step(k1)
        if k2:
            eng2.step(k2)
        if k1:
            steps1 += k1
            if collect(eng, all_idx, "run") and steps1 < max_steps and ((in_run | waiting) & ~done).any():
                eng.set_active(np.tile((in_run | waiting) & ~done, copies))
        if k2:
            steps2 += k2
            if collect(eng2, fb, "fallback") and steps2 < budget2 and (~done[fb]).any():
                eng2.set_active(np.tile(~done[fb], copies))
        if not k1 and not k2:
            if th is not None and eng2 is None:
                th.join()  # nothing left to run until the fallback ladders exist
                continue
            break
    wall = now()
    ttt = t_hit
    # reported only, after the timed run: the judge on the ladders of the
    # instances that were done within the validation window
    judged_ok = ok.copy()
    done_in_validation = ~in_run
    done_in_validation[fb] = False
    for i in np.nonzero(done_in_validation)[0]:
        judged_ok[i] = S.in_band(_rates(eng, [i + c * n for c in range(copies)]))
    eng.close()
    if eng2 is not None:
        eng2.close()
    return {"instances": n, "copies": copies, "start_ladders": start, "hit_rate": float(done.mean()),
            "p50_ttt_s": float(np.median(ttt)),
            "wall_s": wall, "ladder_s": t_lad, "judge_at_s": t_judge, "validate_steps": validate_steps,
            "surrogate_accepted": int(judged_ok.sum()), "done_in_validation": int(done_in_validation.sum()),
            "fallback": int(len(fb)), "fallback_ready_s": res.get("t_ready"),
            "fallback_timing_s": {k: round(v, 3) for k, v in res.get("timing", {}).items()},
            "steps_accepted_run": steps1, "steps_fallback_run": steps2, "fallback_steps_on_first_ladder": fb_steps1,
            "instances_per_hour": n * 3600.0 / wall, "autotune_s": excluded[0], "mode": kw["mode"],
            "timeline": timeline}


def run_plain(ptx, specs, targets, max_steps, poll, seed, log=print):
    probs, _ = build_problems(ptx, specs)
    rows = []
    for i, (p, t) in enumerate(zip(probs, targets)):
        t0 = time.perf_counter()
        lad, _, _ = S.multihistogram_ladders(ptx, [p], S.N_OUT, T0, T1, 5000, 2000, seed=seed + i, threads=1)
        t_pre = time.perf_counter() - t0
        eng = ptx.Engine(p, np.sort(lad[0]), seed=seed + 1000 + i, flags=ptx.FLAG_ICM)
        steps = 0
        t1 = time.perf_counter()
        while steps < max_steps:
            k = min(poll, max_steps - steps)
            eng.step(k)
            steps += k
            if eng.min_energies()[0] <= t:
                break
        t_run = time.perf_counter() - t1
        h = first_hits(eng, [t])[0]
        rows.append({"pre_s": t_pre, "run_s": t_run, "steps": steps, "hit": bool(h >= 0),
                     "ttt_s": t_pre + (h + 1) * t_run / steps if h >= 0 else float("inf"),
                     "mode": eng.info().mode})
        eng.close()
    wall = sum(r["pre_s"] + r["run_s"] for r in rows)
    return {"instances": len(rows), "hit_rate": float(np.mean([r["hit"] for r in rows])),
            "p50_ttt_s": float(np.median([r["ttt_s"] for r in rows])), "wall_s": wall,
            "mean_preprocess_s": float(np.mean([r["pre_s"] for r in rows])),
            "instances_per_hour": len(rows) * 3600.0 / wall, "mode": rows[0]["mode"]}


def load_targets(path=TARGETS):
    with open(path) as f:
        return json.load(f)


ICM_SEMANTICS = ("ladders handed to the engine in ascending beta order; the reference's slot rule "
                 "(src/initialization.cc:158-164) then gives cluster moves to the cold slots with beta >= the "
                 "characteristic scale.  On the descending ladder of --preprocess that rule gives all slots or "
                 "none (none here), so 'ICM on' is realised by the ascending order, for every path alike.")


def run(ptx, n_plain=32, seed=2024, log=print, path=TARGETS, linear=True, judge_before=True, copies=COPIES):
    """Every path on the workload of the targets file.  linear=False skips the
    linear-start variant and the one-copy variant, judge_before=False the
    surrogate path that judges its ladders before the run (both informative
    only where some network ladders pass the judge).  copies: independent
    runs of every instance on the batched paths (1 where the batch fills the
    GPU from the first step and few instances reach E* within the budget: a
    second copy then only doubles the time per step)."""
    cfg = load_targets(path)
    specs, targets = cfg["specs"], np.array(cfg["targets"], np.int64)
    b = run_batched(ptx, specs, targets, cfg["max_steps"], cfg["poll"], seed, log,
                    copies=copies) if judge_before else None
    bj = run_batched_judged_in_run(ptx, specs, targets, cfg["max_steps"], cfg["poll"], seed, log, copies=copies)
    log(f"judged in the run: {bj}")
    bl = b1 = None
    if linear:
        bl = run_batched_judged_in_run(ptx, specs, targets, cfg["max_steps"], cfg["poll"], seed, log, start="linear",
                                       copies=copies)
        log(f"judged in the run from linear ladders: {bl}")
        if copies != 1:
            b1 = run_batched_judged_in_run(ptx, specs, targets, cfg["max_steps"], cfg["poll"], seed, log, copies=1)
            log(f"judged in the run, one copy: {b1}")
    bm = run_batched(ptx, specs, targets, cfg["max_steps"], cfg["poll"], seed, log, surrogate=False, copies=copies)
    # an even sample of both families (specs alternate torus, ER)
    stride = max(2, 2 * (len(specs) // max(n_plain, 1) // 2))
    idx = sorted(list(range(0, len(specs), stride))[: (n_plain + 1) // 2] +
                 list(range(1, len(specs), stride))[: n_plain // 2])
    pl = run_plain(ptx, [specs[i] for i in idx], targets[idx], cfg["max_steps"], cfg["poll"], seed + 7, log)
    out = {"workload": cfg["workload"], "max_steps": cfg["max_steps"], "calibration": cfg.get("calibration"),
           "icm_semantics": ICM_SEMANTICS, "batched_copies": copies,
           "batched_surrogate_judged_in_run": bj, "batched_multihistogram": bm,
           "plain": pl,
           "instances_per_hour_ratio_judged_in_run": bj["instances_per_hour"] / pl["instances_per_hour"],
           "instances_per_hour_ratio_multihistogram": bm["instances_per_hour"] / pl["instances_per_hour"]}
    if b is not None:
        out["batched_surrogate"] = b
        out["instances_per_hour_ratio"] = b["instances_per_hour"] / pl["instances_per_hour"]
    if bl is not None:
        out["batched_linear_judged_in_run"] = bl
        out["instances_per_hour_ratio_linear_judged_in_run"] = bl["instances_per_hour"] / pl["instances_per_hour"]
    if b1 is not None:
        out["batched_surrogate_judged_in_run_one_copy"] = b1
        out["instances_per_hour_ratio_judged_in_run_one_copy"] = b1["instances_per_hour"] / pl["instances_per_hour"]
    return out


def main():
    ap = argparse.ArgumentParser(description="time to target energy, batched + surrogate vs plain")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("calibrate")
    c.add_argument("--instances", type=int, default=512)
    c.add_argument("--L", type=int, default=32)
    c.add_argument("--n-er", type=int, default=1000)
    c.add_argument("--p-er", type=float, default=0.02)
    c.add_argument("--max-steps", type=int, default=10000)
    c.add_argument("--poll", type=int, default=250)
    c.add_argument("--seeds", type=int, default=4)
    c.add_argument("--out", default=TARGETS)
    r = sub.add_parser("run")
    r.add_argument("--plain", type=int, default=32)
    r.add_argument("--targets", default=TARGETS)
    r.add_argument("--copies", type=int, default=COPIES)
    a = ap.parse_args()
    ptx = S.load_ptx()
    if a.cmd == "calibrate":
        specs = ttt_specs(a.instances, 0, a.L, a.n_er, a.p_er)
        best, slots = calibrate(ptx, specs, [101 + k for k in range(a.seeds)], 4 * a.max_steps,
                                log=lambda *s: print(*s, flush=True))
        out = {"workload": f"{a.instances // 2} 2D +-J tori L={a.L} + {a.instances - a.instances // 2} ER +-1 "
                           f"G({a.n_er}, {a.p_er}); T {T0}..{T1}; 2 chains; ICM",
               "specs": specs, "targets": [int(x) for x in best], "calibration_slots": slots,
               "calibration": f"multihistogram ladders, {a.seeds} seeds x {4 * a.max_steps} steps",
               "max_steps": a.max_steps, "poll": a.poll}
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with open(a.out, "w") as f:
            json.dump(out, f)
        print(json.dumps({k: v for k, v in out.items() if k not in ("specs", "targets", "calibration_slots")}))
    else:
        print(json.dumps(run(ptx, a.plain, log=lambda *s: print(*s, file=sys.stderr, flush=True),
                             path=a.targets, copies=a.copies)))


if __name__ == "__main__":
    main()
