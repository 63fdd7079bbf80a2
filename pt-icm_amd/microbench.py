"""Microbenchmarks that pin the roofline peaks used by bench.py.

    python pt-icm_amd/microbench.py lds      # LDS read GB/s (16-byte reads)
    python pt-icm_amd/microbench.py copy     # device copy GB/s (read + write)
    python pt-icm_amd/microbench.py copy --mib 256 --reps 5   # for counter calibration
    python pt-icm_amd/microbench.py batched --steps 32        # 4096 3D EA instances, autotuned launch shape
    python pt-icm_amd/microbench.py batched --mode 1 --waves 1 --steps 32   # split kernels, one wave per replica
    python pt-icm_amd/microbench.py coop --launches 4   # G022 single instance, cooperative kernel, 500 steps/launch
    python pt-icm_amd/microbench.py er --instances 256 --steps 8   # ER +-1 G(5000, 0.004), split kernels
    python pt-icm_amd/microbench.py er --instances 256 --steps 8 --sliced -1   # ... per-replica sweep

The copy kernel (k_copy16) also calibrates the FETCH_SIZE / WRITE_SIZE
counters: a run of it under rocprofv3 --pmc reads back known byte counts.
"""
import argparse
import importlib.util
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def load_ptx():
    spec = importlib.util.spec_from_file_location("ptx", os.path.join(HERE, "ptx.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ptx"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["lds", "copy", "batched", "coop", "er"])
    ap.add_argument("--sliced", type=int, default=0,
                    help="er: the engine's sliced_sweep (0 automatic, -1 per-replica sweep, 1/2/4/8 waves)")
    ap.add_argument("--launches", type=int, default=4, help="coop: timed launches of 500 PT steps")
    ap.add_argument("--icm", action="store_true", help="coop: ascending ladder, so the cold slots run cluster moves")
    ap.add_argument("--short", type=int, default=0,
                    help="coop: afterwards two launches of this many PT steps (launch-length dependence of counters)")
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--warmup", type=int, default=8, help="batched: untimed PT steps first")
    ap.add_argument("--sweeps", type=int, default=1, help="coop: Metropolis sweeps per PT step")
    ap.add_argument("--coop-waves", type=int, default=8, help="coop: waves per workgroup (4, 8, 16)")
    ap.add_argument("--phases", action="store_true",
                    help="coop: afterwards 500 more PT steps with the phase profile on (workgroup 0's split of a step)")
    ap.add_argument("--waves", type=int, default=0, help="waves per workgroup (0 = engine default)")
    ap.add_argument("--mode", type=int, default=-1,
                    help="execution mode (1 split, 2 fused; -1 = the autotuner's choice)")
    a = ap.parse_args()
    ptx = load_ptx()
    if a.what == "batched":
        import time
        sys.path.insert(0, HERE)
        import instances
        _, coos = instances.ea_batch(a.instances, 0)
        probs = [ptx.Problem.from_coo(*c) for c in coos]
        lad = ptx.beta_linspace(0.2, 3.0, 32, True)
        if a.mode < 0:
            kw, _ = ptx.autotune(probs, [lad] * len(probs), seed=0, steps=32, flags=ptx.FLAG_ICM)
        else:
            kw = {"mode": a.mode, "waves_per_group": a.waves}
        eng = ptx.Engine(probs, [lad] * len(probs), seed=0, flags=ptx.FLAG_ICM, **kw)
        tw = time.perf_counter()
        for done in range(0, a.warmup, 250):  # progress on stderr (counter runs serialise every launch)
            eng.step(min(250, a.warmup - done))
            eng.sync()
            print(f"warm-up {done + min(250, a.warmup - done)} steps, {time.perf_counter() - tw:.1f} s",
                  file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        eng.step(a.steps)
        eng.sync()
        dt = time.perf_counter() - t0
        info = eng.info()
        print(json.dumps({"instances": len(probs), "steps": a.steps, "mode": info.mode, "waves": info.waves_per_group,
                          "ms_per_step": dt * 1e3 / a.steps,
                          "spin_updates_per_s": info.spin_updates_per_step * a.steps / dt,
                          "lds_bytes_per_step": info.lds_bytes_per_step,
                          "steps_per_launch": info.steps_per_launch}))
    elif a.what == "er":
        # ER +-1 G(5000, 0.004) instances (8 colour classes of ~625
        # positions), 64 inverse temperatures linear over T 0.5..3.0, 2
        # chains, cluster moves, split kernels: the sweep's time per step
        import time
        import numpy as np
        sys.path.insert(0, HERE)
        import instances
        n_inst = a.instances if a.instances != 4096 else 256
        probs = [ptx.Problem.from_coo(*instances.erdos_renyi(5000, 0.004, seed=i, pm=True)) for i in range(n_inst)]
        lad = np.sort(ptx.beta_linspace(0.5, 3.0, 64, False))
        eng = ptx.Engine(probs, [lad] * n_inst, seed=0, flags=ptx.FLAG_ICM, mode=ptx.MODE_SPLIT, waves_per_group=1,
                         icm_waves=1, sliced_sweep=a.sliced)
        eng.step(a.warmup)
        eng.sync()
        eng.timing(True)
        t0 = time.perf_counter()
        eng.step(a.steps)
        eng.sync()
        dt = time.perf_counter() - t0
        t = eng.timing()
        info = eng.info()
        sliced = sum(eng.instance_info(i).sliced_sweep for i in range(n_inst))
        eng.close()
        print(json.dumps({"instances": n_inst, "sliced_instances": sliced, "steps": a.steps,
                          "ms_per_step": dt * 1e3 / a.steps, "sweep_ms_per_step": t.sweep_ms / max(a.steps, 1),
                          "icm_ms_per_step": t.icm_ms / max(a.steps, 1),
                          "spin_updates_per_s": info.spin_updates_per_step * a.steps / dt}))
    elif a.what == "coop":
        # G022 with a fixed 22-slot ladder over the range the 64-slot
        # preprocess ladder spans (beta 1.667 .. 0.367): descending, as the
        # preprocess returns it, so the slot rule gives no cluster-move slot;
        # --icm hands it over ascending (the cold slots then run cluster moves)
        import numpy as np
        p = ptx.Problem.read(os.path.join(os.path.dirname(HERE), "tests", "golden", "graphs", "G022"))
        lad = np.linspace(1.0 / 0.6, 0.367, 22)
        if a.icm:
            lad = lad[::-1].copy()
        eng = ptx.Engine(p, lad, seed=0, flags=ptx.FLAG_ICM, mode=ptx.MODE_COOP, waves_per_group=a.coop_waves,
                         spins_per_lane=1, steps_per_launch=500, sweeps=a.sweeps)
        eng.step(500)
        eng.sync()
        eng.timing(True)
        eng.timing()
        eng.step(500 * a.launches)
        eng.sync()
        t = eng.timing()
        info = eng.info()
        nc = eng.instance_info().nc
        short = None
        if a.short > 0:
            eng.step(a.short)
            eng.step(a.short)
            ts = eng.timing()  # accumulated since timing was switched on
            k = ts.fused_launches - t.fused_launches
            short = {"steps_per_launch": a.short, "launches": k,
                     "avg_launch_ms": (ts.fused_ms - t.fused_ms) / max(k, 1)}
        phases = None
        if a.phases:
            eng.phase_profile(True)
            eng.step(500)
            eng.sync()
            us = eng.phase_profile()
            eng.phase_profile(False)
            phases = dict(zip(["sweep", "exchange", "icm", "publish_wait"], (us / 500).round(3).tolist()))
        eng.close()
        avg = t.fused_ms / max(t.fused_launches, 1)
        print(json.dumps({"workload": "G022, 22 slots x 2 chains, cooperative kernel, %d waves, %d sweeps per step"
                                      % (a.coop_waves, a.sweeps), "icm_slots": nc,
                          "launches": t.fused_launches, "steps_per_launch": 500, "avg_launch_ms": avg,
                          "us_per_step": avg * 1e3 / 500, "lds_bytes_per_step": info.lds_bytes_per_step,
                          "lds_gbs": info.lds_bytes_per_step * 500 / (avg * 1e-3) / 1e9, "short": short,
                          "phases_us_per_step": phases}),
              flush=True)
    elif a.what == "lds":
        gbs = max(ptx.lds_bandwidth_gbs() for _ in range(3))
        print(json.dumps({"lds_read_gbs": gbs}))
    else:
        nbytes = a.mib << 20
        ms = ptx.copy_ms(nbytes, a.reps)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms, "gbs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}))


if __name__ == "__main__":
    main()
