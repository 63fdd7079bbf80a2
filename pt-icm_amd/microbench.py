"""Microbenchmarks that pin the roofline peaks used by bench.py.

    python pt-icm_amd/microbench.py lds      # LDS read GB/s (16-byte reads)
    python pt-icm_amd/microbench.py copy     # device copy GB/s (read + write)
    python pt-icm_amd/microbench.py copy --mib 256 --reps 5   # for counter calibration
    python pt-icm_amd/microbench.py batched --steps 32        # 4096 3D EA instances, autotuned launch shape
    python pt-icm_amd/microbench.py batched --mode 1 --waves 1 --steps 32   # split kernels, one wave per replica

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
    ap.add_argument("what", choices=["lds", "copy", "batched"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--instances", type=int, default=4096)
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
        eng.step(8)
        eng.sync()
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
    elif a.what == "lds":
        gbs = max(ptx.lds_bandwidth_gbs() for _ in range(3))
        print(json.dumps({"lds_read_gbs": gbs}))
    else:
        nbytes = a.mib << 20
        ms = ptx.copy_ms(nbytes, a.reps)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms, "gbs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}))


if __name__ == "__main__":
    main()
