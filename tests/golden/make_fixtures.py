"""Writes small golden fixtures from data the reference checkout already holds.

Run in a container that has the reference at /root/reference:
    python tests/golden/make_fixtures.py
Outputs (committed): tests/golden/g022_published.json -- the G022 run recorded in
the comment block of scripts/grab_results2csv.py:11-33 (parameters, best state,
its energy and cut, swap counters, the printed beta ladder).
"""
import json
import os
import re
import sys

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = os.path.join(REF, "scripts", "grab_results2csv.py")
    if not os.path.exists(src):
        sys.exit("reference checkout not found")
    lines = [l[2:].rstrip("\n") for l in open(src) if l.startswith("# ")]
    out = {}
    for l in lines:
        if ":" in l:
            k, v = l.split(":", 1)
            out[k.strip()] = v.strip()
    fx = {
        "source": "scripts/grab_results2csv.py:11-33",
        "seed": int(out["Seed"]), "replicas": int(out["Replicas"]), "sweeps": int(out["Sweeps"]),
        "pts": int(out["PT Steps"]), "T0": float(out["T0"]), "T1": float(out["T1"]),
        "inverse": int(out["Inverse"]), "preprocess": int(out["Preprocess"]), "icm": int(out["ICM"]),
        "best_state_energy": int(out["best_state_energy"]), "best_energy": int(out["best_energy"]),
        "best_maxcut": int(out["best_maxcut"]), "num_of_flips": int(out["num_of_flips"]),
        "num_of_swaps": int(out["num_of_swaps"]),
        "swap_attempts": json.loads(out["swap_attempts"]), "swap_successes": json.loads(out["swap_successes"]),
        "swap_rates": json.loads(out["swap_rates"]),
        "best_state": json.loads(out["best_state"]),
    }
    betas = [l for l in lines if l.startswith("[0.5")]
    fx["betas"] = json.loads(betas[0])
    fx["running_time_ms"] = int(re.sub("[^0-9]", "", out["running time"]))
    with open(os.path.join(HERE, "g022_published.json"), "w") as f:
        json.dump(fx, f)
    print("wrote g022_published.json", len(fx["best_state"]), "spins")


if __name__ == "__main__":
    main()
