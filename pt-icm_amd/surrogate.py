"""Learned temperature ladders: a short 8-slot pilot -> features -> the MLP
(BF16 MFMA kernels of libptx) -> a 64-increment inverse-temperature ladder,
with the FP64 multihistogram preprocess as fallback and judge.

The reference computes its ladder with get_beta (src/cae_estimator.cc:
295-428): a 64-slot pilot run (gather_samples, :76-151), a multihistogram
DOS and a search that steps 1/beta until the predicted adjacent swap rate
drops to 0.4.  Here that ladder is the training label; the network sees only
what an 8-slot pilot of ~1/30 the cost shows, and its ladder is accepted
only if a short validation run measures every adjacent swap rate within
TARGET +- TOL; otherwise the instance falls back to the multihistogram
ladder.

Ladders are descending in beta, start at beta_start = max beta and stop
before reaching beta_end (the reference's stopping rule, :413-418).
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

FEAT_SLOTS, FEAT_BINS, N_OUT = 8, 64, 64
N_FEAT = FEAT_SLOTS * FEAT_BINS + 4
DIMS = [N_FEAT, 256, 256, 256, N_OUT]
INC_DIV = 32.0          # increments are predicted in units of (beta_start - beta_end) / 32
TARGET, TOL = 0.4, 0.05
WEIGHTS = os.path.join(HERE, "data", "ladder_mlp.npz")


def load_ptx():
    if "ptx" in sys.modules:
        return sys.modules["ptx"]
    import importlib.util
    spec = importlib.util.spec_from_file_location("ptx", os.path.join(HERE, "ptx.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ptx"] = mod
    spec.loader.exec_module(mod)
    return mod


# --------------------------------------------------------------------------
# instances
# --------------------------------------------------------------------------

def sample_specs(n, seed):
    """C4-style instance specs: half Erdos-Renyi n in [800, 3000], p log-uniform
    in [0.002, 0.02]; half 2D tori L in [20, 100]; weights +1 or +-1."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        pm = bool(rng.integers(0, 2))
        if i % 2 == 0:
            out.append({"family": "er", "n": int(rng.integers(800, 3001)),
                        "p": float(np.exp(rng.uniform(np.log(0.002), np.log(0.02)))), "pm": pm,
                        "seed": int(seed * 1000003 + i)})
        else:
            out.append({"family": "torus", "L": int(rng.integers(20, 101)), "pm": pm,
                        "seed": int(seed * 1000003 + i)})
    return out


def build(spec):
    sys.path.insert(0, HERE)
    import instances
    if spec["family"] == "er":
        return instances.erdos_renyi(spec["n"], spec["p"], seed=spec["seed"], pm=spec["pm"])
    if spec["family"] == "torus":
        return instances.torus_2d(spec["L"], seed=spec["seed"], pm=spec["pm"])
    if spec["family"] == "ea3d":
        return instances.ea_3d(spec["L"], seed=spec["seed"])
    raise ValueError(spec["family"])


def problem_stats(coo):
    n, src, dst, w = coo
    return {"n": int(n), "mean_degree": 2.0 * len(src) / n, "pm": float(np.any(np.asarray(w) < 0))}


# --------------------------------------------------------------------------
# pilots
# --------------------------------------------------------------------------

def pilot(ptx, problems, nt, T0, T1, burn_in, samples, seed, inverse=False):
    """One-chain pilot with per-slot energy histograms (gather_samples,
    src/cae_estimator.cc:76-151) over a batch; returns the engine and ladder."""
    hb, bs, be = ptx.pilot_betas(T0, T1, nt, inverse)
    eng = ptx.Engine(problems, [hb] * len(problems), seed=seed, chains=1,
                     flags=ptx.FLAG_PILOT | ptx.FLAG_HISTOGRAM, hist_start=burn_in)
    eng.step(burn_in + samples)
    eng.sync()
    return eng, hb, bs, be


def features(counts, e_base, stats):
    """8 normalised 64-bin histograms over the pilot's visited energy window,
    then (log n, mean degree, weight type, log window span)."""
    tot = counts.sum(axis=0)
    nz = np.nonzero(tot)[0]
    lo, hi = int(nz[0]), int(nz[-1])
    span = 2 * (hi - lo)
    b = np.arange(lo, hi + 1)
    idx = ((2 * (b - lo)) * FEAT_BINS) // (span + 2)
    f = np.zeros((FEAT_SLOTS, FEAT_BINS))
    for k in range(FEAT_SLOTS):
        np.add.at(f[k], idx, counts[k, lo:hi + 1].astype(np.float64))
        s = f[k].sum()
        if s > 0:
            f[k] *= 8.0 / s
    extra = [np.log(stats["n"]) / 10.0, stats["mean_degree"] / 32.0, stats["pm"], np.log(span + 2.0) / 10.0]
    return np.concatenate([f.ravel(), extra]).astype(np.float32)


def labels(ladder, bs, be):
    """Targets and modes of the 64 outputs for a multihistogram ladder."""
    s = (bs - be) / INC_DIV
    t = np.zeros(N_OUT, np.float32)
    m = np.zeros(N_OUT, np.uint8)
    k = len(ladder)
    inc = -np.diff(ladder) / s
    t[: k - 1] = inc[:N_OUT]
    m[: k - 1] = 1
    if k < N_OUT:  # the next step must cross beta_end (the reference stops there)
        t[k - 1] = (ladder[-1] - be) / s
        m[k - 1] = 2
    return t, m


def ladder_from_increments(y, bs, be, max_slots=N_OUT):
    s = (bs - be) / INC_DIV
    lad = [bs]
    for k in range(max_slots - 1):
        b = lad[-1] - float(y[k]) * s
        if b <= be:
            break
        lad.append(b)
    return np.array(lad)


def multihistogram_ladders(ptx, problems, nt, T0, T1, burn_in, samples, seed, inverse=False, threads=16,
                           timing=None):
    """The reference's --preprocess ladder for every problem (GPU pilot,
    FP64 DOS and ladder search on host threads).  timing (a dict) receives
    the seconds of the GPU pilot and of the host solve."""
    t0 = time.perf_counter()
    eng, hb, bs, be = pilot(ptx, problems, nt, T0, T1, burn_in, samples, seed, inverse)
    hists = [eng.histograms(i) for i in range(len(problems))]
    eng.close()
    t1 = time.perf_counter()

    def one(h):
        counts, e_base = h
        return ptx.ladder_from_histograms(counts, e_base, hb, bs, be, mode=ptx.LADDER_NEWTON)

    with ThreadPoolExecutor(threads) as ex:
        lads = list(ex.map(one, hists))
    if timing is not None:
        timing["pilot_s"] = t1 - t0
        timing["host_s"] = time.perf_counter() - t1
    return lads, hists, (hb, bs, be)


def short_features(ptx, problems, stats, T0, T1, burn_in, samples, seed, inverse=False):
    eng, hb, bs, be = pilot(ptx, problems, FEAT_SLOTS, T0, T1, burn_in, samples, seed, inverse)
    X = np.stack([features(*eng.histograms(i), stats[i]) for i in range(len(problems))])
    return X, bs, be


def measured_rates(ptx, problems, ladders, steps, seed):
    """Adjacent-pair swap acceptance of a 2-chain PT run (no ICM) per problem."""
    eng = ptx.Engine(problems, ladders, seed=seed, chains=2)
    eng.step(steps)
    eng.sync()
    out = []
    for i, lad in enumerate(ladders):
        nt = len(lad)
        a, s, _ = eng.swap_stats(i)
        att = a[: nt - 1].astype(np.float64) + a[nt: 2 * nt - 1]
        suc = s[: nt - 1].astype(np.float64) + s[nt: 2 * nt - 1]
        out.append(suc / np.maximum(att, 1))
    return out


def in_band(rates):
    return bool(len(rates) > 0 and np.all(np.abs(np.asarray(rates) - TARGET) <= TOL))


# --------------------------------------------------------------------------
# the surrogate
# --------------------------------------------------------------------------

class Surrogate:
    def __init__(self, ptx, path=WEIGHTS, max_batch=256):
        z = np.load(path, allow_pickle=False)
        self.net = ptx.Mlp(DIMS, max_batch=max_batch)
        self.net.set_params(z["params"])
        self.max_batch = max_batch

    def close(self):
        self.net.close()

    def predict(self, X):
        out = []
        for i in range(0, len(X), self.max_batch):
            out.append(self.net.forward(X[i:i + self.max_batch]))
        return np.concatenate(out)

    def ladders(self, X, bs, be):
        return [ladder_from_increments(y, bs, be) for y in self.predict(X)]


def choose_ladders(ptx, sur, problems, stats, T0=0.5, T1=3.0, seed=0, short=(500, 1000), validate_steps=2000,
                   long=(5000, 2000), inverse=False):
    """Surrogate ladders with the multihistogram preprocess as fallback.

    Every instance gets the network's ladder from its 8-slot pilot; a
    2-chain validation run of validate_steps measures the adjacent swap
    rates, and an instance whose rates are not all within TARGET +- TOL
    falls back to the 64-slot multihistogram ladder (get_beta's).  Returns
    (ladders, accepted flags, measured surrogate rates, seconds per stage)."""
    t = {}
    t0 = time.perf_counter()
    X, bs, be = short_features(ptx, problems, stats, T0, T1, short[0], short[1], seed, inverse)
    t["pilot8_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    lads = sur.ladders(X, bs, be)
    t["mlp_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    rates = measured_rates(ptx, problems, lads, validate_steps, seed + 1)
    t["validate_s"] = time.perf_counter() - t0
    ok = [in_band(r) for r in rates]
    fb = [i for i, o in enumerate(ok) if not o]
    t0 = time.perf_counter()
    if fb:
        ft = {}
        mh, _, _ = multihistogram_ladders(ptx, [problems[i] for i in fb], N_OUT, T0, T1, long[0], long[1], seed + 2,
                                          inverse, timing=ft)
        t["fallback_pilot_s"], t["fallback_host_s"] = ft["pilot_s"], ft["host_s"]
        for i, lad in zip(fb, mh):
            lads[i] = lad
    t["fallback_s"] = time.perf_counter() - t0
    return lads, ok, rates, t


def chunked(seq, n):
    for i in range(0, len(seq), n):
        yield seq[i:i + n]


def make_dataset(ptx, specs, T0=0.5, T1=3.0, seed=0, chunk=256, long=(5000, 2000), short=(500, 1000), log=print):
    """Features of the 8-slot short pilot and labels from the 64-slot
    multihistogram preprocess for every spec."""
    X, T, Mo, lads, meta = [], [], [], [], []
    order = sorted(range(len(specs)), key=lambda i: (specs[i]["family"], specs[i].get("n", 0), specs[i].get("L", 0)))
    for ci, idxs in enumerate(chunked(order, chunk)):
        t0 = time.time()
        coos = [build(specs[i]) for i in idxs]
        stats = [problem_stats(c) for c in coos]
        probs = [ptx.Problem.from_coo(*c) for c in coos]
        x, bs, be = short_features(ptx, probs, stats, T0, T1, short[0], short[1], seed + 17 * ci)
        lad, _, _ = multihistogram_ladders(ptx, probs, N_OUT, T0, T1, long[0], long[1], seed + 17 * ci + 1)
        for j, i in enumerate(idxs):
            t, m = labels(lad[j], bs, be)
            X.append(x[j]); T.append(t); Mo.append(m); lads.append(lad[j])
            meta.append(dict(specs[i], **stats[j], slots=len(lad[j])))
        log(f"chunk {ci}: {len(idxs)} instances in {time.time() - t0:.1f} s")
    return np.stack(X), np.stack(T), np.stack(Mo), lads, meta, (bs, be)


def train_main(a):
    ptx = load_ptx()
    os.makedirs(a.out, exist_ok=True)
    t0 = time.time()
    specs = sample_specs(a.instances + a.heldout, a.seed)
    X, T, Mo, lads, meta, (bs, be) = make_dataset(ptx, specs, seed=a.seed, chunk=a.chunk,
                                                  log=lambda *s: print(*s, flush=True))
    t_data = time.time() - t0
    rng = np.random.default_rng(a.seed)
    perm = rng.permutation(len(X))
    tr, ho = perm[: a.instances], perm[a.instances:]
    net = ptx.Mlp(DIMS, max_batch=a.batch, seed=a.seed)
    t1 = time.time()
    losses = net.train(X[tr], T[tr], Mo[tr], batch=a.batch, epochs=a.epochs, lr=a.lr, seed=a.seed)
    t_train = time.time() - t1
    params = net.params()
    np.savez(os.path.join(a.out, "ladder_mlp.npz"), params=params, dims=np.array(DIMS, np.int32))
    np.savez(os.path.join(a.out, "dataset.npz"), X=X, T=T, M=Mo, train=tr, heldout=ho)
    with open(os.path.join(a.out, "meta.json"), "w") as f:
        json.dump({"meta": meta, "ladders": [l.tolist() for l in lads], "beta_start": bs, "beta_end": be}, f)
    # held-out evaluation: predicted ladders vs labels, and measured swap rates
    y = np.concatenate([net.forward(X[ho][i:i + a.batch]) for i in range(0, len(ho), a.batch)])
    pred = [ladder_from_increments(yy, bs, be) for yy in y]
    lab = [lads[i] for i in ho]
    probs = [ptx.Problem.from_coo(*build(specs_by_meta(meta[i]))) for i in ho]
    nslots_err = np.array([len(p) - len(l) for p, l in zip(pred, lab)])
    lin = [np.linspace(bs, be, len(p) + 1)[:-1] for p in pred]
    rp = measured_rates(ptx, probs, pred, a.validate_steps, a.seed + 5)
    rl = measured_rates(ptx, probs, lab, a.validate_steps, a.seed + 6)
    rn = measured_rates(ptx, probs, lin, a.validate_steps, a.seed + 7)

    def summ(rr):
        allr = np.concatenate(rr)
        return {"pairs_in_band": float(np.mean(np.abs(allr - TARGET) <= TOL)),
                "instances_all_in_band": float(np.mean([in_band(r) for r in rr])),
                "median_min_rate": float(np.median([r.min() for r in rr])),
                "mean_abs_dev": float(np.mean(np.abs(allr - TARGET)))}
    report = {"instances": a.instances, "heldout": a.heldout, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
              "data_s": t_data, "train_s": t_train, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
              "slot_count_mae": float(np.mean(np.abs(nslots_err))),
              "surrogate": summ(rp), "multihistogram": summ(rl), "linear_same_count": summ(rn)}
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)


def specs_by_meta(m):
    keys = ("family", "n", "p", "L", "pm", "seed")
    return {k: m[k] for k in keys if k in m}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="generate labelled instances on the GPU and train the network")
    t.add_argument("--instances", type=int, default=2048)
    t.add_argument("--heldout", type=int, default=256)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--batch", type=int, default=128)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--chunk", type=int, default=256)
    t.add_argument("--validate-steps", type=int, default=2000)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--out", default="surrogate_out")
    a = ap.parse_args()
    if a.cmd == "train":
        train_main(a)


if __name__ == "__main__":
    main()
