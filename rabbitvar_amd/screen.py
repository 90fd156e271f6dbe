"""Realignment screen: training data from the exact path, training on the GPU with the library's own
kernels (rv_mlp_*), the threshold, the model file, and the call-set comparison of screened mode
against exact mode.

    python -m rabbitvar_amd.screen train --work DIR --out MODEL.json

Labels come only from the realigner itself: a candidate is positive when visiting it changed a count
(rv_screen_dataset).  The reference has no counterpart of any of this.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "screen_model.json")
SYNTH = os.path.join(HERE, "_build", "rvsynth")

# Synthetic genomes for training (seeds disjoint from every test and bench workload): the plain
# 200 kb shape, the 5 Mb shape's mix at 1 Mb (repeats, long deletions with clipped reads), a dense
# indel mix, and tumor-like samples with low-VAF variants.
TRAIN_CASES = [
    ["--len", "200000", "--depth", "30"],
    ["--len", "1000000", "--depth", "60", "--indel-rate", "0.000125", "--indel-max", "30", "--long-dels", "40",
     "--repeat-frac", "0.02"],
    ["--len", "300000", "--depth", "40", "--indel-rate", "0.002", "--indel-max", "40", "--snv-rate", "0.004",
     "--long-dels", "20", "--repeat-frac", "0.05", "--err-rate", "0.01"],
    ["--len", "500000", "--depth", "100", "--somatic-rate", "0.001", "--vaf", "0.01", "0.2", "--indel-rate",
     "0.0002", "--indel-max", "20", "--long-dels", "10", "--repeat-frac", "0.02"],
]
TRAIN_SEEDS = [1001, 1002, 1003]
VALID_SEED = 1101

# The full-scale set (--scale full, about 2e6 candidates): genomes shaped as the three germline and
# somatic workloads the screen serves, at their full sizes, generated, collected and deleted one at
# a time so that disk and memory stay at one genome.  Seeds disjoint from every test, bench and
# validation genome.
FULL_CASES = [
    ("C1 1 Mb 30x", ["--len", "1000000", "--depth", "30"], list(range(2001, 2011))),
    ("C2 5 Mb 60x", ["--len", "5000000", "--depth", "60", "--indel-rate", "0.000125", "--indel-max", "30",
                     "--long-dels", "200", "--repeat-frac", "0.02"], list(range(2101, 2109))),
    ("C3 tumor 10 Mb 200x", ["--len", "10000000", "--depth", "200", "--somatic-rate", "0.00006", "--vaf", "0.01",
                             "0.2", "--indel-rate", "0.0001", "--dup-rate", "0.02"], list(range(2201, 2210))),
]


def synth(out, args, seed):
    os.makedirs(out, exist_ok=True)
    subprocess.run([SYNTH, "--out", out, "--seed", str(seed)] + list(args), check=True, capture_output=True)
    return out


def load_genome(d):
    from . import read_fasta, read_rvb

    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    return name, seq, read_rvb(os.path.join(d, "reads.rvb"))


def collect(engine, dirs):
    """Features, labels and kinds of every realignment candidate of the given genomes (one region
    per contig, exact run)."""
    X, y, k = [], [], []
    for d in dirs:
        name, seq, reads = load_genome(d)
        L = len(seq)
        f, lab, kind = engine.screen_dataset(name, 1, L, reads.select(1, L), name, seq)
        X.append(f)
        y.append(lab)
        k.append(kind)
    return np.concatenate(X), np.concatenate(y), np.concatenate(k)


def calls(text):
    """The call set of a simple-mode text: (chrom, start, end, ref, alt) per line."""
    out = set()
    for ln in text.splitlines():
        f = ln.split("\t")
        out.add((f[2], f[3], f[4], f[5], f[6]))
    return out


def compare(exact_text, screened_text):
    """Recall and precision of the screened call set against the exact one, and the fraction of the
    exact lines reproduced byte for byte."""
    a, b = calls(exact_text), calls(screened_text)
    inter = len(a & b)
    recall = inter / len(a) if a else 1.0
    precision = inter / len(b) if b else 1.0
    la, lb = set(exact_text.splitlines()), set(screened_text.splitlines())
    identical = len(la & lb) / len(la) if la else 1.0
    return {"recall": recall, "precision": precision, "identical_lines": identical, "exact_calls": len(a),
            "screened_calls": len(b)}


def run_screened(engine, params, threshold, genome):
    name, seq, reads = genome
    L = len(seq)
    engine.set_screen(params, threshold)
    try:
        text = engine.run_region(name, 1, L, reads.select(1, L), name, seq)
        st = engine.stats()
    finally:
        engine.set_screen(None, 0)
    return text, st


def choose_threshold(engine, params, genome, exact_text, target=0.999, grid=None):
    """The largest threshold whose call-set recall on the validation genome meets the target."""
    grid = grid or [0.5, 0.4, 0.3, 0.2, 0.15, 0.1, 0.07, 0.05, 0.03, 0.02, 0.01, 0.005, 0.002, 0.001]
    sweep = []
    best = None
    for th in grid:
        text, st = run_screened(engine, params, th, genome)
        m = compare(exact_text, text)
        m.update({"threshold": th, "skipped": st["screen_skipped"], "candidates": st["screen_candidates"]})
        sweep.append(m)
        if best is None and m["recall"] >= target:
            best = th
    return best, sweep


def save_model(path, params, threshold, meta):
    with open(path, "w") as f:
        json.dump({"threshold": threshold, "n_params": int(params.size), "params": [float(x) for x in params],
                   "meta": meta}, f)


def load_model(path=MODEL_PATH):
    with open(path) as f:
        m = json.load(f)
    return np.array(m["params"], np.float32), float(m["threshold"]), m.get("meta", {})


def train_main(args):
    from .engine import Engine, Mlp

    t0 = time.time()
    eng = Engine(sample="S1", fisher=1)
    if args.scale == "full":
        import shutil

        parts, per_case = [], {}
        for label, case, seeds in FULL_CASES:
            for seed in seeds:
                d = synth(os.path.join(args.work, f"full_{seed}"), case, seed)
                Xi, yi, ki = collect(eng, [d])
                shutil.rmtree(d)
                parts.append((Xi, yi, ki))
                per_case[label] = per_case.get(label, 0) + int(len(yi))
                print(f"[screen] {label} seed {seed}: {len(yi)} candidates ({time.time() - t0:.0f} s)", flush=True)
        X = np.concatenate([p[0] for p in parts])
        y = np.concatenate([p[1] for p in parts])
        kinds = np.concatenate([p[2] for p in parts])
        del parts
    else:
        dirs = []
        for ci, case in enumerate(TRAIN_CASES):
            for seed in TRAIN_SEEDS:
                dirs.append(synth(os.path.join(args.work, f"train_{ci}_{seed}"), case, seed))
        print(f"[screen] {len(dirs)} training genomes generated ({time.time() - t0:.1f} s)", flush=True)
        X, y, kinds = collect(eng, dirs)
        per_case = None
    pos = float(y.mean())
    print(f"[screen] {len(y)} candidates, {pos:.4f} positive; per kind "
          + ", ".join(f"{k}: {int((kinds == k).sum())} ({float(y[kinds == k].mean()) if (kinds == k).any() else 0:.3f})"
                      for k in range(4)), flush=True)
    rng = np.random.default_rng(7)
    order = rng.permutation(len(y))
    nval = len(y) // 10
    va, tr = order[:nval], order[nval:]
    pos_weight = float(args.pos_weight) if args.pos_weight else max(1.0, (1 - pos) / max(pos, 1e-6))
    net = Mlp(seed=args.seed)
    net.load(X[tr], y[tr])
    t1 = time.time()
    losses = net.train(args.epochs, batch=args.batch, lr=args.lr, pos_weight=pos_weight, seed=args.seed)
    train_s = time.time() - t1
    print(f"[screen] trained {args.epochs} epochs in {train_s:.1f} s; loss {losses[0]:.4f} -> {losses[-1]:.4f}",
          flush=True)
    params = net.params()
    sv = net.predict(X[va])
    # candidate-level view on the held-out candidates
    posv = y[va] == 1
    cand = {}
    for th in (0.5, 0.2, 0.1, 0.05, 0.02, 0.01):
        keep = sv >= th
        cand[str(th)] = {"recall": float(keep[posv].mean()) if posv.any() else 1.0, "kept": float(keep.mean())}
    # threshold on a fresh validation genome, by call-set recall
    vd = synth(os.path.join(args.work, "valid"), TRAIN_CASES[1], VALID_SEED)
    genome = load_genome(vd)
    name, seq, reads = genome
    exact_text = eng.run_region(name, 1, len(seq), reads.select(1, len(seq)), name, seq)
    _, sweep = choose_threshold(eng, params, genome, exact_text, grid=[float(t) for t in cand])
    for m in sweep:
        print(f"[screen] threshold {m['threshold']}: recall {m['recall']:.5f} precision {m['precision']:.5f} "
              f"skipped {m['skipped']}/{m['candidates']}", flush=True)
    # the largest threshold that keeps both the held-out candidates' and the validation call set's
    # recall at the target
    ok = [m["threshold"] for m in sweep if m["recall"] >= 0.999 and cand[str(m["threshold"])]["recall"] >= 0.999]
    threshold = max(ok) if ok else min(float(t) for t in cand)
    meta = {"samples": int(len(y)), "positive_fraction": pos, "pos_weight": pos_weight, "epochs": args.epochs,
            "batch": args.batch, "lr": args.lr, "seed": args.seed, "epoch_loss": [float(v) for v in losses],
            "train_seconds": train_s, "heldout_candidates": cand, "validation_sweep": sweep, "scale": args.scale,
            "train_cases": ([[lab, case, seeds] for lab, case, seeds in FULL_CASES] if args.scale == "full"
                            else TRAIN_CASES),
            "train_seeds": None if args.scale == "full" else TRAIN_SEEDS, "candidates_per_case": per_case,
            "valid_seed": VALID_SEED,
            "threshold_rule": "largest grid threshold with held-out candidate recall and validation call-set "
                              "recall both >= 0.999"}
    save_model(args.out, params, threshold, meta)
    print(f"[screen] threshold {threshold}; model written to {args.out}", flush=True)
    eng.close()
    net.close()


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m rabbitvar_amd.screen")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--work", required=True)
    t.add_argument("--out", default=MODEL_PATH)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--batch", type=int, default=1024)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--pos-weight", type=float, default=0.0)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--scale", choices=["small", "full"], default="small",
                   help="small: the 12 small genomes; full: the full-size C1/C2/C3-shaped set (about 2e6 "
                        "candidates)")
    args = ap.parse_args(argv)
    if args.cmd == "train":
        train_main(args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
