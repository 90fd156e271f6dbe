"""Pins the oracle's somatic printer (reference src/modes/somaticMode.cpp:130-735).

No fixture of the reference covers somatic mode, so the checks are:
  * every sample block of a somatic line equals the fields the simple-mode printer (pinned by its own
    tests) prints for that sample alone at the same variant: a column read from the wrong variant, the
    wrong sample or the wrong field fails here;
  * the classification agrees with the planted truth of the synthetic pair: every germline variant
    (in both samples) is labelled Germline, every somatic one (tumor only) StrongSomatic or
    LikelySomatic;
  * the three Fisher columns against scipy's two-sided exact test on the printed counts, the log
    ratio columns against numpy float32 / float64 arithmetic on the printed counts and frequencies.
"""
import collections
import os

import numpy as np
import pytest

from conftest import run_oracle, run_oracle_somatic

LABELS = {"StrongSomatic", "SampleSpecific", "LikelyLOH", "Germline", "StrongLOH", "LikelySomatic", "AFDiff"}


def _fields(text):
    return [ln.split("\t") for ln in text.splitlines()]


def _block(f, off):
    """A somatic sample block (18 columns): tcov, cov, rfc, rrc, fwd, rev, genotype, freq, bias, pmean,
    pstd, qual, qstd, mapq, qratio, hifreq, extrafreq, nm."""
    b = f[off:off + 18]
    return b[:17] + [float(b[17])]


def _simple(f):
    """The same fields of a simple-mode line (36 columns, no --fisher)."""
    return f[7:20] + f[20:24] + [float(f[27])]


def test_sample_blocks_equal_simple_mode(dataset):
    d = dataset("som_edge")
    som = _fields(run_oracle_somatic(d, "chrS:1-60000", out="s_nofisher.txt"))
    tumor = {tuple(f[3:7]): f for f in _fields(run_oracle(d, "chrS:1-60000", out="simple_t.txt")[0])}
    normal = {tuple(f[3:7]): f
              for f in _fields(run_oracle(d, "chrS:1-60000", out="simple_n.txt", reads="normal.rvb")[0])}
    seen = collections.Counter()
    for f in som:
        assert len(f) == 56 and f[-1] == ""  # 55 tab-terminated columns without --fisher
        key, label = tuple(f[3:7]), f[49]
        assert label in LABELS
        # lines whose first variant is the tumor's: the tumor block is that variant
        if label in ("StrongSomatic", "LikelySomatic", "Germline", "AFDiff") and key in tumor:
            assert _block(f, 7) == _simple(tumor[key]), (label, key)
            seen["tumor"] += 1
        # lines whose first variant is the normal's (its genotype column repeats the tumor's)
        if label in ("StrongLOH", "LikelyLOH", "Germline") and key in normal:
            a, b = _block(f, 25), _simple(normal[key])
            a[6] = b[6] = ""
            assert a == b, (label, key)
            seen["normal"] += 1
    assert seen["tumor"] > 1000 and seen["normal"] > 500


def test_labels_match_planted_truth(dataset):
    d = dataset("som")
    lines = _fields(run_oracle_somatic(d, "chrS:1-200000", out="s_truth.txt"))
    at = collections.defaultdict(list)
    for f in lines:
        at[int(f[3])].append(f)
    with open(os.path.join(d, "truth.tsv")) as fh:
        truth = [ln.split("\t") for ln in fh.read().splitlines()[1:]]
    got = collections.Counter()
    for pos, kind, _len, alt, hap in truth:
        p = int(pos)
        if kind == "SNV":
            labels = [f[49] for f in at[p] if f[6] == alt]
        else:  # indels print at their anchor base, one before the planted position
            labels = [f[49] for f in at[p] + at[p - 1] if f[50] != "SNV"]
        assert labels, (pos, kind)
        somatic = hap == "2"
        if somatic:
            assert labels[0] in ("StrongSomatic", "LikelySomatic"), (pos, kind, labels)
        else:
            assert labels[0] == "Germline", (pos, kind, labels)
        got[somatic] += 1
    assert got[True] > 50 and got[False] > 150


def test_fisher_columns(dataset):
    from scipy.stats import fisher_exact

    d = dataset("som_edge")
    lines = _fields(run_oracle_somatic(d, "chrS:1-60000", ["--fisher"], out="s_fisher.txt"))
    rng = np.random.default_rng(3)
    pick = rng.choice(len(lines), size=min(400, len(lines)), replace=False)
    for i in pick:
        f = lines[i]
        assert len(f) == 64 and f[-1] == ""  # 63 tab-terminated columns
        t = [int(x) for x in f[7:13]]  # tcov, cov, rfc, rrc, fwd, rev
        n = [int(x) for x in f[27:33]]
        tests = [((t[2], t[3], t[4], t[5]), 25), ((n[2], n[3], n[4], n[5]), 45),
                 ((t[1], max(t[0] - t[1], 0), n[1], max(n[0] - n[1], 0)), 59)]
        for (a, b, c, dd), col in tests:
            p = fisher_exact([[a, b], [c, dd]], alternative="two-sided")[1]
            assert abs(float(f[col]) - min(p, 1.0)) < 2e-6, (i, col, (a, b, c, dd), f[col], p)
            fa, fb, fc, fd = (np.float32(x + 0.5) for x in (a, b, c, dd))
            ad, bc = fa * fd, fb * fc
            ratio = np.log(np.float32(ad / bc + bc / ad))
            assert abs(float(f[col + 1]) - float(ratio)) < 2e-6, (i, col, f[col + 1], ratio)
        tf, nf = float(f[14]), float(f[34])
        # the frequencies are printed to 6 decimals: bound the log's error from that rounding
        tol = 5e-7 / max(tf, 1e-4) + 5e-7 / max(nf, 1e-4) + 1e-6
        assert abs(float(f[61]) - np.log(max(tf, 1e-4) / max(nf, 1e-4))) < tol
        assert abs(float(f[62]) - np.log((t[1] + 0.5) / (n[1] + 0.5))) < 1e-6
