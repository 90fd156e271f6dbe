"""The realignment screen on the GPU: the network's MFMA kernels against the CPU restatement
(oracle/mlp.py) for scores and gradients, training on a learnable rule, candidate collection from
the exact path, and screened mode at its two ends (keep everything: the exact text; skip
everything: no candidate visited)."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp as omlp  # noqa: E402

pytestmark = pytest.mark.gpu

SEGMENTS = [("W1", omlp.W1, omlp.B1), ("b1", omlp.B1, omlp.W2), ("W2", omlp.W2, omlp.B2), ("b2", omlp.B2, omlp.W3),
            ("w3", omlp.W3, omlp.B3), ("b3", omlp.B3, omlp.NPARAM)]


def _params(seed, scale=0.25):
    return np.random.default_rng(seed).normal(0, scale, omlp.NPARAM).astype(np.float32)


def test_forward_matches_oracle():
    from rabbitvar_amd.engine import Mlp

    rng = np.random.default_rng(1)
    P = _params(2)
    X = rng.normal(0, 1, (1031, omlp.F)).astype(np.float32)  # not a multiple of the 16-sample tile
    net = Mlp(params=P)
    got = net.predict(X)
    net.close()
    want = omlp.scores(P, X)
    diff = np.abs(got - want)
    # the kernels sum in FP32 in MFMA order; a sum landing on the other side of a BF16 rounding
    # boundary moves a hidden value by one BF16 step
    assert diff.max() < 5e-3, diff.max()
    assert diff.mean() < 2e-4, diff.mean()


def test_gradient_matches_oracle():
    from rabbitvar_amd.engine import Mlp

    rng = np.random.default_rng(3)
    P = _params(4)
    X = rng.normal(0, 1, (777, omlp.F)).astype(np.float32)
    t = (rng.random(777) < 0.3).astype(np.uint8)
    net = Mlp(params=P)
    net.load(X, t)
    g, loss = net.gradient(pos_weight=3.0)
    net.close()
    wl, wg = omlp.loss_and_grad(P, X, t, pos_weight=3.0)
    assert abs(loss - wl) < 1e-4 * (1 + abs(wl))
    for name, a, b in SEGMENTS:
        err = np.linalg.norm(g[a:b] - wg[a:b])
        assert err <= 2e-2 * np.linalg.norm(wg[a:b]) + 1e-6, (name, err, np.linalg.norm(wg[a:b]))


def test_training_learns_a_rule():
    from rabbitvar_amd.engine import Mlp

    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (20000, omlp.F)).astype(np.float32)
    t = ((X[:, 0] + 0.5 * X[:, 1] - 0.7 * X[:, 5]) > 0.4).astype(np.uint8)
    net = Mlp(seed=3)
    net.load(X, t)
    losses = net.train(15, batch=512, lr=3e-3, pos_weight=1.0, seed=9)
    s = net.predict(X)
    P = net.params()
    net.close()
    assert losses[-1] < 0.25 * losses[0], losses
    assert ((s >= 0.5) == (t == 1)).mean() > 0.97
    # the trained parameters score the same on the CPU restatement
    assert np.abs(s[:2000] - omlp.scores(P, X[:2000])).max() < 5e-3


def _genome(dataset, case):
    from rabbitvar_amd import read_fasta, read_rvb

    d = dataset(case)
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    return name, seq, read_rvb(os.path.join(d, "reads.rvb"))


def test_candidate_collection(dataset):
    from rabbitvar_amd.engine import Engine

    name, seq, reads = _genome(dataset, "c2s")
    L = len(seq)
    eng = Engine(sample="S1", fisher=1)
    exact = eng.run_region(name, 1, L, reads.select(1, L), name, seq)
    X, y, kinds = eng.screen_dataset(name, 1, L, reads.select(1, L), name, seq)
    again = eng.run_region(name, 1, L, reads.select(1, L), name, seq)
    eng.close()
    assert again == exact  # collection leaves exact mode as it was
    assert X.shape == (len(y), 32) and np.isfinite(X).all()
    assert set(np.unique(y)) <= {0, 1} and 0 < y.mean() < 1
    assert set(np.unique(kinds)) <= {0, 1, 2, 3}
    for k in range(4):  # one-hot kind features
        assert (X[kinds == k, k] == 1).all()


def test_screened_mode_extremes(dataset):
    from rabbitvar_amd.engine import Engine

    name, seq, reads = _genome(dataset, "c2s")
    L = len(seq)
    eng = Engine(sample="S1", fisher=1)
    exact = eng.run_region(name, 1, L, reads.select(1, L), name, seq)
    P = np.zeros(omlp.NPARAM, np.float32)
    P[omlp.B3] = 30.0  # every score ~1
    eng.set_screen(P, 0.5)
    keep_all = eng.run_region(name, 1, L, reads.select(1, L), name, seq)
    st = eng.stats()
    assert keep_all == exact
    assert st["screen_candidates"] > 1000 and st["screen_skipped"] == 0
    P[omlp.B3] = -30.0  # every score ~0
    eng.set_screen(P, 0.5)
    eng.run_region(name, 1, L, reads.select(1, L), name, seq)
    st = eng.stats()
    assert st["screen_skipped"] == st["screen_candidates"]
    eng.set_screen(None, 0)
    back = eng.run_region(name, 1, L, reads.select(1, L), name, seq)
    eng.close()
    assert back == exact


def test_screened_somatic(dataset):
    """Screened mode in somatic mode: each sample's candidates are screened after its own pileup.
    Keeping every candidate reproduces the exact text; skipping every one skips both samples' lists;
    the trained model keeps the call set."""
    import os

    from rabbitvar_amd import read_fasta, read_rvb, screen
    from rabbitvar_amd.engine import Engine

    d = dataset("som_edge")
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    L = len(seq)
    tumor = read_rvb(os.path.join(d, "reads.rvb")).select(1, L)
    normal = read_rvb(os.path.join(d, "normal.rvb")).select(1, L)
    eng = Engine(sample="T|N", fisher=1)
    exact = eng.run_region_somatic(name, 1, L, tumor, normal, name, seq)
    P = np.zeros(omlp.NPARAM, np.float32)
    P[omlp.B3] = 30.0
    eng.set_screen(P, 0.5)
    keep_all = eng.run_region_somatic(name, 1, L, tumor, normal, name, seq)
    st = eng.stats()
    assert keep_all == exact
    assert st["screen_candidates"] > 100 and st["screen_skipped"] == 0
    P[omlp.B3] = -30.0
    eng.set_screen(P, 0.5)
    eng.run_region_somatic(name, 1, L, tumor, normal, name, seq)
    st = eng.stats()
    assert st["screen_skipped"] == st["screen_candidates"] > 100
    params, threshold, _ = screen.load_model()
    eng.set_screen(params, threshold)
    screened = eng.run_region_somatic(name, 1, L, tumor, normal, name, seq)
    eng.set_screen(None, 0)
    back = eng.run_region_somatic(name, 1, L, tumor, normal, name, seq)
    eng.close()
    assert back == exact
    m = screen.compare(exact, screened)
    assert m["exact_calls"] > 100
    assert m["recall"] >= 0.99 and m["precision"] >= 0.99
