"""What the trained screen gains, on genomes none of its training or validation seeds produced.

* On a 1 Mb genome of the 5 Mb germline workload's shape, screened mode with the committed model
  keeps the call set (recall and precision >= 99.9% against the exact path) while the realigner
  skips most candidates.  On such genomes one rule looks sufficient: keep every deletion and
  insertion candidate and skip the MNV and clip-cluster ones (built below as a network of the same
  shape); the call set does not tell the two apart there.
* On a tumor genome at 200x (the somatic workload's shape), some MNV visits do change counts.  The
  network, trained on about 2e6 candidates of all three workload shapes, keeps nearly every
  candidate whose visit changes a count; the kind rule drops the MNV ones.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import SYNTH

pytestmark = pytest.mark.gpu


def _kind_rule_params():
    from rabbitvar_amd import _lib

    n = _lib.load().rv_mlp_param_count()
    F, H = 32, 64
    W1, B1 = 0, F * H
    W2 = B1 + H
    B2 = W2 + H * H
    W3 = B2 + H
    B3 = W3 + H
    P = np.zeros(n, np.float32)
    P[W1 + 1 * H + 0] = 1.0  # h1[0] = relu(is_deletion + is_insertion)
    P[W1 + 2 * H + 0] = 1.0
    P[W2 + 0 * H + 0] = 1.0  # h2[0] = h1[0]
    P[W3 + 0] = 20.0
    P[B3] = -10.0
    return P


def test_screen_keeps_the_call_set(tmp_path):
    from rabbitvar_amd import read_fasta, read_rvb, screen
    from rabbitvar_amd.engine import Engine

    d = str(tmp_path / "fresh")
    os.makedirs(d)
    subprocess.run([SYNTH, "--out", d, "--seed", "4242", "--len", "1000000", "--depth", "60", "--indel-rate",
                    "0.000125", "--indel-max", "30", "--long-dels", "40", "--repeat-frac", "0.02"], check=True,
                   capture_output=True)
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb")).select(1, len(seq))
    eng = Engine(sample="S1", fisher=1)
    exact = eng.run_region(name, 1, len(seq), reads, name, seq)
    params, threshold, meta = screen.load_model()
    text, st = screen.run_screened(eng, params, threshold, (name, seq, reads))
    rule_text, rule_st = screen.run_screened(eng, _kind_rule_params(), 0.5, (name, seq, reads))
    eng.close()
    m = screen.compare(exact, text)
    rule = screen.compare(exact, rule_text)
    print(f"network: skipped {st['screen_skipped']}/{st['screen_candidates']}, {m}")
    print(f"kind rule: skipped {rule_st['screen_skipped']}/{rule_st['screen_candidates']}, {rule}")
    assert m["exact_calls"] > 3000
    assert m["recall"] >= 0.999 and m["precision"] >= 0.999
    assert st["screen_skipped"] > 0.8 * st["screen_candidates"]
    assert rule["recall"] >= 0.999  # the kind rule is enough for this genome's call set


def test_screen_keeps_the_visits_that_change_counts(tmp_path):
    from rabbitvar_amd import read_fasta, read_rvb, screen
    from rabbitvar_amd.engine import Engine, Mlp

    d = str(tmp_path / "tumor")
    os.makedirs(d)
    subprocess.run([SYNTH, "--out", d, "--seed", "4343", "--len", "1000000", "--depth", "200", "--somatic-rate",
                    "0.00006", "--vaf", "0.01", "0.2", "--indel-rate", "0.0001", "--dup-rate", "0.02"], check=True,
                   capture_output=True)
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb")).select(1, len(seq))
    eng = Engine(sample="S1", fisher=1)
    X, y, kinds = eng.screen_dataset(name, 1, len(seq), reads, name, seq)
    eng.close()
    params, threshold, _ = screen.load_model()
    net = Mlp(params=params)
    keep = net.predict(X) >= threshold
    net.close()
    pos = y == 1
    rule_keep = (kinds == 1) | (kinds == 2)  # deletion and insertion candidates
    net_recall, rule_recall = float(keep[pos].mean()), float(rule_keep[pos].mean())
    print(f"{len(y)} candidates, {int(pos.sum())} change counts ({int((pos & (kinds == 0)).sum())} MNV); "
          f"network keeps {keep.mean():.3f} with recall {net_recall:.4f}, kind rule keeps {rule_keep.mean():.3f} "
          f"with recall {rule_recall:.4f}")
    assert pos.sum() > 100 and (pos & (kinds == 0)).sum() > 10
    assert net_recall >= 0.99
    assert net_recall > rule_recall
    assert keep.mean() < 0.2
