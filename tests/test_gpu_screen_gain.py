"""What the trained screen gains over realignment without it, on a genome none of its training or
validation seeds produced: screened mode with the committed model keeps the call set (recall and
precision >= 99.9% against the exact path) while the realigner skips most candidates, and it skips
more of them than the rule the training data makes obvious (keep every deletion and insertion
candidate, skip MNV and clip-cluster candidates), built as a network of the same shape."""
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
    assert st["screen_skipped"] > rule_st["screen_skipped"]
