"""realignlgins30's clip-consensus pair scoring (find35match, reference
src/VariationRealigner.cpp:1396-1428) on the device (k_find35, one wavefront per pair) against the
realigner's host find35match, which tests/test_cpu_sim.py pins against the reference's nested loops:
the same (b5, b3, score) for every pair.  And the whole path with the device scoring in use."""
import os
import random
import subprocess

import pytest

from conftest import run_oracle

pytestmark = pytest.mark.gpu


def random_pairs(n, seed):
    rnd = random.Random(seed)
    out = []
    for k in range(n):
        l5, l3 = rnd.randint(5, 160), rnd.randint(5, 160)
        s5 = "".join(rnd.choice("ACGT") for _ in range(l5))
        if k % 2 == 0:
            # built to overlap: the reverse of s3's start is a stretch of s5 with a few substitutions
            a = rnd.randint(0, max(0, l5 - 12))
            piece = list(s5[a:])
            for _ in range(rnd.randint(0, 3)):
                i = rnd.randrange(len(piece))
                piece[i] = rnd.choice("ACGT")
            s3 = "".join(reversed(piece))[: l3] + "".join(rnd.choice("ACGT") for _ in range(max(0, l3 - len(piece))))
        else:
            s3 = "".join(rnd.choice("ACGT") for _ in range(l3))
        out.append((s5, s3))
    return out


def test_device_find35_equals_host(tools, tmp_path):
    from rabbitvar_amd.engine import Engine
    from test_cpu_sim import CSRC, SIM, SIM_DIR
    from conftest import ROOT

    if not os.path.exists(SIM):
        os.makedirs(os.path.dirname(SIM), exist_ok=True)
        srcs = [os.path.join(SIM_DIR, "rv_cpu_sim.cpp")] + [os.path.join(CSRC, f) for f in
                                                            ("rv_host.cpp", "rv_realign.cpp", "rv_format.cpp",
                                                             "rv_order.cpp")]
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", SIM_DIR, "-I", CSRC, "-I", os.path.join(ROOT, "include"),
                        "-o", SIM] + srcs + ["-pthread"], check=True)
    pairs = random_pairs(6000, 17)
    f = tmp_path / "pairs.txt"
    f.write_text("".join(f"{a} {b}\n" for a, b in pairs))
    r = subprocess.run([SIM, "--find35-file", str(f)], capture_output=True, text=True, check=True)
    want = [tuple(int(x) for x in ln.split()) for ln in r.stdout.splitlines()]
    eng = Engine(sample="S1")
    got = eng.find35(pairs)
    eng.close()
    print(f"find35: {len(pairs)} pairs, {sum(1 for w in want if w[2] > 0)} overlapping", flush=True)
    assert sum(1 for w in want if w[2] > 0) > 1000
    assert got == want


@pytest.mark.parametrize("case", ["c2s", "edge"])
def test_calls_with_device_pair_scoring(dataset, case):
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    d = dataset(case)
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    L = len(seq)
    want, _ = run_oracle(d, f"{name}:1-{L}", ["--fisher"], out="f35calls.txt")
    eng = Engine(sample="S1", fisher=1)
    got = eng.run_region(name, 1, L, reads, name, seq)
    st = eng.stats()
    eng.close()
    print(f"{case}: {st['realign_device_pairs']} pairs scored on the device in {st['realign_device_ms']:.3f} ms",
          flush=True)
    assert st["realign_device_pairs"] > 0
    assert got == want
