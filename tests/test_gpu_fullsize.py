"""The calling path at the workloads' real sizes, against the oracle:

* the 5 Mb, 60x germline workload (2.0e6 reads, 200 long deletions with clipped reads, repeats) as
  one region with --fisher: the text identical, lines in the reference's order;
* the deep tumor/normal pair at full depth (tumor 200x, normal 100x, 2% duplicated records) on a
  1 Mb slice, somatic mode with -t and --fisher: the text identical.
"""
import os
import time

import pytest

from conftest import run_oracle, run_oracle_somatic

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[fullsize] {msg}", flush=True)


def test_germline_5mb_60x_full_region(dataset):
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    t0 = time.time()
    d = dataset("c2full")
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    L = len(seq)
    say(f"5 Mb workload: {reads.n} reads generated in {time.time() - t0:.1f} s")
    want, _ = run_oracle(d, f"{name}:1-{L}", ["--fisher"], out="full.txt")
    say(f"oracle done at {time.time() - t0:.1f} s: {len(want.splitlines())} lines")
    eng = Engine(sample="S1", fisher=1)
    got = eng.run_region(name, 1, L, reads, name, seq)
    st = eng.stats()
    eng.close()
    say(f"engine done at {time.time() - t0:.1f} s")
    assert reads.n > 1900000 and L == 5000000
    assert st["walk_errors"] == 0
    assert len(want.splitlines()) > 20000
    assert sorted(got.splitlines()) == sorted(want.splitlines())
    assert got == want


def test_somatic_200x_100x_1mb_slice(dataset):
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    t0 = time.time()
    d = dataset("c3slice")
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    tumor, normal = read_rvb(os.path.join(d, "reads.rvb")), read_rvb(os.path.join(d, "normal.rvb"))
    L = len(seq)
    say(f"1 Mb pair: {tumor.n} tumor + {normal.n} normal reads generated in {time.time() - t0:.1f} s")
    want = run_oracle_somatic(d, f"{name}:1-{L}", ["--fisher", "-t"], out="c3.txt")
    say(f"oracle done at {time.time() - t0:.1f} s: {len(want.splitlines())} lines")
    eng = Engine(sample="T|N", fisher=1, removeDuplicatedReads=1)
    got = eng.run_region_somatic(name, 1, L, tumor, normal, name, seq)
    st = eng.stats()
    eng.close()
    say(f"engine done at {time.time() - t0:.1f} s")
    assert tumor.n > 1300000 and normal.n > 650000
    assert st["walk_errors"] == 0
    assert len(want.splitlines()) > 30000
    assert got == want
