"""Pileup stage on the GPU vs the oracle: every nonInsertionVariants / insertionVariants entry,
refCoverage value, soft-clip aggregate and count table, bit-exact (FP64 sums compared through
%.17g)."""
import os

import pytest

from conftest import run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["c1", "c2s", "edge"])
@pytest.mark.parametrize("k", [1, 0])
def test_pileup_matches_oracle(dataset, case, k):
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    d = dataset(case)
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    L = len(seq)
    region = f"{name}:1-{L}"
    _, want = run_oracle(d, region, ["-k", str(k)], dump=f"pile_{k}.txt", out=f"out_{k}.txt")
    eng = Engine(sample="S1", performLocalRealignment=k)
    got = eng.debug_pileup(name, 1, L, reads.select(1, L), name, seq)
    st = eng.stats()
    eng.close()
    assert st["walk_errors"] == 0
    assert sorted(got.splitlines()) == sorted(want.splitlines())
