"""The whole calling path on the GPU (pileup -> realignment -> scoring -> simple-mode text) against
the oracle: the output text identical, character for character (the 36 fields, 38 with --fisher),
lines in the reference's robin_hood order."""
import os

import pytest

from conftest import run_oracle

pytestmark = pytest.mark.gpu


def _inputs(d):
    from rabbitvar_amd import read_fasta, read_rvb

    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    return name, seq, reads


@pytest.mark.parametrize("case", ["c1", "c2s", "edge"])
@pytest.mark.parametrize("k,fisher", [(1, True), (0, False)])
def test_calls_match_oracle(dataset, case, k, fisher):
    from rabbitvar_amd.engine import Engine

    d = dataset(case)
    name, seq, reads = _inputs(d)
    L = len(seq)
    extra = ["-k", str(k)] + (["--fisher"] if fisher else [])
    want, _ = run_oracle(d, f"{name}:1-{L}", extra, out=f"calls_{k}_{int(fisher)}.txt")
    eng = Engine(sample="S1", performLocalRealignment=k, fisher=int(fisher))
    got = eng.run_region(name, 1, L, reads.select(1, L), name, seq)
    st = eng.stats()
    eng.close()
    assert st["walk_errors"] == 0
    assert len(want.splitlines()) > 100
    assert sorted(got.splitlines()) == sorted(want.splitlines())
    assert got == want  # same lines in the same (robin_hood slot) order


@pytest.mark.parametrize("start,end", [(1, 1), (5000, 5150), (30001, 30500), (59000, 60000)])
def test_subregions_match_oracle(dataset, start, end):
    """Small and ragged regions, including a one-base region and one at the contig end."""
    from rabbitvar_amd.engine import Engine

    d = dataset("edge")
    name, seq, reads = _inputs(d)
    want, _ = run_oracle(d, f"{name}:{start}-{end}", ["--fisher"], out=f"sub_{start}.txt")
    eng = Engine(sample="S1", fisher=1)
    got = eng.run_region(name, start, end, reads.select(start, end), name, seq)
    eng.close()
    assert sorted(got.splitlines()) == sorted(want.splitlines())
    assert got == want  # same lines in the same (robin_hood slot) order


def test_empty_region(dataset):
    """A region no read overlaps prints nothing."""
    from rabbitvar_amd.engine import Engine

    d = dataset("c1")
    name, seq, reads = _inputs(d)
    eng = Engine(sample="S1")
    got = eng.run_region(name, 100, 200, reads.take([]), name, seq)
    eng.close()
    assert got == ""


@pytest.mark.parametrize("samfilter", [0x504, 0])
def test_duplicate_filter_matches_oracle(dataset, samfilter):
    """-t over records with duplicates, unpaired reads and reads flagged unmapped."""
    from rabbitvar_amd.engine import Engine

    d = dataset("dups")
    name, seq, reads = _inputs(d)
    want, _ = run_oracle(d, f"{name}:1-{len(seq)}", ["--fisher", "-t", "-F", f"{samfilter:x}"],
                         out=f"dup_{samfilter}.txt")
    eng = Engine(sample="S1", fisher=1, removeDuplicatedReads=1, samfilter=samfilter)
    got = eng.run_region(name, 1, len(seq), reads.select(1, len(seq)), name, seq)
    eng.close()
    assert len(want.splitlines()) > 500
    assert got == want


@pytest.mark.parametrize("case,start,end", [("c1", 1, 20000), ("edge", 29001, 31000)])
def test_pileup_mode_matches_oracle(dataset, case, start, end):
    """-p: every covered position prints (reference-only ones as their reference variant)."""
    from rabbitvar_amd.engine import Engine

    d = dataset(case)
    name, seq, reads = _inputs(d)
    want, _ = run_oracle(d, f"{name}:{start}-{end}", ["--fisher", "-p"], out=f"pileup_{start}.txt")
    eng = Engine(sample="S1", fisher=1, doPileup=1)
    got = eng.run_region(name, start, end, reads.select(start, end), name, seq)
    eng.close()
    assert len(want.splitlines()) > 0.9 * (end - start + 1)
    assert got == want


def test_structured_records_match_the_text(dataset):
    """rv_callset_get's records (rv_call) carry, in the text's line order, the printed fields of
    each line: positions, counts, the %f values, alleles, genotype, bias, flanks and type."""
    from rabbitvar_amd.engine import Engine

    d = dataset("c1")
    name, seq, reads = _inputs(d)
    L = len(seq)
    want, _ = run_oracle(d, f"{name}:1-{L}", ["--fisher"], out="calls_structured.txt")
    eng = Engine(sample="S1", fisher=1)
    text, calls = eng.run_region_calls(name, 1, L, reads.select(1, L), name, seq)
    eng.close()
    assert text == want
    lines = text.splitlines()
    assert len(calls) == len(lines) > 100

    def f6(x):
        return "%f" % x

    for c, ln in zip(calls, lines):
        f = ln.split("\t")
        assert (int(f[3]), int(f[4])) == (c["start"], c["end"])
        assert (f[5], f[6]) == (c["refallele"], c["varallele"])
        assert [int(x) for x in f[7:13]] == [c["total_pos_coverage"], c["position_coverage"], c["ref_fwd"],
                                             c["ref_rev"], c["var_fwd"], c["var_rev"]]
        assert f[13] == (c["genotype"] or "0")
        assert (f[14], f[15]) == (f6(c["frequency"]), c["bias"])
        assert (f[16], f[18]) == (f6(c["mean_position"]), f6(c["mean_quality"]))
        assert (f[20], f[21]) == (f6(c["fisher_p"]), f6(c["fisher_ratio"]))
        assert f[22] == f6(c["mean_mapq"])
        assert (f[26], f[27], f[28]) == (str(c["shift3"]), f6(c["msi"]), str(c["msint"]))
        assert (f[30], f[31]) == (str(c["hicnt"]), str(c["hicov"]))
        assert (f[32], f[33]) == (c["leftseq"] or "0", c["rightseq"] or "0")
        assert f[35] == c["vartype"]
