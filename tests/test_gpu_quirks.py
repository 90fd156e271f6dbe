"""The reference's rarely taken paths on the GPU: the CIGAR-rewrite known-answer reads (every rule,
the leading D/I and trailing I/D quirks, the negative-length fallback), spliced reads (N operations,
parseCigar.cpp:1733-1753), reads without an NM tag (parseCigar.cpp:525-533), against the oracle
whose rewrite those known answers pin (tests/test_oracle_cigar_kat.py).  Also the capacity error for
a CIGAR longer than the record slot."""
import os

import pytest

import handbuilt as hb
from conftest import run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("extra", [[], ["-k", "0"], ["-F", "0"], ["-p"]])
def test_quirk_reads_match_oracle(tools, tmp_path, extra):
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    d = str(tmp_path / "quirks")
    hb.build_quirks(d)
    want_out, want_dump = run_oracle(d, "chrS:1-600", ["--fisher"] + extra, dump="pile.txt")
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    kw = {}
    if "-k" in extra:
        kw["performLocalRealignment"] = 0
    if "-F" in extra:
        kw["samfilter"] = 0
    if "-p" in extra:
        kw["doPileup"] = 1
    eng = Engine(sample="S1", fisher=1, **kw)
    dump = eng.debug_pileup(name, 1, len(seq), reads, name, seq)
    out = eng.run_region(name, 1, len(seq), reads, name, seq)
    st = eng.stats()
    eng.close()
    print(f"quirks {extra}: {len(want_out.splitlines())} lines, {len(want_dump.splitlines())} pileup records",
          flush=True)
    assert st["walk_errors"] == 0
    assert sorted(dump.splitlines()) == sorted(want_dump.splitlines())
    assert out == want_out
    assert len(want_out.splitlines()) >= 10


def test_long_cigar_fails_with_capacity_error(tools, tmp_path):
    from rabbitvar_amd.engine import Engine, EngineError
    from test_cpu_sim import _long_cigar_records

    ref = hb.reference()
    reads = _long_cigar_records(ref, (5, 11))
    eng = Engine(sample="S1", fisher=1)
    with pytest.raises(EngineError) as ei:
        eng.run_region("chrT", 1, hb.L, reads, "chrT", ref.encode())
    assert "CIGAR longer than" in str(ei.value)
    # the engine stays usable: the same batch without the long reads runs
    ok = eng.run_region("chrT", 1, hb.L, _long_cigar_records(ref, ()), "chrT", ref.encode())
    eng.close()
    assert isinstance(ok, str)


@pytest.mark.parametrize("side,region", [(5, (1, 600)), (5, (251, 600)), (3, (1, 600))])
@pytest.mark.parametrize("k", [0, 1])
def test_softclip_rescue_reads_match_oracle(tools, tmp_path, side, region, k):
    """The soft-clip rescue known-answer reads (tests/test_oracle_softclip_kat.py) on the GPU: the
    pileup (rescued counts, coverage, Sclip aggregates) and the calls equal the oracle's, with and
    without realignment."""
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine
    import test_oracle_softclip_kat as sk

    d = str(tmp_path / f"s{side}")
    sk.build(d, side)
    s, e = region
    want_out, want_dump = run_oracle(d, f"chrS:{s}-{e}", ["--fisher", "-k", str(k)], dump="pile.txt")
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    eng = Engine(sample="S1", fisher=1, performLocalRealignment=k)
    dump = eng.debug_pileup(name, s, e, reads, name, seq)
    out = eng.run_region(name, s, e, reads, name, seq)
    eng.close()
    # with -k 0 the clips survive to Sclip accumulation (with -k 1 the CIGAR rewrite rescues the
    # matching clip bases first); a region starting at 251 leaves the 5' Sclip outside
    assert "S%d\t" % side in want_dump or region[0] > 1 or k == 1
    assert sorted(dump.splitlines()) == sorted(want_dump.splitlines())
    assert out == want_out


@pytest.mark.parametrize("case", ["mnp", "realigndel", "realignins", "lgins30", "caret", "amp"])
def test_realigner_known_answer_reads_match_oracle(tools, tmp_path, case):
    """The realigner known-answer reads (tests/test_oracle_realign_kat.py: adjustMNP taking over an
    MNV's right piece, realigndel and realignins absorbing end mismatches, realignlgins30 pairing two
    soft clips into an insertion) and the deletion-descriptor known answers
    (tests/test_oracle_allele_kat.py) on the GPU: pileup and
    calls equal the oracle's."""
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine
    import test_oracle_allele_kat as ak
    import test_oracle_realign_kat as rk

    d = str(tmp_path / case)
    {"mnp": rk.build, "realigndel": rk.build_realigndel, "realignins": rk.build_realignins,
     "lgins30": rk.build_lgins30, "caret": lambda x: ak.build(x, "caret"), "amp": lambda x: ak.build(x, "amp")}[case](d)
    want_out, want_dump = run_oracle(d, "chrS:1-600", ["--fisher"], dump="pile.txt")
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    eng = Engine(sample="S1", fisher=1)
    dump = eng.debug_pileup(name, 1, 600, reads, name, seq)
    out = eng.run_region(name, 1, 600, reads, name, seq)
    eng.close()
    print(f"realigner known answer {case}: {len(want_out.splitlines())} lines", flush=True)
    assert sorted(dump.splitlines()) == sorted(want_dump.splitlines())
    assert out == want_out
    anchor = {"mnp": "\t301\t302\t", "realigndel": f"\t{rk.P_DEL - 1}\t{rk.P_DEL + 2}\t",
              "realignins": f"\t{rk.P_INS}\t{rk.P_INS}\t", "lgins30": f"\t{rk.Q_LG}\t{rk.Q_LG}\t",
              "caret": f"\t{ak.P_CARET}\t{ak.P_CARET + 1}\t", "amp": f"\t{ak.P_AMP}\t{ak.P_AMP + 2}\t"}[case]
    assert anchor in out
