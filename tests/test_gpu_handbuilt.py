"""The hand-built cases of tests/handbuilt.py through the HIP engine: the pileup state and the
printed call must meet the same first-principles expectations as the oracle, and equal its output."""
import os

import pytest

import handbuilt as hb
from test_oracle_cases import check_common, parse_dump, run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["snv", "del", "ins"])
def test_handbuilt_on_gpu(tools, tmp_path, kind):
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    d = str(tmp_path / kind)
    s = hb.build(kind, d)
    want_out, want_dump = run(d)
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    eng = Engine(sample="S1", fisher=1)
    dump = eng.debug_pileup(name, 1, hb.L, reads, name, seq)
    out = eng.run_region(name, 1, hb.L, reads, name, seq)
    eng.close()
    V, I, C = parse_dump(dump)
    check_common(kind, s, V, C)
    assert sorted(dump.splitlines()) == sorted(want_dump.splitlines())
    assert out == want_out


def test_duplicate_filter_on_gpu(tools, tmp_path):
    """-t on the hand-built duplicate records: the engine's pileup keeps exactly the reads the
    reference's filter keeps, and equals the oracle's."""
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    d = str(tmp_path / "dups")
    s, alt = hb.build_dups(d)
    want_out, want_dump = run(d, ["-t"])
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb"))
    eng = Engine(sample="S1", fisher=1, removeDuplicatedReads=1)
    dump = eng.debug_pileup(name, 1, hb.L, reads, name, seq)
    out = eng.run_region(name, 1, hb.L, reads, name, seq)
    eng.close()
    V, I, C = parse_dump(dump)
    assert C[hb.EV] == len(hb.DUP_KEPT)
    assert V[(hb.EV, alt)][0] == 1
    assert sorted(dump.splitlines()) == sorted(want_dump.splitlines())
    assert out == want_out
