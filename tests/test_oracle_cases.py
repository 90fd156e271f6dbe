"""Pins the oracle's pileup and call arithmetic on hand-built reads (tests/handbuilt.py), where
counts, strand splits, means, coverage and the printed call follow from first principles."""
import os
import subprocess

import pytest

import handbuilt as hb
from conftest import ORACLE


def run(d, extra=()):
    out = os.path.join(d, "out.txt")
    dump = os.path.join(d, "pile.txt")
    subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1", "-R",
                    f"chrT:1-{hb.L}", "--fisher", "--out", out, "--dump-pileup", dump] + list(extra), check=True)
    return open(out).read(), open(dump).read()


def parse_dump(text):
    V, I, C = {}, {}, {}
    for ln in text.splitlines():
        f = ln.split("\t")
        if f[0] in ("V", "I"):
            vals = (int(f[3]), int(f[4]), int(f[5]), int(f[6]), int(f[7]), float(f[8]), float(f[9]), float(f[10]),
                    float(f[11]), int(f[12]), int(f[13]))
            (V if f[0] == "V" else I)[(int(f[1]), f[2])] = vals
        elif f[0] == "C":
            C[int(f[1])] = int(f[2])
    return V, I, C


def check_common(kind, s, V, C):
    rows = hb.expected_ref_rows(kind, s)
    assert len(rows) > 80
    for p, want in rows.items():
        got = V[(p, s[p - 1])]
        assert got == want, (p, got, want)
        assert C[p] == 10


def fields(line):
    f = line.split("\t")
    assert len(f) == 38
    return f


@pytest.fixture
def case(tools, tmp_path):
    def make(kind):
        d = str(tmp_path / kind)
        s = hb.build(kind, d)
        return d, s

    return make


def test_snv(case):
    d, s = case("snv")
    out, dump = run(d)
    V, I, C = parse_dump(dump)
    check_common("snv", s, V, C)
    ref, alt = s[hb.EV - 1], ("T" if s[hb.EV - 1] != "T" else "A")
    i = hb.EV - hb.START - 1
    assert V[(hb.EV, alt)] == (4, 2, 2, 4, 0, 4 * hb.tp(i, 100), 120, 240, 4, 0, 0)
    assert V[(hb.EV, ref)] == (6, 3, 3, 6, 0, 6 * hb.tp(i, 100), 180, 360, 0, 0, 0)
    lines = out.splitlines()
    assert len(lines) == 1
    f = fields(lines[0])
    assert f[:15] == ["S1", "chrT", "chrT", "251", "251", ref, alt, "10", "4", "3", "3", "2", "2", f"{ref}/{alt}",
                      "0.400000"]
    assert f[15:20] == ["2;2", "50.000000", "0", "30.000000", "0"]
    assert f[20:22] == ["1.000000", "1.000000"]  # Fisher on (3, 3, 2, 2); odds ratio 6/6
    assert f[22:26] == ["60.000000", "8.000000", "0.400000", "0.000000"]
    assert f[29:32] == ["1.000000", "4", "10"]
    assert f[32] == s[hb.EV - 21:hb.EV - 1] and f[33] == s[hb.EV:hb.EV + 20]
    assert f[34:] == ["chrT:1-600", "SNV", "0.000000", "0"]


def test_deletion(case):
    d, s = case("del")
    out, dump = run(d)
    V, I, C = parse_dump(dump)
    check_common("del", s, V, C)
    assert V[(hb.EV, "-3")][:5] == (4, 2, 2, 4, 0)
    for p in range(hb.EV, hb.EV + 3):
        assert C[p] == 10  # the deleted span counts toward coverage
    lines = out.splitlines()
    assert len(lines) == 1
    f = fields(lines[0])
    assert f[3:9] == ["250", "253", s[249:253], s[249], "10", "4"]
    assert f[11:15] == ["2", "2", f"{s[hb.EV - 1]}/-3", "0.400000"]
    assert f[35] == "Deletion"


def test_insertion(case):
    d, s = case("ins")
    out, dump = run(d)
    V, I, C = parse_dump(dump)
    check_common("ins", s, V, C)
    assert I[(hb.EV - 1, "+GG")][:5] == (4, 2, 2, 4, 0)
    lines = out.splitlines()
    assert len(lines) == 1
    f = fields(lines[0])
    assert f[3:9] == ["250", "250", s[249], s[249] + "GG", "10", "4"]
    assert f[11:15] == ["2", "2", f"{s[249]}/+2", "0.400000"]
    assert f[35] == "Insertion"


def test_realignment_off_changes_nothing_here(case):
    d, s = case("snv")
    a, _ = run(d)
    b, _ = run(d, ["-k", "0"])
    assert a == b


@pytest.mark.parametrize("dedup", [False, True])
def test_duplicate_filter(tmp_path, tools, dedup):
    """-t (recordPreprocessor.cpp:153-176) on hand-built records: which reads survive follows from the
    key rule and the clear-on-new-start rule (handbuilt.DUP_RECORDS)."""
    d = str(tmp_path / "dups")
    s, alt = hb.build_dups(d)
    _, dump = run(d, ["-t"] if dedup else [])
    V, I, C = parse_dump(dump)
    kept = hb.DUP_KEPT if dedup else [i for i, r in enumerate(hb.DUP_RECORDS) if not r[1] & 0x4]
    recs = [hb.DUP_RECORDS[i] for i in kept]
    # coverage at 251 and at 201 (r6 starts at 202)
    assert C[hb.EV] == len(recs)
    assert C[hb.START + 1] == sum(1 for r in recs if r[0] == 0)
    carriers = [r for r in recs if r[3]]
    assert V[(hb.EV, alt)][0] == len(carriers)
    assert V[(hb.EV, alt)][1:3] == (sum(1 for r in carriers if not r[1] & 0x10),
                                   sum(1 for r in carriers if r[1] & 0x10))
