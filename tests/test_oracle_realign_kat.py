"""Known answers for a realigner decision: adjustMNP taking over a single-base allele that is the
right-hand piece of an MNV (src/VariationRealigner.cpp:334-465, adjCnt include/VariationUtils.h:263-
332), worked out by hand from the reference text.

Ten 100 bp reads at 251-350 on the known-answer reference of test_oracle_cigar_kat.py, Phred 30,
MAPQ 60, alternately forward and reverse:
  * reads 0-4 carry X at 301 and Y at 302 (both differ from the reference): the M-loop joins them into
    the MNV "X&Y" at 301 (two high-quality mismatches in a row, parseCigar.cpp:711-764);
  * reads 5-6 carry only Y at 302: the single-base allele "Y" at 302;
  * reads 7-9 match the reference.
adjustMNP visits "X&Y" (its only candidate).  mnt = "XY"; i = 0: the left piece "X" at 301 does not
exist; the right piece "Y" at 302 has count 2 < 5, so adjCnt adds all of its counts and sums to the
MNV (extracnt += 2, pstd = qstd = true), refCoverage[301] grows by 2, and "Y" is erased from 302.
No soft clip sits at 301 or 303, and no deletion or insertion exists, so nothing else changes.
The printed MNV: refallele R(301..302), varallele "XY" (collectReferenceVariants: the '&' removed,
one reference base appended), start 301, end 302, type Complex (the first bases differ).
"""
import os
import subprocess

import pytest

import test_oracle_cigar_kat as kat
from conftest import ORACLE


def R(k):
    return kat.REF[k - 1]


X = next(b for b in "ACGT" if b != R(301))
Y = next(b for b in "ACGT" if b != R(302))


def build(d):
    from rabbitvar_amd import batch_from_records, write_rvb

    recs = []
    for r in range(10):
        seq = list(kat.REF[250:350])
        nm = 0
        if r < 5:
            seq[50], seq[51] = X, Y
            nm = 2
        elif r < 7:
            seq[51] = Y
            nm = 1
        recs.append({"pos0": 250, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": nm, "seq": "".join(seq),
                     "qual": [30] * 100, "cigar": [(0, 100)]})
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrS\n" + kat.REF + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))


def parse(path):
    V, C = {}, {}
    for ln in open(path).read().splitlines():
        f = ln.split("\t")
        if f[0] == "V":
            V[(int(f[1]), f[2])] = [int(x) for x in f[3:8]] + [float(x) for x in f[8:12]] + [int(x) for x in f[12:15]]
        elif f[0] == "C":
            C[int(f[1])] = int(f[2])
    return V, C


def test_adjust_mnp_takes_over_right_piece(tools, tmp_path):
    d = str(tmp_path / "mnp")
    build(d)
    out, pile, real = (os.path.join(d, x) for x in ("o.txt", "p.txt", "r.txt"))
    subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1", "-R",
                    "chrS:1-600", "--fisher", "--out", out, "--dump-pileup", pile, "--dump-realigned", real],
                   check=True, capture_output=True)
    V0, C0 = parse(pile)
    V1, C1 = parse(real)
    mnv = (301, f"{X}&{Y}")
    # before: the MNV from reads 0-4 (3 forward, 2 reverse), "Y" from reads 5 (reverse) and 6 (forward)
    assert V0[mnv][:5] == [5, 3, 2, 5, 0]
    assert V0[(302, Y)][:5] == [2, 1, 1, 2, 0]
    assert C0[301] == 10 and C0[302] == 10
    # after: every count and sum of "Y" added to the MNV, extracnt 2, pstd = qstd = 1
    want = [a + b for a, b in zip(V0[mnv][:9], V0[(302, Y)][:9])] + [1, 1, V0[mnv][11] + 2]
    assert V1[mnv] == want
    assert (302, Y) not in V1
    assert C1[301] == C0[301] + 2 and C1[302] == C0[302]
    # everything else is unchanged
    assert {k: v for k, v in V1.items() if k != mnv} == {k: v for k, v in V0.items() if k not in (mnv, (302, Y))}
    lines = [ln.split("\t") for ln in open(out).read().splitlines()]
    mline = [f for f in lines if f[3] == "301"]
    assert len(mline) == 1
    f = mline[0]
    assert f[3:7] == ["301", "302", R(301) + R(302), X + Y]
    assert f[7:9] == ["12", "7"]  # refCoverage at 301 after the takeover (10 + 2), the MNV's count
    assert f[35] == "Complex"
    assert not [g for g in lines if g[3] == "302"]
