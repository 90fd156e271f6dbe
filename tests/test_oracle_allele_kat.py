"""Known answers for deletion descriptors that carry more than the deletion, through the walk
(src/parseCigar.cpp:1525-1700, process_deletion), the realigner (which leaves them alone) and the
printed allele (collectReferenceVariants, src/ToVarsBuilder.cpp:780-960; varType
include/Variant.h:167-188), worked out by hand from the reference text.

Six 102-base reads on the known-answer reference of test_oracle_cigar_kat.py, Phred 30, MAPQ 60,
alternately forward and reverse, with a 2 bp deletion of p, p+1 after 50 matched bases.

"-2^XY" (CIGAR "50M2D2I50M", inserted bases X != R(p), Y != R(p+1)):
  * No rewrite rule matches a deletion followed directly by an insertion (every merge rule needs an
    M between two indels, and the end rules need a short M at an end), so the walker sees the CIGAR.
  * process_deletion: the next operation is the insertion, so desc = "-2^XY" (isNextInsertion), and
    the 50M after it matches the reference, so nothing ("&") follows.  The variant sits at p.
  * realigndel: dellen 2, extra "XY", extrains "XY"; findMM3/findMM5 name only alleles at p and p+1
    that no read carries; pe - p = 2 < 5.  Nothing changes.
  * Printed: the '^' keeps the start at p; refallele R(p..p+1) (dellen bases); varallele "^XY" with
    the '^' removed (CARET_ATGNC); end p+1; type Complex (R(p) != X); variant count 6.
"-2&T" (CIGAR "50M2D50M", the first base after the deletion T != R(p+2) and T != R(p)):
  * process_deletion, isNextMatched: the next M's first base mismatches (offset 1) and the following
    bases match, so desc = "-2&T"; the walk resumes at p+3.
  * realigndel: extra "T"; findMM5's first candidate "T" at p+2 is not an allele (the T went into
    the descriptor); nothing else matches; pe - p = 3 < 5.  Nothing changes.
  * Printed: start p; refallele R(p..p+1) then AMP_ATGC appends R(p+2) for the extra base:
    R(p..p+2); varallele "T"; end p+2; type Complex (R(p) != T); variant count 6.
"""
import os
import subprocess

import pytest

import test_oracle_cigar_kat as kat
from conftest import ORACLE


def R(k):
    return kat.REF[k - 1]


def _caret_premises(p):
    x = next((b for b in "ACGT" if b != R(p)), None)
    y = next((b for b in "ACGT" if b != R(p + 1)), None)
    return x + y


def _amp_premises(p):
    return next((b for b in "ACGT" if b != R(p + 2) and b != R(p)), None)


P_CARET = kat.pick(lambda p: True, 200, 400)
XY = _caret_premises(P_CARET)
P_AMP = kat.pick(lambda p: _amp_premises(p) is not None, 250, 400)
T = _amp_premises(P_AMP)


def build(d, case):
    from rabbitvar_amd import batch_from_records, write_rvb

    recs = []
    for r in range(6):
        if case == "caret":
            p = P_CARET
            seq = kat.Rs(p - 50, p - 1) + XY + kat.Rs(p + 2, p + 51)
            cig = [(0, 50), (2, 2), (1, 2), (0, 50)]
        else:
            p = P_AMP
            seq = kat.Rs(p - 50, p - 1) + T + kat.Rs(p + 3, p + 51)
            cig = [(0, 50), (2, 2), (0, 50)]
        recs.append({"pos0": p - 51, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": 4 if case == "caret" else 3,
                     "seq": seq, "qual": [30] * len(seq), "cigar": cig})
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrS\n" + kat.REF + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))


def _run(d):
    out, pile, real = (os.path.join(d, x) for x in ("o.txt", "p.txt", "r.txt"))
    subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1", "-R",
                    "chrS:1-600", "--fisher", "--out", out, "--dump-pileup", pile, "--dump-realigned", real],
                   check=True, capture_output=True)

    def keys(path):
        return {(int(f[1]), f[2]): f[3:] for f in (ln.split("\t") for ln in open(path).read().splitlines())
                if f[0] in ("V", "I")}

    return keys(pile), keys(real), [ln.split("\t") for ln in open(out).read().splitlines()]


def test_deletion_then_insertion(tools, tmp_path):
    p = P_CARET
    d = str(tmp_path / "caret")
    build(d, "caret")
    before, after, lines = _run(d)
    key = (p, "-2^" + XY)
    assert before[key][:3] == ["6", "3", "3"]
    assert after == before  # the realigner changes nothing
    mine = [f for f in lines if f[3] == str(p)]
    assert len(mine) == 1
    f = mine[0]
    assert f[3:7] == [str(p), str(p + 1), kat.Rs(p, p + 1), XY]
    assert f[8] == "6" and f[35] == "Complex"


def test_deletion_then_mismatch(tools, tmp_path):
    p = P_AMP
    d = str(tmp_path / "amp")
    build(d, "amp")
    before, after, lines = _run(d)
    key = (p, "-2&" + T)
    assert before[key][:3] == ["6", "3", "3"]
    assert not [k for k in before if k[0] == p + 2 and k[1] == T]  # the T went into the descriptor
    assert after == before
    mine = [f for f in lines if f[3] == str(p)]
    assert len(mine) == 1
    f = mine[0]
    assert f[3:7] == [str(p), str(p + 2), kat.Rs(p, p + 2), T]
    assert f[8] == "6" and f[35] == "Complex"


def test_premise_cigars_reach_the_walker_unchanged(tools, tmp_path):
    """The rewrite (pinned by its own known answers) leaves both CIGARs as they are."""
    reads = [(P_CARET - 50, "50M2D2I50M", kat.Rs(P_CARET - 50, P_CARET - 1) + XY + kat.Rs(P_CARET + 2, P_CARET + 51), 30),
             (P_AMP - 50, "50M2D50M", kat.Rs(P_AMP - 50, P_AMP - 1) + T + kat.Rs(P_AMP + 3, P_AMP + 51), 30)]
    assert kat.run(tmp_path, reads) == [(0, P_CARET - 50, "50M2D2I50M"), (0, P_AMP - 50, "50M2D50M")]
