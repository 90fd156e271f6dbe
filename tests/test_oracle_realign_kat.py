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

A second case pins a realigndel absorption (src/VariationRealigner.cpp:489-649, findMM3 :1917-1970):
a 3 bp deletion at p (deleted bases p..p+2) carried by six reads, and two reads that end at p+3
without the gap, so that their last four bases (R(p+3..p+6), the sequence after the deletion)
mismatch R(p..p+3).  p is the first position where all four differ.
  * The CIGAR rewrite keeps "100M" for the two reads (four end mismatches, more than three), and the
    M-loop joins the four high-quality mismatches into the MNV "X&YZW" at p, read position 1 from
    the 3' end (tp = rlen1 - rpe = 1), quality 30.
  * adjustMNP finds none of its pieces ("X", "X&Y", "X&YZ" at p; "Y&ZW", "Z&W", "W" at p+1..p+3):
    no read carries them.
  * realigndel visits "-3" at p (dcnt 6, dellen 3, no extra bases).  findMM3(ref, p, R(p+3..))
    skips no match (R(p) != R(p+3)), so Tbp = p, and collects "X", "XY", "XYZ", "XYZW" at p (four
    mismatches, mcnt <= 3); str has four bases, so nm3 = 0.  findMM5's candidates sit at p+2, p+1
    (no key there) and p (no such allele).  "X&YZW" passes every filter: quality 30 >= goodq,
    mean read position 1 <= nm3 + 4, count 2 < dcnt + dellen and 2 / 6 < 8.  mp = p is not > p,
    so adjCnt(vref, tv) without a reference allele: every count and sum of the MNV is added to the
    deletion (extracnt += 2, pstd = qstd = true) and the MNV is erased from p (p keeps "-3").
  * The clip loops find no soft clip, and pe - p = 3 < 5 skips the noPassingReads branch.
The printed deletion: start p - 1, end p + 2, refallele R(p-1..p+2), varallele R(p-1), variant
count 8 (forward 4, reverse 4); no line is left at p for the MNV.

A third case pins a realignins absorption (src/VariationRealigner.cpp:785-1066): a 3 bp insertion
"+ABC" after p carried by six reads ("50M3I50M"), and two reads ("100M" over p-95..p+4) that hold
A, B, C, R(p+1) at p+1..p+4 instead.  The premises: A != R(p+1), B != R(p+2), C != R(p+3),
R(p+1) != R(p+4) (four end mismatches), C != R(p) (adjInsPos shifts nothing), B != R(p-1) and
A != R(p-2) (findMM5 collects three mismatches, so none of its single-mismatch bookkeeping runs).
  * The two reads keep "100M" and carry the MNV "A&BCR" at p+1 (read position 1, quality 30).
  * adjustMNP finds none of its pieces (R(p+1) != R(p+4) keeps "R" off p+4).
  * realignins visits "+ABC" (insertionCount 6, insert1 "ABC", tn "ABC").  findMM3(ref, p+1,
    "ABC" + R(p+1..)) has Tbp = p+1 and collects "A", "AB", "ABC", "ABCR" (nm3 = 0); findMM5's
    candidates "C" at p, "B&C" at p-1, "A&BC" at p-2 carry no reads.  "A&BCR" passes the filters
    (2 < 6 + 3, 2 / 6 < 8); mp = p+1 > p at the 3' end and p holds the reference allele R(p), so
    adjCnt(vref, tv, lref): the MNV's counts and sums move to the insertion (extracnt += 2,
    pstd = qstd = true) and are taken off R(p) at p (no field goes negative, so correctCnt changes
    nothing); the MNV is erased from p+1.
  * No soft clip; first3 = first5 = p+1, so the adjRefFactor step does not run.
The printed insertion: start p, end p, refallele R(p), varallele R(p) + "ABC", variant count 8
(forward 4, reverse 4); no line is left at p+1 for the MNV.

A fourth case pins realignlgins30 pairing a 3' and a 5' soft-clip consensus into an insertion
(src/VariationRealigner.cpp:1119-1330, find35match :1396-1428, findconseq
include/VariationUtils.h:403-509): a 20 bp insertion INS after q that no read aligns.  Four reads
"50M15S" over q-49..q clip INS[0..15) at their 3' end (Sclip at q+1, consensus seq3 = INS[0..15));
four reads "15S50M" from q+1 clip INS[5..20) at their 5' end (Sclip at q+1, read outward:
seq5 = INS[19], INS[18], ..., INS[5]).  Every base has Phred 30, so findconseq returns each clip
whole.  Premises: INS[0] != R(q+1) and INS[19] != R(q) (no clipped base is rescued), INS[1] !=
R(q+2) and INS[18] != R(q-1) (realignins' findMM3/findMM5 each collect more than one mismatch),
INS and both clips are not low-complexity, INS[0..10) != INS[10..20), and no (i, j) that
find35match visits before (5, 1) reaches 9 agreements before its third disagreement.
  * find35match: at i = 5, j = 1 the strings agree on INS[14] down to INS[4] (10 comparisons);
    the 11th compares with the empty string past seq5's end (one mismatch), so totalLength = 11,
    score 10 > 8, 1/11 < 0.1, and seq5 is exhausted: (b5, b3, score) = (5, 1, 10).
  * smscore = 5: ins = seq3[0..10) + reverse(seq5[0..10)) = INS.  p5 = p3 = q+1: the gap is empty,
    INS's halves differ, so ins = "+INS" at bi = q.  Both clips are marked used, pstd = qstd =
    true, refCoverage[q] += 4 (the 5' clip's count), adjCnt(vref, sc3v, R(q) at q) then
    adjCnt(vref, sc5v): the insertion holds the two clips' counts and sums (extracnt 8), and the
    3' clip's are taken off R(q), where correctCnt clamps what goes negative (counts, read
    position, quality, MAPQ; not NM).  p3 - p5 = 0 < 5 skips noPassingReads.
  * The follow-up realignins("+INS") finds no mismatch allele (reads carry only R(q+1) at q+1 and
    R(q) at q), both clips are used, and first3 = first5 = q+1 skips adjRefFactor.
The printed insertion: start q, end q, refallele R(q), varallele R(q) + INS, variant count 8.
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


P_DEL = kat.pick(lambda p: all(R(p + k) != R(p + 3 + k) for k in range(4)), 200, 400)


def build_realigndel(d):
    from rabbitvar_amd import batch_from_records, write_rvb

    p = P_DEL
    recs = []
    for r in range(6):  # "50M3D50M": deleted bases p..p+2
        s0 = p - 50
        seq = kat.REF[s0 - 1:p - 1] + kat.REF[p + 2:p + 52]
        recs.append({"pos0": s0 - 1, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": 3, "seq": seq,
                     "qual": [30] * 100, "cigar": [(0, 50), (2, 3), (0, 50)]})
    for r in range(2):  # "100M" over p-96..p+3, the last four bases from after the deletion
        s0 = p - 96
        seq = kat.REF[s0 - 1:p - 1] + kat.REF[p + 2:p + 6]
        assert len(seq) == 100
        recs.append({"pos0": s0 - 1, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": 4, "seq": seq,
                     "qual": [30] * 100, "cigar": [(0, 100)]})
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrS\n" + kat.REF + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))


def test_realigndel_absorbs_end_mismatches(tools, tmp_path):
    p = P_DEL
    d = str(tmp_path / "del")
    build_realigndel(d)
    out, pile, real = (os.path.join(d, x) for x in ("o.txt", "p.txt", "r.txt"))
    subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1", "-R",
                    "chrS:1-600", "--fisher", "--out", out, "--dump-pileup", pile, "--dump-realigned", real],
                   check=True, capture_output=True)
    V0, _ = parse(pile)
    V1, _ = parse(real)
    dl = (p, "-3")
    mnv = (p, R(p + 3) + "&" + R(p + 4) + R(p + 5) + R(p + 6))
    # before: six deletion reads (3 forward, 3 reverse), two MNV reads (1 and 1), all high quality
    assert V0[dl][:5] == [6, 3, 3, 6, 0]
    assert V0[mnv][:5] == [2, 1, 1, 2, 0]
    assert V0[mnv][5] == 2.0 and V0[mnv][6] == 60.0  # read position 1 and quality 30, per read
    # after: the MNV's counts and sums added to the deletion, extracnt 2, pstd = qstd = 1
    want = [a + b for a, b in zip(V0[dl][:9], V0[mnv][:9])] + [1, 1, V0[dl][11] + 2]
    assert V1[dl] == want
    assert mnv not in V1
    assert {k: v for k, v in V1.items() if k != dl} == {k: v for k, v in V0.items() if k not in (dl, mnv)}
    lines = [ln.split("\t") for ln in open(out).read().splitlines()]
    dline = [f for f in lines if f[3] == str(p - 1)]
    assert len(dline) == 1
    f = dline[0]
    assert f[3:7] == [str(p - 1), str(p + 2), kat.Rs(p - 1, p + 2), R(p - 1)]
    assert f[8] == "8" and f[11:13] == ["4", "4"]
    assert not [g for g in lines if g[3] == str(p)]



def parse_all(path):
    V, I = {}, {}
    for ln in open(path).read().splitlines():
        f = ln.split("\t")
        if f[0] in ("V", "I"):
            (V if f[0] == "V" else I)[(int(f[1]), f[2])] = ([int(x) for x in f[3:8]] + [float(x) for x in f[8:12]] +
                                                             [int(x) for x in f[12:15]])
    return V, I


def _ins_premises(p):
    a = next((b for b in "ACGT" if b != R(p + 1) and b != R(p - 2)), None)
    b = next((x for x in "ACGT" if x != R(p + 2) and x != R(p - 1)), None)
    c = next((x for x in "ACGT" if x != R(p + 3) and x != R(p)), None)
    if None in (a, b, c) or R(p + 1) == R(p + 4):
        return None
    return a + b + c


P_INS = kat.pick(lambda p: _ins_premises(p) is not None, 200, 400)
INS = _ins_premises(P_INS)


def build_realignins(d):
    from rabbitvar_amd import batch_from_records, write_rvb

    p = P_INS
    recs = []
    for r in range(6):  # "50M3I50M": the insertion after p
        s0 = p - 49
        seq = kat.Rs(s0, p) + INS + kat.Rs(p + 1, p + 50)
        recs.append({"pos0": s0 - 1, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": 3, "seq": seq,
                     "qual": [30] * 103, "cigar": [(0, 50), (1, 3), (0, 50)]})
    for r in range(2):  # "100M" over p-95..p+4
        s0 = p - 95
        seq = kat.Rs(s0, p) + INS + R(p + 1)
        assert len(seq) == 100
        recs.append({"pos0": s0 - 1, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": 4, "seq": seq,
                     "qual": [30] * 100, "cigar": [(0, 100)]})
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrS\n" + kat.REF + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))


def test_realignins_absorbs_end_mismatches(tools, tmp_path):
    p = P_INS
    d = str(tmp_path / "ins")
    build_realignins(d)
    out, pile, real = (os.path.join(d, x) for x in ("o.txt", "p.txt", "r.txt"))
    subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1", "-R",
                    "chrS:1-600", "--fisher", "--out", out, "--dump-pileup", pile, "--dump-realigned", real],
                   check=True, capture_output=True)
    V0, I0 = parse_all(pile)
    V1, I1 = parse_all(real)
    ins = (p, "+" + INS)
    mnv = (p + 1, INS[0] + "&" + INS[1:] + R(p + 1))
    rf = (p, R(p))
    assert I0[ins][:5] == [6, 3, 3, 6, 0]
    assert V0[mnv][:5] == [2, 1, 1, 2, 0]
    assert V0[mnv][5] == 2.0 and V0[mnv][6] == 60.0
    assert all(a >= b for a, b in zip(V0[rf][:9], V0[mnv][:9]))  # nothing for correctCnt to clamp
    assert I1[ins] == [a + b for a, b in zip(I0[ins][:9], V0[mnv][:9])] + [1, 1, I0[ins][11] + 2]
    assert V1[rf] == [a - b for a, b in zip(V0[rf][:9], V0[mnv][:9])] + V0[rf][9:]
    assert mnv not in V1
    assert {k: v for k, v in V1.items() if k != rf} == {k: v for k, v in V0.items() if k not in (rf, mnv)}
    assert {k: v for k, v in I1.items() if k != ins} == {k: v for k, v in I0.items() if k != ins}
    lines = [ln.split("\t") for ln in open(out).read().splitlines()]
    iline = [f for f in lines if f[3] == str(p) and f[6] == R(p) + INS]
    assert len(iline) == 1
    f = iline[0]
    assert f[3:7] == [str(p), str(p), R(p), R(p) + INS]
    assert f[8] == "8" and f[11:13] == ["4", "4"]
    assert not [g for g in lines if g[3] == str(p + 1)]



def _agreements_before_third_miss(seq5, seq3, i, j):
    """find35match's walk from (i, j) (seq3 read backwards from its j-th last base, seq5 forwards
    from i), counting agreements until the third disagreement or either string's end; a position
    past seq5's end counts as a disagreement."""
    agree = miss = t = 0
    while t + j <= len(seq3) and i + t <= len(seq5):
        a = seq3[len(seq3) - j - t]
        b = seq5[i + t] if i + t < len(seq5) else ""
        if a != b:
            miss += 1
            if miss > 2:
                break
        else:
            agree += 1
        t += 1
    return agree


def _lowcomplex(x):
    return max(x.count(b) for b in "ACGT") / len(x) > 0.75 or sum(1 for b in "ACGT" if b in x) < 3


def _lgins_case():
    import random

    rnd = random.Random(35)
    for q in range(200, 400):
        for _ in range(20):
            ins = "".join(rnd.choice("ACGT") for _ in range(20))
            seq3, seq5 = ins[:15], ins[5:][::-1]
            if ins[0] == R(q + 1) or ins[19] == R(q) or ins[1] == R(q + 2) or ins[18] == R(q - 1):
                continue
            if _lowcomplex(ins) or _lowcomplex(seq3) or _lowcomplex(seq5) or ins[:10] == ins[10:]:
                continue
            if any(_agreements_before_third_miss(seq5, seq3, i, j) > 8 for i in range(5) for j in range(1, 7)):
                continue
            assert _agreements_before_third_miss(seq5, seq3, 5, 1) == 10
            return q, ins
    raise AssertionError("no case satisfies the premises")


Q_LG, INS_LG = _lgins_case()


def build_lgins30(d):
    from rabbitvar_amd import batch_from_records, write_rvb

    q, ins = Q_LG, INS_LG
    recs = []
    for r in range(4):  # "50M15S" over q-49..q, clipping INS[0..15)
        recs.append({"pos0": q - 50, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": 0,
                     "seq": kat.Rs(q - 49, q) + ins[:15], "qual": [30] * 65, "cigar": [(0, 50), (4, 15)]})
    for r in range(4):  # "15S50M" from q+1, clipping INS[5..20)
        recs.append({"pos0": q, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": 0,
                     "seq": ins[5:] + kat.Rs(q + 1, q + 50), "qual": [30] * 65, "cigar": [(4, 15), (0, 50)]})
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrS\n" + kat.REF + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))


def parse_clips(path):
    S = {}
    for ln in open(path).read().splitlines():
        f = ln.split("\t")
        if f[0] in ("S5", "S3"):
            S[(f[0], int(f[1]))] = [int(x) for x in f[2:7]] + [float(x) for x in f[7:11]]
    return S


def test_realignlgins30_pairs_clips_into_insertion(tools, tmp_path):
    q, ins = Q_LG, INS_LG
    d = str(tmp_path / "lgins")
    build_lgins30(d)
    out, pile, real = (os.path.join(d, x) for x in ("o.txt", "p.txt", "r.txt"))
    subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1", "-R",
                    "chrS:1-600", "--fisher", "--out", out, "--dump-pileup", pile, "--dump-realigned", real],
                   check=True, capture_output=True)
    V0, I0 = parse_all(pile)
    V1, I1 = parse_all(real)
    _, C0 = parse(pile)
    _, C1 = parse(real)
    S = parse_clips(pile)
    s3, s5 = S[("S3", q + 1)], S[("S5", q + 1)]
    assert s3[:5] == [4, 2, 2, 4, 0] and s5[:5] == [4, 2, 2, 4, 0]
    assert not I0
    key = (q, "+" + ins)
    assert I1 == {key: [a + b for a, b in zip(s3, s5)] + [1, 1, 8]}
    rf = (q, R(q))
    # R(q) at q comes from the four "50M15S" reads' last aligned base (read position 1 from the
    # 3' end, Phred 30, MAPQ 60); subtracting the 3' clip's sums drives its read-position sum
    # negative, and correctCnt clamps counts, read position, quality and MAPQ at 0 (not NM)
    assert V0[rf] == [4, 2, 2, 4, 0, 4.0, 120.0, 240.0, 0.0, 0, 0, 0]
    assert s3[5] > V0[rf][5]
    want = [max(0, a - b) for a, b in zip(V0[rf][:8], s3[:8])] + [V0[rf][8] - s3[8]] + V0[rf][9:]
    assert V1[rf] == want
    assert {k: v for k, v in V1.items() if k != rf} == {k: v for k, v in V0.items() if k != rf}
    assert C1[q] == C0[q] + 4 and {k: v for k, v in C1.items() if k != q} == {k: v for k, v in C0.items() if k != q}
    lines = [ln.split("\t") for ln in open(out).read().splitlines()]
    iline = [f for f in lines if f[3] == str(q) and f[6] == R(q) + ins]
    assert len(iline) == 1
    f = iline[0]
    assert f[3:7] == [str(q), str(q), R(q), R(q) + ins]
    assert f[8] == "8"
