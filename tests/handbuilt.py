"""Hand-built inputs whose pileup and calls follow from first principles (used by the oracle tests
and the matching GPU tests).

A 600 bp contig chrT; ten 100 bp reads aligned at 201-300 (alternately forward and reverse, Phred
30, MAPQ 60); four of them carry one event at 251:
  snv  - a substitution at 251                       (CIGAR 100M, NM 1)
  del  - a 3 bp deletion of 251-253                  (50M3D47M, NM 3)
  ins  - a 2 bp insertion after 250                  (50M2I50M, NM 2)
"""
import os
import random

from rabbitvar_amd import batch_from_records, write_rvb

L = 600
START = 200  # pos0 of every read (1-based 201)
RLEN = 100
EV = 251


def reference():
    rnd = random.Random(42)
    s = [rnd.choice("ACGT") for _ in range(L)]
    # no short repeats around the event so that nothing shifts the alleles
    s[EV - 2:EV + 4] = list("ACGTCA")  # positions 250..255 = A C G T C A
    return "".join(s)


def tp(i, rlen1):
    """The reference's read-position statistic for read index i (parseCigar: min(i + 1, rlen1 - i))."""
    return i + 1 if i < rlen1 - i else rlen1 - i


def build(kind, d):
    s = reference()
    recs = []
    for r in range(10):
        flag = 16 if r % 2 else 0
        seq = s[START:START + RLEN]
        cigar = [(0, RLEN)]
        nm = 0
        if r < 4:
            if kind == "snv":
                alt = "T" if s[EV - 1] != "T" else "A"
                seq = seq[:EV - 1 - START] + alt + seq[EV - START:]
                nm = 1
            elif kind == "del":
                seq = s[START:EV - 1] + s[EV + 2:START + RLEN]
                cigar = [(0, EV - 1 - START), (2, 3), (0, RLEN - (EV - 1 - START) - 3)]
                nm = 3
            elif kind == "ins":
                seq = s[START:EV - 1] + "GG" + s[EV - 1:START + RLEN]
                cigar = [(0, EV - 1 - START), (1, 2), (0, RLEN - (EV - 1 - START))]
                nm = 2
        recs.append({"pos0": START, "flag": flag, "mapq": 60, "nm": nm, "seq": seq, "qual": [30] * len(seq),
                     "cigar": cigar})
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrT\n")
        for i in range(0, L, 60):
            f.write(s[i:i + 60] + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))
    return s


def read_layout(kind):
    """Per read: (rlen1 = M + I length, {ref position: read index} over its M bases, NM - indel length)."""
    out = []
    for r in range(10):
        carrier = r < 4
        m = {}
        if not carrier or kind == "snv":
            for i in range(RLEN):
                m[START + 1 + i] = i
            out.append((RLEN, m, 1 if (carrier and kind == "snv") else 0))
        elif kind == "del":
            a = EV - 1 - START
            for i in range(a):
                m[START + 1 + i] = i
            for i in range(a, RLEN - 3):
                m[START + 1 + i + 3] = i
            out.append((RLEN - 3, m, 0))
        else:  # ins
            a = EV - 1 - START
            for i in range(a):
                m[START + 1 + i] = i
            for i in range(a, RLEN):
                m[START + 1 + i] = i + 2
            out.append((RLEN + 2, m, 0))
    return out


def expected_ref_rows(kind, s):
    """position -> (cnt, fwd, rev, hicnt, locnt, pmean, qmean, mqmean, nm, pstd, qstd) of the reference allele
    at positions where every covering read matches the reference.  Each read contributes its
    read-position statistic tp, quality 30, MAPQ 60 and its NM minus its inserted and deleted bases
    (parseCigar's nm - idlen)."""
    lay = read_layout(kind)
    rows = {}
    for p in range(START + 1, START + RLEN + 1):
        if kind == "snv" and p == EV:
            continue
        if kind == "del" and EV <= p <= EV + 2:
            continue
        if kind == "ins" and p == EV - 1:
            continue  # the insertion's anchor base carries the insertion bookkeeping
        cov = [(r, rlen1, m[p], nm) for r, (rlen1, m, nm) in enumerate(lay) if p in m]
        if len(cov) != 10:
            continue
        fwd = sum(1 for r, *_ in cov if r % 2 == 0)
        tps = [tp(i, rl) for _, rl, i, _ in cov]
        # pstd: contributions at two different read positions; qstd: two different qualities
        rows[p] = (10, fwd, 10 - fwd, 10, 0, sum(tps), 300, 600, sum(nm for *_, nm in cov),
                   int(len(set(tps)) > 1), 0)
    return rows


# Duplicate filter (-t).  Records in iterator order, as (start offset, flag, mate pos0, SNV carrier):
# the reference drops a read whose key (start, mate contig, mate start) for a mate starting before
# base 10 was already seen since the last change of alignment start.
DUP_RECORDS = [
    (0, 0x0, -1, True),     # r0 unpaired: key 201-*-0, kept
    (0, 0x10, -1, True),    # r1 unpaired: same key, dropped
    (0, 0x1, 500, False),   # r2 paired, mate at 501: no key, kept
    (0, 0x11, 3, False),    # r3 paired, mate at 4: key 201-=-4, kept
    (0, 0x1, 3, False),     # r4 same key as r3, dropped
    (0, 0x0, -1, True),     # r5 key of r0, still the same start: dropped
    (1, 0x0, -1, False),    # r6 starts at 202: the set is cleared; key 202-*-0, kept
    (0, 0x10, -1, False),   # r7 back at 201 (a second file's records): cleared again, kept
    (0, 0x5, 3, True),      # r8 flagged unmapped: removed by -F 0x504 before the duplicate filter
]
DUP_KEPT = [0, 2, 3, 6, 7]


def build_dups(d):
    s = reference()
    alt = "T" if s[EV - 1] != "T" else "A"
    recs = []
    for off, flag, mpos0, carrier in DUP_RECORDS:
        p0 = START + off
        seq = s[p0:p0 + RLEN]
        nm = 0
        if carrier:
            seq = seq[:EV - 1 - p0] + alt + seq[EV - p0:]
            nm = 1
        recs.append({"pos0": p0, "flag": flag, "mapq": 60, "nm": nm, "seq": seq, "qual": [30] * RLEN,
                     "cigar": [(0, RLEN)], "mpos0": mpos0, "tid": 0, "mtid": 0 if flag & 1 else -1})
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrT\n")
        for i in range(0, L, 60):
            f.write(s[i:i + 60] + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))
    return s, alt


def build_quirks(d, copies=3):
    """Reads that exercise the reference's rarely taken paths, on the known-answer reference of
    test_oracle_cigar_kat.py (chrS, 600 bp):
      * every CIGAR-rewrite known-answer read, `copies` times each (forward and reverse);
      * spliced reads (N operations, parseCigar.cpp:1733-1753);
      * reads without an NM tag (parseCigar.cpp:525-533), mapped and flagged unmapped;
      * 30 plain 100M reads for background coverage.
    Records are written in alignment-start order."""
    import test_oracle_cigar_kat as kat

    ref = kat.REF
    recs = []

    def ops_of(cig):
        ops, num = [], ""
        for ch in cig:
            if ch.isdigit():
                num += ch
            else:
                ops.append(("MIDNSHP=X".index(ch), int(num)))
                num = ""
        return ops

    for name in kat.CASES:
        P, cig, seq, q, _ = kat.CASES[name]()
        ops = ops_of(cig)
        quals = [int(x) for x in q.split(",")] if isinstance(q, str) else [q] * len(seq)
        idlen = sum(n for o, n in ops if o in (1, 2))
        for c in range(copies):
            recs.append({"pos0": P - 1, "flag": 16 if c % 2 else 0, "mapq": 60, "nm": idlen, "seq": seq,
                         "qual": quals, "cigar": ops})
    rnd = random.Random(77)
    for i in range(30):
        p0 = 140 + i * 10
        recs.append({"pos0": p0, "flag": 16 if i % 2 else 0, "mapq": 60, "nm": 0, "seq": ref[p0:p0 + 100],
                     "qual": [rnd.randint(20, 40) for _ in range(100)], "cigar": [(0, 100)]})
    for i in range(6):  # spliced: 40M 100N 60M, 35M 2I 20M 150N 43M
        p0 = 150 + 7 * i
        if i % 2 == 0:
            seq = ref[p0:p0 + 40] + ref[p0 + 140:p0 + 200]
            cig = [(0, 40), (3, 100), (0, 60)]
        else:
            seq = ref[p0:p0 + 35] + "AC" + ref[p0 + 35:p0 + 55] + ref[p0 + 205:p0 + 248]
            cig = [(0, 35), (1, 2), (0, 20), (3, 150), (0, 43)]
        recs.append({"pos0": p0, "flag": 0, "mapq": 60, "nm": 2 if i % 2 else 0, "seq": seq,
                     "qual": [30] * len(seq), "cigar": cig})
    for i in range(4):  # no NM tag: mapped reads are walked (NM counted as 0), unmapped ones dropped
        p0 = 260 + 5 * i
        seq = ref[p0:p0 + 50] + ("A" if ref[p0 + 50] != "A" else "C") + ref[p0 + 51:p0 + 100]
        rec = {"pos0": p0, "flag": 0x8 if i < 2 else 0x4, "mapq": 60, "seq": seq, "qual": [30] * 100,
               "cigar": [(0, 100)]}
        recs.append(rec)
    recs.sort(key=lambda r: r["pos0"])
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrS\n")
        for i in range(0, len(ref), 60):
            f.write(ref[i:i + 60] + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))
    return ref
