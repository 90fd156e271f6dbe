"""Known answers for the soft-clip rescue and Sclip accumulation of parseCigar's process_softclip
(src/parseCigar.cpp:1112-1301, addCnt :300-325), worked out by hand from the reference text.

-k 0 keeps modifyCigar (whose tails would rescue the same bases first) out of the way, so the
rescue loops of process_softclip itself are what runs.  Ten 100 bp reads on the known-answer
reference of test_oracle_cigar_kat.py, alternately forward and reverse, MAPQ 60, NM 0, Phred 30
except one clipped base of Phred 5:

5' case, "5S95M" at P = 251, clipped bases equal R(246..250), the third (read index 2) at Phred 5:
  * the rescue walks back from the last clipped base while it equals the reference with Phred > 10:
    index 4 -> position 250 (addCnt with read position 5), index 3 -> 249 (read position 4); index
    2 (Phred 5) stops it.  addCnt adds count, strand, read position, quality, MAPQ and NM; it does
    not touch pstd/qstd.  refCoverage of 250 and 249 grows by one per read.
  * the 3 clipped bases left (Phred 5, 30, 30 from index 2 down): one low-quality base, not more
    than one, so all three count: 3 high-quality bases (3 > 1 low), sum 65.  The Sclip sits at 249
    (the start after the rescue): nt[0] = read index 2 = R(248), nt[1] = R(247), nt[2] = R(246), each
    seen 10 times; the cluster gets addCnt(read position 3, quality 65/3.0): 10 reads, Σ position
    30, ten additions of 65/3.0, all low quality (21.67 < 22.5).
  * with the region starting at 251, the rescue still counts 250 and 249 (it has no region check),
    but the Sclip is not made (its start, 249, is outside the region).
3' case, "95M5S" at P = 251, clipped bases equal R(346..350), read index 97 at Phred 5:
  * rescue forward from read index 95: 346 (read position 100 - 95 = 5), 347 (4); index 97 stops it.
  * the Sclip at 348: nt[0] = R(348), nt[1] = R(349), nt[2] = R(350); addCnt(read position 3,
    quality 65/3.0) per read.
"""
import os
import subprocess

import pytest

import test_oracle_cigar_kat as kat
from conftest import ORACLE

P = 251


def R(k):
    return kat.REF[k - 1]


def build(d, side):
    from rabbitvar_amd import batch_from_records, write_rvb

    recs = []
    for r in range(10):
        if side == 5:
            seq = kat.REF[P - 6:P - 1] + kat.REF[P - 1:P + 94]
            cig = [(4, 5), (0, 95)]
            q = [30] * 100
            q[2] = 5
        else:
            seq = kat.REF[P - 1:P + 94] + kat.REF[P + 94:P + 99]
            cig = [(0, 95), (4, 5)]
            q = [30] * 100
            q[97] = 5
        recs.append({"pos0": P - 1, "flag": 16 if r % 2 else 0, "mapq": 60, "nm": 0, "seq": seq, "qual": q,
                     "cigar": cig})
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chrS\n" + kat.REF + "\n")
    write_rvb(os.path.join(d, "reads.rvb"), batch_from_records(recs))


def dump(d, region):
    out, pile = os.path.join(d, "o.txt"), os.path.join(d, "p.txt")
    subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1", "-R",
                    region, "-k", "0", "--out", out, "--dump-pileup", pile], check=True, capture_output=True)
    V, C, S, SN = {}, {}, {}, {}
    for ln in open(pile).read().splitlines():
        f = ln.split("\t")
        if f[0] == "V":
            V[(int(f[1]), f[2])] = tuple(int(x) for x in f[3:8]) + tuple(float(x) for x in f[8:12]) + (int(f[12]),
                                                                                                        int(f[13]))
        elif f[0] == "C":
            C[int(f[1])] = int(f[2])
        elif f[0] in ("S5", "S3"):
            S[(f[0], int(f[1]))] = tuple(int(x) for x in f[2:7]) + tuple(float(x) for x in f[7:11]) + (int(f[11]),)
        elif f[0] in ("S5n", "S3n"):
            SN[(f[0], int(f[1]), int(f[2]))] = f[3:]
    return V, C, S, SN


def ten_thirds():
    s = 0.0
    for _ in range(10):
        s += 65 / 3.0
    return s


@pytest.mark.parametrize("region", ["chrS:1-600", "chrS:251-600"])
def test_five_prime_rescue_and_sclip(tools, tmp_path, region):
    d = str(tmp_path / "s5")
    build(d, 5)
    V, C, S, SN = dump(d, region)
    # (cnt, fwd, rev, hi, lo, Σ read position, Σ quality, Σ MAPQ, Σ NM, pstd, qstd)
    assert V[(250, R(250))] == (10, 5, 5, 10, 0, 50.0, 300.0, 600.0, 0.0, 0, 0)
    assert V[(249, R(249))] == (10, 5, 5, 10, 0, 40.0, 300.0, 600.0, 0.0, 0, 0)
    assert (248, R(248)) not in V
    assert C[250] == 10 and C[249] == 10 and 248 not in C
    if region == "chrS:1-600":
        assert S[("S5", 249)] == (10, 5, 5, 0, 10, 30.0, ten_thirds(), 600.0, 0.0, 3)
        assert SN[("S5n", 249, 0)] == [f"{R(248)}:10"]
        assert SN[("S5n", 249, 1)] == [f"{R(247)}:10"]
        assert SN[("S5n", 249, 2)] == [f"{R(246)}:10"]
    else:
        assert not S


def test_three_prime_rescue_and_sclip(tools, tmp_path):
    d = str(tmp_path / "s3")
    build(d, 3)
    V, C, S, SN = dump(d, "chrS:1-600")
    assert V[(346, R(346))] == (10, 5, 5, 10, 0, 50.0, 300.0, 600.0, 0.0, 0, 0)
    assert V[(347, R(347))] == (10, 5, 5, 10, 0, 40.0, 300.0, 600.0, 0.0, 0, 0)
    assert (348, R(348)) not in V
    assert C[346] == 10 and C[347] == 10 and 348 not in C
    assert S[("S3", 348)] == (10, 5, 5, 0, 10, 30.0, ten_thirds(), 600.0, 0.0, 3)
    assert SN[("S3n", 348, 0)] == [f"{R(348)}:10"]
    assert SN[("S3n", 348, 1)] == [f"{R(349)}:10"]
    assert SN[("S3n", 348, 2)] == [f"{R(350)}:10"]
