"""Known-answer tests for the oracle's CIGAR rewrite (CigarModifier::modifyCigar,
src/cigarModifier.cpp:219-410 and its helpers :412-1000, rule texts include/patterns.h:52-153),
driven one read at a time through `rvcpu --cigar-kat`.

Every expected answer below was worked out by hand from the reference text, rule by rule, and is
written as a literal (positions relative to the read's alignment start P).  The derivation is in
each case's comment.  Conventions used by the derivations:
  * R(k) is the reference base at 1-based position k; reads are placed at P.
  * Clipped and inserted bases are 'N' unless stated: 'N' never equals a reference base, so the
    rescue comparisons (isHasAndEquals) on them are false.  Every base has quality 30 (> 10,
    CONF_LOWQUAL) unless stated.
  * A case's premises about the reference (for example R(P+4) != R(P+2)) are asserted; the position
    P is chosen as the first one in a range where they hold.
  * The answer is what parseCigar then walks: modifyCigar's words read through the record's own
    CIGAR pointer with the count it returns (parseCigar.cpp:548-574), after cleanupCigar
    (parseCigar.cpp:1858-1892).  "changed" is modifyCigar's return value.
"""
import os
import random
import subprocess

import pytest

from conftest import ORACLE

_rnd = random.Random(2024)
REF = "".join(_rnd.choice("ACGT") for _ in range(600))


def R(k):
    return REF[k - 1]


def Rs(a, b):
    """R(a..b) inclusive."""
    return REF[a - 1:b]


def pick(pred, lo=150, hi=400):
    for p in range(lo, hi):
        if pred(p):
            return p
    raise AssertionError("no position satisfies the premises")


def run(tmp_path, reads):
    inp, out = tmp_path / "kat_in.txt", tmp_path / "kat_out.txt"
    lines = ["REF " + REF]
    for p, cig, seq, q in reads:
        qs = q if isinstance(q, str) else str(q)
        lines.append(f"{p} {cig} {seq} {qs}")
    inp.write_text("\n".join(lines) + "\n")
    subprocess.run([ORACLE, "--cigar-kat", str(inp), str(out)], check=True)
    res = []
    for ln in out.read_text().splitlines():
        c, p, cg = ln.split()
        res.append((int(c), int(p), cg))
    return res


CASES = {}


def case(fn):
    CASES[fn.__name__] = fn
    return fn


# ------------------------------------------------------------------ the indel loop (:261-363)

@case
def s_then_i_at_start():
    # BEGIN_NUMBER_S_NUMBER_IorD "5S3I" -> tslen = 5 + 3 -> "8S92M", position unchanged.  Tails: the
    # clip is all N, the M part matches: nothing more.
    P = 201
    return P, "5S3I92M", "N" * 8 + Rs(P, P + 91), 30, (1, P, "8S92M")


@case
def s_then_d_at_start():
    # "5S3D" -> "5S", position += 3.  combineDigSDigM: query[4] = N, no rescue; forward walk matches.
    P = 201
    return P, "5S3D95M", "N" * 5 + Rs(P + 3, P + 97), 30, (1, P + 3, "5S95M")


@case
def i_then_s_at_end():
    # NUMBER_IorD_NUMBER_S_END "3I5S" -> 5 + 3 = "8S": "92M8S".  captureMisSoftlyMS: query[92] = N.
    P = 201
    return P, "92M3I5S", Rs(P, P + 91) + "N" * 8, 30, (1, P, "92M8S")


@case
def d_then_s_at_end():
    # "3D5S" -> "5S" (a deletion adds nothing to the clip): "92M5S".
    P = 201
    return P, "92M3D5S", Rs(P, P + 91) + "N" * 5, 30, (1, P, "92M5S")


@case
def s_short_m_then_i():
    # BEGIN_NUMBER_S_NUMBER_M_NUMBER_IorD "5S8M2I", 8 <= 10: tslen = 5 + 8 + 2 = 15, position += 8:
    # "15S85M" at P+8.  combineDigSDigM: query[14] (inserted N) vs R(P+7) false, query[13] N false,
    # the forward walk from P+8 matches: unchanged.
    P = 201
    seq = "N" * 5 + Rs(P, P + 7) + "NN" + Rs(P + 8, P + 92)
    return P, "5S8M2I85M", seq, 30, (1, P + 8, "15S85M")


@case
def s_short_m_then_d():
    # "5S8M2D": tslen = 5 + 8 = 13, position += 8 + 2: "13S85M" at P+10.  combineDigSDigM: query[12]
    # = R(P+7) vs R(P+9): different (premise); query[11] = R(P+6) vs R(P+8): different (premise);
    # query[12] vs R(P+9) again false; forward walk matches.
    P = pick(lambda p: R(p + 9) != R(p + 7) and R(p + 8) != R(p + 6))
    seq = "N" * 5 + Rs(P, P + 7) + Rs(P + 10, P + 94)
    return P, "5S8M2D85M", seq, 30, (1, P + 10, "13S85M")


@case
def s_long_m_then_i_unchanged():
    # "5S12M2I": 12 > 10, the rule does not apply; no other rule matches; the tails change nothing:
    # modifyCigar returns false.
    P = 201
    seq = "N" * 5 + Rs(P, P + 11) + "NN" + Rs(P + 12, P + 92)
    return P, "5S12M2I81M", seq, 30, (0, P, "5S12M2I81M")


@case
def i_short_m_then_s():
    # NUMBER_IorD_NUMBER_M_NUMBER_S_END "2I8M5S", 8 <= 10: 5 + 8 + 2 = 15: "85M15S".
    P = 201
    seq = Rs(P, P + 84) + "NN" + Rs(P + 85, P + 92) + "N" * 5
    return P, "85M2I8M5S", seq, 30, (1, P, "85M15S")


@case
def leading_digit_m_then_d():
    # BEGIN_DIGIT_M_NUMBER_IorD_NUMBER_M "3M2D95M": tslen = 3, position += 3 + 2 = P+5; query[3] =
    # R(P+5) equals the reference there, so tn = 0: "3S95M" at P+5.  combineDigSDigM: query[2] =
    # R(P+2) vs R(P+4) and query[1] = R(P+1) vs R(P+3) differ (premises): unchanged.
    P = pick(lambda p: R(p + 4) != R(p + 2) and R(p + 3) != R(p + 1))
    seq = Rs(P, P + 2) + Rs(P + 5, P + 99)
    return P, "3M2D95M", seq, 30, (1, P + 5, "3S95M")


@case
def leading_digit_m_then_i():
    # "4M2I94M": tslen = 4 + 2 = 6, position += 4; query[6] = R(P+4): tn = 0: "6S94M" at P+4.
    P = 201
    seq = Rs(P, P + 3) + "NN" + Rs(P + 4, P + 97)
    return P, "4M2I94M", seq, 30, (1, P + 4, "6S94M")


@case
def leading_digit_m_then_d_with_mismatches():
    # "4M2D94M" whose next two bases mismatch (N): tslen = 4, position P+6, then tn = 2 (query[4],
    # query[5] differ from R(P+6), R(P+7)), query[6] = R(P+8) stops it: "6S92M" at P+8.
    P = 201
    seq = Rs(P, P + 3) + "NN" + Rs(P + 8, P + 99)
    return P, "4M2D94M", seq, 30, (1, P + 8, "6S92M")


@case
def trailing_i_then_digit_m():
    # NUMBER_IorD_DIGIT_M_END "2I4M": 4 + 2 = 6: "94M6S".  captureMisSoftlyMS: query[94] = N: no
    # rescue; the backward walk matches.
    P = 201
    seq = Rs(P, P + 93) + "NN" + Rs(P + 94, P + 97)
    return P, "94M2I4M", seq, 30, (1, P, "94M6S")


@case
def trailing_d_then_digit_m():
    # "2D4M" -> "4S": "94M4S".  captureMisSoftlyMS: query[94] = R(P+96) vs R(P+94) and query[95] =
    # R(P+97) vs R(P+95) differ (premises): no rescue.
    P = pick(lambda p: R(p + 94) != R(p + 96) and R(p + 95) != R(p + 97))
    seq = Rs(P, P + 93) + Rs(P + 96, P + 99)
    return P, "94M2D4M", seq, 30, (1, P, "94M4S")


@case
def two_deletions_merge():
    # DIG_D_DIG_D "2D3D" -> "5D".
    P = 201
    seq = Rs(P, P + 49) + Rs(P + 55, P + 102)
    return P, "50M2D3D48M", seq, 30, (1, P, "50M5D48M")


@case
def two_insertions_merge():
    # DIG_I_DIG_I "2I3I" -> "5I".
    P = 201
    seq = Rs(P, P + 49) + "N" * 5 + Rs(P + 50, P + 94)
    return P, "50M2I3I45M", seq, 30, (1, P, "50M5I45M")


@case
def close_deletions_to_complex():
    # DIG_D_DIG_M_DIG_DI_DIGI "2D5M3D", 5 <= 15: dlen = 2 + 5 + 3 = 10, ilen = 5: "10D5I".
    P = 201
    seq = Rs(P, P + 39) + Rs(P + 42, P + 46) + Rs(P + 50, P + 102)
    return P, "40M2D5M3D53M", seq, 30, (1, P, "40M10D5I53M")


@case
def close_insertion_deletion_to_complex():
    # NOTDIG_DIG_I_DIG_M_DIG_DI_DIGI "M2I5M3D" (the character before is M, neither D nor H), 5 <= 15:
    # dlen = 5 + 3 = 8, ilen = 2 + 5 = 7; the first DIG_I_DIG_M_DIG_DI_DIGI match is replaced: "8D7I".
    P = 201
    seq = Rs(P, P + 39) + "NN" + Rs(P + 40, P + 44) + Rs(P + 48, P + 100)
    return P, "40M2I5M3D53M", seq, 30, (1, P, "40M8D7I53M")


@case
def deletion_insertion_deletion_to_complex():
    # D_M_D_DD_M_D_I_D_M_D_DD "30M2D4M3I5M2D55M": tslen = 4 + 3 + 5 = 12, dlen = 2 + 4 + 5 + 2 = 13,
    # mid = 9; rn = 0 because query[30] = R(P+32) differs from R(P+30) (premise): "30M13D12I55M".
    P = pick(lambda p: R(p + 32) != R(p + 30))
    seq = Rs(P, P + 29) + Rs(P + 32, P + 35) + "NNN" + Rs(P + 36, P + 40) + Rs(P + 43, P + 97)
    return P, "30M2D4M3I5M2D55M", seq, 30, (1, P, "30M13D12I55M")


@case
def three_deletions_to_complex():
    # threeDeletionsPattern "30M2D4M3D5M2D54M": tslen = 4 + 5 = 9, dlen = 2 + 4 + 3 + 5 + 2 = 16,
    # mid = 9; rn = 0 (premise R(P+32) != R(P+30)): "30M16D9I54M".
    P = pick(lambda p: R(p + 32) != R(p + 30))
    seq = Rs(P, P + 29) + Rs(P + 32, P + 35) + Rs(P + 39, P + 43) + Rs(P + 46, P + 99)
    return P, "30M2D4M3D5M2D54M", seq, 30, (1, P, "30M16D9I54M")


@case
def three_indels_to_complex():
    # threeIndelsPattern "30M2I4M3D5M1I54M": mid = 9, tslen = 9 + 2 + 1 = 12, dlen = 9 + 3 = 12;
    # rn = 0 (query[30] = N): "30M12D12I54M".
    P = 201
    seq = Rs(P, P + 29) + "NN" + Rs(P + 30, P + 33) + Rs(P + 37, P + 41) + "N" + Rs(P + 42, P + 95)
    return P, "30M2I4M3D5M1I54M", seq, 30, (1, P, "30M12D12I54M")


@case
def negative_length_restores_the_original():
    # A record whose sequence (54 bases) is longer than its CIGAR's 52.  threeIndels on
    # "30M1I4M3D5M3D12M": mid = 9, tslen = 10, dlen = 15; the bases after the first M match the
    # reference for rn = 23 positions (query[53] = R(P+58) differs from R(P+53), premise), so
    # RDOFF = 53, dlen = -8, tslen = -13 -> dlen = 5, rm = 12 - 13 = -1: "53M5D-1M".  The tails leave
    # it (query[53] equals R(P+58)); cigarstr_2cigar meets -1, restores the original text and
    # modifyCigar returns false.
    P = pick(lambda p: R(p + 58) != R(p + 53))
    seq = Rs(P, P + 52) + R(P + 58)
    return P, "30M1I4M3D5M3D12M", seq, 30, (0, P, "30M1I4M3D5M3D12M")


# ------------------------------------------------------------------ record quirks (:247-258)

@case
def leading_deletion():
    # cigar[0] is D: position += 3 and the CIGAR pointer advances without decrementing n_cigar, so the
    # text is "97M" plus the first SEQ word read as an operation.  The read's second base is not G
    # (premise), so that word is not a soft clip and no rule fires.  The rewrite writes back at the
    # advanced pointer and returns true with n_cigar = 2; parseCigar reads from the record's own
    # pointer: the leading 3D and the 97M, at P+3.
    P = pick(lambda p: R(p + 4) != "G")
    seq = Rs(P + 3, P + 99)
    return P, "3D97M", seq, 30, (1, P + 3, "3D97M")


@case
def leading_insertion_becomes_deletion():
    # cigar[0] is I: rewritten in place to 2D: "2D98M".  No rule fires in the loop.
    # captureMisSoftly3Mismatches walks back from R(P+99) against query[97] = R(P+95) and so on; the
    # premises make the first four comparisons differ, so rn > 3 and nothing is clipped.  The text
    # differs from "2I98M": true, walked as 2D 98M at P.
    P = pick(lambda p: all(R(p + 99 - k) != R(p + 95 - k) for k in range(4)))
    seq = "NN" + Rs(P, P + 97)
    return P, "2I98M", seq, 30, (1, P, "2D98M")


@case
def trailing_insertion_is_not_converted():
    # The trailing check compares the whole word with BAM_CINS (1), so 2I stays; nothing fires and
    # modifyCigar returns false.  cleanupCigar turns the trailing I into S.
    P = 201
    seq = Rs(P, P + 97) + "NN"
    return P, "98M2I", seq, 30, (0, P, "98M2S")


@case
def trailing_deletion_is_kept():
    # The trailing check compares the whole word with BAM_CDEL (2): 3D stays; unchanged.
    P = 201
    return P, "97M3D", Rs(P, P + 96), 30, (0, P, "97M3D")


# ------------------------------------------------------------------ tails (:365-377, :412-668)

@case
def end_mismatches_become_clip():
    # captureMisSoftly3Mismatches: query[99], query[98] = N differ, then three matches: rn = 2 <= 3:
    # "98M2S".
    P = 201
    return P, "100M", Rs(P, P + 97) + "NN", 30, (1, P, "98M2S")


@case
def four_end_mismatches_stay():
    # rn = 4 > 3: unchanged.
    P = 201
    return P, "100M", Rs(P, P + 95) + "NNNN", 30, (0, P, "100M")


@case
def end_mismatch_pattern_mmM():
    # query[99] = N (rn = 1), query[98] matches, query[97] = N (rn = 3), then three matches: "97M3S".
    P = 201
    seq = Rs(P, P + 96) + "N" + R(P + 98) + "N"
    return P, "100M", seq, 30, (1, P, "97M3S")


@case
def end_mismatch_after_deletion():
    # "50M2D48M", last base N: captureMisSoftly3Mismatches with prefix "50M2D": refoff = P + 48 + 52,
    # rdoff = 48 + 50; rn = 1: DIG_M_END "48M" -> "47M1S".
    P = 201
    seq = Rs(P, P + 49) + Rs(P + 52, P + 98) + "N"
    return P, "50M2D48M", seq, 30, (1, P, "50M2D47M1S")


@case
def begin_mismatch_becomes_clip():
    # combineBeginDigM: query[0] = N differs (rn = 1), then three matches: "1S99M" at P+1.
    P = 201
    return P, "100M", "N" + Rs(P + 1, P + 99), 30, (1, P + 1, "1S99M")


@case
def five_prime_clip_rescued():
    # combineDigSDigM: the 5 clipped bases equal R(P-5..P-1) with quality 30: rn = 5, soft = 0:
    # "100M" at P-5.
    P = 201
    return P, "5S95M", Rs(P - 5, P + 94), 30, (1, P - 5, "100M")


@case
def five_prime_clip_low_quality_base_absorbed():
    # query[2] has quality 5: the first loop stops at rn = 2 ("3S97M" at P-2); the second loop
    # matches query[1], query[0] (rn = 2) and, since query[2] equals R(P-3) (the check ignores
    # quality), rn + 1 = 3 more bases join: "100M" at P-5.
    P = 201
    q = ",".join(["30", "30", "5"] + ["30"] * 97)
    return P, "5S95M", Rs(P - 5, P + 94), q, (1, P - 5, "100M")


@case
def five_prime_clip_partly_rescued():
    # query[2] = N: first loop rn = 2 -> "3S97M" at P-2; second loop rn = 2 but query[2] != R(P-3)
    # and rn <= 4: no more; rn != 0: no forward walk.
    P = 201
    seq = Rs(P - 5, P - 4) + "N" + Rs(P - 2, P + 94)
    return P, "5S95M", seq, 30, (1, P - 2, "3S97M")


@case
def three_prime_clip_rescued():
    # captureMisSoftlyMS: the 5 clipped bases equal R(P+95..P+99): "100M".
    P = 201
    return P, "95M5S", Rs(P, P + 99), 30, (1, P, "100M")


@case
def three_prime_clip_partly_rescued_with_stale_offsets():
    # query[97] = N: first loop rn = 2 -> "97M3S".  The second loop keeps the old offsets
    # (refoff = P+95, rdoff = 95): query[96] matches (rn = 1), query[97] does not; rn <= 4 and
    # rn != 0: no further change.
    P = 201
    seq = Rs(P, P + 96) + "N" + Rs(P + 98, P + 99)
    return P, "95M5S", seq, 30, (1, P, "97M3S")


@case
def three_prime_backward_mismatches():
    # captureMisSoftlyMS: clip all N, no rescue; backward walk: query[94], query[93] = N (rn = 2),
    # then three matches: soft = 7, mch = 93: "93M7S".
    P = 201
    seq = Rs(P, P + 92) + "NN" + "N" * 5
    return P, "95M5S", seq, 30, (1, P, "93M7S")


@pytest.mark.parametrize("name", list(CASES))
def test_cigar_rewrite_known_answer(tools, tmp_path, name):
    P, cig, seq, q, want = CASES[name]()
    ((c, p, cg),) = run(tmp_path, [(P, cig, seq, q)])
    assert (c, p, cg) == want, (name, P, cig)


def test_all_cases_in_one_batch(tools, tmp_path):
    """The rewrite keeps no state across reads."""
    rows = [CASES[n]() for n in CASES]
    got = run(tmp_path, [(P, c, s, q) for P, c, s, q, _ in rows])
    assert got == [w for *_, w in rows]
