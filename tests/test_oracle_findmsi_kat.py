"""Known answers for findMSI (src/ToVarsBuilder.cpp:656-722), the microsatellite count and 3' shift
behind the printed msi / msint / shift3 columns.  Each answer is worked out by hand from the
reference text:

  * curmsi_from_end(left + tseq1, unit) starts at 1 and counts the units that precede the last one,
    but only while the index i > nmsi, so the first unit of the string is never counted;
  * the tseq2 loop counts leading units while i + nmsi < len(tseq2), so a unit that ends exactly at
    the end of tseq2 is not counted;
  * the unit with the largest count wins (ties keep the shorter unit);
  * shift3 counts s while tseq[s] == tseq2[s], where tseq = tseq1 + tseq2.

The same inputs go through the oracle (rvcpu --findmsi) and through the scoring kernels' find_msi
compiled for the host (the CPU harness)."""
import os
import subprocess

import pytest

from conftest import ORACLE

CASES = [
    # deletion of AA from a run: left "CCGTAAA" + "AA" ends with 5 A's counted (the run reaches back
    # to index 4), tseq2 adds 3 (AAA before C): 8, unit A; shift3: AA vs AAAC -> 3
    ("AA", "AAACTGCATG", "CCGTAAA", (8.0, 3, "A")),
    # CA repeat: five CA units at the end of tseq1 (the G G T prefix stops it), three more at the
    # start of tseq2 (CACACA before GG): 8, unit CA; shift3: G vs C -> 0
    ("GGTCACACACACA", "CACACAGG", "", (8.0, 0, "CA")),
    # +CAG in a CAG repeat: from the end of left + tseq1 the units at 9 and 6 count, the one at 3 is
    # stopped by i > nmsi: 3; tseq2 adds CAG at 0 and 3 (the third would end past index 8): 5;
    # shift3: CAGCAGCAGTT vs CAGCAGTT agree on 6
    ("CAG", "CAGCAGTT", "ACGCAGCAG", (5.0, 6, "CAG")),
    # both boundary rules at once: AAAA counts 3 (index 1 is not > 1), C adds nothing
    ("AAAA", "C", "", (3.0, 0, "A")),
    # empty tseq1: no unit; tseq equals tseq2, so shift3 runs to its end
    ("", "ACGT", "", (0.0, 4, "")),
    # a single-base tseq1 against a homopolymer tseq2: 1 + 4 (AAAAA: the fifth A ends the string)
    ("A", "AAAAA", "", (5.0, 5, "A")),
]


def parse(text):
    f = text.rstrip("\n").split(" ")
    return float(f[0]), int(f[1]), f[2] if len(f) > 2 else ""


@pytest.mark.parametrize("t1,t2,left,want", CASES)
def test_findmsi_known_answer_oracle(tools, t1, t2, left, want):
    r = subprocess.run([ORACLE, "--findmsi", t1 or "-", t2 or "-", left or "-"], capture_output=True, text=True,
                       check=True)
    assert parse(r.stdout) == want


@pytest.mark.parametrize("t1,t2,left,want", CASES)
def test_findmsi_known_answer_engine(t1, t2, left, want):
    from test_cpu_sim import SIM

    if not os.path.exists(SIM):
        pytest.skip("CPU harness not built (tests/test_cpu_sim.py builds it)")
    r = subprocess.run([SIM, "--findmsi", t1 or "-", t2 or "-", left or "-"], capture_output=True, text=True,
                       check=True)
    assert parse(r.stdout) == want
