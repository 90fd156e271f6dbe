"""The reference is built with -O3 -ffast-math (CMakeLists.txt:28); the oracle and the engine are
built without it.  This builds the oracle a second time with -O3 -ffast-math and requires every
printed byte to be the same on C1 (200 kb, 30x), a C2-shaped slice (300 kb, 60x, long deletions,
repeats), the dense edge-case mix, and a deep 250x/120x tumor/normal pair (simple mode on the tumor
and somatic mode on the pair, with -t), so the flags change nothing these workloads print.

What -ffast-math could change in the reference's own build: createVariant divides four sums by the
same count (ToVarsBuilder.cpp:487-522); -freciprocal-math may rewrite such divisions as products
with one reciprocal, which moves a quotient by an ulp.  roundHalfEven is the identity
(util.h:181-184), so the quotient reaches printf's %f directly, where an ulp matters only on an
exact 7th-decimal tie.  See the design notes for why this build keeps true divisions.
"""
import os
import subprocess

import pytest

from conftest import ROOT, run_oracle, run_oracle_somatic

FM = os.path.join(ROOT, "oracle", "_build", "rvcpu_fastmath")


@pytest.fixture(scope="module")
def fastmath(tools):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "fastmath"], check=True)
    return FM


def contig_len(d):
    with open(os.path.join(d, "ref.fa")) as f:
        return sum(len(x.strip()) for x in f if not x.startswith(">"))


@pytest.mark.parametrize("case", ["c1", "c2s", "edge", "deep"])
def test_fastmath_build_prints_the_same(fastmath, dataset, case):
    d = dataset(case)
    region = f"chrS:1-{contig_len(d)}"
    want, _ = run_oracle(d, region, ["--fisher"], out="fm_plain.txt")
    out = os.path.join(d, "fm_fast.txt")
    subprocess.run([fastmath, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1", "-R",
                    region, "--fisher", "--out", out], check=True, capture_output=True)
    assert len(want.splitlines()) > 300
    assert open(out).read() == want


def test_fastmath_build_prints_the_same_somatic(fastmath, dataset):
    d = dataset("deep")
    region = f"chrS:1-{contig_len(d)}"
    want = run_oracle_somatic(d, region, ["--fisher", "-t"], out="fm_som_plain.txt")
    out = os.path.join(d, "fm_som_fast.txt")
    bams = os.path.join(d, "reads.rvb") + "|" + os.path.join(d, "normal.rvb")
    subprocess.run([fastmath, "-G", os.path.join(d, "ref.fa"), "-b", bams, "-N", "T|N", "-R", region, "--fisher", "-t",
                    "--out", out], check=True, capture_output=True)
    assert len(want.splitlines()) > 1000
    assert open(out).read() == want
