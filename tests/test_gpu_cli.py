"""The command line end to end on the GPU: BAM records decoded by the library, regions from -R and
from a BED file, simple and somatic mode, against the oracle on the same records (RVB)."""
import os
import subprocess
import sys

import pytest

from conftest import ORACLE, ROOT, SYNTH

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bams(tmp_path_factory, tools):
    d = str(tmp_path_factory.mktemp("cli"))
    subprocess.run([SYNTH, "--out", d, "--seed", "21", "--len", "120000", "--depth", "50", "--indel-rate", "0.0005",
                    "--indel-max", "30", "--long-dels", "6", "--repeat-frac", "0.03", "--normal-depth", "30",
                    "--somatic-rate", "0.001", "--vaf", "0.05", "0.4", "--bam"], check=True, capture_output=True)
    return d


def cli(args):
    r = subprocess.run([sys.executable, "-m", "rabbitvar_amd"] + args, cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def oracle(args, out):
    subprocess.run([ORACLE] + args + ["--out", out], check=True, capture_output=True)
    return open(out).read()


def test_region_from_bam(bams):
    d = bams
    got = cli(["-G", f"{d}/ref.fa", "-b", f"{d}/reads.bam", "-N", "S1", "-R", "chrS:1-120000", "--fisher"])
    want = oracle(["-G", f"{d}/ref.fa", "-b", f"{d}/reads.rvb", "-N", "S1", "-R", "chrS:1-120000", "--fisher"],
                  f"{d}/o1.txt")
    assert len(want.splitlines()) > 200
    assert got == want


def test_bed_regions_from_bam(bams):
    d = bams
    bed = f"{d}/panel.bed"
    with open(bed, "w") as f:
        f.write("chrS\t1001\t5000\tgeneA\nchrS\t40001\t41500\tgeneB\nchrS\t90000\t119000\tgeneC\n")
    args = ["-G", f"{d}/ref.fa", "-N", "S1", "-i", bed, "-c", "1", "-S", "2", "-E", "3", "-g", "4", "--fisher", "-t"]
    got = cli(args[:2] + ["-b", f"{d}/reads.bam"] + args[2:])
    want = oracle(args[:2] + ["-b", f"{d}/reads.rvb"] + args[2:], f"{d}/o2.txt")
    assert len(want.splitlines()) > 100
    assert {ln.split("\t")[1] for ln in want.splitlines()} == {"geneA", "geneB", "geneC"}
    assert got == want


def test_somatic_from_bams(bams):
    d = bams
    got = cli(["-G", f"{d}/ref.fa", "-b", f"{d}/reads.bam|{d}/normal.bam", "-N", "T|N", "-R", "chrS:2001-118000",
               "--fisher"])
    want = oracle(["-G", f"{d}/ref.fa", "-b", f"{d}/reads.rvb|{d}/normal.rvb", "-N", "T|N", "-R",
                   "chrS:2001-118000", "--fisher"], f"{d}/o3.txt")
    assert len(want.splitlines()) > 200
    assert got == want


def test_screened_cli_runs(bams):
    d = bams
    got = cli(["-G", f"{d}/ref.fa", "-b", f"{d}/reads.bam", "-N", "S1", "-R", "chrS:1-120000", "--fisher",
               "--screen"])
    assert len(got.splitlines()) > 200
