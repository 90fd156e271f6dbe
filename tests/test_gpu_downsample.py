"""-Z downsampling through the C-ABI against the oracle: the library draws the process's rand() for
each record that passes -F (recordPreprocessor.cpp:133), so the test seeds it as a process that never
calls srand() starts (seed 1) before each run.  Simple mode with the duplicate filter, somatic mode
(the first sample's draws before the second's), and the engine pool over a region list (draws in
region order whatever the number of engines)."""
import ctypes
import os
import subprocess

import pytest

from conftest import ORACLE, SYNTH, run_oracle, run_oracle_somatic

pytestmark = pytest.mark.gpu


def reseed():
    ctypes.CDLL(None).srand(ctypes.c_uint(1))


def test_simple_with_dedup(dataset):
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    d = dataset("dups")
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    reads = read_rvb(os.path.join(d, "reads.rvb")).select(1, 60000)
    want, _ = run_oracle(d, f"{name}:1-60000", ["--fisher", "-t", "-Z", "0.3"], out="gpu_ds.txt")
    full, _ = run_oracle(d, f"{name}:1-60000", ["--fisher", "-t"], out="gpu_ds_full.txt")
    eng = Engine(sample="S1", fisher=1, removeDuplicatedReads=1, isDownsampling=1, downsampling=0.3)
    reseed()
    got = eng.run_region(name, 1, 60000, reads, name, seq)
    eng.close()
    assert len(want.splitlines()) > 300 and want != full
    assert got == want


def test_somatic(dataset):
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import Engine

    d = dataset("som_edge")
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    tumor = read_rvb(os.path.join(d, "reads.rvb")).select(1, len(seq))
    normal = read_rvb(os.path.join(d, "normal.rvb")).select(1, len(seq))
    want = run_oracle_somatic(d, f"{name}:1-{len(seq)}", ["--fisher", "-Z", "0.25"], out="gpu_ds_som.txt")
    eng = Engine(sample="T|N", fisher=1, isDownsampling=1, downsampling=0.25)
    reseed()
    got = eng.run_region_somatic(name, 1, len(seq), tumor, normal, name, seq)
    eng.close()
    assert len(want.splitlines()) > 100
    assert got == want


@pytest.fixture(scope="module")
def panel(tmp_path_factory, tools):
    d = str(tmp_path_factory.mktemp("ds_panel"))
    subprocess.run([SYNTH, "--out", d, "--seed", "9", "--len", "400000", "--panel", "16", "--panel-len", "7.0", "1.0",
                    "--panel-depth", "200", "800", "--snv-rate", "0.002", "--indel-rate", "0.0005", "--bam"],
                   check=True, capture_output=True)
    subprocess.run([ORACLE, "-G", f"{d}/ref.fa", "-b", f"{d}/reads.rvb", "-N", "S1", "-i", f"{d}/panel.bed", "-c", "1",
                    "-S", "2", "-E", "3", "-g", "4", "--fisher", "-Z", "0.2", "--out", f"{d}/oracle.txt"], check=True,
                   capture_output=True)
    return d


@pytest.mark.parametrize("workers", [1, 4])
def test_pool_region_order(panel, workers):
    from rabbitvar_amd import BamFile, read_fasta
    from rabbitvar_amd.engine import EnginePool

    d = panel
    regions = []
    for ln in open(f"{d}/panel.bed"):
        c, s, e, g = ln.rstrip("\n").split("\t")
        regions.append((c, int(s), int(e), g))
    name, seq = read_fasta(f"{d}/ref.fa")
    bam = BamFile(f"{d}/reads.bam", threads=8)
    pool = EnginePool(workers=workers, sample="S1", fisher=1, isDownsampling=1, downsampling=0.2)
    reseed()
    texts = pool.run_regions(regions, bam, {name: seq})
    pool.close()
    want = open(f"{d}/oracle.txt").read()
    assert len(want.splitlines()) > 30
    assert "".join(texts) == want
