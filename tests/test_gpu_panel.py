"""A region list (amplicon-style panel: reads start at the amplicons' ends, depth 300-1500) through the
engine pool, against the oracle's BED run: the texts in region order, identical, with one engine and
with several engines working at once."""
import os
import subprocess

import pytest

from conftest import ORACLE, SYNTH

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def panel(tmp_path_factory, tools):
    d = str(tmp_path_factory.mktemp("panel"))
    subprocess.run([SYNTH, "--out", d, "--seed", "8", "--len", "1000000", "--panel", "30", "--panel-len", "7.0", "1.0",
                    "--panel-depth", "300", "1500", "--snv-rate", "0.002", "--indel-rate", "0.0005", "--bam"],
                   check=True, capture_output=True)
    subprocess.run([ORACLE, "-G", f"{d}/ref.fa", "-b", f"{d}/reads.rvb", "-N", "S1", "-i", f"{d}/panel.bed", "-c", "1",
                    "-S", "2", "-E", "3", "-g", "4", "--fisher", "--out", f"{d}/oracle.txt"], check=True,
                   capture_output=True)
    return d


def _regions(d):
    out = []
    for ln in open(f"{d}/panel.bed"):
        c, s, e, g = ln.rstrip("\n").split("\t")
        out.append((c, int(s), int(e), g))
    return out


@pytest.mark.parametrize("workers", [1, 4])
def test_pool_matches_oracle(panel, workers):
    from rabbitvar_amd import BamFile, read_fasta
    from rabbitvar_amd.engine import EnginePool

    d = panel
    regions = _regions(d)
    assert len(regions) >= 20
    name, seq = read_fasta(f"{d}/ref.fa")
    bam = BamFile(f"{d}/reads.bam", threads=8)
    pool = EnginePool(workers=workers, sample="S1", fisher=1)
    texts = pool.run_regions(regions, bam, {name: seq})
    pool.close()
    want = open(f"{d}/oracle.txt").read()
    assert len(want.splitlines()) > 100
    assert "".join(texts) == want
    assert sum(1 for t in texts if t) > len(regions) // 2
