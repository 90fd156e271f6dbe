"""Somatic mode on the GPU (tumor and normal piled up, realigned and scored over one reference
window, then classified and printed) against the oracle's somatic printer: the text identical,
character for character (63 tab-terminated columns with --fisher), lines in the reference's
position order (reference src/modes/somaticMode.cpp:83-128, :311-735)."""
import os

import pytest

from conftest import run_oracle_somatic

pytestmark = pytest.mark.gpu


def _inputs(d):
    from rabbitvar_amd import read_fasta, read_rvb

    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    return name, seq, read_rvb(os.path.join(d, "reads.rvb")), read_rvb(os.path.join(d, "normal.rvb"))


def _run(d, start, end, extra, out, **cfg):
    from rabbitvar_amd.engine import Engine

    name, seq, tumor, normal = _inputs(d)
    end = min(end, len(seq))
    want = run_oracle_somatic(d, f"{name}:{start}-{end}", extra, out=out)
    eng = Engine(sample="T|N", **cfg)
    got = eng.run_region_somatic(name, start, end, tumor.select(start, end), normal.select(start, end), name, seq)
    st = eng.stats()
    eng.close()
    assert st["walk_errors"] == 0
    return got, want


@pytest.mark.parametrize("case", ["som", "som_edge"])
@pytest.mark.parametrize("fisher", [True, False])
def test_somatic_matches_oracle(dataset, case, fisher):
    d = dataset(case)
    extra = ["--fisher"] if fisher else []
    got, want = _run(d, 1, 10**9, extra, f"som_{int(fisher)}.txt", fisher=int(fisher))
    known = {"StrongSomatic", "SampleSpecific", "LikelyLOH", "Germline", "StrongLOH", "LikelySomatic", "AFDiff"}
    labels = {f for ln in want.splitlines() for f in ln.split("\t") if f in known}
    assert len(want.splitlines()) > 500
    assert {"StrongSomatic", "Germline", "StrongLOH", "LikelySomatic"} <= labels
    assert sorted(got.splitlines()) == sorted(want.splitlines())
    assert got == want


@pytest.mark.parametrize("start,end", [(1, 1), (2001, 2150), (30001, 30500), (59000, 60000)])
def test_somatic_subregions_match_oracle(dataset, start, end):
    """Small and ragged regions, including a one-base region and one at the contig end."""
    d = dataset("som_edge")
    got, want = _run(d, start, end, ["--fisher"], f"som_sub_{start}.txt", fisher=1)
    assert got == want


def test_somatic_without_realignment(dataset):
    d = dataset("som_edge")
    got, want = _run(d, 1, 10**9, ["--fisher", "-k", "0"], "som_k0.txt", fisher=1, performLocalRealignment=0)
    assert got == want
