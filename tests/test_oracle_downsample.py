"""-Z downsampling in the oracle (recordPreprocessor.cpp:133): each record that passes -F draws the C
library's rand() in iterator order and is dropped when the draw is <= RAND_MAX * fraction.

The pin: the oracle run with -Z over all records equals the oracle run without -Z over the records
that survive the draws, where the draws are made here with the C library's own rand() (seeded 1, the
state of a process that never seeds it).  Drawing for records that -F drops, or after the MAPQ or
duplicate filters, or skipping the draw for some records, changes which records survive and fails
the comparison."""
import ctypes
import os

import numpy as np
import pytest

from conftest import run_oracle

RAND_MAX = 2147483647  # glibc


def libc():
    c = ctypes.CDLL(None)
    c.rand.restype = ctypes.c_int
    c.srand.argtypes = [ctypes.c_uint]
    return c


def expected_drops(flags, samfilter, fraction):
    c = libc()
    c.srand(1)
    limit = RAND_MAX * fraction
    drop = np.zeros(len(flags), bool)
    drawn = 0
    for i, f in enumerate(flags):
        if int(f) & samfilter:
            continue
        drawn += 1
        drop[i] = c.rand() <= limit
    return drop, drawn


@pytest.mark.parametrize("fraction,dedup", [(0.3, False), (0.3, True), (0.05, True)])
def test_downsample_equals_prefiltered_reads(dataset, fraction, dedup):
    from rabbitvar_amd import read_rvb
    from rabbitvar_amd.io import write_rvb

    d = dataset("dups")
    region = "chrS:1-60000"
    reads = read_rvb(os.path.join(d, "reads.rvb")).select(1, 60000)
    assert int(((reads.flag & 0x504) != 0).sum()) > 100  # records the flag filter drops take no draw
    drop, drawn = expected_drops(reads.flag, 0x504, fraction)
    assert abs(drop.sum() / drawn - fraction) < 0.02
    kept = reads.take(np.nonzero(~drop)[0])
    tag = f"{fraction}_{int(dedup)}"
    write_rvb(os.path.join(d, f"kept_{tag}.rvb"), kept)
    extra = ["--fisher"] + (["-t"] if dedup else [])
    got, _ = run_oracle(d, region, extra + ["-Z", str(fraction)], out=f"ds_{tag}.txt")
    want, _ = run_oracle(d, region, extra, out=f"ds_kept_{tag}.txt", reads=f"kept_{tag}.rvb")
    full, _ = run_oracle(d, region, extra, out=f"ds_full_{tag}.txt")
    assert len(want.splitlines()) > 300
    assert got == want
    assert got != full


def test_downsample_all_and_none(dataset):
    d = dataset("dups")
    region = "chrS:1-60000"
    everything, _ = run_oracle(d, region, ["-Z", "1"], out="ds_all.txt")
    assert everything == ""  # every draw is <= RAND_MAX
    full, _ = run_oracle(d, region, [], out="ds_none_full.txt")
    # a draw of exactly 0 is the only drop at fraction 0; glibc's first draws from seed 1 are not 0
    c = libc()
    c.srand(1)
    assert all(c.rand() != 0 for _ in range(200000))
    none, _ = run_oracle(d, region, ["-Z", "0"], out="ds_none.txt")
    assert none == full
