"""The device's robin_hood slot order (rv_slotorder.hip: growth by growth, runs simulated one
wavefront each) against the oracle's robin_hood clone (oracle/rh_map.h via rhmap_probe), on
insert and insert-then-erase sequences of distinct keys: plain random keys, positions in a hashed
order as alignedVariants receives them, and keys crowded into a few home ranges so that runs grow
long and the info increment halves (the clone's final increment is checked to be halved)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

PROBE = os.path.join(ROOT, "oracle", "_build", "rhmap_probe")
M64 = (1 << 64) - 1


def hash_int(k):
    p = (int(k) & M64) * 0xDE5FB9D2630458E9
    return ((p >> 64) + (p & M64)) & M64


def clone_order(keys, transient):
    text = "".join((f"i {k}\ne {k}\n" if t else f"i {k}\n") for k, t in zip(keys.tolist(), transient.tolist()))
    r = subprocess.run([PROBE], input=text, capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    inc = int(lines[0].split()[3])
    return [int(ln.split()[1]) for ln in lines[1:-1]], inc


def sequence(seed, n, crowd):
    rnd = np.random.default_rng(seed)
    if seed % 3 == 0:
        keys = rnd.choice(np.arange(-5000, 4 * n + 5000), size=n, replace=False)
    else:
        pos = np.arange(1, n + 1)
        h = np.array([hash_int(p) for p in pos], dtype=object)
        keys = pos[np.argsort([x & ((1 << 22) - 1) for x in h], kind="stable")]
    keys = keys.astype(np.int64)
    if crowd:
        bits = int(rnd.integers(10, 17))
        band = (1 << bits) // int(rnd.integers(8, 32))
        crowded = np.array([(hash_int(k) & ((1 << bits) - 1)) < band for k in keys])
        keep = crowded | (rnd.random(len(keys)) < 0.4)
        keys = keys[keep]
    u = rnd.random(len(keys))
    frac, tfrac = float(rnd.uniform(0.05, 0.9)), float(rnd.uniform(0.0, 0.25))
    sel = u < frac + tfrac
    return keys[sel].astype(np.int32), (u[sel] >= frac).astype(np.uint8)


@pytest.fixture(scope="module")
def engine(tools):
    from rabbitvar_amd.engine import Engine

    if not os.path.exists(PROBE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    e = Engine(sample="S1")
    yield e
    e.close()


@pytest.mark.parametrize("seed,n,crowd", [(1, 20, False), (2, 3000, False), (3, 60000, False), (4, 400000, False),
                                          (5, 60000, True), (6, 300000, True), (7, 900000, True), (8, 150000, True)])
def test_device_slot_order_matches_clone(engine, seed, n, crowd):
    keys, tr = sequence(seed, n, crowd)
    print(f"slot order seed {seed}: {len(keys)} operations, {int(tr.sum())} transient", flush=True)
    want, inc = clone_order(keys, tr)
    got = engine.slot_order(keys, tr)
    print(f"  final info increment {inc}, device {'declined' if got is None else 'applied'}", flush=True)
    if got is None:
        assert crowd  # only sequences with crowded homes may need the host
        return
    assert got.tolist() == want


def test_device_slot_order_reaches_halved_levels(engine):
    """Among crowded sequences the device path covers ones whose table ended with a halved info
    increment (two or more halvings)."""
    covered = 0
    for seed in range(20, 26):
        keys, tr = sequence(seed, 200000, True)
        want, inc = clone_order(keys, tr)
        got = engine.slot_order(keys, tr)
        if got is not None:
            assert got.tolist() == want
            if inc <= 8:
                covered += 1
    assert covered > 0


def test_empty_and_all_transient(engine):
    assert engine.slot_order(np.zeros(0, np.int32), np.zeros(0, np.uint8)).tolist() == []
    keys = np.arange(100, dtype=np.int32)
    assert engine.slot_order(keys, np.ones(100, np.uint8)).tolist() == []
