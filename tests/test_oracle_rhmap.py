"""Pins the oracle's robin_hood clone (oracle/rh_map.h), whose iteration order fixes the reference's
output line order and soft-clip consensus tie-breaks.

Checks, on random insert/erase sequences:
  * membership: iteration visits exactly the live keys, once each;
  * bookkeeping: every key sits at or after its home slot, with the info byte equal to
    inc * (displacement + 1) + top hash bits;
  * order: with the initial info width in force, the slot order equals the robin-hood order
    characterised independently of the insertion procedure: home slot ascending, then top hash
    bits descending, then insertion time ascending.
"""
import random
import subprocess

import pytest

from conftest import ROOT

import os

PROBE = os.path.join(ROOT, "oracle", "_build", "rhmap_probe")
M64 = (1 << 64) - 1


def hash_int(k):
    # robin_hood 3.4.3 hash_int on x86-64: 128-bit product with the mixing constant, halves summed
    p = (k & M64) * 0xDE5FB9D2630458E9
    return ((p >> 64) + (p & M64)) & M64


def run_probe(ops):
    r = subprocess.run([PROBE], input="".join(f"{o} {k}\n" for o, k in ops), capture_output=True, text=True,
                       check=True)
    lines = r.stdout.splitlines()
    head = lines[0].split()
    mask, inc, shift = int(head[1]), int(head[3]), int(head[5])
    slots = [tuple(int(x) for x in ln.split()) for ln in lines[1:-1]]
    chars = [int(x) for x in lines[-1].split()[1:]]
    return mask, inc, shift, slots, chars


@pytest.fixture(scope="module")
def probe(tools):
    if not os.path.exists(PROBE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return run_probe


@pytest.mark.parametrize("seed", range(12))
def test_random_sequences(probe, seed):
    rnd = random.Random(seed)
    n = rnd.choice([5, 40, 300, 2500])
    span = rnd.choice([n * 2, 10**6])
    ops, live, when, t = [], set(), {}, 0
    for _ in range(n):
        k = rnd.randrange(1, span)
        if live and rnd.random() < (0.25 if seed % 2 else 0.0):
            e = rnd.choice(sorted(live))
            ops.append(("e", e))
            live.discard(e)
            continue
        ops.append(("i", k))
        if k not in live:
            live.add(k)
            when[k] = t
            t += 1
    mask, inc, shift, slots, _ = probe(ops)
    keys = [k for _, k, _ in slots]
    assert sorted(keys) == sorted(live) and len(keys) == len(set(keys))
    for slot, k, info in slots:
        h = hash_int(k)
        home = h & mask
        assert slot >= home
        assert info == inc * (slot - home + 1) + (h >> shift)
    if inc == 32:
        model = sorted(keys, key=lambda k: (hash_int(k) & mask, -(hash_int(k) >> shift), when[k]))
        assert keys == model


def test_lone_key_and_growth(probe):
    mask, inc, shift, slots, _ = probe([("i", 12345)])
    assert mask == 7 and slots[0][0] == hash_int(12345) & 7
    # 8 initial slots, doubled when the 80% load would be exceeded: 6 keys fit, the 7th grows
    mask6, *_ = probe([("i", k) for k in range(1, 7)])
    mask7, *_ = probe([("i", k) for k in range(1, 8)])
    assert (mask6, mask7) == (7, 15)


def test_erase_then_reinsert_goes_last_among_equals(probe):
    keys = list(range(1, 200))
    ops = [("i", k) for k in keys] + [("e", 17), ("i", 17)]
    mask, inc, shift, slots, _ = probe(ops)
    assert 17 in [k for _, k, _ in slots]
