"""Pins the robin_hood order models against the reference's own robin_hood.h.

oracle/_ref/rh_ref_probe is built (`make -C oracle ref`) from the reference's include/robin_hood.h,
unmodified, with the reference's flags.  It applies insert / erase sequences to
robin_hood::unordered_map<int,int> and <char,int> and prints the iteration orders.  The same
sequences go through the oracle's clone (oracle/rh_map.h, via rhmap_probe) and through the
engine's slot-order model (rabbitvar_amd/csrc/rv_rhorder.h and the <char,int> replay, via the CPU
harness), and all three orders must be identical.

The sequences reach 1e6 operations, so the table grows many times, the info increment is halved
(try_increase_info, robin_hood.h:1972-1995) and erasure shifts runs down (robin_hood.h:1695-1713).
These tests need the reference tree and are skipped where it is absent.
"""
import os
import random
import subprocess

import pytest

from conftest import ROOT

REFERENCE = "/root/reference"
REF_PROBE = os.path.join(ROOT, "oracle", "_ref", "rh_ref_probe")
CLONE = os.path.join(ROOT, "oracle", "_build", "rhmap_probe")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE, "include", "robin_hood.h")),
                                reason="the reference tree is not present")


@pytest.fixture(scope="module")
def ref_probe(tools):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "all"], check=True)
    return REF_PROBE


def ref_orders(text):
    r = subprocess.run([REF_PROBE], input=text, capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    assert lines[0] == "keys"
    return [int(x) for x in lines[1:-1]], [int(x) for x in lines[-1].split()[1:]]


def clone_orders(text):
    r = subprocess.run([CLONE], input=text, capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    return [int(ln.split()[1]) for ln in lines[1:-1]], [int(x) for x in lines[-1].split()[1:]]


def op_sequence(seed, n, erase_frac, keyspace):
    rnd = random.Random(seed)
    ops, live = [], []
    for i in range(n):
        if live and rnd.random() < erase_frac:
            k = live.pop(rnd.randrange(len(live)))
            ops.append(f"e {k}\n")
        else:
            if keyspace == "dense":  # genomic positions: nearly consecutive, the pileup's key shape
                k = 1000 + i * 2 + rnd.randrange(3)
            elif keyspace == "wide":
                k = rnd.randrange(-(1 << 31), 1 << 31)
            else:
                k = rnd.randrange(0, 50000)
            live.append(k)
            ops.append(f"i {k}\n")
    return "".join(ops)


CASES = [
    (0, 50, 0.0, "dense"),
    (1, 3000, 0.1, "small"),
    (2, 70000, 0.0, "dense"),
    (3, 70000, 0.2, "wide"),
    (4, 400000, 0.05, "dense"),
    (5, 1000000, 0.1, "dense"),
    (6, 1000000, 0.0, "wide"),
    (7, 250000, 0.45, "small"),
]


@pytest.mark.parametrize("seed,n,erase_frac,keyspace", CASES)
def test_int_map_order_equals_reference(ref_probe, seed, n, erase_frac, keyspace):
    text = op_sequence(seed, n, erase_frac, keyspace)
    want, _ = ref_orders(text)
    got, _ = clone_orders(text)
    assert len(want) > 0
    assert got == want


@pytest.mark.parametrize("seed,n,erase_frac,keyspace", CASES)
def test_engine_slot_model_equals_reference(ref_probe, seed, n, erase_frac, keyspace):
    from test_cpu_sim import SIM, SIM_DIR, CSRC  # noqa: F401

    if not os.path.exists(SIM):
        pytest.skip("CPU harness not built (tests/test_cpu_sim.py builds it)")
    text = op_sequence(seed, n, erase_frac, keyspace)
    want, _ = ref_orders(text)
    g = subprocess.run([SIM, "--rh-ints"], input=text, capture_output=True, text=True, check=True)
    assert [int(x) for x in g.stdout.split()] == want


def test_char_map_order_equals_reference(ref_probe):
    """robin_hood<char,int> iteration order (findconseq's ties, VariationUtils.h:431-443) for random
    insertion sequences with repeats, against the clone and the engine's replay."""
    from test_cpu_sim import SIM

    rnd = random.Random(9)
    alphabet = [ord(c) for c in "ACGTN"] + list(range(0, 256, 13))
    for t in range(300):
        seq = [rnd.choice(alphabet) for _ in range(rnd.randint(1, 40))]
        text = "".join(f"c {b}\n" for b in seq)
        _, want = ref_orders(text)
        _, got = clone_orders(text)
        assert got == want, seq
        if os.path.exists(SIM):
            distinct = list(dict.fromkeys(seq))
            g = subprocess.run([SIM, "--rh-chars"] + [str(b) for b in distinct], capture_output=True, text=True,
                               check=True)
            assert [int(x) for x in g.stdout.split()] == want, seq
