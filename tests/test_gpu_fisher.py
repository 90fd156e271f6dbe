"""The device's FP64 Fisher test (the scoring kernels' fisher_two_sided, through rv_debug_fisher)
against the oracle's restatement of htslib's kt_fisher_exact, to 1e-12 relative on every output
(two-sided p, left and right tails).

Tables: the 2,196 golden tables (tests/golden/fisher_bruteforce.json, pinned to exact rationals in
test_oracle_fisher.py), plus tables shaped like the deep workloads' tests: strand tables
(ref fwd, ref rev, alt fwd, alt rev) at 200-5000x and tumor/normal tables (alt, ref - alt) at
200x/100x, with low and high allele fractions."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ORACLE, ROOT

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden", "fisher_bruteforce.json")


def oracle_fisher(tables, tmp_path):
    inp = tmp_path / "tables.txt"
    out = tmp_path / "p.txt"
    inp.write_text("".join(f"{a} {b} {c} {d}\n" for a, b, c, d in tables))
    subprocess.run([ORACLE, "--fisher-table", str(inp), str(out)], check=True)
    return np.array([[float(x) for x in line.split()] for line in out.read_text().splitlines()])


def deep_tables(seed=5):
    rnd = np.random.default_rng(seed)
    out = []
    for depth in (200, 500, 1000, 2500, 5000):
        for _ in range(400):
            cov = rnd.integers(depth // 2, depth * 3 // 2)
            vaf = rnd.choice([0.005, 0.02, 0.1, 0.3, 0.5])
            alt = rnd.binomial(cov, vaf)
            f_ref = rnd.binomial(cov - alt, 0.5)
            f_alt = rnd.binomial(alt, rnd.choice([0.5, 0.1, 0.9]))
            out.append([f_ref, cov - alt - f_ref, f_alt, alt - f_alt])
    for _ in range(1000):  # tumor vs normal: (tumor alt, tumor ref - alt, normal alt, normal ref - alt)
        t, n = rnd.integers(120, 300), rnd.integers(60, 150)
        ta, na = rnd.binomial(t, rnd.uniform(0.01, 0.2)), rnd.binomial(n, rnd.choice([0.0, 0.002, 0.3]))
        out.append([ta, t - ta, na, n - na])
    return out


@pytest.fixture(scope="module")
def engine(tools):
    from rabbitvar_amd.engine import Engine

    e = Engine(device=0, sample="S1", fisher=1)
    yield e
    e.close()


def check(engine, tables, tmp_path):
    got = engine.fisher(np.array(tables))
    want = oracle_fisher(tables, tmp_path)
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-300)
    worst = np.unravel_index(np.argmax(rel), rel.shape)
    print(f"fisher: {len(tables)} tables, max relative difference {rel.max():.3e} at {tables[worst[0]]}", flush=True)
    assert np.all(rel <= 1e-12), (rel.max(), tables[worst[0]], got[worst[0]], want[worst[0]])
    # the printed form (%f) is the same
    assert [f"{x:f}" for x in got[:, 0]] == [f"{x:f}" for x in want[:, 0]]
    return got


def test_golden_tables(engine, tmp_path):
    with open(GOLDEN) as f:
        tables = [g["t"] for g in json.load(f)]
    check(engine, tables, tmp_path)


def test_deep_workload_tables(engine, tmp_path):
    got = check(engine, deep_tables(), tmp_path)
    assert (got[:, 0] < 1e-6).sum() > 50  # p-values that print as 0.000000 are compared too


def test_degenerate_tables(engine, tmp_path):
    tables = [[0, 0, 0, 0], [0, 0, 5, 7], [3, 0, 0, 0], [0, 4, 0, 9], [1, 0, 0, 1], [5000, 0, 0, 5000]]
    got = check(engine, tables, tmp_path)
    np.testing.assert_array_equal(got[:4, 0], [1.0, 1.0, 1.0, 1.0])
