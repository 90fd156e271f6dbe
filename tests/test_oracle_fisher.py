"""Pins the oracle's two-sided Fisher exact test (htslib kt_fisher_exact, restated in
oracle/score.cpp) against exact rational arithmetic and against the mathematics it must obey."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ORACLE, ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "fisher_bruteforce.json")


def oracle_fisher(tables, tmp_path):
    inp = tmp_path / "tables.txt"
    out = tmp_path / "p.txt"
    inp.write_text("".join(f"{a} {b} {c} {d}\n" for a, b, c, d in tables))
    subprocess.run([ORACLE, "--fisher-table", str(inp), str(out)], check=True)
    return np.array([[float(x) for x in line.split()] for line in out.read_text().splitlines()])


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_two_sided_matches_exact_rationals(tools, golden, tmp_path):
    tables = [g["t"] for g in golden]
    got = oracle_fisher(tables, tmp_path)[:, 0]
    want = np.array([g["two_f"] for g in golden])
    near = np.array([g["near_tie"] for g in golden])
    rel = np.abs(got - want) / np.maximum(want, 1e-300)
    # htslib's method (Lanczos log-gamma, then a ratio recurrence) loses accuracy as the table total n
    # grows: the log-probabilities it exponentiates are sums of log-gammas of size ~n ln n, each
    # carrying a relative rounding error near 1e-16.  Below n = 500 it stays within 1e-12 relative
    # of the exact rational value; on the golden set's largest tables (n up to ~2000) within 5e-12.
    n = np.array([sum(t) for t in tables])
    small = ~near & (n < 500)
    assert np.all(rel[small] < 1e-12), (np.max(rel[small]), tables[int(np.argmax(np.where(small, rel, 0)))])
    assert np.all(rel[~near] < 5e-12), (np.max(rel[~near]), tables[int(np.argmax(np.where(near, 0, rel)))])
    assert small.sum() > 2000 and (~near & (n >= 500)).sum() > 100


def test_properties(tools, tmp_path):
    rnd = np.random.default_rng(7)
    t = rnd.integers(0, 80, size=(400, 4))
    base = oracle_fisher(t.tolist(), tmp_path)
    # swapping rows or columns of the table leaves the two-sided p-value unchanged
    rows = oracle_fisher(t[:, [2, 3, 0, 1]].tolist(), tmp_path)
    cols = oracle_fisher(t[:, [1, 0, 3, 2]].tolist(), tmp_path)
    trans = oracle_fisher(t[:, [0, 2, 1, 3]].tolist(), tmp_path)
    for other in (rows, cols, trans):
        np.testing.assert_allclose(other[:, 0], base[:, 0], rtol=1e-9)
    assert np.all(base[:, 0] <= 1.0) and np.all(base[:, 0] > 0)
    # one-sided tails: left(n11) + right(n11) = 1 + P(n11) >= 1
    assert np.all(base[:, 1] + base[:, 2] >= 1 - 1e-12)
    # degenerate margins: a single possible table has p = 1
    deg = oracle_fisher([[0, 0, 5, 7], [3, 0, 0, 0], [0, 4, 0, 9]], tmp_path)
    np.testing.assert_array_equal(deg[:, 0], [1.0, 1.0, 1.0])


def test_against_scipy_where_no_ties(tools, tmp_path):
    """scipy's fisher_exact sums the tables no more probable than the observed one (relative
    tolerance 1e-7); with no probabilities near that boundary both rules give the same sum."""
    stats = pytest.importorskip("scipy.stats")
    rnd = np.random.default_rng(11)
    t = rnd.integers(0, 40, size=(150, 4))
    got = oracle_fisher(t.tolist(), tmp_path)[:, 0]
    for row, g in zip(t.tolist(), got):
        want = stats.fisher_exact([[row[0], row[1]], [row[2], row[3]]]).pvalue
        assert abs(g - want) <= 1e-9 * max(want, 1e-300) or abs(g - want) < 1e-12, (row, g, want)
