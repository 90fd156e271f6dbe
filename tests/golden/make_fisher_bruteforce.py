"""Writes fisher_bruteforce.json: 2x2 tables with the two-sided p-value of htslib's kt_fisher_exact
rule evaluated in exact rational arithmetic.

The probabilities of all tables with the observed margins are exact hypergeometric terms
(math.comb); the two tails are walked inward from both ends while p < 0.99999999 * p_obs, the table
where each walk stops is added when p < 1.00000001 * p_obs, and the sum is capped at 1.  Tables for
which some term lies within 1e-9 (relative) of either threshold are marked `near_tie`: a rounding
difference may legitimately move them across.

    python tests/golden/make_fisher_bruteforce.py
"""
import itertools
import json
import math
import os
import random
from fractions import Fraction

A = Fraction(99999999, 100000000)
B = Fraction(100000001, 100000000)


def exact(n11, n12, n21, n22):
    n1_, n_1, n = n11 + n12, n11 + n21, n11 + n12 + n21 + n22
    hi = min(n_1, n1_)
    lo = max(0, n1_ + n_1 - n)
    if lo == hi:
        return Fraction(1), False
    tot = math.comb(n, n_1)
    P = {i: Fraction(math.comb(n1_, i) * math.comb(n - n1_, n_1 - i), tot) for i in range(lo, hi + 1)}
    q = P[n11]
    near = any(abs(p - A * q) <= Fraction(1, 10**9) * q or abs(p - B * q) <= Fraction(1, 10**9) * q for p in P.values())
    left = Fraction(0)
    i = lo + 1
    p = P[lo]
    while p < A * q and i <= hi:
        left += p
        p = P[i]
        i += 1
    i -= 1
    if p < B * q:
        left += p
    right = Fraction(0)
    j = hi - 1
    p = P[hi]
    while p < A * q and j >= 0:
        right += p
        p = P[j] if j >= lo else Fraction(0)
        j -= 1
    j += 1
    if p < B * q:
        right += p
    two = min(Fraction(1), left + right)
    return two, near


def main():
    rnd = random.Random(12345)
    tables = [t for t in itertools.product(range(6), repeat=4)]
    for _ in range(600):
        n = rnd.choice([20, 60, 150, 400, 1000])
        t = [rnd.randint(0, n // 2) for _ in range(4)]
        tables.append(tuple(t))
    # strand tables typical of calls: ref fwd/rev vs alt fwd/rev
    for _ in range(300):
        rf, rr = rnd.randint(0, 120), rnd.randint(0, 120)
        af, ar = rnd.randint(0, 60), rnd.randint(0, 60)
        tables.append((rf, rr, af, ar))
    out = []
    for t in tables:
        two, near = exact(*t)
        out.append({"t": list(t), "two": [two.numerator, two.denominator] if two.denominator < 10**300 else None,
                    "two_f": float(two), "near_tie": near})
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "fisher_bruteforce.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(len(out), "tables")


if __name__ == "__main__":
    main()
