"""Pins the CPU restatement of the realignment screen's network (oracle/mlp.py): its backward
pass against central finite differences of its loss (float64, no BF16 rounding), its BF16 rounding
against the bit patterns rounding to nearest even gives, and the BF16 forward pass against the
float64 one within BF16 precision."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp  # noqa: E402


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, mlp.NPARAM)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    X = rng.normal(0, 1, (40, mlp.F))
    t = (rng.random(40) < 0.3).astype(np.float64)
    P = _params(2)
    loss, g = mlp.loss_and_grad(P, X, t, pos_weight=3.0, bf=False)
    idx = np.concatenate([rng.choice(np.arange(mlp.W1, mlp.B1), 40), rng.choice(np.arange(mlp.B1, mlp.W2), 10),
                          rng.choice(np.arange(mlp.W2, mlp.B2), 40), rng.choice(np.arange(mlp.B2, mlp.W3), 10),
                          np.arange(mlp.W3, mlp.B3)[:10], [mlp.B3]])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (mlp.loss_and_grad(Pp, X, t, 3.0, bf=False)[0] - mlp.loss_and_grad(Pm, X, t, 3.0, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    # BF16 keeps 7 fraction bits: the spacing just above 1 is 2^-7
    vals = np.array([1.0, 1 + 2**-8, 1 + 1.5 * 2**-8, 1 + 3 * 2**-8, 1 + 2**-9, -2.5, 0.0], np.float32)
    got = mlp.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                      # halfway, ties to the even mantissa (0)
    assert got[2] == np.float32(1 + 2**-7)    # above halfway
    assert got[3] == np.float32(1 + 2**-6)    # halfway between mantissas 1 and 2: even is 2
    assert got[4] == 1.0                      # below halfway
    assert got[5] == -2.5 and got[6] == 0.0
    assert np.all(got.view(np.uint32) & 0xFFFF == 0)


def test_bf16_forward_close_to_exact():
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (200, mlp.F))
    P = _params(6, 0.2)
    a = mlp.scores(P, X, bf=True)
    b = mlp.scores(P, X, bf=False)
    assert np.max(np.abs(a - b)) < 0.02
    assert np.max(np.abs(a - b)) > 0  # the rounding is applied
