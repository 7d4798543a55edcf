"""Pins of the oracle's r2r transforms (CPU).

The oracle's fast path (2N-point complex FFT of the even extension, DST by the
sign/reversal identities) is checked against
  - the O(N^2) direct evaluation of FFTW's published definitions,
  - SciPy golden vectors committed in tests/golden/r2r_golden.npz,
  - closed forms: single-mode inputs, constant input, inverse pairs (2N),
  - 3D separability and the 8 N^3 round trip that sets quasi.cpp:219's scale.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "r2r_golden.npz")
KINDS = (4, 5, 8, 9)


def test_fast_matches_direct(oracle_mod):
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 5, 8, 12, 16, 24, 48, 64, 96, 100, 128):
        x = rng.standard_normal(n)
        for kind in KINDS:
            a = oracle_mod.r2r_1d(x, kind)
            b = oracle_mod.r2r_1d(x, kind, direct=True)
            assert rel_l2(a, b) < 1e-14, (n, kind)


def test_golden_vectors(oracle_mod):
    g = np.load(GOLDEN, allow_pickle=False)
    for n in g["lengths"]:
        x = g[f"x_{n}"]
        for kind in KINDS:
            ref = g[f"y_{n}_{kind}"]
            assert rel_l2(oracle_mod.r2r_1d(x, kind), ref) < 1e-14, (n, kind)
            if n <= 128:
                assert rel_l2(oracle_mod.r2r_1d(x, kind, direct=True), ref) < 1e-14


@pytest.mark.parametrize("n", (8, 12, 64))
def test_single_mode_closed_forms(oracle_mod, n):
    j = np.arange(n)
    for m in range(n):
        e = np.zeros(n)
        e[m] = 1.0
        # REDFT10 of a unit impulse at j=m: Y_k = 2 cos(pi (m+1/2) k / N)
        assert np.allclose(oracle_mod.r2r_1d(e, 5), 2 * np.cos(np.pi * (m + 0.5) * j / n), atol=1e-13)
        # REDFT01: Y_k = (m==0 ? 1 : 2 cos(pi m (k+1/2)/N))
        ref = np.ones(n) if m == 0 else 2 * np.cos(np.pi * m * (j + 0.5) / n)
        assert np.allclose(oracle_mod.r2r_1d(e, 4), ref, atol=1e-13)
        # RODFT10: Y_k = 2 sin(pi (m+1/2)(k+1)/N)
        assert np.allclose(oracle_mod.r2r_1d(e, 9), 2 * np.sin(np.pi * (m + 0.5) * (j + 1) / n), atol=1e-13)
        # RODFT01: Y_k = (m==N-1 ? (-1)^k : 2 sin(pi (m+1)(k+1/2)/N))
        ref = (-1.0) ** j if m == n - 1 else 2 * np.sin(np.pi * (m + 1) * (j + 0.5) / n)
        assert np.allclose(oracle_mod.r2r_1d(e, 8), ref, atol=1e-13)


def test_constant_and_inverse_pairs(oracle_mod):
    n = 48
    y = oracle_mod.r2r_1d(np.ones(n), 5)
    assert abs(y[0] - 2 * n) < 1e-12 and np.max(np.abs(y[1:])) < 1e-12
    x = np.random.default_rng(3).standard_normal(n)
    for fwd, inv in ((5, 4), (9, 8)):
        back = oracle_mod.r2r_1d(oracle_mod.r2r_1d(x, fwd), inv)
        assert rel_l2(back, 2 * n * x) < 1e-14
        back = oracle_mod.r2r_1d(oracle_mod.r2r_1d(x, inv), fwd)
        assert rel_l2(back, 2 * n * x) < 1e-14


def test_3d_separable_and_roundtrip(oracle_mod):
    rng = np.random.default_rng(11)
    a = rng.standard_normal((6, 10, 12))
    kinds = (9, 4, 8)
    got = oracle_mod.r2r_3d(a, kinds)
    ref = a.copy()
    for ax, k in enumerate(kinds):
        ref = np.apply_along_axis(lambda v: oracle_mod.r2r_1d(v, k, direct=True), ax, ref)
    assert rel_l2(got, ref) < 1e-13
    fwd = oracle_mod.r2r_3d(a, (5, 9, 5))
    back = oracle_mod.r2r_3d(fwd, (4, 8, 4))
    assert rel_l2(back, 8 * 6 * 10 * 12 * a) < 1e-14
