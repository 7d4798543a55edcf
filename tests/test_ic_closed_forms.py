"""Closed-form checks of the initial-condition restatements in ic_ref (CPU)."""
import numpy as np
import pytest

import ic_ref


def test_focal_conic_ic_closed_forms():
    """Structure the restatement must have: zero from k = bE + 1 up; the flat
    layers A cos(q0 zs dz) on the circle ds = mid (the reference's outer
    branch takes ds >= mid) and in the corners outside it; zero on the
    cylinder axis (ds - mid = -mid < -aE: outside the tube); the tube fills
    a good part of the layers below bE."""
    shape, fc = (32, 32, 24), 10
    psi = ic_ref.focal_conic_psi(shape, 0.5, 1.0, 8, fc)
    dz = 2 * np.pi / 8
    assert np.all(psi[:, :, fc + 1:] == 0)
    for k in range(fc):
        assert psi[0, 16, k] == pytest.approx(0.5 * np.cos(k * dz), abs=1e-15)  # ds = mid
        assert psi[0, 0, k] == pytest.approx(0.5 * np.cos(k * dz), abs=1e-15)   # corner, ds > mid
        assert psi[16, 16, k] == 0.0                                            # axis
    assert np.count_nonzero(psi[:, :, :fc]) > 0.3 * psi[:, :, :fc].size
