"""Independent numpy restatements of the reference's initial conditions,
used by the tests as the input the oracle starts from."""
import numpy as np


def focal_conic_psi(shape, amp, q0, nw, fc):
    """The focal-conic initial condition of quasi.cpp:543-589 (aE = bE = fc,
    mid = Nx/2 in integer division, z from the bottom), written with numpy
    array operations: layers A cos(q0 dz r'), r' = sqrt(((bE/aE)(ds - mid))^2
    + zs^2), inside the quarter-ellipse tube around the circle ds = mid,
    flat layers A cos(q0 zs dz) outside that circle for zs < bE, and 0 above
    k = bE + 1 and in the rest of the cylinder."""
    nx, ny, nz = shape
    dz = 2 * np.pi / (q0 * nw)
    mid = float(nx // 2)
    a = b = float(fc)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    xs, ys, zs = i - mid, j - mid, k.astype(float)
    ds = np.sqrt(xs * xs + ys * ys)
    inside = ds < mid
    tube = np.sqrt(((ds - mid) / a) ** 2 + (zs / b) ** 2) <= 1
    psi = np.where(inside & tube, amp * np.cos(q0 * dz * np.sqrt(((b / a) * (ds - mid)) ** 2 + zs ** 2)),
                   0.0)
    psi = np.where(~inside & (np.abs(zs) < b), amp * np.cos(q0 * zs * dz), psi)
    return np.where(k < b + 1, psi, 0.0)
