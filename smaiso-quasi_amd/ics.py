"""Run configurations and initial conditions (host-side input preparation).

Each named workload fixes the reference parameters (quasi.cpp:159-219 names)
and builds psi(x) on the host in FP64, as the reference's main() does before
its setup transforms (quasi.cpp:504-589, quasi-stb.cpp:529-552).

Parameter mapping of the workloads onto the reference constants:
  eps      the workloads' epsilon = 0.5 is taken with the reference's sign,
           eps = -0.001*argv[2] <= 0 (quasi.cpp:186-187): eps = -0.5
  rho      rho = kp sqrt(q0^2 psi^2 + |grad psi|^2) + rho0 (quasi.cpp:1243-1247)
           with rho_i = rho0 = 0.9 in the isotropic phase and
           rho_s = kp A + rho0 = 1.0 for A = 0.5, so kp = 0.2
  rho_m    the density of the implicit linear operator (C1, C2 of
           quasi.cpp:883-884).  The reference sets rho_m = rho_s/2 with
           rho0 = 0.005, i.e. the midpoint of its density range [rho0, rho_s];
           the same midpoint here is (rho_s + rho_i)/2 = 0.95.  (rho_m = 0.5
           leaves (rho - rho_m) aLin, the explicitly treated part, too stiff
           for dt = 0.01: the step blows up within 20 steps.)
  nu       argv[1] of the reference, not fixed by the workloads: 1.0
  dt       0.01; nw = 8 points per wavelength (dx = 2 pi/8), q0 = 1
  z        the reference samples its ICs at k*dz (quasi.cpp:576)
"""
import math

import numpy as np

BASE = dict(
    variant=0, compat_ranks=0, nw=8, divv_switch=200, hydro=1,
    q0=1.0, alpha=1.0, beta=2.0, gamma=1.0, eps=-0.5, nu=1.0,
    kp=0.2, rho0=0.9, amp=0.5, rho_m=0.95, dt=0.01, filter=1.57 * 1.57 * 8, growth_q=0.0,
)

WORKLOADS = {
    # plain method, 64^3, FP64, 100 steps; the reference ran with mpirun -np 4
    "interface64": dict(shape=(64, 64, 64), steps=100, compat_ranks=4, seed=1, ic="interface"),
    # FP64 drop-in path, 256^3, 1000 steps
    "interface256": dict(shape=(256, 256, 256), steps=1000, compat_ranks=0, seed=1, ic="interface"),
    # mixed vs FP64 at 512^3: tilted, undulated slab
    "tilted512": dict(shape=(512, 512, 512), steps=1000, compat_ranks=0, seed=3, ic="tilted"),
    # large mixed run: spherical smectic domain, 1024 x 1024 x 512
    "sphere1024": dict(shape=(1024, 1024, 512), steps=500, compat_ranks=0, seed=5, ic="sphere",
                       rho0=0.85, kp=0.3),
}


def params_dict(name, **over):
    w = WORKLOADS[name]
    d = dict(BASE)
    d["nx"], d["ny"], d["nz"] = w["shape"]
    d["compat_ranks"] = w["compat_ranks"]
    for k in ("rho0", "kp"):
        if k in w:
            d[k] = w[k]
    d["rho_m"] = (d["kp"] * d["amp"] + 2 * d["rho0"]) / 2
    d.update(over)
    return d


def _axis(n, nw=8, q0=1.0):
    dx = 2 * math.pi / (q0 * nw)
    return np.arange(n) * dx, n * dx


def interface_psi(shape, amp=0.5, q0=1.0, nw=8, noise=1e-4, seed=1):
    """Flat smectic slab below Lz/2 blended into the isotropic phase:
    psi = A cos(q0 z) (1 - tanh((z - Lz/2)/2))/2 + N(0, noise^2)."""
    nx, ny, nz = shape
    z, lz = _axis(nz, nw, q0)
    prof = 0.5 * (1.0 - np.tanh((z - lz / 2) / 2.0))
    line = amp * np.cos(q0 * z) * prof
    psi = np.broadcast_to(line, shape).copy()
    if noise:
        psi += noise * np.random.default_rng(seed).standard_normal(shape)
    return psi


def tilted_psi(shape, amp=0.5, q0=1.0, nw=8, noise=1e-4, seed=3, tilt_deg=15.0):
    """Smectic slab tilted by 15 degrees about x: layer normal
    n = (0, sin t, cos t); the interface plane has the same normal, passes
    through the box centre and is displaced along n by 2 dx sin(2 pi x/(Lx/4))."""
    nx, ny, nz = shape
    x, lx = _axis(nx, nw, q0)
    y, ly = _axis(ny, nw, q0)
    z, lz = _axis(nz, nw, q0)
    dx = x[1] - x[0] if nx > 1 else 2 * math.pi / (q0 * nw)
    t = math.radians(tilt_deg)
    u = (z[None, None, :] * math.cos(t) + y[None, :, None] * math.sin(t))
    u0 = (lz / 2) * math.cos(t) + (ly / 2) * math.sin(t)
    shift = 2 * dx * np.sin(2 * math.pi * x / (lx / 4))
    d = u - u0 - shift[:, None, None]
    psi = amp * np.cos(q0 * u) * 0.5 * (1.0 - np.tanh(d / 2.0))
    psi = np.ascontiguousarray(np.broadcast_to(psi, shape))
    if noise:
        psi = psi + noise * np.random.default_rng(seed).standard_normal(shape)
    return psi


def sphere_psi(shape, amp=0.5, q0=1.0, nw=8, radius_pts=150, noise=1e-4, seed=5):
    """Spherical smectic domain of radius 150 dx with radial lamellae
    psi = A cos(q0 r), blended by tanh into the isotropic phase."""
    nx, ny, nz = shape
    x, lx = _axis(nx, nw, q0)
    y, ly = _axis(ny, nw, q0)
    z, lz = _axis(nz, nw, q0)
    dx = 2 * math.pi / (q0 * nw)
    out = np.empty(shape)
    r0 = radius_pts * dx
    for i in range(nx):
        r = np.sqrt((x[i] - lx / 2) ** 2 + (y[:, None] - ly / 2) ** 2 + (z[None, :] - lz / 2) ** 2)
        out[i] = amp * np.cos(q0 * r) * 0.5 * (1.0 - np.tanh((r - r0) / 2.0))
    if noise:
        out += noise * np.random.default_rng(seed).standard_normal(shape)
    return out


def workload_psi(name, shape=None, **kw):
    w = WORKLOADS[name]
    shape = shape or w["shape"]
    ic = w["ic"]
    if ic == "interface":
        return interface_psi(shape, seed=w["seed"], **kw)
    if ic == "tilted":
        return tilted_psi(shape, seed=w["seed"], **kw)
    if ic == "sphere":
        return sphere_psi(shape, seed=w["seed"], **kw)
    raise KeyError(ic)


def stb_psi(shape, amp=math.sqrt(0.12), q0=1.0, nw=8, qi=0.5, phi=0.1):
    """quasi-stb.cpp:546-547: psi = A cos(q0 k dx) + (A phi/2) cos(q0 k dx + Qi i dx)
    - (A phi/2) cos(q0 k dx - Qi i dx), sampled at integer indices."""
    nx, ny, nz = shape
    dx = 2 * math.pi / (q0 * nw)
    i = np.arange(nx)[:, None, None]
    k = np.arange(nz)[None, None, :]
    psi = (amp * np.cos(q0 * k * dx) + (amp * phi / 2) * np.cos(q0 * k * dx + qi * i * dx)
           - (amp * phi / 2) * np.cos(q0 * k * dx - qi * i * dx))
    return np.ascontiguousarray(np.broadcast_to(psi, shape))


# stability ensemble (quasi-stb.cpp workload): pure smectic, CN/AB2 step,
# constants of quasi-stb.cpp:149-220 with eps near coexistence
# (eps_c = -0.16875 beta^2/gamma = -0.009 for beta = 0.4, gamma = 3, where the
# reference's Amp = sqrt(0.9 beta/gamma) = sqrt(0.12) is the one-mode amplitude)
STB_BASE = dict(
    variant=1, compat_ranks=0, nw=8, divv_switch=200, hydro=1,
    q0=1.0, alpha=1.0, beta=0.4, gamma=3.0, eps=-0.009, nu=0.1,
    kp=0.5, rho0=0.1, amp=math.sqrt(0.12), rho_m=0.5, dt=0.0005, filter=1.57 * 1.57 * 8,
    growth_q=0.5,
)
ENSEMBLE_EDGES = (32, 48, 64, 96, 128)


def ensemble_members(count=96, seed=4, pert_amp=1e-3):
    """`count` instances: each grid edge drawn uniformly from ENSEMBLE_EDGES
    per axis, perturbation wavenumber Qi ~ U(0.05, 1.0), perturbation amplitude
    A phi / 2 = pert_amp.  Returns [(params dict, psi)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx, ny, nz = (int(v) for v in rng.choice(ENSEMBLE_EDGES, size=3))
        qi = float(rng.uniform(0.05, 1.0))
        d = dict(STB_BASE)
        d.update(nx=nx, ny=ny, nz=nz, growth_q=qi)
        phi = 2 * pert_amp / d["amp"]
        out.append((d, stb_psi((nx, ny, nz), amp=d["amp"], qi=qi, phi=phi)))
    return out
