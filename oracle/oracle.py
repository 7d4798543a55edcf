"""ctypes view of the CPU oracle (oracle/qsm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Nothing in the product library
imports it.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqsm_oracle.so")

REDFT01, REDFT10, RODFT01, RODFT10 = 4, 5, 8, 9

FIELDS = [
    "PSI", "RHO", "VX", "VY", "VZ", "P", "DIVV", "PSIHAT", "F", "G", "GX", "GY", "GZ",
    "RX", "RY", "RZ", "LX", "LY", "LZ", "RHOHAT", "NHAT", "PSIHAT_OLD", "PSI_OLD",
    "SX", "SY", "SZ", "DHAT", "MU", "DTPSI", "RHO_OLD",
]
FIELD = {n: i for i, n in enumerate(FIELDS)}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("variant", ctypes.c_int), ("compat_ranks", ctypes.c_int), ("nw", ctypes.c_int),
        ("divv_switch", ctypes.c_int), ("hydro", ctypes.c_int),
        ("q0", ctypes.c_double), ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
        ("gamma", ctypes.c_double), ("eps", ctypes.c_double), ("nu", ctypes.c_double),
        ("kp", ctypes.c_double), ("rho0", ctypes.c_double), ("amp", ctypes.c_double),
        ("rho_m", ctypes.c_double), ("dt", ctypes.c_double), ("filter", ctypes.c_double),
        ("growth_q", ctypes.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def load_library(path):
    """Load an oracle build (this directory's, or a test's variant of it) and
    declare its signatures."""
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    pp = ctypes.POINTER(OrcParams)
    sz = ctypes.c_size_t
    L.orc_r2r_1d.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
    L.orc_r2r_1d_direct.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
    L.orc_r2r_3d.argtypes = [ctypes.c_int] * 6 + [dp, dp]
    L.orc_wavenumbers.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
    L.orc_spec_grad.argtypes = [ctypes.c_int] * 3 + [dp, dp, dp, ctypes.c_int, dp, dp]
    L.orc_halo_sx_slab.argtypes = [ctypes.c_int] * 3
    L.orc_project.argtypes = [pp] + [dp] * 8
    L.orc_density_law.argtypes = [pp, sz] + [dp] * 5
    L.orc_density_filter.argtypes = [pp, dp, dp, dp]
    L.orc_divergence.argtypes = [pp, sz] + [dp] * 9
    L.orc_stress.argtypes = [sz] + [dp] * 12
    L.orc_nonlinear.argtypes = [pp, sz, ctypes.c_int] + [dp] * 15
    L.orc_stb_stress.argtypes = [pp, sz] + [dp] * 9
    L.orc_stb_nonlinear.argtypes = [pp, sz] + [dp] * 10
    L.orc_stb_growth_mode.argtypes = [pp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.orc_create.argtypes = [pp]
    L.orc_create.restype = ctypes.c_void_p
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_set_psi.argtypes = [ctypes.c_void_p, dp]
    L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    L.orc_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    L.orc_diag.argtypes = [ctypes.c_void_p, dp]
    L.orc_get_step.argtypes = [ctypes.c_void_p]
    L.orc_get_growth.argtypes = [ctypes.c_void_p]
    L.orc_get_growth.restype = ctypes.c_double
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load_library(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using_library(path):
    """Run the enclosed calls against another build of the oracle (the
    mutation tests plant a mistake in a copy of qsm_oracle.c and check that a
    pin catches it)."""
    global _lib
    saved = lib()
    _lib = load_library(path)
    try:
        yield _lib
    finally:
        _lib = saved


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _opt(a):
    return _p(a) if a is not None else ctypes.POINTER(ctypes.c_double)()


def r2r_1d(x, kind, direct=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    f = lib().orc_r2r_1d_direct if direct else lib().orc_r2r_1d
    if f(x.size, kind, _p(x), _p(y)) != 0:
        raise ValueError("bad kind")
    return y


def r2r_3d(a, kinds):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty_like(a)
    nx, ny, nz = a.shape
    lib().orc_r2r_3d(nx, ny, nz, kinds[0], kinds[1], kinds[2], _p(a), _p(out))
    return out


def wavenumbers(n, length):
    q = np.empty(n)
    s = np.empty(n)
    lib().orc_wavenumbers(n, length, _p(q), _p(s))
    return q, s


def spec_grad(fhat, s_tables, axis):
    fhat = np.ascontiguousarray(fhat, dtype=np.float64)
    out = np.empty_like(fhat)
    nx, ny, nz = fhat.shape
    sx, sy, sz = [np.ascontiguousarray(t, dtype=np.float64) for t in s_tables]
    lib().orc_spec_grad(nx, ny, nz, _p(sx), _p(sy), _p(sz), axis, _p(fhat), _p(out))
    return out


def halo_sx_slab(nx, compat_ranks, i):
    return lib().orc_halo_sx_slab(nx, compat_ranks, i)


def make_params(**kw):
    p = OrcParams()
    for name, _ in OrcParams._fields_:
        if name in kw:
            setattr(p, name, kw[name])
    return p


def project(params, Sx, Sy, Sz, dhat):
    shp = Sx.shape
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (Sx, Sy, Sz)]
    dh = None if dhat is None else np.ascontiguousarray(dhat, dtype=np.float64)
    outs = [np.empty(shp) for _ in range(4)]
    lib().orc_project(ctypes.byref(params), *[_p(a) for a in arrs], _opt(dh),
                      *[_p(o) for o in outs])
    return outs  # phat, vxh, vyh, vzh


def _arr(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else np.ascontiguousarray(np.broadcast_to(a, shape))


def _fields(shape, *arrs):
    return [_arr(a, shape) for a in arrs]


def density_law(params, psi, g):
    """rho* of quasi.cpp:1237-1247 from psi and grad psi (three arrays)."""
    psi, gx, gy, gz = _fields(np.shape(psi), psi, *g)
    out = np.empty_like(psi)
    lib().orc_density_law(ctypes.byref(params), psi.size, _p(psi), _p(gx), _p(gy), _p(gz), _p(out))
    return out


def density_filter(params, rho_star):
    """(rhohat, rho) of quasi.cpp:1249-1268."""
    rs = _arr(rho_star)
    rh, rho = np.empty_like(rs), np.empty_like(rs)
    lib().orc_density_filter(ctypes.byref(params), _p(rs), _p(rh), _p(rho))
    return rh, rho


def divergence(params, rho, rho_old, v, r):
    """D of quasi.cpp:1273-1287 (real space, before its transform)."""
    shape = np.shape(rho)
    a = _fields(shape, rho, rho_old, *v, *r)
    out = np.empty(shape)
    lib().orc_divergence(ctypes.byref(params), out.size, *[_p(x) for x in a], _p(out))
    return out


def stress(rho, G, F, g, l):
    """(S_x, S_y, S_z, G <- rho G) of quasi.cpp:1298-1317."""
    shape = np.shape(rho)
    rho, F, gx, gy, gz, lx, ly, lz = _fields(shape, rho, F, *g, *l)
    Gc = _arr(G, shape).copy()
    S = [np.empty(shape) for _ in range(3)]
    lib().orc_stress(rho.size, _p(rho), _p(Gc), _p(F), _p(gx), _p(gy), _p(gz), _p(lx), _p(ly),
                     _p(lz), *[_p(x) for x in S])
    return S, Gc


def nonlinear(params, step, rho, psi, F, G, v, g, P, r):
    """N of quasi.cpp:1905-1931 at step `step` (real space)."""
    shape = np.shape(rho)
    a = _fields(shape, rho, psi, F, G, *v, *g, P, *r)
    out = np.empty(shape)
    lib().orc_nonlinear(ctypes.byref(params), out.size, step, *[_p(x) for x in a], _p(out))
    return out


def stb_stress(params, sh, psi, g):
    """(mu, [S_x, S_y, S_z]) of quasi-stb.cpp:1284-1307."""
    shape = np.shape(sh)
    a = _fields(shape, sh, psi, *g)
    mu = np.empty(shape)
    S = [np.empty(shape) for _ in range(3)]
    lib().orc_stb_stress(ctypes.byref(params), mu.size, *[_p(x) for x in a], _p(mu),
                         *[_p(x) for x in S])
    return mu, S


def stb_nonlinear(params, psi, mu, v, g):
    """(N, dtpsi) of quasi-stb.cpp:1893-1923."""
    shape = np.shape(psi)
    a = _fields(shape, psi, mu, *v, *g)
    N, dt = np.empty(shape), np.empty(shape)
    lib().orc_stb_nonlinear(ctypes.byref(params), N.size, *[_p(x) for x in a], _p(N), _p(dt))
    return N, dt


def stb_growth_mode(params):
    """(i*, k*) of quasi-stb.cpp:2147-2150, or None off the grid."""
    i, k = ctypes.c_int(), ctypes.c_int()
    rc = lib().orc_stb_growth_mode(ctypes.byref(params), ctypes.byref(i), ctypes.byref(k))
    return (i.value, k.value) if rc == 0 else None


class State:
    """One oracle run: create, set_psi (setup), step(n), fields, diagnostics."""

    def __init__(self, params):
        self.params = params
        self.shape = (params.nx, params.ny, params.nz)
        self.h = lib().orc_create(ctypes.byref(params))
        if not self.h:
            raise MemoryError("orc_create failed")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_psi(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(self.shape)
        lib().orc_set_psi(self.h, _p(psi))

    def step(self, n=1):
        lib().orc_step(self.h, n)

    @property
    def nstep(self):
        return lib().orc_get_step(self.h)

    @property
    def growth(self):
        return lib().orc_get_growth(self.h)

    def field(self, name):
        out = np.empty(self.shape)
        lib().orc_get_field(self.h, FIELD[name], _p(out))
        return out

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        lib().orc_set_field(self.h, FIELD[name], _p(a))

    def diag(self):
        out = np.empty(3)
        lib().orc_diag(self.h, _p(out))
        return {"L1": out[0], "mass": out[1], "energy": out[2]}
