"""ctypes view of the CPU oracle (oracle/qsm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Nothing in the product library
imports it.
"""
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
    "SX", "SY", "SZ", "DHAT", "MU",
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


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_r2r_1d.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_r2r_1d_direct.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_r2r_3d.argtypes = [ctypes.c_int] * 6 + [dp, dp]
        L.orc_wavenumbers.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
        L.orc_spec_grad.argtypes = [ctypes.c_int] * 3 + [dp, dp, dp, ctypes.c_int, dp, dp]
        L.orc_halo_sx_slab.argtypes = [ctypes.c_int] * 3
        L.orc_project.argtypes = [ctypes.POINTER(OrcParams)] + [dp] * 8
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
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
        _lib = L
    return _lib


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
