"""Python view of libqsm.so (include/qsm.h) through ctypes.

This is the host-side mirror used by the tests, bench.py and scripts.  It has
no compute of its own: every call goes to the HIP library, and loading fails
loudly when libqsm.so has not been built.

The names follow the reference program quasi/quasi.cpp: a run is a set of
parameters (quasi.cpp:159-219), an initial psi (quasi.cpp:504-589), setup
(quasi.cpp:875-951), steps (quasi.cpp:1014-1965) and diagnostics
(quasi.cpp:1969-2031).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqsm.so")

QSM_OK = 0
ERRORS = {-1: "ARG", -2: "UNSUPPORTED", -3: "HIP", -4: "NONFINITE", -5: "STATE", -6: "NOMEM"}

REDFT01, REDFT10, RODFT01, RODFT10 = 4, 5, 8, 9
FP64, MIXED = 0, 1
PLAN_MEASURE = 1  # QSM_PLAN_MEASURE
QUASI, STB = 0, 1

# names of the fused load / store operator codes (qsm_ops.hpp LoadCode,
# StoreCode) as the timing entry points report them
LOAD_OPS = ["plain", "zero", "coef_f", "coef_g", "grad_x", "grad_y", "grad_z", "lapgrad_x",
            "lapgrad_y", "lapgrad_z", "sh", "rhostar", "divv",
            "stress", "pressure", "vel0", "vel1", "vel2", "scale", "nonlin", "nonlin_p",
            "stress_stb", "nonlin_stb", "nonlin_stb0", "stress3", "stress_stb3", "stress3_divv",
            "nonlin_py"]
STORE_OPS = ["plain", "filter", "update", "update_stb", "filter_x", "update_x", "update_stb_x"]

FIELDS = [
    "PSI", "RHO", "VX", "VY", "VZ", "P", "DIVV", "PSIHAT", "F", "G", "GX", "GY", "GZ",
    "RX", "RY", "RZ", "LX", "LY", "LZ", "RHOHAT", "NHAT", "PSIHAT_OLD", "PSI_OLD",
    "SX", "SY", "SZ", "DHAT", "DTPSI",
]
FIELD = {n: i for i, n in enumerate(FIELDS)}


class QsmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qsm error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(ctypes.Structure):
    """qsm_params; field order and meaning follow include/qsm.h."""
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

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Diag(ctypes.Structure):
    _fields_ = [
        ("L1", ctypes.c_double), ("mass", ctypes.c_double), ("energy", ctypes.c_double),
        ("interface_z", ctypes.c_double), ("growth", ctypes.c_double),
        ("step", ctypes.c_longlong),
    ]


# exported symbols and their signatures (checked by the CPU tests against include/qsm.h)
_VP, _DP, _IP = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
SIGNATURES = {
    "qsm_version": ([], ctypes.c_char_p),
    "qsm_last_error": ([], ctypes.c_char_p),
    "qsm_default_params": ([_IP, ctypes.POINTER(Params)], _IP),
    "qsm_r2r3d_plan": ([_IP] * 7 + [ctypes.c_uint, ctypes.POINTER(_VP)], _IP),
    "qsm_r2r3d_execute": ([_VP, _VP, _VP, _VP], _IP),
    "qsm_r2r3d_execute_host": ([_VP, _VP, _VP], _IP),
    "qsm_r2r3d_execute_host_noalloc": ([_VP, _VP, _VP], _IP),
    "qsm_r2r3d_destroy": ([_VP], _IP),
    "qsm_create": ([ctypes.POINTER(Params), _IP, ctypes.POINTER(_VP)], _IP),
    "qsm_set_psi": ([_VP, _DP], _IP),
    "qsm_step": ([_VP, _IP], _IP),
    "qsm_set_graphs": ([_VP, _IP], _IP),
    "qsm_run_l1": ([_VP, _IP, _IP, ctypes.c_double, ctypes.POINTER(_IP)], _IP),
    "qsm_get_field": ([_VP, _IP, _DP], _IP),
    "qsm_get_diag": ([_VP, ctypes.POINTER(Diag)], _IP),
    "qsm_destroy": ([_VP], _IP),
    "qsm_checkpoint_write": ([_VP, ctypes.c_char_p], _IP),
    "qsm_checkpoint_read": ([_VP, ctypes.c_char_p], _IP),
    "qsm_ens_create": ([ctypes.POINTER(Params), _IP, _IP, ctypes.POINTER(_VP)], _IP),
    "qsm_ens_set_psi": ([_VP, _IP, _DP], _IP),
    "qsm_ens_step": ([_VP, _IP], _IP),
    "qsm_ens_growth": ([_VP, _DP], _IP),
    "qsm_ens_get_field": ([_VP, _IP, _IP, _DP], _IP),
    "qsm_ens_get_diag": ([_VP, _IP, ctypes.POINTER(Diag)], _IP),
    "qsm_ens_set_graphs": ([_VP, _IP], _IP),
    "qsm_ens_set_graph_steps": ([_VP, _IP], _IP),
    "qsm_ens_timing_enable": ([_VP, _IP], _IP),
    "qsm_ens_timing_kernel": ([_VP, _IP] + [ctypes.POINTER(_IP)] * 6 +
                              [_DP, ctypes.POINTER(ctypes.c_longlong), _DP], _IP),
    "qsm_r2r3d_set_variant": ([_VP, _IP, _IP], _IP),
    "qsm_ens_destroy": ([_VP], _IP),
    "qsm_autotune": ([_VP, ctypes.c_char_p], _IP),
    "qsm_ens_autotune": ([_VP, ctypes.c_char_p], _IP),
    "qsm_set_variant": ([_VP, _IP, _IP, _IP], _IP),
    "qsm_timing_enable": ([_VP, _IP], _IP),
    "qsm_timing_select": ([_VP] + [_IP] * 5, _IP),
    "qsm_timing_get": ([_VP, _IP, _DP, ctypes.POINTER(ctypes.c_longlong), _DP], _IP),
    "qsm_timing_kernel": ([_VP, _IP] + [ctypes.POINTER(_IP)] * 5 +
                          [_DP, ctypes.POINTER(ctypes.c_longlong), _DP], _IP),
}

_lib = None


def lib():
    """Load libqsm.so; raises if it is missing (there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built: run `make -C {HERE}` "
                          "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != QSM_OK:
        raise QsmError(rc, lib().qsm_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_params(variant=QUASI, **overrides):
    p = Params()
    _check(lib().qsm_default_params(variant, ctypes.byref(p)))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Plan:
    """3D r2r plan: FFTW semantics (unnormalised, one kind per axis)."""

    def __init__(self, shape, kinds, prec=FP64, measure=False):
        self.shape = tuple(shape)
        self.prec = prec
        self.dtype = np.float64 if prec == FP64 else np.float32
        h = ctypes.c_void_p()
        flags = PLAN_MEASURE if measure else 0
        _check(lib().qsm_r2r3d_plan(*self.shape, *kinds, prec, flags, ctypes.byref(h)))
        self.h = h

    def set_variant(self, axis, variant):
        _check(lib().qsm_r2r3d_set_variant(self.h, axis, variant))

    def __call__(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if x.shape != self.shape:
            raise ValueError(f"shape {x.shape} != plan shape {self.shape}")
        out = np.empty_like(x)
        _check(lib().qsm_r2r3d_execute_host(self.h, x.ctypes.data_as(ctypes.c_void_p),
                                            out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def noalloc(self, x):
        """The transform through qsm_r2r3d_execute_host_noalloc (staging in
        the library's static device buffer, no device allocation)."""
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if x.shape != self.shape:
            raise ValueError(f"shape {x.shape} != plan shape {self.shape}")
        out = np.empty_like(x)
        _check(lib().qsm_r2r3d_execute_host_noalloc(self.h, x.ctypes.data_as(ctypes.c_void_p),
                                                    out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def close(self):
        if self.h:
            lib().qsm_r2r3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sim:
    """One simulation handle (qsm_create ... qsm_destroy)."""

    def __init__(self, params, prec=FP64):
        self.params = params
        self.prec = prec
        self.shape = (params.nx, params.ny, params.nz)
        h = ctypes.c_void_p()
        _check(lib().qsm_create(ctypes.byref(params), prec, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().qsm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_psi(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(self.shape)
        _check(lib().qsm_set_psi(self.h, _dp(psi)))

    def step(self, n=1):
        _check(lib().qsm_step(self.h, n))

    def graphs(self, enable=True):
        _check(lib().qsm_set_graphs(self.h, 1 if enable else 0))

    def run_l1(self, max_steps, step_l1=50, lim_l1=1e-6):
        done = ctypes.c_int()
        _check(lib().qsm_run_l1(self.h, max_steps, step_l1, lim_l1, ctypes.byref(done)))
        return done.value

    def field(self, name):
        out = np.empty(self.shape)
        _check(lib().qsm_get_field(self.h, FIELD[name], _dp(out)))
        return out

    def diag(self):
        d = Diag()
        _check(lib().qsm_get_diag(self.h, ctypes.byref(d)))
        return {n: getattr(d, n) for n, _ in Diag._fields_}

    def autotune(self, cache_path=None):
        _check(lib().qsm_autotune(self.h, os.fsencode(cache_path) if cache_path else None))

    def set_variant(self, axis, inv, variant):
        _check(lib().qsm_set_variant(self.h, axis, inv, variant))

    def checkpoint_write(self, path):
        _check(lib().qsm_checkpoint_write(self.h, os.fsencode(path)))

    def checkpoint_read(self, path):
        _check(lib().qsm_checkpoint_read(self.h, os.fsencode(path)))

    def timing(self, enable=True):
        _check(lib().qsm_timing_enable(self.h, 1 if enable else 0))

    def timing_select(self, axis, inv, ld, st, variant=0):
        """Time only the launches of one kernel (the fields of timing_kernel)."""
        _check(lib().qsm_timing_select(self.h, axis, inv, ld, st, variant))

    def timing_get(self, which=0):
        ms = ctypes.c_double()
        cnt = ctypes.c_longlong()
        by = ctypes.c_double()
        _check(lib().qsm_timing_get(self.h, which, ctypes.byref(ms), ctypes.byref(cnt),
                                    ctypes.byref(by)))
        return {"ms": ms.value, "launches": cnt.value, "bytes": by.value}

    def timing_kernel(self, rank=0):
        """rank-th kernel by device time: {axis, inv, ld, st, ms, launches, bytes}"""
        ax, inv, ld, st, var = (ctypes.c_int() for _ in range(5))
        ms, cnt, by = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        _check(lib().qsm_timing_kernel(self.h, rank, ctypes.byref(ax), ctypes.byref(inv),
                                       ctypes.byref(ld), ctypes.byref(st), ctypes.byref(var),
                                       ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(by)))
        return {"axis": ax.value, "inv": inv.value, "ld": ld.value, "st": st.value,
                "variant": var.value, "ms": ms.value, "launches": cnt.value, "bytes": by.value}


class Ensemble:
    """Independent instances advanced together (qsm_ens_*): one batched
    kernel per pencil length per launch of the step, one graph per step."""

    def __init__(self, params_list, prec=FP64):
        self.params = list(params_list)
        self.prec = prec
        arr = (Params * len(self.params))(*self.params)
        h = ctypes.c_void_p()
        _check(lib().qsm_ens_create(arr, len(self.params), prec, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().qsm_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, i):
        p = self.params[i]
        return (p.nx, p.ny, p.nz)

    def set_psi(self, i, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(self.shape(i))
        _check(lib().qsm_ens_set_psi(self.h, i, _dp(psi)))

    def step(self, n=1):
        _check(lib().qsm_ens_step(self.h, n))

    def graphs(self, enable=True):
        _check(lib().qsm_ens_set_graphs(self.h, 1 if enable else 0))

    def graph_steps(self, steps):
        _check(lib().qsm_ens_set_graph_steps(self.h, steps))

    def timing(self, enable=True):
        _check(lib().qsm_ens_timing_enable(self.h, 1 if enable else 0))

    def timing_kernel(self, rank=0):
        """The rank-th timed batched kernel: its pass fields, pencil length n,
        total ms, launches and algorithmic bytes."""
        iv = [ctypes.c_int() for _ in range(6)]
        ms, by = ctypes.c_double(), ctypes.c_double()
        cnt = ctypes.c_longlong()
        _check(lib().qsm_ens_timing_kernel(self.h, rank, *[ctypes.byref(x) for x in iv],
                                           ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(by)))
        keys = ("axis", "inv", "ld", "st", "variant", "n")
        out = {k: v.value for k, v in zip(keys, iv)}
        out.update(ms=ms.value, launches=cnt.value, bytes=by.value)
        return out

    def autotune(self, cache_path=None):
        _check(lib().qsm_ens_autotune(self.h, os.fsencode(cache_path) if cache_path else None))

    def growth(self):
        out = np.empty(len(self.params))
        _check(lib().qsm_ens_growth(self.h, _dp(out)))
        return out

    def field(self, i, name):
        out = np.empty(self.shape(i))
        _check(lib().qsm_ens_get_field(self.h, i, FIELD[name], _dp(out)))
        return out

    def diag(self, i):
        d = Diag()
        _check(lib().qsm_ens_get_diag(self.h, i, ctypes.byref(d)))
        return {n: getattr(d, n) for n, _ in Diag._fields_}
