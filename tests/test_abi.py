"""C ABI checks that need no GPU: libqsm.so loads, exports every function
include/qsm.h declares, keeps FFTW's kind values, and rejects bad arguments
before touching the device.  Also checks that the product library neither
links nor embeds the oracle."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qsm.h")
LIB = os.path.join(ROOT, "smaiso-quasi_amd", "libqsm.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qsm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for n in ("qsm_r2r3d_plan", "qsm_r2r3d_execute", "qsm_r2r3d_destroy", "qsm_create",
              "qsm_set_psi", "qsm_step", "qsm_get_field", "qsm_get_diag", "qsm_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol(qsm_mod):
    L = ctypes.CDLL(LIB)
    for n in declared():
        assert hasattr(L, n), n
    assert set(declared()) == set(qsm_mod.SIGNATURES), "binding and header disagree"


def test_no_oracle_in_product():
    out = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert "orc_" not in out
    deps = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "oracle" not in deps


def test_kind_values_are_fftw(qsm_mod):
    text = open(HEADER).read()
    vals = dict(re.findall(r"#define (QSM_R[EO]DFT\d\d) (\d+)", text))
    assert vals == {"QSM_REDFT01": "4", "QSM_REDFT10": "5", "QSM_RODFT01": "8", "QSM_RODFT10": "9"}
    assert (qsm_mod.REDFT01, qsm_mod.REDFT10, qsm_mod.RODFT01, qsm_mod.RODFT10) == (4, 5, 8, 9)


def test_default_params(qsm_mod):
    p = qsm_mod.default_params(qsm_mod.QUASI)
    assert (p.nx, p.ny, p.nz) == (256, 256, 256)
    assert p.divv_switch == 200 and p.dt == 0.001 and p.beta == 2.0 and p.gamma == 1.0
    assert abs(p.filter - 1.57 * 1.57 * 8) < 1e-15
    s = qsm_mod.default_params(qsm_mod.STB)
    assert (s.beta, s.gamma, s.dt, s.nz) == (0.4, 3.0, 0.0005, 512)
    assert qsm_mod.lib().qsm_version().startswith(b"qsm")


def test_argument_errors_before_device(qsm_mod):
    with pytest.raises(qsm_mod.QsmError) as e:
        qsm_mod.Plan((8, 8, 8), (5, 5, 3))  # REDFT00 is not a kind the path uses
    assert e.value.code == -1
    with pytest.raises(qsm_mod.QsmError) as e:
        qsm_mod.Plan((8, 8, 100), (5, 5, 5))  # no pass kernel for 100
    assert e.value.code == -2
    p = qsm_mod.default_params(qsm_mod.QUASI, nx=64, ny=64, nz=64, compat_ranks=3)
    with pytest.raises(qsm_mod.QsmError) as e:
        qsm_mod.Sim(p)
    assert e.value.code == -1 and "compat_ranks" in str(e.value)
    p = qsm_mod.default_params(qsm_mod.QUASI, nx=64, ny=64, nz=64, dt=0.0)
    with pytest.raises(qsm_mod.QsmError):
        qsm_mod.Sim(p)
    with pytest.raises(qsm_mod.QsmError):
        qsm_mod.Sim(p, prec=7)
