"""GPU tests of the host side around the step: full-state checkpoint resume
(bit for bit) and the qsm_run driver's reference-format outputs."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "smaiso-quasi_amd", "qsm_run")


@pytest.mark.parametrize("prec", (0, 1))
def test_checkpoint_resume_is_bit_exact(qsm_mod, tmp_path, prec):
    import ics
    d = ics.params_dict("interface64", divv_switch=4)
    d["nx"], d["ny"], d["nz"] = (32, 32, 48)
    qp = qsm_mod.Params(**{k: d[k] for k, _ in qsm_mod.Params._fields_})
    a = qsm_mod.Sim(qp, prec)
    a.set_psi(ics.workload_psi("interface64", shape=(32, 32, 48)))
    a.step(5)
    path = str(tmp_path / "state.ckpt")
    a.checkpoint_write(path)
    a.step(4)
    b = qsm_mod.Sim(qp, prec)
    b.checkpoint_read(path)
    assert b.diag()["step"] == 5
    b.step(4)
    for f in ("PSI", "PSIHAT", "PSIHAT_OLD", "RHO", "VX", "VY", "VZ", "NHAT", "P"):
        assert np.array_equal(a.field(f), b.field(f)), f
    assert a.diag() == b.diag()
    other = qsm_mod.Params(**{k: d[k] for k, _ in qsm_mod.Params._fields_})
    other.dt = 0.02
    c = qsm_mod.Sim(other, prec)
    with pytest.raises(qsm_mod.QsmError):
        c.checkpoint_read(path)


def test_driver_reference_outputs(qsm_mod, tmp_path):
    out = tmp_path / "run"
    cmd = [RUN, "nx=32", "ny=32", "nz=32", "eps=-0.5", "beta=2", "gamma=1", "kp=0.2",
           "rho0=0.9", "amp=0.5", "rho_m=0.95", "dt=0.01", "nu=1", "compat_ranks=4",
           "steps=20", "step_l1=5", "step_save=2", f"out={out}", "ic=interface",
           f"ckpt_out={tmp_path / 'end.ckpt'}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    l1 = np.loadtxt(out / "L1.dat")
    mass = np.loadtxt(out / "mass.dat")
    assert l1.shape == (4,) and mass.shape == (4,) and np.all(np.isfinite(l1))
    # two saves of the j = Ny/2 slice, [i][k] order, appended
    assert np.loadtxt(out / "psiMid.dat").shape == (2 * 32 * 32,)
    for r_ in range(4):  # psi<rank>.dat of the 4-rank layout, 8 x-slabs each
        assert np.loadtxt(out / f"psi{r_}.dat").shape == (8 * 32 * 32,)
    # restart from the reference's text files (6 digits) reproduces psi to that precision
    r2 = subprocess.run([RUN, "nx=32", "ny=32", "nz=32", "eps=-0.5", "beta=2", "gamma=1", "kp=0.2",
                         "rho0=0.9", "amp=0.5", "rho_m=0.95", "dt=0.01", "nu=1", "compat_ranks=4",
                         "ic=load", f"load={out}", "ranks=4", "steps=1", "step_l1=1",
                         f"out={tmp_path / 'run2'}"], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    # the binary checkpoint written at the end resumes in the Python binding
    import ics
    d = ics.params_dict("interface64")
    d.update(nx=32, ny=32, nz=32, compat_ranks=4, rho_m=0.95)
    sim = qsm_mod.Sim(qsm_mod.Params(**{k: d[k] for k, _ in qsm_mod.Params._fields_}), 0)
    sim.checkpoint_read(str(tmp_path / "end.ckpt"))
    assert sim.diag()["step"] == 20
    assert abs(sim.diag()["L1"] - l1[-1]) <= 1e-5 * abs(l1[-1])


@pytest.mark.parametrize("prec", (0, 1))
def test_timing_select_records_one_kernel(qsm_mod, prec):
    """qsm_timing_select: the counters hold that kernel's launches only, as
    many per step as full timing counts for it, and the step's results do
    not depend on which launches are timed."""
    import ics
    d = ics.params_dict("interface64")
    d["nx"], d["ny"], d["nz"] = (32, 32, 64)
    qp = qsm_mod.Params(**{k: d[k] for k, _ in qsm_mod.Params._fields_})
    psi = ics.workload_psi("interface64", shape=(32, 32, 64))
    a = qsm_mod.Sim(qp, prec)
    a.set_psi(psi)
    a.step(2)
    a.timing(True)
    a.step(3)
    full = a.timing_kernel(0)
    second = a.timing_kernel(1)
    a.timing_select(full["axis"], full["inv"], full["ld"], full["st"], full["variant"])
    a.step(3)
    sel = a.timing_kernel(0)
    for k in ("axis", "inv", "ld", "st", "variant", "launches"):
        assert sel[k] == full[k], (k, sel, full)
    assert sel["bytes"] == full["bytes"]
    assert 0 < sel["ms"]
    with pytest.raises(qsm_mod.QsmError):
        a.timing_kernel(1)  # nothing else was timed
    assert second["launches"] > 0
    a.timing(False)
    b = qsm_mod.Sim(qp, prec)
    b.set_psi(psi)
    b.step(8)
    for f in ("PSI", "RHO"):
        assert np.array_equal(a.field(f), b.field(f)), f
    with pytest.raises(qsm_mod.QsmError):
        a.timing_select(3, 0, 0, 0, 0)
