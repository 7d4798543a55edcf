"""Each pin of tests/pins.py catches a mistake in the stage it pins (CPU).

A mutation is one slip planted in a copy of oracle/qsm_oracle.c, the kind a
restatement of quasi.cpp / quasi-stb.cpp could make: a sign, a power, a
swapped field or component, a missing factor, a wrong step gate.  The copy
is compiled into the test's temporary directory and the named pin must fail
against it.  The product library is not involved.
"""
import os
import subprocess

import pytest

import pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "qsm_oracle.c")

# (name, text in qsm_oracle.c, replacement, pin that must fail)
MUTATIONS = [
    # S3 filter and density law (quasi.cpp:1243-1262)
    ("filter_sqrt_k2", "exp(-p->filter * mq2)", "exp(-p->filter * sqrt(mq2))", "check_filter"),
    ("filter_scale", "double sc = 0.125 / ((double)n);", "double sc = 0.25 / ((double)n);",
     "check_filter"),
    ("density_q0_unsquared", "q02 * pow(psi[t], 2)", "p->q0 * pow(psi[t], 2)", "check_density_law"),
    ("density_gz_as_gy", "pow(gz[t], 2)) + p->rho0", "pow(gy[t], 2)) + p->rho0",
     "check_density_law"),
    # S4 divergence (quasi.cpp:1280-1285) and its data flow
    ("divv_rho_swapped", "-((rho[t] - rho_old[t]) / p->dt", "-((rho_old[t] - rho[t]) / p->dt",
     "check_divergence"),
    ("divv_denominator", "(0.5 * (rho[t] + rho_old[t]))", "(0.5 * (rho[t] + rho[t]))",
     "check_divergence"),
    ("divv_pairing", "vy[t] * ry[t]", "vy[t] * rz[t]", "check_divergence"),
    ("divv_rho_old_is_current", "orc_divergence(p, s->n, rho, f[ORC_RHO_OLD],",
     "orc_divergence(p, s->n, rho, f[ORC_RHO],", "check_rho_old_is_previous_rho"),
    ("grad_rho_not_lagged", "orc_spec_grad(nx, ny, nz, sx, sy, sz, c, f[ORC_RHOHAT], f[ORC_RX + c]);",
     "orc_spec_grad(nx, ny, nz, sx, sy, sz, c, f[ORC_PSIHAT], f[ORC_RX + c]);",
     "check_grad_rho_is_lagged"),
    ("divv_reads_new_v", "  if (n > p->divv_switch)\n    orc_divergence(p, s->n, rho, f[ORC_RHO_OLD], f[ORC_VX], f[ORC_VY], f[ORC_VZ],",
     "  if (n > p->divv_switch)\n    orc_divergence(p, s->n, rho, f[ORC_RHO_OLD], f[ORC_SX], f[ORC_VY], f[ORC_VZ],",
     "check_divergence_reads_previous_v"),
    # S5 stress (quasi.cpp:1305-1317)
    ("stress_sign", "Sz[t] = g * gz[t] - rf * lz[t];", "Sz[t] = g * gz[t] + rf * lz[t];",
     "check_stress_single_mode"),
    ("stress_rho_f_missing", "double rf = rho[t] * F[t];", "double rf = F[t];",
     "check_stress_single_mode"),
    ("stress_g_not_kept", "    G[t] = g;\n", "    G[t] = G[t] * 1.0;\n", "check_stress_single_mode"),
    ("stress_pairing", "Sz[t] = g * gz[t] - rf * lz[t];", "Sz[t] = g * lz[t] - rf * gz[t];",
     "check_stress_two_modes"),
    ("coef_f_sign", "double fh = p->alpha * (q02 - mq2) * f[ORC_PSIHAT][id];",
     "double fh = p->alpha * (mq2 - q02) * f[ORC_PSIHAT][id];", "check_stress_two_modes"),
    # S9 N (quasi.cpp:1914-1930)
    ("nonlin_gamma_sign", "- p->gamma * pow(ps, 5)) -", "+ p->gamma * pow(ps, 5)) -",
     "check_nonlinear_harmonics"),
    ("nonlin_beta_power", "p->beta * pow(ps, 3) - p->gamma * pow(ps, 5)) -",
     "p->beta * pow(ps, 2) - p->gamma * pow(ps, 5)) -", "check_nonlinear_harmonics"),
    ("nonlin_eps_sign", "rho[t] * (p->eps * ps - F[t]", "rho[t] * (-p->eps * ps - F[t]",
     "check_nonlinear_harmonics"),
    ("nonlin_f_sign", "rho[t] * (p->eps * ps - F[t]", "rho[t] * (p->eps * ps + F[t]",
     "check_nonlinear_off_ring"),
    ("nonlin_g_sign", "               G[t] - vx[t] * gx[t]", "             - G[t] - vx[t] * gx[t]",
     "check_nonlinear_off_ring"),
    ("nonlin_g_times_rho", "               G[t] - vx[t] * gx[t]", "           rho[t] * G[t] - vx[t] * gx[t]",
     "check_nonlinear_advection"),
    ("nonlin_adv_pairing", "vy[t] * gy[t] - vz[t] * gz[t];", "vy[t] * gz[t] - vz[t] * gz[t];",
     "check_nonlinear_advection"),
    ("nonlin_rho_s_unsquared", "double rs2 = pow(p->kp * p->amp + p->rho0, 2);",
     "double rs2 = pow(p->kp * p->amp + p->rho0, 1);", "check_nonlinear_pressure"),
    ("nonlin_p_gate", "int with_p = step > p->divv_switch + 1;", "int with_p = step > p->divv_switch;",
     "check_nonlinear_pressure"),
    ("nonlin_p_rho_s_from_rho0", "double rs2 = pow(p->kp * p->amp + p->rho0, 2);",
     "double rs2 = pow(p->rho0, 2);", "check_nonlinear_pressure"),
    # stb mu-stress, N, dtpsi, growth (quasi-stb.cpp:1284-1307, 1893-1923, 2147-2158)
    ("stb_mu_beta_sign", "sh[t] - p->beta * pow(psi[t], 3)", "sh[t] + p->beta * pow(psi[t], 3)",
     "check_stb_dtpsi"),
    ("stb_mu_sh_sign", "double mu = sh[t] - p->beta", "double mu = -sh[t] - p->beta", "check_stb_dtpsi"),
    ("stb_stress_pairing", "Sy[t] = rho * mu * gy[t];", "Sy[t] = rho * mu * gz[t];",
     "check_stb_local_mu"),
    ("stb_nonlin_gamma_sign", "double nl = p->beta * pow(psi[t], 3) - p->gamma * pow(psi[t], 5);",
     "double nl = p->beta * pow(psi[t], 3) + p->gamma * pow(psi[t], 5);", "check_stb_harmonics"),
    ("stb_dtpsi_sign", "double dp = -rho * mu[t];", "double dp = rho * mu[t];", "check_stb_dtpsi"),
    ("stb_growth_kstar", "*ks = (int)lround(p->q0 / qz1);", "*ks = (int)lround(p->q0 / qx1);",
     "check_stb_growth_mode"),
    ("stb_growth_denominator", "s->growth = (ph[id] - pho[id]) / (p->dt * pho[id]);",
     "s->growth = (ph[id] - pho[id]) / (p->dt * ph[id]);", "check_stb_growth_mode"),
]

CHECKS = {c.__name__: c for c in pins.ALL_CHECKS}


def test_every_pin_is_exercised():
    """Every pin is the target of at least one mutation."""
    assert {m[3] for m in MUTATIONS} == set(CHECKS)


@pytest.mark.parametrize("name,old,new,check", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pin_catches_mutation(oracle_mod, tmp_path, name, old, new, check):
    src = open(SRC).read()
    assert src.count(old) == 1, f"mutation site of {name} is not unique in qsm_oracle.c"
    mutant = tmp_path / f"{name}.c"
    mutant.write_text(src.replace(old, new))
    so = tmp_path / f"lib{name}.so"
    subprocess.run(["gcc", "-O1", "-fopenmp", "-fPIC", "-std=gnu11", "-shared", "-o", str(so),
                    str(mutant), "-lm"], check=True)
    with oracle_mod.using_library(str(so)):
        with pytest.raises(AssertionError):
            CHECKS[check](oracle_mod)
