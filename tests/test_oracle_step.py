"""Closed-form pins of the oracle's step stages (CPU).

Each check is chosen so that a plausible restatement mistake fails it:
  gradient    shifted DST stencil == exact derivative of a cosine mode
              (wrong shift direction, q instead of s, or a sign flip fails)
  laplacian   l_c == -K^2 d_c psi for the same mode
  density     single-mode psi at q0 gives constant rho = kp A q0 + rho0
  projection  S = grad(phi), D = 0   -> v = 0 and p = phi
              S = 0, D given         -> spectral div v = D on interior modes
              P-rank compat          -> v_y, v_z change exactly at slabs r*n0
  update      linear regime (kp = 0): per-mode two-step recurrence with
              C1 = 2 - dt rho_m aLin, C2 = 1.5 - dt rho_m aLin, AB2 with
              N_old = N at step 1
  stb update  Crank-Nicolson ratio (1 + dt aLin/2)/(1 - dt aLin/2) when
              beta = gamma = 0
  diagnostics closed-form sums of cos^2, cos^4, cos^6 on the half-sample grid
  halo map    emulation of the reference's flat-index halo copies
"""
import math

import numpy as np

from conftest import rel_l2

NW = 8
DX = 2 * math.pi / NW  # q0 = 1


def grid(n):
    return (np.arange(n) + 0.5) * DX, n * DX


def waves(oracle_mod, n):
    return oracle_mod.wavenumbers(n, n * DX)


def base_params(oracle_mod, shape, **over):
    d = dict(nx=shape[0], ny=shape[1], nz=shape[2], variant=0, compat_ranks=0, nw=NW,
             divv_switch=1000, hydro=1, q0=1.0, alpha=1.0, beta=2.0, gamma=1.0, eps=-0.5,
             nu=1.0, kp=0.2, rho0=0.9, amp=0.5, rho_m=0.95, dt=0.01,
             filter=1.57 * 1.57 * 8, growth_q=0.0)
    d.update(over)
    return oracle_mod.make_params(**d), d


def cos_mode(shape, modes):
    (x, lx), (y, ly), (z, lz) = grid(shape[0]), grid(shape[1]), grid(shape[2])
    a, b, c = modes
    qa, qb, qc = math.pi * a / lx, math.pi * b / ly, math.pi * c / lz
    cx, cy, cz = np.cos(qa * x), np.cos(qb * y), np.cos(qc * z)
    sx, sy, sz = np.sin(qa * x), np.sin(qb * y), np.sin(qc * z)
    psi = cx[:, None, None] * cy[None, :, None] * cz[None, None, :]
    d = [-qa * sx[:, None, None] * cy[None, :, None] * cz[None, None, :],
         -qb * cx[:, None, None] * sy[None, :, None] * cz[None, None, :],
         -qc * cx[:, None, None] * cy[None, :, None] * sz[None, None, :]]
    return psi, d, qa * qa + qb * qb + qc * qc


def hat(oracle_mod, f):
    n = f.size
    return oracle_mod.r2r_3d(f, (5, 5, 5)) * (0.125 / n)


INV = {0: (8, 4, 4), 1: (4, 8, 4), 2: (4, 4, 8)}
FWD = {0: (9, 5, 5), 1: (5, 9, 5), 2: (5, 5, 9)}


def test_gradient_and_laplacian_closed_form(oracle_mod):
    shape = (16, 12, 32)
    s_tabs = [waves(oracle_mod, n)[1] for n in shape]
    psi1, d1, k1 = cos_mode(shape, (3, 2, 5))
    psi2, d2, k2 = cos_mode(shape, (7, 0, 11))
    psi = psi1 - 0.25 * psi2
    ph = hat(oracle_mod, psi)
    for c in range(3):
        gh = oracle_mod.spec_grad(ph, s_tabs, c)
        g = oracle_mod.r2r_3d(gh, INV[c])
        assert rel_l2(g, d1[c] - 0.25 * d2[c]) < 1e-13, c
        # laplacian of the gradient: the DST index m carries s[m] along c
        qs = [waves(oracle_mod, n) for n in shape]
        tabs = [qs[a][1] if a == c else qs[a][0] for a in range(3)]
        k2c = tabs[0][:, None, None] ** 2 + tabs[1][None, :, None] ** 2 + tabs[2][None, None, :] ** 2
        lap = oracle_mod.r2r_3d(-k2c * gh, INV[c])
        assert rel_l2(lap, -k1 * d1[c] + 0.25 * k2 * d2[c]) < 1e-13, c


def test_density_law_single_mode(oracle_mod):
    shape = (8, 8, 32)
    amp, kp, rho0 = 0.5, 0.2, 0.9
    p, _ = base_params(oracle_mod, shape, kp=kp, rho0=rho0, amp=amp)
    z, lz = grid(shape[2])
    psi = np.broadcast_to(amp * np.cos(z), shape)  # q0 = 1 is mode nz/4
    st = oracle_mod.State(p)
    st.set_psi(psi)
    st.step(1)
    rho = st.field("RHO")
    assert np.max(np.abs(rho - (kp * amp + rho0))) < 1e-13
    assert np.max(np.abs(st.field("GX"))) < 1e-14 and np.max(np.abs(st.field("GY"))) < 1e-14
    assert rel_l2(st.field("GZ"), np.broadcast_to(-amp * np.sin(z), shape)) < 1e-13


def _gradient_S(oracle_mod, shape, phi_hat):
    s_tabs = [waves(oracle_mod, n)[1] for n in shape]
    S = []
    for c in range(3):
        g = oracle_mod.r2r_3d(oracle_mod.spec_grad(phi_hat, s_tabs, c), INV[c])
        S.append(oracle_mod.r2r_3d(g, FWD[c]))
    return S


def _smooth_hat(shape, seed):
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return rng.standard_normal(shape) * np.exp(-0.05 * (i * i + j * j + k * k))


def test_projection_of_gradient(oracle_mod):
    shape = (16, 12, 20)
    p, _ = base_params(oracle_mod, shape, nu=0.7)
    phi = _smooth_hat(shape, 5)
    Sx, Sy, Sz = _gradient_S(oracle_mod, shape, phi)
    ph, vx, vy, vz = oracle_mod.project(p, Sx, Sy, Sz, None)
    ref = phi.copy()
    ref[0, 0, 0] = 0.0
    assert rel_l2(ph, ref) < 1e-13
    scale = 0.125 / np.prod(shape)
    mag = scale * max(np.max(np.abs(a)) for a in (Sx, Sy, Sz))
    for v in (vx, vy, vz):
        assert np.max(np.abs(v)) < 1e-12 * mag


def test_projection_compat_slabs(oracle_mod):
    shape = (16, 12, 20)
    phi = _smooth_hat(shape, 6)
    Sx, Sy, Sz = _gradient_S(oracle_mod, shape, phi)
    p4, _ = base_params(oracle_mod, shape, compat_ranks=4)
    _, vx, vy, vz = oracle_mod.project(p4, Sx, Sy, Sz, None)
    scale = 0.125 / np.prod(shape)
    mag = scale * max(np.max(np.abs(a)) for a in (Sx, Sy, Sz))
    assert np.max(np.abs(vx)) < 1e-12 * mag
    for v in (vy, vz):
        bad = np.max(np.abs(v), axis=(1, 2)) > 1e-9 * mag
        assert list(np.nonzero(bad)[0]) == [4, 8, 12]


def test_projection_divergence(oracle_mod):
    shape = (12, 16, 10)
    p, _ = base_params(oracle_mod, shape, nu=1.3)
    dh = _smooth_hat(shape, 7)
    z = np.zeros(shape)
    ph, vx, vy, vz = oracle_mod.project(p, z, z, z, dh)
    assert np.max(np.abs(ph)) == 0.0
    (qx, sx), (qy, sy), (qz, sz) = [waves(oracle_mod, n) for n in shape]
    div = np.zeros(shape)
    div[1:, :, :] += sx[:-1, None, None] * vx[:-1, :, :]
    div[:, 1:, :] += sy[None, :-1, None] * vy[:, :-1, :]
    div[:, :, 1:] += sz[None, None, :-1] * vz[:, :, :-1]
    scale = 0.125 / np.prod(shape)
    inner = (slice(None), slice(0, shape[1] - 1), slice(0, shape[2] - 1))
    ref = scale * dh
    ref[0, 0, 0] = 0.0
    assert rel_l2(div[inner], ref[inner]) < 1e-13


def test_linear_regime_recurrence(oracle_mod):
    shape = (16, 8, 32)
    p, d = base_params(oracle_mod, shape, kp=0.0, rho0=0.9, rho_m=0.95, eps=-0.3, dt=0.02)
    A = 1e-8
    psi, _, k2 = cos_mode(shape, (3, 0, 9))
    st = oracle_mod.State(p)
    st.set_psi(A * psi)
    idx = (3, 0, 9)
    x0 = st.field("PSIHAT")[idx]
    alin = d["eps"] - d["alpha"] * (1.0 - k2) ** 2
    a = d["rho0"] * alin  # q0 = 1: rho(eps - F - G) = rho aLin psi
    c1 = 2.0 - d["dt"] * d["rho_m"] * alin
    c2 = 1.5 - d["dt"] * d["rho_m"] * alin
    dt = d["dt"]
    xs = [x0, (c1 * x0 - 0.5 * x0 + dt / 2 * (2 * a * x0)) / c2]
    for _ in range(8):
        xm, xn = xs[-2], xs[-1]
        xs.append((c1 * xn - 0.5 * xm + dt / 2 * (3 * a * xn - a * xm)) / c2)
    for n in range(1, 10):
        st.step(1)
        got = st.field("PSIHAT")[idx]
        assert abs(got - xs[n]) <= 1e-10 * abs(xs[n]), (n, got, xs[n])


def test_stb_crank_nicolson_ratio(oracle_mod):
    shape = (8, 8, 32)
    p, d = base_params(oracle_mod, shape, variant=1, beta=0.0, gamma=0.0, eps=-0.2, dt=0.01)
    psi, _, k2 = cos_mode(shape, (2, 0, 6))
    st = oracle_mod.State(p)
    st.set_psi(1e-3 * psi)
    alin = d["eps"] - d["alpha"] * (1.0 - k2) ** 2
    ratio = (1 + d["dt"] / 2 * alin) / (1 - d["dt"] / 2 * alin)
    prev = st.field("PSIHAT")[2, 0, 6]
    for _ in range(3):
        st.step(1)
        cur = st.field("PSIHAT")[2, 0, 6]
        assert abs(cur / prev - ratio) < 1e-12
        prev = cur


def test_diagnostics_closed_form(oracle_mod):
    shape = (4, 6, 32)
    p, d = base_params(oracle_mod, shape)
    st = oracle_mod.State(p)
    z, _ = grid(shape[2])
    A, rc = 0.4, 0.97
    psi = np.broadcast_to(A * np.cos(z), shape)
    st.set_field("PSI", psi)
    st.set_field("PSI_OLD", 0.9 * psi)
    st.set_field("RHO", np.full(shape, rc))
    st.set_field("F", 0.3 * psi)
    g = st.diag()
    nx, ny, nz = shape
    dx3 = DX ** 3
    e = rc * dx3 * nx * ny * (-0.5 * d["eps"] * A ** 2 * nz / 2 + 0.5 * d["alpha"] * 0.09 * A ** 2 * nz / 2
                             - 0.25 * d["beta"] * A ** 4 * 3 * nz / 8 + d["gamma"] / 6 * A ** 6 * 5 * nz / 16)
    assert abs(g["mass"] - rc * nx * ny * nz * dx3) < 1e-12 * g["mass"]
    assert abs(g["energy"] - e) < 1e-12 * abs(e)
    assert abs(g["L1"] - 0.1 / d["dt"]) < 1e-12 * g["L1"]


def test_halo_map_matches_reference_copies(oracle_mod):
    """Emulate quasi.cpp:1639-1690 on flat per-rank arrays: which global x-slab
    of S_x lands in the v_y stencil at each i."""
    nx, ny, nz = 16, 3, 2
    for P in (1, 2, 4, 8, 16):
        n0 = nx // P
        glob = np.repeat(np.arange(nx, dtype=float), ny * nz)  # S_x[i,j,k] = i
        local = [glob[r * n0 * ny * nz:(r + 1) * n0 * ny * nz] for r in range(P)]
        front = []
        for r in range(P):  # psi_front[index] = Sx_local[index], index = j*Nz + k
            buf = np.zeros(ny * nz)
            for j in range(ny):
                for k in range(nz):
                    buf[j * nz + k] = local[r][j * nz + k]
            front.append(buf)
        for r in range(P):
            back = np.zeros(ny * nz) if r == 0 else front[r - 1]  # Recv from rank-1 (rank 0: zeros)
            for il in range(n0):
                i = r * n0 + il
                j, k = 0, 0
                if il == 0:
                    seen = back[(j + 1) * nz + k] if r > 0 else -1
                else:
                    seen = local[r][((il - 1) * ny + j + 1) * nz + k]
                want = oracle_mod.halo_sx_slab(nx, P, i)
                assert (seen if i > 0 else -1) == want, (P, i, seen, want)
