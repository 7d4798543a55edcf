"""Closed-form pins of the oracle's pointwise stages and of their composition
in the step (CPU).  Each check takes the oracle module and raises
AssertionError on a mismatch; tests/test_oracle_pins.py runs them against the
oracle, tests/test_oracle_mutations.py against copies of the oracle with one
planted mistake each, where the named check must fail.

None of the expected values below is the oracle's formula evaluated again:
each comes from a property the stage must have, chosen so that the plausible
slips of a restatement (a sign, a power, a swapped field or component, a
missing factor, a wrong step gate) change the answer.

  filter      DCT of a long cosine mode on the half-sample grid: one
              coefficient, scaled by exp(-1.57*1.57*8 K^2) (quasi.cpp:1259-1262)
  density     sqrt of Pythagorean sums: |(q0 psi, g)| exactly (quasi.cpp:1243-1247)
  divergence  continuity: uniform growth rho ~ e^{lambda t} gives
              D = -(2/dt) tanh(lambda dt/2); a linear profile advected with
              the velocity gives D = 0 (quasi.cpp:1280-1285)
  stress      a single Fourier mode has G g_c = F l_c, so S = 0 for any rho;
              two modes give the cross term in closed form (quasi.cpp:1305-1317)
  N           Chebyshev harmonics of eps psi + beta psi^3 - gamma psi^5 for a
              mode at q0; -F - rho G for a mode off q0 with constant rho;
              orthogonal / parallel v and grad psi; the pressure term with a
              Pythagorean |grad rho| and rho_s^2, from step divvSwitch + 2
              (quasi.cpp:1905-1931)
  stb         a mode at q0 with eps = 0 has a local mu, so S = grad Phi(psi)
              and v = 0; harmonics of beta psi^3 - gamma psi^5; scale DCT(dtpsi)
              is the update's increment times (1 - dt aLin/2)/dt (quasi-stb.cpp:1284-1307,
              1893-1923, 1944-1957); growth rate (r - 1)/dt at (i*, 0, k*)
              (quasi-stb.cpp:2147-2158)
  data flow   rho_old is the previous step's rho; grad rho is that of the
              previous step's rhohat; D reads the previous step's v
"""
import math

import numpy as np

# the reference's filter constant, quasi.cpp:1261: exp(-1.57*1.57*mq2c*8)
REF_FILTER = 1.57 * 1.57 * 8


def params(orc, shape, **over):
    d = dict(nx=shape[0], ny=shape[1], nz=shape[2], variant=0, compat_ranks=0, nw=8,
             divv_switch=1000, hydro=1, q0=1.0, alpha=1.0, beta=2.0, gamma=1.0, eps=-0.5,
             nu=1.0, kp=0.2, rho0=0.9, amp=0.5, rho_m=0.95, dt=0.01, filter=REF_FILTER,
             growth_q=0.0)
    d.update(over)
    return orc.make_params(**d), d


def axis(n, d):
    """Half-sample grid x_m = (m + 1/2) dx and box length, dx = 2 pi/(q0 nw)."""
    dx = 2 * math.pi / (d["q0"] * d["nw"])
    return (np.arange(n) + 0.5) * dx, n * dx


def wavenumber(m, n, d):
    """q[m] = pi m / L (quasi.cpp:397-431)."""
    return math.pi * m / axis(n, d)[1]


def mode(shape, d, m):
    """cos(q_x x) cos(q_y y) cos(q_z z) of mode m = (mx, my, mz), with its
    derivatives and K^2."""
    xs = [axis(n, d)[0] for n in shape]
    qs = [wavenumber(m[a], shape[a], d) for a in range(3)]
    c = [np.cos(qs[a] * xs[a]) for a in range(3)]
    s = [np.sin(qs[a] * xs[a]) for a in range(3)]

    def outer(a, b, e):
        return a[:, None, None] * b[None, :, None] * e[None, None, :]
    f = outer(c[0], c[1], c[2])
    grad = [-qs[0] * outer(s[0], c[1], c[2]), -qs[1] * outer(c[0], s[1], c[2]),
            -qs[2] * outer(c[0], c[1], s[2])]
    return f, grad, sum(q * q for q in qs)


def close(got, want, tol, what):
    got, want = np.asarray(got, dtype=np.float64), np.asarray(want, dtype=np.float64)
    scale = max(float(np.max(np.abs(want))), 1e-300)
    err = float(np.max(np.abs(got - want))) / scale
    assert err <= tol, f"{what}: max error {err:.3e} of {scale:.3e} (tolerance {tol:.1e})"


# ---------------------------------------------------------------- S3 filter

def check_filter(orc):
    """rho* = c0 + a cos cos cos (mode (1,1,1)) + b cos(q_z[2] z): the filter
    keeps c0 at (0,0,0), puts a/8 e^{-f K1^2} at (1,1,1) and b/2 e^{-f K2^2}
    at (0,0,2) (DCT-II of a cosine on the half-sample grid: N per axis, 2N
    for the constant; scale = 1/(8 NxNyNz)), and rho = IDCT of that.  The
    modes are long enough (K^2 ~ 0.07-0.09) that e^{-f K^2} is O(0.2)."""
    shape = (32, 24, 40)
    p, d = params(orc, shape, q0=1.3)
    assert abs(d["filter"] - 19.7192) < 1e-12
    c0, a, b = 0.93, 0.5, -0.3
    m1, _, k1 = mode(shape, d, (1, 1, 1))
    m2, _, k2 = mode(shape, d, (0, 0, 2))
    f1, f2 = math.exp(-REF_FILTER * k1), math.exp(-REF_FILTER * k2)
    assert 0.05 < f1 < 0.5 and 0.05 < f2 < 0.5
    rh, rho = orc.density_filter(p, c0 + a * m1 + b * m2)
    for idx, want in (((0, 0, 0), c0), ((1, 1, 1), a / 8 * f1), ((0, 0, 2), b / 2 * f2)):
        assert abs(rh[idx] - want) <= 1e-12 * abs(want), f"filtered rhohat{idx}: {rh[idx]} != {want}"
    rest = rh.copy()
    rest[0, 0, 0] = rest[1, 1, 1] = rest[0, 0, 2] = 0.0
    assert float(np.max(np.abs(rest))) < 1e-14, "filtered rhohat outside the three modes"
    close(rho, c0 + a * f1 * m1 + b * f2 * m2, 1e-13, "filtered rho")


# ---------------------------------------------------------------- S3 density law

def check_density_law(orc):
    """rho* = kp |(q0 psi, g)| + rho0 on Pythagorean tuples (q0 = 1.7):
    (3,4,0) -> 5, (0,5,12) -> 13, (2,3,6) -> 7, and q0 psi with g_x = 4/3 q0 psi -> 5/3 q0 |psi|."""
    shape = (2, 3, 8)
    p, d = params(orc, shape, q0=1.7, kp=0.35, rho0=0.6)
    kp, r0, q0 = d["kp"], d["rho0"], d["q0"]
    t = np.linspace(-1.0, 1.0, 48).reshape(shape)
    z = np.zeros(shape)
    for g, norm in (((3, 4, 0), 5), ((0, 5, 12), 13), ((2, 3, 6), 7)):
        got = orc.density_law(p, z, [g[0] * t, g[1] * t, g[2] * t])
        close(got, kp * norm * np.abs(t) + r0, 1e-14, f"density law, g = {g} t")
    got = orc.density_law(p, t, [z, z, z])
    close(got, kp * q0 * np.abs(t) + r0, 1e-14, "density law, psi alone")
    got = orc.density_law(p, t, [4.0 / 3.0 * q0 * t, z, z])
    close(got, kp * 5.0 / 3.0 * q0 * np.abs(t) + r0, 1e-14, "density law, psi and g_x")


# ---------------------------------------------------------------- S4 divergence

def check_divergence(orc):
    """Continuity, div v = -(D rho/Dt)/rho, in the reference's discrete form:
    (a) v = 0, rho = rho_old e^{lambda dt}: D = -(2/dt) tanh(lambda dt/2),
        whatever rho_old is;
    (b) rho = rho_a + G.(x - u t) carried by v = u: D = 0."""
    shape = (4, 5, 6)
    p, d = params(orc, shape, dt=0.02)
    dt = d["dt"]
    rng = np.random.default_rng(11)
    rho_old = 0.8 + 0.4 * rng.random(shape)
    lam = 40.0
    z = np.zeros(shape)
    D = orc.divergence(p, rho_old * math.exp(lam * dt), rho_old, [z, z, z], [z, z, z])
    close(D, np.full(shape, -(2.0 / dt) * math.tanh(lam * dt / 2)), 1e-12, "uniform growth")
    xs = np.meshgrid(*[np.arange(n) * 0.3 for n in shape], indexing="ij")
    u = (0.7, -1.1, 0.4)
    G = (0.05, 0.02, -0.03)
    rho_new = 1.0 + sum(G[c] * (xs[c] - u[c] * dt) for c in range(3))
    rho_then = 1.0 + sum(G[c] * xs[c] for c in range(3))
    D = orc.divergence(p, rho_new, rho_then, [np.full(shape, c) for c in u],
                       [np.full(shape, c) for c in G])
    assert float(np.max(np.abs(D))) < 1e-11, f"advected linear profile: D = {np.max(np.abs(D)):.3e}"


# ---------------------------------------------------------------- S5 stress

def check_stress_single_mode(orc):
    """One Fourier mode: G = -K^2 F and l_c = -K^2 g_c, so G g_c = F l_c and
    S = rho (G g_c - F l_c) = 0 for any rho; G comes back multiplied by rho."""
    shape = (8, 10, 12)
    p, d = params(orc, shape, q0=1.2)
    f, g, k2 = mode(shape, d, (2, 3, 5))
    F = d["alpha"] * (d["q0"] ** 2 - k2) * 0.4 * f
    G = -k2 * F
    gr = [0.4 * x for x in g]
    lp = [-k2 * x for x in gr]
    rho = 0.9 + 0.3 * np.random.default_rng(3).random(shape)
    S, Gout = orc.stress(rho, G, F, gr, lp)
    mag = float(np.max(np.abs(G * gr[2])))
    for c in range(3):
        assert float(np.max(np.abs(S[c]))) < 1e-13 * mag, f"single-mode stress S_{'xyz'[c]} != 0"
    close(Gout, rho * G, 1e-15, "G <- rho G")


def check_stress_two_modes(orc):
    """psi = A cos(a z) + B cos(b z) through one oracle step: with
    F = fa cos(az) + fb cos(bz), fa = alpha (q0^2 - a^2) A (quasi.cpp:1042),
    G = lap F, g = d_z psi, l = lap g, the stress is
    S_z = rho (a^2 - b^2) (b B fa cos(az) sin(bz) - a A fb sin(az) cos(bz)),
    S_x = S_y = 0."""
    shape = (4, 4, 32)
    p, d = params(orc, shape, kp=0.3, rho0=0.8)
    z, _ = axis(32, d)
    A, B = 0.3, 0.2
    a, b = wavenumber(3, 32, d), wavenumber(7, 32, d)
    psi = np.broadcast_to(A * np.cos(a * z) + B * np.cos(b * z), shape)
    st = orc.State(p)
    st.set_psi(psi)
    st.step(1)
    rho = st.field("RHO")
    fa, fb = d["alpha"] * (d["q0"] ** 2 - a * a) * A, d["alpha"] * (d["q0"] ** 2 - b * b) * B
    want = rho * (a * a - b * b) * np.broadcast_to(
        b * B * fa * np.cos(a * z) * np.sin(b * z) - a * A * fb * np.sin(a * z) * np.cos(b * z), shape)
    n8 = 8.0 * np.prod(shape)
    Sz = orc.r2r_3d(st.field("SZ"), (4, 4, 8)) / n8
    close(Sz, want, 1e-12, "two-mode S_z")
    for c in ("SX", "SY"):
        assert float(np.max(np.abs(st.field(c)))) < 1e-12 * float(np.max(np.abs(want))) * n8, c


# ---------------------------------------------------------------- S9 N

def _harmonics(A, eps, beta, gamma):
    """eps psi + beta psi^3 - gamma psi^5 for psi = A cos: coefficients of
    cos, cos 3, cos 5 (cos^3 = (3 cos + cos 3)/4, cos^5 = (10 cos + 5 cos 3 + cos 5)/16)."""
    return (eps * A + 0.75 * beta * A ** 3 - 10 / 16 * gamma * A ** 5,
            0.25 * beta * A ** 3 - 5 / 16 * gamma * A ** 5,
            -gamma * A ** 5 / 16)


def check_nonlinear_harmonics(orc):
    """psi = A cos(q0 z), nw = 16 (no aliasing up to 5 q0): F = G = 0,
    rho = kp A q0 + rho0 = rho_s, S = 0 so v = 0, and N-hat of step 1 holds
    rho_s times the harmonics of eps psi + beta psi^3 - gamma psi^5 at modes
    m, 3m, 5m (m = Nz/8), times 4 Nx Ny Nz (unnormalised DCT-II)."""
    shape = (4, 4, 32)
    p, d = params(orc, shape, nw=16, amp=0.7, kp=0.2, rho0=0.9, eps=-0.5, beta=2.0, gamma=1.0)
    z, _ = axis(32, d)
    A = d["amp"]
    st = orc.State(p)
    st.set_psi(np.broadcast_to(A * np.cos(d["q0"] * z), shape))
    st.step(1)
    rs = d["kp"] * A * d["q0"] + d["rho0"]
    close(st.field("RHO"), np.full(shape, rs), 1e-13, "rho of a mode at q0")
    want = np.zeros(shape)
    for h, c in zip((4, 12, 20), _harmonics(A, d["eps"], d["beta"], d["gamma"])):
        want[0, 0, h] = 4 * np.prod(shape) * rs * c
    close(st.field("NHAT"), want, 1e-11, "N-hat harmonics")


def check_nonlinear_off_ring(orc):
    """kp = 0 (rho = rho0 = 0.9 everywhere), q0 = 1.25 and psi = A cos(q z),
    q = 0.75 q0: F = alpha (q0^2 - q^2) psi, G = -q^2 F enters N as rho G, v = 0,
    so the first harmonic of N is rho0 ((eps - alpha (q0^2 - q^2)(1 - q^2)) A
    + 3/4 beta A^3 - 10/16 gamma A^5)."""
    shape = (4, 4, 32)
    p, d = params(orc, shape, nw=16, q0=1.25, kp=0.0, rho0=0.9, eps=-0.3, beta=1.5, gamma=0.8,
                  alpha=1.1)
    z, _ = axis(32, d)
    A = 0.6
    q = wavenumber(3, 32, d)
    assert abs(q - 0.75 * d["q0"]) < 1e-14
    st = orc.State(p)
    st.set_psi(np.broadcast_to(A * np.cos(q * z), shape))
    st.step(1)
    r0, al, q02 = d["rho0"], d["alpha"], d["q0"] ** 2
    lin = d["eps"] - al * (q02 - q * q) * (1 - q * q)
    h1, h3, h5 = _harmonics(A, lin, d["beta"], d["gamma"])
    want = np.zeros(shape)
    for m, c in zip((3, 9, 15), (h1, h3, h5)):
        want[0, 0, m] = 4 * np.prod(shape) * r0 * c
    close(st.field("NHAT"), want, 1e-10, "N-hat off the ring")


def check_nonlinear_advection(orc):
    """With rho = psi = F = 0: N = -G - v.g; v perpendicular to g adds
    nothing, v parallel to g subtracts |v||g| (|(1,2,2)| = 3, |(2,4,4)| = 6);
    G enters once, not multiplied by rho again."""
    shape = (2, 2, 4)
    p, d = params(orc, shape)
    z = np.zeros(shape)
    one = np.ones(shape)
    v = [1 * one, 2 * one, 2 * one]
    for g, want in (((2, 1, -2), 0.0), ((2, 4, 4), -18.0)):
        N = orc.nonlinear(p, 1, z, z, z, z, v, [c * one for c in g], z, [z, z, z])
        assert float(np.max(np.abs(N - want))) < 1e-14, f"advection term, g = {g}"
    N = orc.nonlinear(p, 1, 2.0 * one, z, z, 0.25 * one, [z, z, z], [z, z, z], z, [z, z, z])
    assert float(np.max(np.abs(N + 0.25))) < 1e-15, "N holds -G (G already multiplied by rho)"


def check_nonlinear_pressure(orc):
    """From step divvSwitch + 2 only, N gains p kp |grad rho| / rho_s^2 with
    rho_s = kp Amp + rho0 = 0.2 * 1.5 + 0.9 = 1.2 (|(3,4,12)| = 13)."""
    shape = (2, 2, 4)
    p, d = params(orc, shape, divv_switch=7, kp=0.2, amp=1.5, rho0=0.9)
    z = np.zeros(shape)
    t = np.linspace(0.1, 0.8, 16).reshape(shape)
    P = 0.45 * np.ones(shape)
    r = [3 * t, 4 * t, 12 * t]
    rs2 = (0.2 * 1.5 + 0.9) ** 2
    for step, on in ((8, False), (9, True), (40, True)):
        N = orc.nonlinear(p, step, z, z, z, z, [z, z, z], [z, z, z], P, r)
        want = 0.45 * 0.2 * 13 * t / rs2 if on else np.zeros(shape)
        assert float(np.max(np.abs(N - want))) < 1e-14, f"pressure term at step {step}"


# ---------------------------------------------------------------- stb

def stb_params(orc, shape, **over):
    base = dict(variant=1, beta=0.4, gamma=3.0, eps=-0.009, nu=0.1, amp=math.sqrt(0.12),
                dt=0.0005, kp=0.5, rho0=0.1, rho_m=0.5, growth_q=0.5)
    base.update(over)
    return params(orc, shape, **base)


def check_stb_local_mu(orc):
    """stb, eps = 0, mode (0,3,4) with q_y^2 + q_z^2 = q0^2 (nw = 16, 40^2
    points across, modes up to 6x resolved): sh = -aLin psi = 0, so mu is a
    function of psi and S = mu grad psi = grad Phi(psi) projects to v = 0."""
    shape = (4, 40, 40)
    p, d = stb_params(orc, shape, nw=16, eps=0.0, beta=0.4, gamma=3.0)
    f, _, k2 = mode(shape, d, (0, 3, 4))
    assert abs(k2 - d["q0"] ** 2) < 1e-14
    st = orc.State(p)
    st.set_psi(0.5 * f)
    st.step(1)
    smag = max(float(np.max(np.abs(st.field(c)))) for c in ("SX", "SY", "SZ"))
    assert smag > 1e-3
    for c in ("VX", "VY", "VZ"):
        vmax = float(np.max(np.abs(st.field(c))))
        assert vmax < 1e-13, f"stb local mu: {c} = {vmax:.3e}"


def check_stb_harmonics(orc):
    """stb, psi = A cos(q0 z) (nw = 16): v = 0 as above, so N-hat holds the
    harmonics of beta psi^3 - gamma psi^5 (quasi-stb.cpp:1914)."""
    shape = (4, 4, 32)
    p, d = stb_params(orc, shape, nw=16, eps=-0.2)
    z, _ = axis(32, d)
    A = 0.5
    st = orc.State(p)
    st.set_psi(np.broadcast_to(A * np.cos(d["q0"] * z), shape))
    st.step(1)
    want = np.zeros(shape)
    for h, c in zip((4, 12, 20), _harmonics(A, 0.0, d["beta"], d["gamma"])):
        want[0, 0, h] = 4 * np.prod(shape) * c
    close(st.field("NHAT"), want, 1e-11, "stb N-hat harmonics")


def check_stb_dtpsi(orc):
    """stb step 1 (N_old = N): the update gives
    (psihat_new - psihat)(1 - dt aLin/2)/dt = aLin psihat + scale N-hat, and
    dtpsi = -mu - v.g = aLin psi + N, so scale DCT(dtpsi) equals
    (psihat_new - psihat_old)(1 - dt aLin/2)/dt mode by mode."""
    shape = (8, 8, 16)
    p, d = stb_params(orc, shape, eps=-0.05, dt=0.002)
    rng = np.random.default_rng(5)
    psi = 0.3 * np.cos(axis(16, d)[0])[None, None, :] + 0.05 * rng.standard_normal(shape)
    st = orc.State(p)
    st.set_psi(psi)
    st.step(1)
    q = [np.array([wavenumber(m, n, d) for m in range(n)]) for n in shape]
    k2 = q[0][:, None, None] ** 2 + q[1][None, :, None] ** 2 + q[2][None, None, :] ** 2
    alin = d["eps"] - d["alpha"] * (d["q0"] ** 2 - k2) ** 2
    scale = 0.125 / np.prod(shape)
    lhs = scale * orc.r2r_3d(st.field("DTPSI"), (5, 5, 5))
    rhs = (st.field("PSIHAT") - st.field("PSIHAT_OLD")) * (1 - d["dt"] * alin / 2) / d["dt"]
    close(lhs, rhs, 1e-10, "dtpsi against the update's increment")


def check_stb_growth_mode(orc):
    """beta = gamma = 0, one small mode at (i*, 0, k*) with Qi = q_x[3] on a
    non-cubic grid: i* = 3, k* = Nz/4 (q0 = 1, nw = 8), and the growth rate of
    the last step is (r - 1)/dt with the Crank-Nicolson ratio
    r = (1 + dt aLin/2)/(1 - dt aLin/2)."""
    shape = (24, 8, 32)
    p, d = stb_params(orc, shape, beta=0.0, gamma=0.0, eps=-0.1, dt=0.01)
    qi = wavenumber(3, 24, d)
    p.growth_q = qi
    assert orc.stb_growth_mode(p) == (3, 8)
    f, _, k2 = mode(shape, d, (3, 0, 8))
    st = orc.State(p)
    st.set_psi(1e-7 * f)
    st.step(3)
    alin = d["eps"] - d["alpha"] * (d["q0"] ** 2 - k2) ** 2
    r = (1 + d["dt"] * alin / 2) / (1 - d["dt"] * alin / 2)
    assert abs(st.growth - (r - 1) / d["dt"]) < 1e-8 * abs((r - 1) / d["dt"]), (st.growth, (r - 1) / d["dt"])


# ---------------------------------------------------------------- data flow of the step

def modulated_layers(shape, d):
    """Layers cos(q0 z) with an amplitude modulated at mode 1 of every axis,
    slow enough to survive the density filter, so that rho and grad rho vary."""
    f = [0.0] * 3
    for a in range(3):
        f[a] = np.cos(wavenumber(1, shape[a], d) * axis(shape[a], d)[0])
    amp = 0.4 * (1 + 0.3 * f[0][:, None, None] + 0.2 * f[1][None, :, None] + 0.25 * f[2][None, None, :])
    return np.ascontiguousarray(amp * np.cos(d["q0"] * axis(shape[2], d)[0])[None, None, :])


def check_rho_old_is_previous_rho(orc):
    """Uniform psi: rho uniform, grad rho = 0, v = 0, so past the switch
    D = -(2/dt)(rho_n - rho_{n-1})/(rho_n + rho_{n-1}) with rho_{n-1} the
    density the previous step formed (the D output is IDCT(scale D-hat))."""
    shape = (4, 4, 8)
    p, d = params(orc, shape, divv_switch=2, dt=0.05, eps=-0.4)
    st = orc.State(p)
    st.set_psi(np.full(shape, 0.6))
    rhos = []
    for n in range(1, 6):
        st.step(1)
        rhos.append(float(np.mean(st.field("RHO"))))
        if n > d["divv_switch"]:
            a, b = rhos[-1], rhos[-2]
            want = -(2.0 / d["dt"]) * (a - b) / (a + b)
            assert abs(want) > 1e-3
            close(st.field("DIVV"), np.full(shape, want), 1e-10, f"D of step {n}")


def check_grad_rho_is_lagged(orc):
    """The grad rho a step reads is that of the previous step's rhohat:
    r_c(n) = IDST_c(-s_c rhohat_{n-1}[+1 along c])."""
    shape = (8, 8, 16)
    p, d = params(orc, shape, divv_switch=1)
    psi = modulated_layers(shape, d)
    st = orc.State(p)
    st.set_psi(psi)
    st.step(2)
    prev = st.field("RHOHAT")
    st.step(1)
    assert float(np.max(np.abs(st.field("RZ")))) > 1e-4
    s_tabs = [np.array([wavenumber(m + 1, n, d) for m in range(n)]) for n in shape]
    inv = {0: (8, 4, 4), 1: (4, 8, 4), 2: (4, 4, 8)}
    for c, name in enumerate(("RX", "RY", "RZ")):
        want = orc.r2r_3d(orc.spec_grad(prev, s_tabs, c), inv[c])
        close(st.field(name), want, 1e-12, f"lagged {name}")


def check_divergence_reads_previous_v(orc):
    """Past the switch, D of step n is the continuity form over rho_n, the
    previous step's rho and v, and this step's lagged grad rho."""
    shape = (8, 8, 16)
    p, d = params(orc, shape, divv_switch=2)
    psi = modulated_layers(shape, d)
    st = orc.State(p)
    st.set_psi(psi)
    st.step(3)
    rho_prev = st.field("RHO")
    v_prev = [st.field(c) for c in ("VX", "VY", "VZ")]
    st.step(1)
    rho = st.field("RHO")
    r = [st.field(c) for c in ("RX", "RY", "RZ")]
    adv = sum(v_prev[c] * r[c] for c in range(3))
    assert float(np.max(np.abs(adv))) > 1e-7
    want = -((rho - rho_prev) / d["dt"] + adv) / (0.5 * (rho + rho_prev))
    close(st.field("DIVV"), want, 1e-10, "D from the previous step's rho and v")


STAGE_CHECKS = [check_filter, check_density_law, check_divergence, check_stress_single_mode,
                check_nonlinear_advection, check_nonlinear_pressure]
STEP_CHECKS = [check_stress_two_modes, check_nonlinear_harmonics, check_nonlinear_off_ring,
               check_stb_local_mu, check_stb_harmonics, check_stb_dtpsi, check_stb_growth_mode,
               check_rho_old_is_previous_rho, check_grad_rho_is_lagged,
               check_divergence_reads_previous_v]
ALL_CHECKS = STAGE_CHECKS + STEP_CHECKS
