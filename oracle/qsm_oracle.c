/*
 * qsm_oracle.c -- CPU FP64 restatement of the reference time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only to compare against or to time it beside the GPU path.  The
 * product library never links, loads or calls anything in oracle/.
 *
 * What it restates (reference paths relative to the reference repo root):
 *   - FFTW r2r kinds REDFT10/REDFT01/RODFT10/RODFT01, unnormalised, as used
 *     by the 8 plans of quasi/quasi.cpp:449-495 (FFTW 3.3.x, third party,
 *     not vendored; its published definitions are the spec).
 *   - wavenumber tables              quasi/quasi.cpp:397-431
 *   - coefficients C1/C2/CM1/CM2     quasi/quasi.cpp:875-902, quasi-stb.cpp:840-871
 *   - setup psi -> psihat -> psi     quasi/quasi.cpp:905-951
 *   - the quasi step S1..S10         quasi/quasi.cpp:1014-1965
 *   - the stb step deltas            quasi/quasi-stb.cpp:1003-1038, 1218-1308,
 *                                    1369-1404, 1893-1968
 *   - diagnostics L1/mass/energy     quasi/quasi.cpp:1969-2031
 *   - stb growth rate                quasi/quasi-stb.cpp:2147-2158
 *   - the rank-count-dependent halo slab of quasi/quasi.cpp:1639-1650 (the
 *     velocity stencils of v_y, v_z take S_x of slab 0 of the previous rank
 *     instead of its last slab), reproduced when compat_ranks = P > 0.
 *
 * Pins (tests/test_oracle_transforms.py, test_oracle_step.py, test_oracle_pins.py
 * with the checks in tests/pins.py; tests/test_oracle_mutations.py plants one
 * mistake per stage in a copy of this file and shows that its pin fails):
 *   r2r fast path      vs O(N^2) direct definitions and SciPy golden vectors
 *   3D r2r             separability vs 1D, round trip 2N per axis
 *   shifted gradient   closed-form derivative of the cosine interpolant
 *   density law        Pythagorean |(q0 psi, g)|; a mode at q0 gives constant rho
 *   density filter     one long cosine mode scaled by exp(-1.57*1.57*8 K^2)
 *   divergence         continuity: uniform growth -(2/dt) tanh(lambda dt/2),
 *                      advected linear profile 0; rho_old, grad rho and v lags
 *   stress             single mode S = 0; two-mode closed form; G <- rho G
 *   nonlinear N        Chebyshev harmonics at q0; -F - rho G off q0; v.g for
 *                      orthogonal/parallel vectors; pressure term and its gate
 *   stb                local mu -> v = 0; harmonics; dtpsi vs the update;
 *                      growth-rate mode and Crank-Nicolson ratio
 *   projection         gradient S -> v = 0, p = phi; solenoidal S -> p = 0;
 *                      S = 0 -> div v = D on interior modes
 *   update             linear-regime two-root recurrence (closed form)
 *   diagnostics        closed-form sums of cos^2, cos^4, cos^6
 *   halo compat        slab map checked by brute force against a rank loop
 * Every public function below has at least one of these checks.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* FFTW's fftw_r2r_kind values (fftw3.h), kept identical for familiarity. */
enum { ORC_REDFT01 = 4, ORC_REDFT10 = 5, ORC_RODFT01 = 8, ORC_RODFT10 = 9 };

typedef struct {
  int nx, ny, nz;
  int variant;      /* 0 = quasi (quasi.cpp), 1 = stb (quasi-stb.cpp)          */
  int compat_ranks; /* 0 = exact stencils, P = reproduce P-rank slab halo map   */
  int nw;           /* points per wavelength; L = N*2*pi/(q0*nw)               */
  int divv_switch;  /* quasi.cpp:161 (200)                                     */
  int hydro;        /* stb advection switch, quasi-stb.cpp:155                 */
  double q0, alpha, beta, gamma, eps, nu, kp, rho0, amp, rho_m, dt;
  double filter;    /* 1.57*1.57*8 in quasi.cpp:1261                           */
  double growth_q;  /* stb perturbation wavenumber Qi, quasi-stb.cpp:157       */
} orc_params;

/* ------------------------------------------------------------------ */
/* complex FFT (mixed radix, recursive) used by the fast r2r path      */
/* ------------------------------------------------------------------ */

typedef double complex cplx;

typedef struct {
  int n;        /* transform length 2N */
  cplx *tw;     /* tw[t] = exp(-2 pi i t / n) */
  cplx *post;   /* post[k] = exp(-i pi k / n), k < n/2 (the half-sample shift) */
  int *rev;     /* bit reversal permutation when n is a power of two, else NULL */
} fft_plan;

static int smallest_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

/* out[k] = sum_m in[m*stride] exp(sign*2 pi i k m / n); tw is for length ntot,
 * and twstep = ntot / n.  scratch holds at least n entries. */
static void fft_rec(int n, int stride, const cplx *in, cplx *out, const cplx *tw,
                    int twstep, int sign, cplx *scratch) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = smallest_factor(n), m = n / p;
  /* sub-transforms of length m on the p decimated sequences */
  for (int r = 0; r < p; ++r)
    fft_rec(m, stride * p, in + (size_t)r * stride, scratch + (size_t)r * m, tw,
            twstep * p, sign, out);
  int ntot = n * twstep;
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < p; ++q) {
      int kk = k + q * m;
      cplx acc = 0;
      for (int r = 0; r < p; ++r) {
        long e = ((long)r * kk * twstep) % ntot;
        cplx w = tw[e];
        if (sign > 0) w = conj(w);
        acc += w * scratch[(size_t)r * m + k];
      }
      out[kk] = acc;
    }
  }
}

/* Iterative radix-2 decimation-in-time FFT (power-of-two n): out = DFT(in). */
static void fft_pow2(const fft_plan *pl, const cplx *in, cplx *out, int sign) {
  int n = pl->n;
  for (int t = 0; t < n; ++t) out[pl->rev[t]] = in[t];
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int b = 0; b < n; b += len)
      for (int t = 0; t < half; ++t) {
        cplx w = pl->tw[t * step];
        if (sign > 0) w = conj(w);
        cplx u = out[b + t], v = w * out[b + t + half];
        out[b + t] = u + v;
        out[b + t + half] = u - v;
      }
  }
}

static void fft_any(const fft_plan *pl, const cplx *in, cplx *out, int sign, cplx *scratch) {
  if (pl->rev) fft_pow2(pl, in, out, sign);
  else fft_rec(pl->n, 1, in, out, pl->tw, 1, sign, scratch);
}

static fft_plan *plan_cache[8];
static int plan_cache_n = 0;

static const fft_plan *get_plan(int n) {
  for (int i = 0; i < plan_cache_n; ++i)
    if (plan_cache[i]->n == n) return plan_cache[i];
  fft_plan *p = (fft_plan *)malloc(sizeof(fft_plan));
  p->n = n;
  p->tw = (cplx *)malloc(sizeof(cplx) * n);
  p->post = (cplx *)malloc(sizeof(cplx) * n);
  for (int t = 0; t < n; ++t) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L * t / n;
    p->tw[t] = (double)cosl(a) + I * (double)sinl(a);
    long double b = -3.141592653589793238462643383279502884L * t / n;
    p->post[t] = (double)cosl(b) + I * (double)sinl(b);
  }
  p->rev = NULL;
  if ((n & (n - 1)) == 0) {
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    p->rev = (int *)malloc(sizeof(int) * n);
    for (int t = 0; t < n; ++t) {
      int r = 0;
      for (int b = 0; b < lg; ++b) if (t & (1 << b)) r |= 1 << (lg - 1 - b);
      p->rev[t] = r;
    }
  }
  if (plan_cache_n < 8) plan_cache[plan_cache_n++] = p;
  return p;
}

/* Plans must exist before threads use them. */
static void prepare_plans(const orc_params *pr) {
  get_plan(2 * pr->nx); get_plan(2 * pr->ny); get_plan(2 * pr->nz);
}

/* ------------------------------------------------------------------ */
/* 1D r2r transforms                                                    */
/* ------------------------------------------------------------------ */

static const long double PI_L = 3.141592653589793238462643383279502884L;

/* O(N^2) evaluation of FFTW's r2r definitions (FFTW manual, "1d Real-even
 * DFTs (DCTs)" and "1d Real-odd DFTs (DSTs)"). */
int orc_r2r_1d_direct(int n, int kind, const double *in, double *out) {
  for (int k = 0; k < n; ++k) {
    long double acc = 0;
    switch (kind) {
    case ORC_REDFT10:
      for (int j = 0; j < n; ++j) acc += 2.0L * in[j] * cosl(PI_L * (j + 0.5L) * k / n);
      break;
    case ORC_REDFT01:
      acc = in[0];
      for (int j = 1; j < n; ++j) acc += 2.0L * in[j] * cosl(PI_L * j * (k + 0.5L) / n);
      break;
    case ORC_RODFT10:
      for (int j = 0; j < n; ++j) acc += 2.0L * in[j] * sinl(PI_L * (j + 0.5L) * (k + 1) / n);
      break;
    case ORC_RODFT01:
      acc = ((k & 1) ? -1.0L : 1.0L) * in[n - 1];
      for (int j = 0; j < n - 1; ++j) acc += 2.0L * in[j] * sinl(PI_L * (j + 1) * (k + 0.5L) / n);
      break;
    default:
      return -1;
    }
    out[k] = (double)acc;
  }
  return 0;
}

/* Fast path: DCT-II/III through a 2N-point complex DFT of the even extension;
 * DST-II/III through  DST2(x)[k] = DCT2((-1)^n x)[N-1-k]  and
 * DST3(X)[n] = (-1)^n DCT3(rev X)[n].  work holds >= 6N complex entries. */
static void r2r_fast(int n, int kind, const double *in, int istride, double *out,
                     int ostride, cplx *work) {
  const fft_plan *pl = get_plan(2 * n);
  cplx *a = work, *b = work + 2 * n, *s = work + 4 * n;
  int dst = (kind == ORC_RODFT10 || kind == ORC_RODFT01);
  if (kind == ORC_REDFT10 || kind == ORC_RODFT10) {
    for (int j = 0; j < n; ++j) {
      double x = in[(size_t)j * istride];
      if (dst && (j & 1)) x = -x;
      a[j] = x;
      a[2 * n - 1 - j] = x;
    }
    fft_any(pl, a, b, -1, s);
    for (int k = 0; k < n; ++k) {
      /* Y_k = Re(exp(-i pi k / 2N) U_k) */
      double y = creal(pl->post[k] * b[k]);
      int ko = dst ? (n - 1 - k) : k;
      out[(size_t)ko * ostride] = y;
    }
  } else {
    for (int k = 0; k < n; ++k) {
      double x = dst ? in[(size_t)(n - 1 - k) * istride] : in[(size_t)k * istride];
      double w = k ? 2.0 : 1.0;
      a[k] = w * x * conj(pl->post[k]);
    }
    for (int k = n; k < 2 * n; ++k) a[k] = 0;
    fft_any(pl, a, b, +1, s);
    for (int j = 0; j < n; ++j) {
      double y = creal(b[j]);
      if (dst && (j & 1)) y = -y;
      out[(size_t)j * ostride] = y;
    }
  }
}

int orc_r2r_1d(int n, int kind, const double *in, double *out) {
  if (n < 1) return -1;
  if (kind != ORC_REDFT10 && kind != ORC_REDFT01 && kind != ORC_RODFT10 && kind != ORC_RODFT01)
    return -1;
  cplx *work = (cplx *)malloc(sizeof(cplx) * 6 * (size_t)n);
  double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
  r2r_fast(n, kind, in, 1, tmp, 1, work);
  memcpy(out, tmp, sizeof(double) * (size_t)n);
  free(tmp);
  free(work);
  return 0;
}

/* In-place separable 3D r2r on row-major [nx][ny][nz] (z fastest); the three
 * axis kinds are independent, as in fftw_mpi_plan_r2r_3d (quasi.cpp:449-495). */
static void r2r3d_inplace(int nx, int ny, int nz, int kx, int ky, int kz, double *a) {
  get_plan(2 * nx); get_plan(2 * ny); get_plan(2 * nz);
  int nmax = nx > ny ? nx : ny;
  if (nz > nmax) nmax = nz;
#pragma omp parallel
  {
    cplx *work = (cplx *)malloc(sizeof(cplx) * 6 * (size_t)nmax);
    double *line = (double *)malloc(sizeof(double) * (size_t)nmax);
    /* z */
#pragma omp for schedule(static)
    for (long r = 0; r < (long)nx * ny; ++r) {
      double *p = a + (size_t)r * nz;
      r2r_fast(nz, kz, p, 1, line, 1, work);
      memcpy(p, line, sizeof(double) * (size_t)nz);
    }
    /* y */
#pragma omp for schedule(static)
    for (long r = 0; r < (long)nx * nz; ++r) {
      long i = r / nz, k = r % nz;
      double *p = a + (size_t)i * ny * nz + k;
      r2r_fast(ny, ky, p, nz, line, 1, work);
      for (int j = 0; j < ny; ++j) p[(size_t)j * nz] = line[j];
    }
    /* x */
#pragma omp for schedule(static)
    for (long r = 0; r < (long)ny * nz; ++r) {
      double *p = a + r;
      r2r_fast(nx, kx, p, ny * nz, line, 1, work);
      for (int i = 0; i < nx; ++i) p[(size_t)i * ny * nz] = line[i];
    }
    free(line);
    free(work);
  }
}

int orc_r2r_3d(int nx, int ny, int nz, int kx, int ky, int kz, const double *in, double *out) {
  size_t n = (size_t)nx * ny * nz;
  if (out != in) memcpy(out, in, n * sizeof(double));
  r2r3d_inplace(nx, ny, nz, kx, ky, kz, out);
  return 0;
}

/* ------------------------------------------------------------------ */
/* wavenumbers and shifted spectral stencils                            */
/* ------------------------------------------------------------------ */

/* quasi.cpp:397-431: q[m] = pi m / L (DCT modes), s[m] = pi (m+1) / L (DST). */
void orc_wavenumbers(int n, double len, double *q, double *s) {
  double d = len / n;
  for (int m = 0; m < n; ++m) {
    q[m] = M_PI * m / (d * n);
    s[m] = M_PI * (m + 1) / (d * n);
  }
}

static double box_len(const orc_params *p, int n) { return n * 2.0 * M_PI / (p->q0 * p->nw); }

#define IDX(i, j, k) (((size_t)(i) * ny + (j)) * nz + (k))

/* quasi.cpp:1062-1175: ghat_c[..m..] = -s_c[m] * fhat[..m+1..], 0 at m = N_c-1.
 * With DST index m carrying wavenumber s[m] = q[m+1], this is d/dc of the
 * cosine interpolant. */
void orc_spec_grad(int nx, int ny, int nz, const double *qs_x, const double *qs_y,
                   const double *qs_z, int axis, const double *fhat, double *out) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        double v = 0;
        if (axis == 0 && i < nx - 1) v = -qs_x[i] * fhat[IDX(i + 1, j, k)];
        if (axis == 1 && j < ny - 1) v = -qs_y[j] * fhat[IDX(i, j + 1, k)];
        if (axis == 2 && k < nz - 1) v = -qs_z[k] * fhat[IDX(i, j, k + 1)];
        out[IDX(i, j, k)] = v;
      }
}

/* Slab that the P-rank reference reads for "S_x at x-mode i-1" in the v_y and
 * v_z stencils.  quasi.cpp:1639-1650 fills the halo from S_x[index] with
 * index = j*Nz+k, i.e. slab i_local = 0 of the sending rank, so at every rank
 * boundary i = r*n0 (r >= 1) the neighbour is (r-1)*n0 instead of i-1.
 * Returns -1 for "no neighbour" (i = 0). */
int orc_halo_sx_slab(int nx, int compat_ranks, int i) {
  if (i == 0) return -1;
  if (compat_ranks <= 0) return i - 1;
  int n0 = nx / compat_ranks;
  if (i % n0 == 0) return i - n0;
  return i - 1;
}

/* Pressure (quasi.cpp:1385-1415) and velocity projection (quasi.cpp:1425-1872).
 * Sx, Sy, Sz: forward DST_c outputs; dhat: forward DCT of D (raw).
 * Outputs phat (scaled, ready for DCT-III), vx/vy/vz hat (ready for the
 * mixed inverse transforms).  Any of the outputs may be NULL. */
void orc_project(const orc_params *pr, const double *Sx, const double *Sy, const double *Sz,
                 const double *dhat, double *phat, double *vxh, double *vyh, double *vzh) {
  int nx = pr->nx, ny = pr->ny, nz = pr->nz;
  double *qx = malloc(sizeof(double) * nx), *sx = malloc(sizeof(double) * nx);
  double *qy = malloc(sizeof(double) * ny), *sy = malloc(sizeof(double) * ny);
  double *qz = malloc(sizeof(double) * nz), *sz = malloc(sizeof(double) * nz);
  orc_wavenumbers(nx, box_len(pr, nx), qx, sx);
  orc_wavenumbers(ny, box_len(pr, ny), qy, sy);
  orc_wavenumbers(nz, box_len(pr, nz), qz, sz);
  double scale = 0.125 / ((double)nx * ny * nz);
  double nu = pr->nu;
  int stb = pr->variant == 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    int ih = orc_halo_sx_slab(nx, pr->compat_ranks, i);
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t id = IDX(i, j, k);
        double mq2 = qx[i] * qx[i] + qy[j] * qy[j] + qz[k] * qz[k];
        if (phat) {
          /* quasi.cpp:1387-1415; stb (quasi-stb.cpp:1394-1402) uses i*j*k > 0
           * and adds (4/3)*nu*D with integer 4/3 = 1 */
          double a = i > 0 ? Sx[IDX(i - 1, j, k)] : 0.0;
          double b = j > 0 ? Sy[IDX(i, j - 1, k)] : 0.0;
          double c = k > 0 ? Sz[IDX(i, j, k - 1)] : 0.0;
          int on = stb ? ((long)i * j * k > 0) : (i + j + k != 0);
          if (on) {
            double v = -(qx[i] * a + qy[j] * b + qz[k] * c) / mq2;
            if (stb) v += 1.0 * nu * (dhat ? dhat[id] : 0.0);
            phat[id] = scale * v;
          } else {
            phat[id] = 0;
          }
        }
        if (vxh) {
          /* quasi.cpp:1522-1634 */
          double v = 0;
          if (i < nx - 1) {
            double m2 = sx[i] * sx[i] + qy[j] * qy[j] + qz[k] * qz[k];
            double cm1 = scale / (nu * m2), cm2 = sx[i] / m2;
            double syv = j > 0 ? Sy[IDX(i + 1, j - 1, k)] : 0.0;
            double szv = k > 0 ? Sz[IDX(i + 1, j, k - 1)] : 0.0;
            double dot = sx[i] * Sx[id] + qy[j] * syv + qz[k] * szv;
            double dd = dhat ? dhat[IDX(i + 1, j, k)] : 0.0;
            v = cm1 * (Sx[id] - cm2 * dot + nu * sx[i] * dd);
          }
          vxh[id] = v;
        }
        if (vyh || vzh) {
          /* quasi.cpp:1681-1872: v_y, v_z are zero at j = Ny-1 or k = Nz-1 */
          double vy = 0, vz = 0;
          if (j < ny - 1 && k < nz - 1) {
            double sxa = ih >= 0 ? Sx[IDX(ih, j + 1, k)] : 0.0;
            double szv = k > 0 ? Sz[IDX(i, j + 1, k - 1)] : 0.0;
            double m2 = qx[i] * qx[i] + sy[j] * sy[j] + qz[k] * qz[k];
            double dot = qx[i] * sxa + sy[j] * Sy[id] + qz[k] * szv;
            double dd = dhat ? dhat[IDX(i, j + 1, k)] : 0.0;
            vy = (scale / (nu * m2)) * (Sy[id] - (sy[j] / m2) * dot + nu * sy[j] * dd);

            double sxb = ih >= 0 ? Sx[IDX(ih, j, k + 1)] : 0.0;
            double syv = j > 0 ? Sy[IDX(i, j - 1, k + 1)] : 0.0;
            m2 = qx[i] * qx[i] + qy[j] * qy[j] + sz[k] * sz[k];
            dot = qx[i] * sxb + qy[j] * syv + sz[k] * Sz[id];
            dd = dhat ? dhat[IDX(i, j, k + 1)] : 0.0;
            vz = (scale / (nu * m2)) * (Sz[id] - (sz[k] / m2) * dot + nu * sz[k] * dd);
          }
          if (vyh) vyh[id] = vy;
          if (vzh) vzh[id] = vz;
        }
      }
  }
  free(qx); free(sx); free(qy); free(sy); free(qz); free(sz);
}

/* ------------------------------------------------------------------ */
/* pointwise stages of the step (each exported so that the tests can    */
/* feed it prescribed fields; the step below composes them in the       */
/* reference's order)                                                   */
/* ------------------------------------------------------------------ */

/* quasi.cpp:1237-1247: rho* = kp sqrt(q0^2 psi^2 + |grad psi|^2) + rho0 */
void orc_density_law(const orc_params *p, size_t n, const double *psi, const double *gx,
                     const double *gy, const double *gz, double *rho_star) {
  const double q02 = p->q0 * p->q0;
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < n; ++t)
    rho_star[t] = p->kp * sqrt(q02 * pow(psi[t], 2) + pow(gx[t], 2) + pow(gy[t], 2) +
                               pow(gz[t], 2)) + p->rho0;
}

/* quasi.cpp:1249-1268: rhohat = scale DCT(rho*) exp(-filter K^2), rho = IDCT rhohat.
 * rho_star is left unchanged. */
void orc_density_filter(const orc_params *p, const double *rho_star, double *rhohat,
                        double *rho) {
  int nx = p->nx, ny = p->ny, nz = p->nz;
  size_t n = (size_t)nx * ny * nz;
  double sc = 0.125 / ((double)n);
  double *qx = malloc(sizeof(double) * nx), *sx = malloc(sizeof(double) * nx);
  double *qy = malloc(sizeof(double) * ny), *sy = malloc(sizeof(double) * ny);
  double *qz = malloc(sizeof(double) * nz), *sz = malloc(sizeof(double) * nz);
  orc_wavenumbers(nx, nx * 2.0 * M_PI / (p->q0 * p->nw), qx, sx);
  orc_wavenumbers(ny, ny * 2.0 * M_PI / (p->q0 * p->nw), qy, sy);
  orc_wavenumbers(nz, nz * 2.0 * M_PI / (p->q0 * p->nw), qz, sz);
  memcpy(rhohat, rho_star, n * sizeof(double));
  r2r3d_inplace(nx, ny, nz, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10, rhohat);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
    size_t id = IDX(i, j, k);
    double mq2 = pow(qx[i], 2) + pow(qy[j], 2) + pow(qz[k], 2);
    rhohat[id] = sc * rhohat[id] * exp(-p->filter * mq2);
  }
  memcpy(rho, rhohat, n * sizeof(double));
  r2r3d_inplace(nx, ny, nz, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01, rho);
  free(qx); free(sx); free(qy); free(sy); free(qz); free(sz);
}

/* quasi.cpp:1273-1287: D = -((rho - rho_old)/dt + v.r)/(0.5 (rho + rho_old)) */
void orc_divergence(const orc_params *p, size_t n, const double *rho, const double *rho_old,
                    const double *vx, const double *vy, const double *vz, const double *rx,
                    const double *ry, const double *rz, double *D) {
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < n; ++t)
    D[t] = -((rho[t] - rho_old[t]) / p->dt + vx[t] * rx[t] + vy[t] * ry[t] + vz[t] * rz[t]) /
           (0.5 * (rho[t] + rho_old[t]));
}

/* quasi.cpp:1298-1317: G <- rho G (kept: N reads it), S_c = G g_c - rho F l_c */
void orc_stress(size_t n, const double *rho, double *G, const double *F, const double *gx,
                const double *gy, const double *gz, const double *lx, const double *ly,
                const double *lz, double *Sx, double *Sy, double *Sz) {
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < n; ++t) {
    double g = rho[t] * G[t];
    G[t] = g;
    double rf = rho[t] * F[t];
    Sx[t] = g * gx[t] - rf * lx[t];
    Sy[t] = g * gy[t] - rf * ly[t];
    Sz[t] = g * gz[t] - rf * lz[t];
  }
}

/* quasi.cpp:1905-1931: N = rho (eps psi - F + beta psi^3 - gamma psi^5) - G - v.g,
 * plus p kp |grad rho| / rho_s^2 (rho_s = kp Amp + rho0) from step divvSwitch + 2 */
void orc_nonlinear(const orc_params *p, size_t n, int step, const double *rho,
                   const double *psi, const double *F, const double *G, const double *vx,
                   const double *vy, const double *vz, const double *gx, const double *gy,
                   const double *gz, const double *P, const double *rx, const double *ry,
                   const double *rz, double *N) {
  double rs2 = pow(p->kp * p->amp + p->rho0, 2);
  int with_p = step > p->divv_switch + 1;
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < n; ++t) {
    double ps = psi[t];
    double v = rho[t] * (p->eps * ps - F[t] + p->beta * pow(ps, 3) - p->gamma * pow(ps, 5)) -
               G[t] - vx[t] * gx[t] - vy[t] * gy[t] - vz[t] * gz[t];
    if (with_p) v += P[t] * p->kp * sqrt(pow(rx[t], 2) + pow(ry[t], 2) + pow(rz[t], 2)) / rs2;
    N[t] = v;
  }
}

/* quasi-stb.cpp:1284-1307: mu = sh - beta psi^3 + gamma psi^5, S_c = rho mu g_c (rho = 1) */
void orc_stb_stress(const orc_params *p, size_t n, const double *sh, const double *psi,
                    const double *gx, const double *gy, const double *gz, double *mu_out,
                    double *Sx, double *Sy, double *Sz) {
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < n; ++t) {
    double rho = 1;
    double mu = sh[t] - p->beta * pow(psi[t], 3) + p->gamma * pow(psi[t], 5);
    mu_out[t] = mu;
    Sx[t] = rho * mu * gx[t];
    Sy[t] = rho * mu * gy[t];
    Sz[t] = rho * mu * gz[t];
  }
}

/* quasi-stb.cpp:1893-1923: N = beta psi^3 - gamma psi^5 [- v.g if hydro == 1] and the
 * time derivative written to dtpsi.dat, dtpsi = -rho mu [- v.g] (rho = 1) */
void orc_stb_nonlinear(const orc_params *p, size_t n, const double *psi, const double *mu,
                       const double *vx, const double *vy, const double *vz, const double *gx,
                       const double *gy, const double *gz, double *N, double *dtpsi) {
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < n; ++t) {
    double rho = 1;
    double adv = vx[t] * gx[t] + vy[t] * gy[t] + vz[t] * gz[t];
    double nl = p->beta * pow(psi[t], 3) - p->gamma * pow(psi[t], 5);
    double dp = -rho * mu[t];
    if (p->hydro == 1) { nl -= adv; dp -= adv; }
    N[t] = nl;
    dtpsi[t] = dp;
  }
}

/* quasi-stb.cpp:2147-2150: the mode (i*, 0, k*) of the growth rate, i* =
 * round(Qi / q[1]), k* = round(q0 / q_z[1]).  The reference takes q_z[1] for
 * both (its grids are cubic); here i* uses the x table, the same value on a
 * cubic grid.  Returns 0 when the mode lies on the grid. */
int orc_stb_growth_mode(const orc_params *p, int *is, int *ks) {
  double qx1 = M_PI / (p->nx * 2.0 * M_PI / (p->q0 * p->nw));
  double qz1 = M_PI / (p->nz * 2.0 * M_PI / (p->q0 * p->nw));
  *is = (int)lround(p->growth_q / qx1);
  *ks = (int)lround(p->q0 / qz1);
  return (*is < p->nx && *ks < p->nz) ? 0 : -1;
}

/* ------------------------------------------------------------------ */
/* state and step                                                       */
/* ------------------------------------------------------------------ */

enum {
  ORC_PSI = 0, ORC_RHO, ORC_VX, ORC_VY, ORC_VZ, ORC_P, ORC_DIVV, ORC_PSIHAT,
  ORC_F, ORC_G, ORC_GX, ORC_GY, ORC_GZ, ORC_RX, ORC_RY, ORC_RZ, ORC_LX, ORC_LY, ORC_LZ,
  ORC_RHOHAT, ORC_NHAT, ORC_PSIHAT_OLD, ORC_PSI_OLD, ORC_SX, ORC_SY, ORC_SZ, ORC_DHAT,
  ORC_MU, ORC_DTPSI, ORC_RHO_OLD, ORC_NFIELDS
};

typedef struct {
  orc_params p;
  size_t n;
  int step;
  double *qx, *sx, *qy, *sy, *qz, *sz;
  double *f[ORC_NFIELDS];
  double growth; /* stb growth of the last step */
} orc_state;

orc_state *orc_create(const orc_params *p) {
  orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
  s->p = *p;
  s->n = (size_t)p->nx * p->ny * p->nz;
  for (int f = 0; f < ORC_NFIELDS; ++f) s->f[f] = (double *)calloc(s->n, sizeof(double));
  s->qx = malloc(sizeof(double) * p->nx); s->sx = malloc(sizeof(double) * p->nx);
  s->qy = malloc(sizeof(double) * p->ny); s->sy = malloc(sizeof(double) * p->ny);
  s->qz = malloc(sizeof(double) * p->nz); s->sz = malloc(sizeof(double) * p->nz);
  orc_wavenumbers(p->nx, box_len(p, p->nx), s->qx, s->sx);
  orc_wavenumbers(p->ny, box_len(p, p->ny), s->qy, s->sy);
  orc_wavenumbers(p->nz, box_len(p, p->nz), s->qz, s->sz);
  prepare_plans(p);
  return s;
}

void orc_destroy(orc_state *s) {
  if (!s) return;
  for (int f = 0; f < ORC_NFIELDS; ++f) free(s->f[f]);
  free(s->qx); free(s->sx); free(s->qy); free(s->sy); free(s->qz); free(s->sz);
  free(s);
}

static double scale_of(const orc_params *p) { return 0.125 / ((double)p->nx * p->ny * p->nz); }

/* quasi.cpp:905-951 (stb 874-920): psihat = scale*DCT(psi), psi = IDCT(psihat),
 * psihat_old = psihat, psi_old = psi; rhohat, rho_old, v, D start at zero. */
void orc_set_psi(orc_state *s, const double *psi) {
  const orc_params *p = &s->p;
  int nx = p->nx, ny = p->ny, nz = p->nz;
  double sc = scale_of(p);
  double *ph = s->f[ORC_PSIHAT];
  orc_r2r_3d(nx, ny, nz, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10, psi, ph);
  for (size_t t = 0; t < s->n; ++t) ph[t] *= sc;
  orc_r2r_3d(nx, ny, nz, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01, ph, s->f[ORC_PSI]);
  memcpy(s->f[ORC_PSIHAT_OLD], ph, s->n * sizeof(double));
  memcpy(s->f[ORC_PSI_OLD], s->f[ORC_PSI], s->n * sizeof(double));
  /* setup N (quasi.cpp:929-947) is overwritten at step 1 before it is read */
  int zero[] = {ORC_RHOHAT, ORC_RHO, ORC_VX, ORC_VY, ORC_VZ, ORC_DIVV, ORC_DHAT, ORC_P,
                ORC_MU, ORC_DTPSI, ORC_RHO_OLD};
  for (unsigned t = 0; t < sizeof(zero) / sizeof(zero[0]); ++t)
    memset(s->f[zero[t]], 0, s->n * sizeof(double));
  double *nh = s->f[ORC_NHAT];
  const double *ps = s->f[ORC_PSI];
  for (size_t t = 0; t < s->n; ++t)
    nh[t] = p->beta * pow(ps[t], 3) - p->gamma * pow(ps[t], 5);
  r2r3d_inplace(nx, ny, nz, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10, nh);
  s->step = 0;
  s->growth = 0;
}

static void inv3(const orc_params *p, int axis_dst, double *a) {
  int kx = ORC_REDFT01, ky = ORC_REDFT01, kz = ORC_REDFT01;
  if (axis_dst == 0) kx = ORC_RODFT01;
  if (axis_dst == 1) ky = ORC_RODFT01;
  if (axis_dst == 2) kz = ORC_RODFT01;
  r2r3d_inplace(p->nx, p->ny, p->nz, kx, ky, kz, a);
}

static void fwd3(const orc_params *p, int axis_dst, double *a) {
  int kx = ORC_REDFT10, ky = ORC_REDFT10, kz = ORC_REDFT10;
  if (axis_dst == 0) kx = ORC_RODFT10;
  if (axis_dst == 1) ky = ORC_RODFT10;
  if (axis_dst == 2) kz = ORC_RODFT10;
  r2r3d_inplace(p->nx, p->ny, p->nz, kx, ky, kz, a);
}

/* S10 (quasi.cpp:1944-1965; stb 1944-1957): psihat_new from psihat, psihat_old,
 * N and N_old; psihat_old <- psihat; psi_old <- psi; psi = IDCT psihat_new. */
static void update(orc_state *s, const double *nhat_old) {
  const orc_params *p = &s->p;
  int nx = p->nx, ny = p->ny, nz = p->nz;
  const double *qx = s->qx, *qy = s->qy, *qz = s->qz;
  double sc = scale_of(p), q02 = p->q0 * p->q0;
  int stb = p->variant == 1;
  double *ph = s->f[ORC_PSIHAT], *pho = s->f[ORC_PSIHAT_OLD];
  const double *nh = s->f[ORC_NHAT];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
    size_t id = IDX(i, j, k);
    double mq2 = pow(qx[i], 2) + pow(qy[j], 2) + pow(qz[k], 2);
    double alin = p->eps - p->alpha * pow(q02 - mq2, 2);
    double old = ph[id];
    if (stb) {
      /* quasi-stb.cpp:855-856, 1954-1956: Crank-Nicolson, no psihat_old term */
      double c1 = 1.0 + (p->dt / 2) * alin, c2 = 1.0 - (p->dt / 2) * alin;
      ph[id] = (c1 * old + (p->dt / 2) * sc * (3.0 * nh[id] - nhat_old[id])) / c2;
    } else {
      /* quasi.cpp:883-884, 1949-1951 */
      double c1 = 2.0 - p->dt * p->rho_m * alin, c2 = 1.5 - p->dt * p->rho_m * alin;
      ph[id] = (c1 * old - 0.5 * pho[id] + (p->dt / 2) * sc * (3.0 * nh[id] - nhat_old[id])) / c2;
    }
    pho[id] = old;
  }
  memcpy(s->f[ORC_PSI_OLD], s->f[ORC_PSI], s->n * sizeof(double));
  memcpy(s->f[ORC_PSI], ph, s->n * sizeof(double));
  inv3(p, -1, s->f[ORC_PSI]);
}

/* One step of quasi.cpp:1017-1965. */
static void quasi_step(orc_state *s) {
  const orc_params *p = &s->p;
  int nx = p->nx, ny = p->ny, nz = p->nz;
  const double *qx = s->qx, *qy = s->qy, *qz = s->qz, *sx = s->sx, *sy = s->sy, *sz = s->sz;
  double **f = s->f;
  double sc = scale_of(p), q02 = p->q0 * p->q0;
  int n = ++s->step;
  double *nhat_old = (double *)malloc(s->n * sizeof(double));
  memcpy(nhat_old, f[ORC_NHAT], s->n * sizeof(double)); /* quasi.cpp:1032 */

  /* S1: F, G (quasi.cpp:1034-1057) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
    size_t id = IDX(i, j, k);
    double mq2 = pow(qx[i], 2) + pow(qy[j], 2) + pow(qz[k], 2);
    double fh = p->alpha * (q02 - mq2) * f[ORC_PSIHAT][id];
    f[ORC_F][id] = fh;
    f[ORC_G][id] = -mq2 * fh;
  }
  inv3(p, -1, f[ORC_F]);
  inv3(p, -1, f[ORC_G]);

  /* S2: gradients of psi, lagged rho (from the previous step's rhohat, before
   * S3 rewrites it), laplacians of grad psi (quasi.cpp:1060-1232) */
  for (int c = 0; c < 3; ++c) {
    orc_spec_grad(nx, ny, nz, sx, sy, sz, c, f[ORC_PSIHAT], f[ORC_GX + c]);
    orc_spec_grad(nx, ny, nz, sx, sy, sz, c, f[ORC_RHOHAT], f[ORC_RX + c]);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
    size_t id = IDX(i, j, k);
    double m;
    m = pow(sx[i], 2) + pow(qy[j], 2) + pow(qz[k], 2);
    f[ORC_LX][id] = -m * f[ORC_GX][id];
    m = pow(qx[i], 2) + pow(sy[j], 2) + pow(qz[k], 2);
    f[ORC_LY][id] = -m * f[ORC_GY][id];
    m = pow(qx[i], 2) + pow(qy[j], 2) + pow(sz[k], 2);
    f[ORC_LZ][id] = -m * f[ORC_GZ][id];
  }
  for (int c = 0; c < 3; ++c) {
    inv3(p, c, f[ORC_GX + c]);
    inv3(p, c, f[ORC_RX + c]);
    inv3(p, c, f[ORC_LX + c]);
  }

  /* S3: density law and filter (quasi.cpp:1237-1268); rho* goes through the
   * D field's array, rewritten below before anything reads it */
  double *rho = f[ORC_RHO], *dv = f[ORC_DIVV];
  double *rstar = (double *)malloc(s->n * sizeof(double));
  orc_density_law(p, s->n, f[ORC_PSI], f[ORC_GX], f[ORC_GY], f[ORC_GZ], rstar);
  orc_density_filter(p, rstar, f[ORC_RHOHAT], rho);
  free(rstar);

  /* S4: divergence (quasi.cpp:1273-1293).  D keeps its last value (0 from
   * set-up) up to divvSwitch; rho_old <- rho afterwards. */
  if (n > p->divv_switch)
    orc_divergence(p, s->n, rho, f[ORC_RHO_OLD], f[ORC_VX], f[ORC_VY], f[ORC_VZ], f[ORC_RX],
                   f[ORC_RY], f[ORC_RZ], dv);
  memcpy(f[ORC_RHO_OLD], rho, s->n * sizeof(double));
  memcpy(f[ORC_DHAT], dv, s->n * sizeof(double));
  fwd3(p, -1, f[ORC_DHAT]);

  /* S5: stress (quasi.cpp:1298-1330); G <- rho G */
  orc_stress(s->n, rho, f[ORC_G], f[ORC_F], f[ORC_GX], f[ORC_GY], f[ORC_GZ], f[ORC_LX],
             f[ORC_LY], f[ORC_LZ], f[ORC_SX], f[ORC_SY], f[ORC_SZ]);
  for (int c = 0; c < 3; ++c) fwd3(p, c, f[ORC_SX + c]);

  /* S6, S7: pressure and projection (quasi.cpp:1336-1884) */
  orc_project(p, f[ORC_SX], f[ORC_SY], f[ORC_SZ], f[ORC_DHAT], f[ORC_P], f[ORC_VX], f[ORC_VY],
              f[ORC_VZ]);
  inv3(p, -1, f[ORC_P]);
  for (int c = 0; c < 3; ++c) inv3(p, c, f[ORC_VX + c]);

  /* S8: divv output (quasi.cpp:1888-1900) */
  for (size_t t = 0; t < s->n; ++t) dv[t] = sc * f[ORC_DHAT][t];
  inv3(p, -1, dv);

  /* S9: nonlinear term (quasi.cpp:1905-1939) */
  orc_nonlinear(p, s->n, n, rho, f[ORC_PSI], f[ORC_F], f[ORC_G], f[ORC_VX], f[ORC_VY],
                f[ORC_VZ], f[ORC_GX], f[ORC_GY], f[ORC_GZ], f[ORC_P], f[ORC_RX], f[ORC_RY],
                f[ORC_RZ], f[ORC_NHAT]);
  fwd3(p, -1, f[ORC_NHAT]);
  if (n == 1) memcpy(nhat_old, f[ORC_NHAT], s->n * sizeof(double));

  /* S10: update (quasi.cpp:1944-1965) */
  update(s, nhat_old);
  free(nhat_old);
}

/* One step of quasi-stb.cpp:986-1968 (live path; dead transforms dropped
 * because their results are never read). */
static void stb_step(orc_state *s) {
  const orc_params *p = &s->p;
  int nx = p->nx, ny = p->ny, nz = p->nz;
  const double *qx = s->qx, *qy = s->qy, *qz = s->qz, *sx = s->sx, *sy = s->sy, *sz = s->sz;
  double **f = s->f;
  double q02 = p->q0 * p->q0;
  int n = ++s->step;
  double *nhat_old = (double *)malloc(s->n * sizeof(double));
  memcpy(nhat_old, f[ORC_NHAT], s->n * sizeof(double));

  /* sh = IDCT(-aLin psihat) (quasi-stb.cpp:1015-1038), stored in ORC_F */
  double *sh = f[ORC_F];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
    size_t id = IDX(i, j, k);
    double mq2 = pow(qx[i], 2) + pow(qy[j], 2) + pow(qz[k], 2);
    double alin = p->eps - p->alpha * pow(q02 - mq2, 2);
    sh[id] = -alin * f[ORC_PSIHAT][id];
  }
  inv3(p, -1, sh);
  for (int c = 0; c < 3; ++c) {
    orc_spec_grad(nx, ny, nz, sx, sy, sz, c, f[ORC_PSIHAT], f[ORC_GX + c]);
    inv3(p, c, f[ORC_GX + c]);
  }
  /* rho = 1 (quasi-stb.cpp:1232); mu and stress (1284-1308) */
  for (size_t t = 0; t < s->n; ++t) f[ORC_RHO][t] = 1;
  orc_stb_stress(p, s->n, sh, f[ORC_PSI], f[ORC_GX], f[ORC_GY], f[ORC_GZ], f[ORC_MU],
                 f[ORC_SX], f[ORC_SY], f[ORC_SZ]);
  for (int c = 0; c < 3; ++c) fwd3(p, c, f[ORC_SX + c]);
  /* divv_local stays 0 in stb, so Dhat = 0 */
  memset(f[ORC_DHAT], 0, s->n * sizeof(double));
  orc_project(p, f[ORC_SX], f[ORC_SY], f[ORC_SZ], f[ORC_DHAT], NULL, f[ORC_VX], f[ORC_VY],
              f[ORC_VZ]);
  for (int c = 0; c < 3; ++c) inv3(p, c, f[ORC_VX + c]);

  /* N and dtpsi (quasi-stb.cpp:1893-1923) */
  orc_stb_nonlinear(p, s->n, f[ORC_PSI], f[ORC_MU], f[ORC_VX], f[ORC_VY], f[ORC_VZ], f[ORC_GX],
                    f[ORC_GY], f[ORC_GZ], f[ORC_NHAT], f[ORC_DTPSI]);
  fwd3(p, -1, f[ORC_NHAT]);
  if (n == 1) memcpy(nhat_old, f[ORC_NHAT], s->n * sizeof(double));

  /* CN update (quasi-stb.cpp:1944-1957) */
  update(s, nhat_old);

  /* growth rate at mode (i*, 0, k*) (quasi-stb.cpp:2147-2158) */
  int is, ks;
  if (n > 1 && orc_stb_growth_mode(p, &is, &ks) == 0) {
    size_t id = IDX(is, 0, ks);
    const double *ph = f[ORC_PSIHAT], *pho = f[ORC_PSIHAT_OLD];
    s->growth = (ph[id] - pho[id]) / (p->dt * pho[id]);
  }
  free(nhat_old);
}

int orc_step(orc_state *s, int nsteps) {
  for (int t = 0; t < nsteps; ++t) {
    if (s->p.variant == 1) stb_step(s);
    else quasi_step(s);
  }
  return 0;
}

int orc_get_step(const orc_state *s) { return s->step; }
double orc_get_growth(const orc_state *s) { return s->growth; }

int orc_get_field(const orc_state *s, int which, double *out) {
  if (which < 0 || which >= ORC_NFIELDS) return -1;
  memcpy(out, s->f[which], s->n * sizeof(double));
  return 0;
}

int orc_set_field(orc_state *s, int which, const double *in) {
  if (which < 0 || which >= ORC_NFIELDS) return -1;
  memcpy(s->f[which], in, s->n * sizeof(double));
  return 0;
}

/* quasi.cpp:1976-2029: out = {L1, mass, energy}.  Plain FP64 sums in index
 * order, like each rank's loop (for P = 1). */
void orc_diag(const orc_state *s, double *out) {
  const orc_params *p = &s->p;
  const double *psi = s->f[ORC_PSI], *po = s->f[ORC_PSI_OLD], *rho = s->f[ORC_RHO],
               *F = s->f[ORC_F];
  double a = 0, b = 0, m = 0, e = 0;
  double dx = 2.0 * M_PI / (p->q0 * p->nw);
  for (size_t t = 0; t < s->n; ++t) {
    a += fabs(psi[t] - po[t]);
    b += fabs(psi[t]);
    m += rho[t];
    e += rho[t] * (-0.5 * p->eps * pow(psi[t], 2) + 0.5 * p->alpha * pow(F[t], 2) -
                   0.25 * p->beta * pow(psi[t], 4) + (p->gamma / 6) * pow(psi[t], 6));
  }
  out[0] = a / (p->dt * b);
  out[1] = m * dx * dx * dx;
  out[2] = e * dx * dx * dx;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
