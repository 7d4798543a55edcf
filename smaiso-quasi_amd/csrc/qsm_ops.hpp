// qsm_ops.hpp -- argument blocks of one r2r pass and the pointwise operators
// fused into the first and last passes of the step's transforms.
//
// Every pointwise loop of the reference step that sits between two
// fftw_execute calls becomes either a LOAD operator (evaluated while the first
// pass of the following transform reads its tile) or a STORE operator
// (evaluated while the last pass of the preceding transform writes its tile).
// Coefficients are evaluated in FP64 in registers from the 1-D wavenumber
// tables; none of the reference's coefficient arrays is materialised.
#pragma once
#include <hip/hip_runtime.h>

namespace qsm {

// geometry of the [nx][ny][nz] field (z fastest)
struct Geo {
  int nx, ny, nz;
  int pad_;
};

// FP64 wavenumber tables, quasi.cpp:397-431: q[m] = pi m/L, s[m] = pi (m+1)/L,
// and the density filter exp(-filter K^2) = fx[i] fy[j] fz[k] with
// f_c[m] = exp(-filter q_c[m]^2) (quasi.cpp:1261 is separable in K^2).
struct Waves {
  const double *qx, *sx, *qy, *sy, *qz, *sz;
  const double *fx, *fy, *fz;
};

// physical constants of the step (quasi.cpp:159-219, quasi-stb.cpp:149-220)
struct Phys {
  double q02, alpha, beta, gamma, eps, nu, kp, rho0, dt, scale, filter, rho_s2, rho_m;
  int hydro;
  int pad_;
};

enum LoadCode : int {
  LD_PLAIN = 0,     // a0[idx] (field type T)
  LD_PLAIN_D,       // (T) a0[idx], a0 is double (psihat)
  LD_ZERO,          // 0 (divergence before divv_switch: the reference transforms zeros)
  LD_COEF_F,        // alpha (q0^2 - K^2) psihat                          quasi.cpp:1042
  LD_COEF_G,        // -K^2 alpha (q0^2 - K^2) psihat                     quasi.cpp:1044-1045
  LD_GRAD,          // -s_c[m] src[m+1 along c], 0 at m = N_c-1 (src: T) quasi.cpp:1124-1171
  LD_GRAD_D,        // same with a double source (psihat)
  LD_LAPGRAD,       // -K_c^2 (-s_c[m] psihat[m+1])                       quasi.cpp:1176-1191
  LD_SH,            // -aLin psihat (stb)                                 quasi-stb.cpp:1023
  LD_RHOSTAR,       // kp sqrt(q0^2 psi^2 + |g|^2) + rho0                 quasi.cpp:1243-1247
  LD_DIVV,          // -((rho-rho_old)/dt + v.r)/(0.5(rho+rho_old))       quasi.cpp:1280-1285
  LD_STRESS,        // (rho G) g_c - (rho F) l_c                          quasi.cpp:1305-1317
  LD_PRESSURE,      // scale (-(q.S_shifted)/K^2)                         quasi.cpp:1385-1415
  LD_VEL0,          // Stokes/Leray projection of S, component x          quasi.cpp:1522-1634
  LD_VEL1,          // component y                                        quasi.cpp:1681-1872
  LD_VEL2,          // component z                                        quasi.cpp:1681-1872
  LD_SCALE,         // scale a0[idx]                                      quasi.cpp:1891-1896
  LD_NONLIN,        // N of quasi.cpp:1914-1923 (steps <= divv_switch+1)
  LD_NONLIN_P,      // N with the pressure term of quasi.cpp:1925-1930
  LD_STRESS_STB,    // rho mu g_c, mu = sh - beta psi^3 + gamma psi^5     quasi-stb.cpp:1295-1307
  LD_NONLIN_STB,    // beta psi^3 - gamma psi^5 - v.g (hydro = 1)         quasi-stb.cpp:1903-1914
  LD_NONLIN_STB0,   // beta psi^3 - gamma psi^5 (hydro = 0)               quasi-stb.cpp:1915-1919
};

enum StoreCode : int {
  ST_PLAIN = 0,     // d0[idx] = v
  ST_FILTER,        // d0[idx] = scale v exp(-filter K^2)                 quasi.cpp:1260-1261
  ST_UPDATE,        // semi-implicit AB2/CN update of psihat              quasi.cpp:1949-1951
  ST_UPDATE_STB,    // CN update of psihat (stb)                          quasi-stb.cpp:1954-1956
};

struct LoadOp {
  int code;
  int axis;      // component c for GRAD/GRAD_D/LAPGRAD
  int flag;      // unused (reserved)
  int n0;        // VEL0-2: slab size of the P-rank halo compat map (0 = exact)
  const void* a[14];
};

struct StoreOp {
  int code;
  int first;           // UPDATE: first step (N_old = N)
  void* d[3];          // outputs
  const void* s[2];    // extra inputs
  const double* dc;    // mixed precision: FP64 sum of the forward input (DC override), or null
  double* dc_out;      // mixed precision: FP64 (0,0,0) mode written here, or null
  const double* dc_old;// UPDATE: FP64 N_old(0,0,0) in mixed precision, or null
};

__device__ __forceinline__ double sq(double x) { return x * x; }

template <typename T>
__device__ __forceinline__ double rd(const void* p, long idx) {
  return (double)((const T*)p)[idx];
}

// p[idx + off] if ok, else 0, without a branch: the load always reads a valid
// address (idx itself when !ok) and the result is selected afterwards, so
// the loads of one element can be issued together.
template <typename T>
__device__ __forceinline__ double rd_if(const void* p, long idx, long off, bool ok) {
  const double v = (double)((const T*)p)[ok ? idx + off : idx];
  return ok ? v : 0.0;
}

// Wavenumber values at one element of an x pass.  An x-pass thread keeps one
// (j, k) for all its elements, so the y and z entries are read once per
// thread; xi is the launch's i-table at the element's i (itab() below).
struct Kv {
  double xi;
  double qy, sy, qz, sz, fy, fz;
};

// The one i-indexed table a launch's operator reads (uniform per launch),
// or null: q_x or s_x for the spectral prologues, q_x for the update and the
// filter factors f_x for the density filter.
template <int LD, int ST>
__device__ __forceinline__ const double* itab(const LoadOp& L, const Waves& W) {
  if constexpr (LD == LD_COEF_F || LD == LD_COEF_G || LD == LD_SH || LD == LD_PRESSURE ||
                LD == LD_VEL1 || LD == LD_VEL2) {
    return W.qx;
  } else if constexpr (LD == LD_VEL0 || LD == LD_GRAD || LD == LD_GRAD_D) {
    return W.sx;
  } else if constexpr (LD == LD_LAPGRAD) {
    return L.axis == 0 ? W.sx : W.qx;
  } else if constexpr (ST == ST_UPDATE || ST == ST_UPDATE_STB) {
    return W.qx;
  } else if constexpr (ST == ST_FILTER) {
    return W.fx;
  } else {
    return nullptr;
  }
}

template <int LD, int ST>
constexpr bool uses_kv() {
  return LD == LD_COEF_F || LD == LD_COEF_G || LD == LD_SH || LD == LD_PRESSURE ||
         LD == LD_VEL0 || LD == LD_VEL1 || LD == LD_VEL2 || LD == LD_GRAD ||
         LD == LD_GRAD_D || LD == LD_LAPGRAD || ST == ST_UPDATE || ST == ST_UPDATE_STB ||
         ST == ST_FILTER;
}

__device__ __forceinline__ Kv kv_thread(const Waves& W, int j, int k) {
  return Kv{0.0, W.qy[j], W.sy[j], W.qz[k], W.sz[k], W.fy[j], W.fz[k]};
}

// Shifted read along axis c used by the gradient stencils: -s_c[m] src[m+1],
// 0 at m = N_c - 1.  The axis is uniform for a launch, so the per-element
// choices below are selects, not branches.
template <typename S>
__device__ __forceinline__ double grad_hat(const LoadOp& L, const Geo& G, const Kv& K, int i,
                                           int j, int k, long idx) {
  const long nyz = (long)G.ny * G.nz;
  const int c = L.axis;
  const int m = c == 0 ? i : (c == 1 ? j : k);
  const int nc = c == 0 ? G.nx : (c == 1 ? G.ny : G.nz);
  const long o = c == 0 ? nyz : (c == 1 ? (long)G.nz : 1L);
  const double st = c == 0 ? K.xi : (c == 1 ? K.sy : K.sz);
  const bool ok = m < nc - 1;
  return -st * rd_if<S>(L.a[0], idx, o, ok);
}

// K carries the element's wavenumbers for the x-pass spectral operators.
template <typename T, int LD>
__device__ __forceinline__ T load_eval(const LoadOp& L, const Phys& P, const Kv& K,
                                       const Geo& G, int i, int j, int k, long idx) {
  if constexpr (LD == LD_PLAIN) {
    return ((const T*)L.a[0])[idx];
  } else if constexpr (LD == LD_PLAIN_D) {
    return (T)((const double*)L.a[0])[idx];
  } else if constexpr (LD == LD_ZERO) {
    return (T)0;
  } else if constexpr (LD == LD_COEF_F || LD == LD_COEF_G) {
    const double mq2 = sq(K.xi) + sq(K.qy) + sq(K.qz);
    const double fh = P.alpha * (P.q02 - mq2) * ((const double*)L.a[0])[idx];
    return (T)(LD == LD_COEF_F ? fh : -mq2 * fh);
  } else if constexpr (LD == LD_GRAD) {
    return (T)grad_hat<T>(L, G, K, i, j, k, idx);
  } else if constexpr (LD == LD_GRAD_D) {
    return (T)grad_hat<double>(L, G, K, i, j, k, idx);
  } else if constexpr (LD == LD_LAPGRAD) {
    // i-table: s_x when c = 0, else q_x
    const int c = L.axis;
    const double ay = c == 1 ? K.sy : K.qy;
    const double az = c == 2 ? K.sz : K.qz;
    const double m = sq(K.xi) + sq(ay) + sq(az);
    return (T)(-m * grad_hat<double>(L, G, K, i, j, k, idx));
  } else if constexpr (LD == LD_SH) {
    const double mq2 = sq(K.xi) + sq(K.qy) + sq(K.qz);
    const double alin = P.eps - P.alpha * sq(P.q02 - mq2);
    return (T)(-alin * ((const double*)L.a[0])[idx]);
  } else if constexpr (LD == LD_RHOSTAR) {
    const T ps = ((const T*)L.a[0])[idx], gx = ((const T*)L.a[1])[idx],
            gy = ((const T*)L.a[2])[idx], gz = ((const T*)L.a[3])[idx];
    return (T)P.kp * sqrt((T)P.q02 * (ps * ps) + gx * gx + gy * gy + gz * gz) + (T)P.rho0;
  } else if constexpr (LD == LD_DIVV) {
    // a: rho, rho_old, vx, vy, vz, rx, ry, rz
    const T r = ((const T*)L.a[0])[idx], ro = ((const T*)L.a[1])[idx];
    T t = (r - ro) / (T)P.dt + ((const T*)L.a[2])[idx] * ((const T*)L.a[5])[idx];
    t = t + ((const T*)L.a[3])[idx] * ((const T*)L.a[6])[idx];
    t = t + ((const T*)L.a[4])[idx] * ((const T*)L.a[7])[idx];
    return -t / ((T)0.5 * (r + ro));
  } else if constexpr (LD == LD_STRESS) {
    // a: rho, G, F, g_c, l_c
    const T r = ((const T*)L.a[0])[idx];
    const T gg = r * ((const T*)L.a[1])[idx];
    const T rf = r * ((const T*)L.a[2])[idx];
    return gg * ((const T*)L.a[3])[idx] - rf * ((const T*)L.a[4])[idx];
  } else if constexpr (LD == LD_PRESSURE) {
    // a: Sx, Sy, Sz (forward DST outputs); i-table q_x
    const long nyz = (long)G.ny * G.nz;
    const double a = rd_if<T>(L.a[0], idx, -nyz, i > 0);
    const double b = rd_if<T>(L.a[1], idx, -(long)G.nz, j > 0);
    const double c = rd_if<T>(L.a[2], idx, -1, k > 0);
    const double mq2 = sq(K.xi) + sq(K.qy) + sq(K.qz);
    const double pv = P.scale * (-(K.xi * a + K.qy * b + K.qz * c) / mq2);
    return (T)((i + j + k == 0) ? 0.0 : pv);
  } else if constexpr (LD == LD_VEL0) {
    // a: Sx, Sy, Sz, Dhat (may be null); i-table s_x.  Global-index form of
    // quasi.cpp:1522-1634.
    const long nyz = (long)G.ny * G.nz;
    const bool in = i < G.nx - 1;  // v_x = 0 at i = Nx-1
    const double nu = P.nu;
    const double rm2 = 1.0 / (sq(K.xi) + sq(K.qy) + sq(K.qz));
    const double cm1 = (P.scale / nu) * rm2, cm2 = K.xi * rm2;
    const double sx = rd<T>(L.a[0], idx);
    const double syv = rd_if<T>(L.a[1], idx, nyz - G.nz, in && j > 0);
    const double szv = rd_if<T>(L.a[2], idx, nyz - 1, in && k > 0);
    const double dot = K.xi * sx + K.qy * syv + K.qz * szv;
    const double dd = L.a[3] ? rd_if<T>(L.a[3], idx, nyz, in) : 0.0;
    const double vx = cm1 * (sx - cm2 * dot + nu * K.xi * dd);
    return (T)(in ? vx : 0.0);
  } else if constexpr (LD == LD_VEL1 || LD == LD_VEL2) {
    // quasi.cpp:1681-1872: zero at j = Ny-1 or k = Nz-1; i-table q_x.  The
    // S_x neighbour slab is i-1, or the P-rank halo slab (quasi.cpp:1639-1650).
    const long nyz = (long)G.ny * G.nz;
    const bool in = j < G.ny - 1 && k < G.nz - 1;
    const double nu = P.nu;
    int ih = i - 1;
    if (L.n0 > 0 && i > 0 && (i % L.n0) == 0) ih = i - L.n0;
    const long dx = (long)(ih - i) * nyz;
    double vv;
    if constexpr (LD == LD_VEL1) {
      const double rm2 = 1.0 / (sq(K.xi) + sq(K.sy) + sq(K.qz));
      const double sy = rd<T>(L.a[1], idx);
      const double sxa = rd_if<T>(L.a[0], idx, dx + G.nz, in && i > 0);
      const double szv = rd_if<T>(L.a[2], idx, G.nz - 1, in && k > 0);
      const double dot = K.xi * sxa + K.sy * sy + K.qz * szv;
      const double dd = L.a[3] ? rd_if<T>(L.a[3], idx, G.nz, in) : 0.0;
      vv = ((P.scale / nu) * rm2) * (sy - (K.sy * rm2) * dot + nu * K.sy * dd);
    } else {
      const double rm2 = 1.0 / (sq(K.xi) + sq(K.qy) + sq(K.sz));
      const double sz = rd<T>(L.a[2], idx);
      const double sxb = rd_if<T>(L.a[0], idx, dx + 1, in && i > 0);
      const double syv = rd_if<T>(L.a[1], idx, 1 - (long)G.nz, in && j > 0);
      const double dot = K.xi * sxb + K.qy * syv + K.sz * sz;
      const double dd = L.a[3] ? rd_if<T>(L.a[3], idx, 1, in) : 0.0;
      vv = ((P.scale / nu) * rm2) * (sz - (K.sz * rm2) * dot + nu * K.sz * dd);
    }
    return (T)(in ? vv : 0.0);
  } else if constexpr (LD == LD_SCALE) {
    return (T)(P.scale * rd<T>(L.a[0], idx));
  } else if constexpr (LD == LD_NONLIN || LD == LD_NONLIN_P) {
    // a: rho, psi, F, G, vx, vy, vz, gx, gy, gz, p, rx, ry, rz
    const T r = ((const T*)L.a[0])[idx], ps = ((const T*)L.a[1])[idx];
    const T ps2 = ps * ps, ps3 = ps2 * ps, ps5 = ps3 * ps2;
    T v = r * ((T)P.eps * ps - ((const T*)L.a[2])[idx] + (T)P.beta * ps3 - (T)P.gamma * ps5) -
          r * ((const T*)L.a[3])[idx];
    v = v - ((const T*)L.a[4])[idx] * ((const T*)L.a[7])[idx];
    v = v - ((const T*)L.a[5])[idx] * ((const T*)L.a[8])[idx];
    v = v - ((const T*)L.a[6])[idx] * ((const T*)L.a[9])[idx];
    if constexpr (LD == LD_NONLIN_P) {
      const T rx = ((const T*)L.a[11])[idx], ry = ((const T*)L.a[12])[idx],
              rz = ((const T*)L.a[13])[idx];
      v = v + ((const T*)L.a[10])[idx] * (T)P.kp * sqrt(rx * rx + ry * ry + rz * rz) /
                  (T)P.rho_s2;
    }
    return v;
  } else if constexpr (LD == LD_STRESS_STB) {
    // a: sh, psi, g_c ; rho = 1
    const T ps = ((const T*)L.a[1])[idx];
    const T ps2 = ps * ps, ps3 = ps2 * ps, ps5 = ps3 * ps2;
    const T mu = ((const T*)L.a[0])[idx] - (T)P.beta * ps3 + (T)P.gamma * ps5;
    return mu * ((const T*)L.a[2])[idx];
  } else if constexpr (LD == LD_NONLIN_STB || LD == LD_NONLIN_STB0) {
    // a: psi, vx, vy, vz, gx, gy, gz
    const T ps = ((const T*)L.a[0])[idx];
    const T ps2 = ps * ps, ps3 = ps2 * ps, ps5 = ps3 * ps2;
    T v = (T)P.beta * ps3 - (T)P.gamma * ps5;
    if constexpr (LD == LD_NONLIN_STB) {
      v = v - ((const T*)L.a[1])[idx] * ((const T*)L.a[4])[idx];
      v = v - ((const T*)L.a[2])[idx] * ((const T*)L.a[5])[idx];
      v = v - ((const T*)L.a[3])[idx] * ((const T*)L.a[6])[idx];
    }
    return v;
  } else {
    static_assert(LD == LD_PLAIN, "unknown load operator");
    return (T)0;
  }
}

// Store operators that read fields besides the pass output (the update reads
// psihat, psihat_old, N_old) split into a load phase and an apply phase, so a
// caller can issue the loads of many elements before any store: the compiler
// cannot move a load above an earlier store to a possibly aliasing array.
struct StIn {
  double ph, pho, nold;
};

template <int ST>
constexpr bool store_reads() {
  return ST == ST_UPDATE || ST == ST_UPDATE_STB || ST == ST_FILTER;
}

template <typename T, int ST>
__device__ __forceinline__ StIn store_inputs(const StoreOp& S, const Kv& K, long idx) {
  StIn in{0.0, 0.0, 0.0};
  if constexpr (ST == ST_UPDATE || ST == ST_UPDATE_STB) {
    in.ph = ((const double*)S.s[0])[idx];
    in.pho = ((const double*)S.d[0])[idx];
    in.nold = (double)((const T*)S.d[1])[idx];
  } else if constexpr (ST == ST_FILTER) {
    in.ph = K.xi * K.fy * K.fz;  // i-table f_x
  }
  return in;
}

// d0: psihat_old (double, in) -> psihat_new (out); d1: N_old (T, in) -> N (out);
// s0: psihat (double, in).  quasi.cpp:1949-1951, quasi-stb.cpp:1954-1956.
template <typename T, int ST>
__device__ __forceinline__ void store_apply(const StoreOp& S, const Phys& P, const Kv& K, long idx,
                                            T v, const StIn& in) {
  if constexpr (ST == ST_FILTER) {
    // quasi.cpp:1260-1261 with exp(-filter K^2) = fx[i] fy[j] fz[k] (in.ph)
    double raw = (double)v;
    if (S.dc && idx == 0) raw = 8.0 * S.dc[0];
    const double r = P.scale * raw * in.ph;
    ((T*)S.d[0])[idx] = (T)r;
    if (S.dc_out && idx == 0) S.dc_out[0] = r;
    return;
  }
  const double mq2 = sq(K.xi) + sq(K.qy) + sq(K.qz);  // i-table q_x
  const double alin = P.eps - P.alpha * sq(P.q02 - mq2);
  double nh = (double)v;
  if (S.dc && idx == 0) nh = 8.0 * S.dc[0];
  double nold = in.nold;
  if (S.first) nold = nh;
  else if (S.dc_old && idx == 0) nold = S.dc_old[0];
  double nw;
  if constexpr (ST == ST_UPDATE) {
    const double c1 = 2.0 - P.dt * P.rho_m * alin, c2 = 1.5 - P.dt * P.rho_m * alin;
    nw = (c1 * in.ph - 0.5 * in.pho + (P.dt / 2) * P.scale * (3.0 * nh - nold)) / c2;
  } else {
    const double c1 = 1.0 + (P.dt / 2) * alin, c2 = 1.0 - (P.dt / 2) * alin;
    nw = (c1 * in.ph + (P.dt / 2) * P.scale * (3.0 * nh - nold)) / c2;
  }
  ((double*)S.d[0])[idx] = nw;
  ((T*)S.d[1])[idx] = (T)nh;
  if (S.dc_out && idx == 0) S.dc_out[0] = nh;
}

// Load operators that read many fields or do FP64 projection arithmetic are
// evaluated in chunks of 8 elements (register pressure); the rest in one go.
template <int LD>
constexpr bool load_heavy() {
  // the x-pass projection operators: FP64 arithmetic on four fields in a
  // 512-thread block (128-VGPR budget); the real-space operators of the
  // z pass run in 256-thread blocks and stay fully unrolled
  return LD == LD_VEL0 || LD == LD_VEL1 || LD == LD_VEL2 || LD == LD_PRESSURE;
}

// Plain stores; the operators that read other fields go through
// store_inputs/store_apply.
template <typename T, int ST>
__device__ __forceinline__ void store_eval(const StoreOp& S, long idx, T v) {
  static_assert(ST == ST_PLAIN, "reading store operators use store_inputs/store_apply");
  ((T*)S.d[0])[idx] = v;
}

}  // namespace qsm
