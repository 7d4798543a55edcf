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

// Every source is an array of the field type T.  The spectral prologues
// read psihat in the field precision: the FP64 state itself, or in mixed
// precision its FP32 copy (rounded once by the update that writes it).
enum LoadCode : int {
  LD_PLAIN = 0,     // a0[idx]
  LD_ZERO,          // 0 (the reference transforms zeros: D before divv_switch)
  LD_COEF_F,        // alpha (q0^2 - K^2) psihat                          quasi.cpp:1042
  LD_COEF_G,        // -K^2 alpha (q0^2 - K^2) psihat                     quasi.cpp:1044-1045
  LD_GRAD_X,        // -s_c[m] src[m+1 along c], 0 at m = N_c-1, c = x   quasi.cpp:1124-1171
  LD_GRAD_Y,        //   c = y
  LD_GRAD_Z,        //   c = z
  LD_LAPGRAD_X,     // -K_c^2 (-s_c[m] psihat[m+1]), c = x               quasi.cpp:1176-1191
  LD_LAPGRAD_Y,     //   c = y
  LD_LAPGRAD_Z,     //   c = z
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
  // multi-output operators (three values per element, one per component c)
  LD_STRESS3,       // LD_STRESS for c = x, y, z; a: rho, G, F, g_x,y,z, l_x,y,z
  LD_STRESS_STB3,   // LD_STRESS_STB for c = x, y, z; a: sh, psi, g_x,y,z
  LD_STRESS3_DIVV,  // LD_STRESS3, then LD_DIVV as a fourth value; a: as LD_STRESS3, then
                    //   rho_old, v_x,y,z, r_x,y,z (rho is a0; in the multi-output pass
                    //   with lz = 2, a15 is rho's y-passed field, r_z formed from it)
  LD_NONLIN_PY,     // LD_NONLIN_P with p (a10) y-passed and |grad rho|^2 (a11) formed by
                    //   the stress pass: the pass completes p's inverse z transform and
                    //   keeps the pressure term
};

enum StoreCode : int {
  ST_PLAIN = 0,     // d0[idx] = v
  ST_FILTER,        // d0[idx] = scale v exp(-filter K^2)                 quasi.cpp:1260-1261
  ST_UPDATE,        // semi-implicit AB2/CN update of psihat              quasi.cpp:1949-1951
  ST_UPDATE_STB,    // CN update of psihat (stb)                          quasi-stb.cpp:1954-1956
  // The same stores, then the inverse x transform (DCT-III) of the stored
  // spectral values into d3: the field's first inverse pass, run on the tile
  // the forward x pass holds (rho's x-passed field after the density filter;
  // psi's after either update, from psihat_new, its FP32 copy in mixed
  // precision)
  ST_FILTER_X,
  ST_UPDATE_X,
  ST_UPDATE_STB_X,
};

// the store operator an X code extends, and whether a code has the x inverse
__host__ __device__ constexpr int st_base(int st) {
  return st == ST_FILTER_X ? ST_FILTER
         : st == ST_UPDATE_X ? ST_UPDATE
         : st == ST_UPDATE_STB_X ? ST_UPDATE_STB : st;
}
__host__ __device__ constexpr bool st_inverts(int st) { return st_base(st) != st; }

struct LoadOp {
  int code;
  int axis;      // component c (informative: STRESS)
  int flag;      // unused (reserved)
  int n0;        // VEL0-2: slab size of the P-rank halo compat map (0 = exact)
  const void* a[16];
};

struct StoreOp {
  int code;
  int first;           // UPDATE: first step (N_old = N)
  void* d[4];          // outputs (d3: the x-inverse output of ST_FILTER_X / ST_UPDATE_X)
  const void* s[2];    // extra inputs
  const double* dc;    // mixed precision: FP64 sum of the forward input (DC override), or null
  double* dc_out;      // mixed precision: FP64 (0,0,0) mode written here, or null
  const double* dc_old;// UPDATE: FP64 N_old(0,0,0) in mixed precision, or null
};

__device__ __forceinline__ double sq(double x) { return x * x; }

// Element index into a field.  Fields hold at most 2^29 elements (checked
// on the host), so the byte offset of any FP64 element fits in 32 bits and
// every access is a uniform base pointer plus a 32-bit unsigned offset (the
// scalar-base addressing form, one vector add per access).
using Ix = unsigned;

template <typename T>
__device__ __forceinline__ T ld(const void* base, Ix idx) {
  const Ix off = idx * (Ix)sizeof(T);
  return *(const T*)((const char*)base + off);
}
template <typename T>
__device__ __forceinline__ void st(void* base, Ix idx, T v) {
  const Ix off = idx * (Ix)sizeof(T);
  *(T*)((char*)base + off) = v;
}

// Split addressing: element thr + uni of a field, with thr a per-thread
// offset (>= 0) and uni an offset uniform across the wavefront (may be
// negative; thr + uni >= 0, or within a padded field's padding).  The
// uniform part is added to the base pointer with scalar instructions and the
// per-thread part is the 32-bit vector offset, so the accesses of a walk that
// share thr share one offset register.
template <typename T>
__device__ __forceinline__ T ld_at(const void* base, Ix thr, int uni) {
  const char* b = (const char*)base + (long)uni * (long)sizeof(T);
  return *(const T*)(b + thr * (Ix)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st_at(void* base, Ix thr, int uni, T v) {
  char* b = (char*)base + (long)uni * (long)sizeof(T);
  *(T*)(b + thr * (Ix)sizeof(T)) = v;
}

// non-temporal form of st_at (streaming store, low L2 retention)
template <typename T>
__device__ __forceinline__ void st_at_nt(void* base, Ix thr, int uni, T v) {
  char* b = (char*)base + (long)uni * (long)sizeof(T);
  __builtin_nontemporal_store(v, (T*)(b + thr * (Ix)sizeof(T)));
}

template <typename T>
__device__ __forceinline__ double rd(const void* p, Ix idx) {
  return (double)ld<T>(p, idx);
}

// p[idx + off] if ok, else 0, without a branch: the load always reads
// idx + off, inside the field or its padding (a step's fields carry one y-z
// slab of padding on either side, and |off| <= ny nz), at a uniform offset,
// and the value is selected afterwards.
template <typename T>
__device__ __forceinline__ double rd_if(const void* p, Ix idx, int off, bool ok) {
  const double v = (double)ld_at<T>(p, idx, off);
  return ok ? v : 0.0;
}

// Wavenumbers at one element: q_c = pi m_c/L_c, s_c = pi (m_c+1)/L_c and the
// density filter factor f_c = exp(-filter q_c^2) per axis (quasi.cpp:397-431,
// 1261).  A pass fills the entries of its own axis per element (from a table
// it stages) and the others once per thread or per row.
struct Kv {
  double qx, sx, fx, qy, sy, fy, qz, sz, fz;
};

template <int LD, int STX>
constexpr bool uses_kv() {
  constexpr int ST = st_base(STX);
  return LD == LD_COEF_F || LD == LD_COEF_G || LD == LD_SH || LD == LD_PRESSURE ||
         LD == LD_VEL0 || LD == LD_VEL1 || LD == LD_VEL2 || (LD >= LD_GRAD_X && LD <= LD_LAPGRAD_Z) ||
         ST == ST_UPDATE || ST == ST_UPDATE_STB || ST == ST_FILTER;
}

// Which wavenumbers of axis c an operator reads: q_c, s_c, f_c.
template <int LD, int STX>
constexpr bool uses_q(int c) {
  constexpr int ST = st_base(STX);
  return LD == LD_COEF_F || LD == LD_COEF_G || LD == LD_SH || LD == LD_PRESSURE ||
         ST == ST_UPDATE || ST == ST_UPDATE_STB ||
         (LD == LD_VEL0 && c != 0) || (LD == LD_VEL1 && c != 1) || (LD == LD_VEL2 && c != 2) ||
         (LD >= LD_LAPGRAD_X && LD <= LD_LAPGRAD_Z && c != LD - LD_LAPGRAD_X);
}
template <int LD, int ST>
constexpr bool uses_s(int c) {
  return (LD == LD_VEL0 && c == 0) || (LD == LD_VEL1 && c == 1) || (LD == LD_VEL2 && c == 2) ||
         (LD >= LD_GRAD_X && LD <= LD_GRAD_Z && c == LD - LD_GRAD_X) ||
         (LD >= LD_LAPGRAD_X && LD <= LD_LAPGRAD_Z && c == LD - LD_LAPGRAD_X);
}
template <int LD, int ST>
constexpr bool uses_f(int) {
  return st_base(ST) == ST_FILTER;
}
// pass-axis tables staged for an operator (bit 0: q, 1: s, 2: f)
template <int AX, int LD, int ST>
constexpr int kv_tables() {
  return (uses_q<LD, ST>(AX) ? 1 : 0) | (uses_s<LD, ST>(AX) ? 2 : 0) | (uses_f<LD, ST>(AX) ? 4 : 0);
}
constexpr int kv_count(int mask) { return (mask & 1) + ((mask >> 1) & 1) + ((mask >> 2) & 1); }

// Shifted read along axis C used by the gradient stencils: -s_C[m] src[m+1],
// 0 at m = N_C - 1.
template <typename S, int C>
__device__ __forceinline__ double grad_hat(const LoadOp& L, const Geo& G, const Kv& K, int i,
                                           int j, int k, Ix idx) {
  const int m = C == 0 ? i : (C == 1 ? j : k);
  const int nc = C == 0 ? G.nx : (C == 1 ? G.ny : G.nz);
  const int o = C == 0 ? G.ny * G.nz : (C == 1 ? G.nz : 1);
  const double st = C == 0 ? K.sx : (C == 1 ? K.sy : K.sz);
  return -st * rd_if<S>(L.a[0], idx, o, m < nc - 1);
}

// The projection prologues (pressure and the three velocity components) read
// shifted neighbours of three or four fields and do FP64 arithmetic with one
// division.  They run in two phases: fetch issues every load of an element
// (raw values at the shifted positions, inside the field or its padding: the
// shifts stay within one y-z slab), finish selects and evaluates.  A pass
// fetches a chunk of elements before finishing any, so the chunk's loads are
// in flight together.  The offsets are uniform (the P-rank halo slab of
// v_y and v_z aside), so a load costs no vector address arithmetic.
template <int LD>
constexpr bool load_split() {
  return LD == LD_PRESSURE || LD == LD_VEL0 || LD == LD_VEL1 || LD == LD_VEL2;
}

template <typename T>
struct Raw4 {
  T a, b, c, d;
};

// branch-free logical and (both operands are evaluated)
__device__ __forceinline__ bool band(bool a, bool b) { return (int)a & (int)b; }

// flags and offsets of the shifted reads, shared by fetch and finish
struct Shifts {
  bool fa, fb, fc, fd;   // read a/b/c/d is live (else its value is 0)
  int oa, ob, oc, od;    // index offsets of the reads
};

template <int LD>
__device__ __forceinline__ Shifts shifts(const LoadOp& L, const Geo& G, int i, int j, int k) {
  const int nyz = G.ny * G.nz;
  const bool hasd = L.a[3] != nullptr;  // D-hat is null while the divergence term is off
  if constexpr (LD == LD_PRESSURE) {
    // a: S_x[i-1], b: S_y[j-1], c: S_z[k-1]
    return Shifts{i > 0, j > 0, k > 0, false, -nyz, -G.nz, -1, 0};
  } else if constexpr (LD == LD_VEL0) {
    // a: S_x, b: S_y[i+1][j-1], c: S_z[i+1][k-1], d: D[i+1]; zero at i = Nx-1
    const bool in = i < G.nx - 1;
    return Shifts{true, band(in, j > 0), band(in, k > 0), band(in, hasd), 0, nyz - G.nz, nyz - 1, nyz};
  } else {
    // zero at j = Ny-1 or k = Nz-1.  The S_x neighbour slab is i-1 here
    // (exact stencil); fetch reads the P-rank halo slab instead where that
    // map applies (halo_slab).
    const bool in = (j < G.ny - 1) & (k < G.nz - 1);
    if constexpr (LD == LD_VEL1)  // a: S_y, b: S_x[ih][j+1], c: S_z[j+1][k-1], d: D[j+1]
      return Shifts{true, band(in, i > 0), band(in, k > 0), band(in, hasd), 0, G.nz - nyz, G.nz - 1, G.nz};
    else  // a: S_z, b: S_x[ih][k+1], c: S_y[j-1][k+1], d: D[k+1]
      return Shifts{true, band(in, i > 0), band(in, j > 0), band(in, hasd), 0, 1 - nyz, 1 - G.nz, 1};
  }
}

// The P-rank halo map of v_y and v_z (quasi.cpp:1639-1650): at the first slab
// of a rank's slab range (i mod n0 = 0, i > 0) the S_x neighbour is slab
// i - n0 instead of i - 1.  i mod n0 without an integer division:
// floor((i + 1/2)/n0) in FP32 is exact for i < 2^22 (relative error 2^-23
// against a margin of 1/(2 n0)).
__device__ __forceinline__ bool halo_slab(const LoadOp& L, int i) {
  const int qn = (int)(((float)i + 0.5f) * (1.0f / (float)L.n0));
  return (i > 0) & (i == qn * L.n0);
}

template <typename T, int LD>
__device__ __forceinline__ Raw4<T> fetch(const LoadOp& L, const Geo& G, int i, int j, int k,
                                         Ix idx) {
  const Shifts h = shifts<LD>(L, G, i, j, k);
  // field of each read: PRESSURE S_x, S_y, S_z; VEL_c own component first
  const void* pa = LD == LD_VEL1 ? L.a[1] : (LD == LD_VEL2 ? L.a[2] : L.a[0]);
  const void* pb = (LD == LD_PRESSURE || LD == LD_VEL0) ? L.a[1] : L.a[0];
  const void* pc = LD == LD_VEL2 ? L.a[1] : L.a[2];
  const void* pd = L.a[3] ? L.a[3] : L.a[0];  // a valid address when D-hat is null
  Raw4<T> r;
  r.a = ld_at<T>(pa, idx, h.oa);
  r.b = ld_at<T>(pb, idx, h.ob);
  r.c = ld_at<T>(pc, idx, h.oc);
  if constexpr (LD != LD_PRESSURE) r.d = ld_at<T>(pd, idx, h.od);
  else r.d = (T)0;
  if constexpr (LD == LD_VEL1 || LD == LD_VEL2) {
    if (L.n0 > 0) {  // uniform: the P-rank halo map
      if (halo_slab(L, i)) r.b = ld_at<T>(pb, idx, h.ob - (L.n0 - 1) * G.ny * G.nz);
    }
  }
  return r;
}

// 1/x in FP64 from the hardware reciprocal and one Newton step (x > 0)
__device__ __forceinline__ double rcp64(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

// Mixed precision (T = float): the projection reads FP32 fields and writes
// an FP32 field.  Its coefficients (1/K^2 with the scale, C_M1 = scale/(nu
// K^2), C_M2 = s_c/K^2, nu s_c) are formed in FP64 and rounded once; the
// products with the FP32 data are FP32.
template <int LD>
__device__ __forceinline__ float finish32(const Raw4<float>& r, const LoadOp& L, const Phys& P,
                                          const Kv& K, const Geo& G, int i, int j, int k) {
  const Shifts h = shifts<LD>(L, G, i, j, k);
  const float a = h.fa ? r.a : 0.0f, b = h.fb ? r.b : 0.0f, c = h.fc ? r.c : 0.0f,
              dd = h.fd ? r.d : 0.0f;
  // these operators run on the x pass: the y, z terms are per thread
  if constexpr (LD == LD_PRESSURE) {
    const float cp = (float)(-P.scale * rcp64(sq(K.qx) + (sq(K.qy) + sq(K.qz))));
    const float pv = cp * ((float)K.qx * a + (float)K.qy * b + (float)K.qz * c);
    return (i + j + k == 0) ? 0.0f : pv;
  } else {
    // own component: wavenumber s_c; the others q
    constexpr int cc = LD == LD_VEL0 ? 0 : (LD == LD_VEL1 ? 1 : 2);
    const double kx = cc == 0 ? K.sx : K.qx, ky = cc == 1 ? K.sy : K.qy, kz = cc == 2 ? K.sz : K.qz;
    const double ks = cc == 0 ? kx : (cc == 1 ? ky : kz);
    const double rm2 = rcp64(sq(kx) + (sq(ky) + sq(kz)));
    const float cm1 = (float)((P.scale / P.nu) * rm2), cm2 = (float)(ks * rm2),
                cnu = (float)(P.nu * ks);
    const float fx = (float)kx, fy = (float)ky, fz = (float)kz;
    // read a is the own component; b, c the others in axis order
    const float dot = cc == 0 ? fx * a + fy * b + fz * c
                              : (cc == 1 ? fx * b + fy * a + fz * c : fx * b + fy * c + fz * a);
    const float vv = cm1 * (a - cm2 * dot + cnu * dd);
    const bool in = cc == 0 ? i < G.nx - 1 : ((j < G.ny - 1) & (k < G.nz - 1));
    return in ? vv : 0.0f;
  }
}

template <typename T, int LD>
__device__ __forceinline__ T finish(const Raw4<T>& r, const LoadOp& L, const Phys& P, const Kv& K,
                                    const Geo& G, int i, int j, int k) {
  if constexpr (sizeof(T) == sizeof(float)) return finish32<LD>(r, L, P, K, G, i, j, k);
  const Shifts h = shifts<LD>(L, G, i, j, k);
  const double a = h.fa ? (double)r.a : 0.0, b = h.fb ? (double)r.b : 0.0,
               c = h.fc ? (double)r.c : 0.0, dd = h.fd ? (double)r.d : 0.0;
  const double nu = P.nu;
  if constexpr (LD == LD_PRESSURE) {
    // quasi.cpp:1385-1415
    const double mq2 = sq(K.qx) + sq(K.qy) + sq(K.qz);
    const double pv = P.scale * (-(K.qx * a + K.qy * b + K.qz * c) / mq2);
    return (T)((i + j + k == 0) ? 0.0 : pv);
  } else if constexpr (LD == LD_VEL0) {
    // global-index form of quasi.cpp:1522-1634
    const bool in = i < G.nx - 1;
    const double rm2 = 1.0 / (sq(K.sx) + sq(K.qy) + sq(K.qz));
    const double cm1 = (P.scale / nu) * rm2, cm2 = K.sx * rm2;
    const double dot = K.sx * a + K.qy * b + K.qz * c;
    const double vx = cm1 * (a - cm2 * dot + nu * K.sx * dd);
    return (T)(in ? vx : 0.0);
  } else if constexpr (LD == LD_VEL1) {
    // quasi.cpp:1681-1872
    const bool in = (j < G.ny - 1) & (k < G.nz - 1);
    const double rm2 = 1.0 / (sq(K.qx) + sq(K.sy) + sq(K.qz));
    const double dot = K.qx * b + K.sy * a + K.qz * c;
    const double vv = ((P.scale / nu) * rm2) * (a - (K.sy * rm2) * dot + nu * K.sy * dd);
    return (T)(in ? vv : 0.0);
  } else {
    const bool in = (j < G.ny - 1) & (k < G.nz - 1);
    const double rm2 = 1.0 / (sq(K.qx) + sq(K.qy) + sq(K.sz));
    const double dot = K.qx * b + K.qy * c + K.sz * a;
    const double vv = ((P.scale / nu) * rm2) * (a - (K.sz * rm2) * dot + nu * K.sz * dd);
    return (T)(in ? vv : 0.0);
  }
}

// Real-space operators of the forward z passes as functions of their input
// values: in[f] is field a[f] at the element.  load_eval reads the inputs
// and applies the same function, so element-wise and vectorized loads give
// the same bits.
template <int LD>
constexpr int op_nin() {
  return LD == LD_PLAIN ? 1 : LD == LD_RHOSTAR ? 4 : LD == LD_DIVV ? 8 : LD == LD_STRESS ? 5
       : LD == LD_NONLIN ? 10 : LD == LD_NONLIN_P ? 14 : LD == LD_NONLIN_PY ? 11
       : LD == LD_STRESS_STB ? 3
       : LD == LD_NONLIN_STB ? 7 : LD == LD_NONLIN_STB0 ? 1 : 0;
}
template <typename T, int LD>
__device__ __forceinline__ T op_apply(const T* in, const Phys& P) {
  if constexpr (LD == LD_PLAIN) {
    return in[0];
  } else if constexpr (LD == LD_RHOSTAR) {
    const T ps = in[0], gx = in[1], gy = in[2], gz = in[3];
    return (T)P.kp * sqrt((T)P.q02 * (ps * ps) + gx * gx + gy * gy + gz * gz) + (T)P.rho0;
  } else if constexpr (LD == LD_DIVV) {
    // rho, rho_old, vx, vy, vz, rx, ry, rz
    const T r = in[0], ro = in[1];
    T t = (r - ro) / (T)P.dt + in[2] * in[5];
    t = t + in[3] * in[6];
    t = t + in[4] * in[7];
    return -t / ((T)0.5 * (r + ro));
  } else if constexpr (LD == LD_STRESS) {
    // rho, G, F, g_c, l_c
    const T r = in[0];
    const T gg = r * in[1];
    const T rf = r * in[2];
    return gg * in[3] - rf * in[4];
  } else if constexpr (LD == LD_NONLIN || LD == LD_NONLIN_P || LD == LD_NONLIN_PY) {
    // rho, psi, F, G, vx, vy, vz, gx, gy, gz, p, rx, ry, rz; LD_NONLIN_PY:
    // in[10] is the pressure term itself (pressure_term, formed by the pass)
    const T r = in[0], ps = in[1];
    const T ps2 = ps * ps, ps3 = ps2 * ps, ps5 = ps3 * ps2;
    T v = r * ((T)P.eps * ps - in[2] + (T)P.beta * ps3 - (T)P.gamma * ps5) - r * in[3];
    v = v - in[4] * in[7];
    v = v - in[5] * in[8];
    v = v - in[6] * in[9];
    if constexpr (LD == LD_NONLIN_P) {
      const T rx = in[11], ry = in[12], rz = in[13];
      v = v + in[10] * (T)P.kp * sqrt(rx * rx + ry * ry + rz * rz) / (T)P.rho_s2;
    } else if constexpr (LD == LD_NONLIN_PY) {
      v = v + in[10];
    }
    return v;
  } else if constexpr (LD == LD_STRESS_STB) {
    // sh, psi, g_c ; rho = 1
    const T ps = in[1];
    const T ps2 = ps * ps, ps3 = ps2 * ps, ps5 = ps3 * ps2;
    const T mu = in[0] - (T)P.beta * ps3 + (T)P.gamma * ps5;
    return mu * in[2];
  } else if constexpr (LD == LD_NONLIN_STB || LD == LD_NONLIN_STB0) {
    // psi, vx, vy, vz, gx, gy, gz
    const T ps = in[0];
    const T ps2 = ps * ps, ps3 = ps2 * ps, ps5 = ps3 * ps2;
    T v = (T)P.beta * ps3 - (T)P.gamma * ps5;
    if constexpr (LD == LD_NONLIN_STB) {
      v = v - in[1] * in[4];
      v = v - in[2] * in[5];
      v = v - in[3] * in[6];
    }
    return v;
  } else {
    static_assert(op_nin<LD>() > 0, "not a real-space operator");
    return (T)0;
  }
}

// K carries the element's wavenumbers for the x-pass spectral operators.
// The pressure term p kp |grad rho| / rho_s^2 of quasi.cpp:1925-1930, from
// |grad rho|^2 accumulated as rx^2, + ry^2, + rz^2 (the order of LD_NONLIN_P).
template <typename T>
__device__ __forceinline__ T pressure_term(T p, T r2, const Phys& P) {
  return p * (T)P.kp * sqrt(r2) / (T)P.rho_s2;
}

// held: a value the pass formed itself (LD_NONLIN_PY: the pressure term).
template <typename T, int LD>
__device__ __forceinline__ T load_eval(const LoadOp& L, const Phys& P, const Kv& K,
                                       const Geo& G, int i, int j, int k, Ix idx,
                                       T held = (T)0) {
  constexpr bool MIX = sizeof(T) == sizeof(float);  // FP32 fields: mixed precision
  if constexpr (LD == LD_PLAIN) {
    return ld<T>(L.a[0], idx);
  } else if constexpr (LD == LD_ZERO) {
    return (T)0;
  } else if constexpr (MIX && (LD == LD_COEF_F || LD == LD_COEF_G || LD == LD_SH)) {
    // mixed precision: the coefficient in FP64 (q0^2 - K^2 cancels near the
    // ring), rounded once, times the FP32 input
    const double mq2 = sq(K.qx) + (sq(K.qy) + sq(K.qz));
    const double d = P.q02 - mq2;
    const T x = ld<T>(L.a[0], idx);
    if constexpr (LD == LD_SH) return (T)(P.alpha * sq(d) - P.eps) * x;
    const T fh = (T)(P.alpha * d) * x;
    return LD == LD_COEF_F ? fh : (T)(-mq2) * fh;
  } else if constexpr (MIX && LD >= LD_GRAD_X && LD <= LD_LAPGRAD_Z) {
    // mixed precision: one FP64 factor per element, rounded once, times the
    // FP32 input (grad: -s_c; lap grad: K_c^2 s_c with K_c the wavenumbers
    // of the shifted component)
    constexpr int c = (LD - LD_GRAD_X) % 3;
    const double st = c == 0 ? K.sx : (c == 1 ? K.sy : K.sz);
    double f = -st;
    if constexpr (LD >= LD_LAPGRAD_X) {
      const double ax = c == 0 ? K.sx : K.qx, ay = c == 1 ? K.sy : K.qy, az = c == 2 ? K.sz : K.qz;
      // these operators run on the x pass: the y, z terms are per thread
      const double m = sq(ax) + (sq(ay) + sq(az));
      f = m * st;
    }
    const int m = c == 0 ? i : (c == 1 ? j : k);
    const int nc = c == 0 ? G.nx : (c == 1 ? G.ny : G.nz);
    const int o = c == 0 ? G.ny * G.nz : (c == 1 ? G.nz : 1);
    const bool ok = m < nc - 1;
    const T x = ld_at<T>(L.a[0], idx, o);  // inside the field or its padding
    return ok ? (T)f * x : (T)0;
  } else if constexpr (LD == LD_COEF_F || LD == LD_COEF_G) {
    const double mq2 = sq(K.qx) + sq(K.qy) + sq(K.qz);
    const double fh = P.alpha * (P.q02 - mq2) * (double)ld<T>(L.a[0], idx);
    return (T)(LD == LD_COEF_F ? fh : -mq2 * fh);
  } else if constexpr (LD >= LD_GRAD_X && LD <= LD_GRAD_Z) {
    return (T)grad_hat<T, LD - LD_GRAD_X>(L, G, K, i, j, k, idx);
  } else if constexpr (LD >= LD_LAPGRAD_X && LD <= LD_LAPGRAD_Z) {
    constexpr int c = LD - LD_LAPGRAD_X;
    const double ax = c == 0 ? K.sx : K.qx;
    const double ay = c == 1 ? K.sy : K.qy;
    const double az = c == 2 ? K.sz : K.qz;
    const double m = sq(ax) + sq(ay) + sq(az);
    return (T)(-m * grad_hat<T, c>(L, G, K, i, j, k, idx));
  } else if constexpr (LD == LD_SH) {
    const double mq2 = sq(K.qx) + sq(K.qy) + sq(K.qz);
    const double alin = P.eps - P.alpha * sq(P.q02 - mq2);
    return (T)(-alin * (double)ld<T>(L.a[0], idx));
  } else if constexpr (LD == LD_RHOSTAR || LD == LD_DIVV || LD == LD_STRESS) {
    T in[op_nin<LD>()];
#pragma unroll
    for (int f = 0; f < op_nin<LD>(); ++f) in[f] = ld<T>(L.a[f], idx);
    return op_apply<T, LD>(in, P);
  } else if constexpr (LD == LD_PRESSURE || LD == LD_VEL0 || LD == LD_VEL1 || LD == LD_VEL2) {
    return finish<T, LD>(fetch<T, LD>(L, G, i, j, k, idx), L, P, K, G, i, j, k);
  } else if constexpr (LD == LD_SCALE) {
    return (T)(P.scale * rd<T>(L.a[0], idx));
  } else if constexpr (LD == LD_NONLIN || LD == LD_NONLIN_P || LD == LD_STRESS_STB ||
                       LD == LD_NONLIN_STB || LD == LD_NONLIN_STB0) {
    T in[op_nin<LD>()];
#pragma unroll
    for (int f = 0; f < op_nin<LD>(); ++f) in[f] = ld<T>(L.a[f], idx);
    return op_apply<T, LD>(in, P);
  } else if constexpr (LD == LD_NONLIN_PY) {
    T in[11];
#pragma unroll
    for (int f = 0; f < 10; ++f) in[f] = ld<T>(L.a[f], idx);
    in[10] = held;
    return op_apply<T, LD>(in, P);
  } else {
    static_assert(LD == LD_PLAIN, "unknown load operator");
    return (T)0;
  }
}

// Multi-output operators: the three components of the stress at one
// element, reading the shared inputs once.  Each value is the expression of
// the single-output operator (LD_STRESS, LD_STRESS_STB) term for term.
template <int LD>
constexpr int multi_count() {
  return LD == LD_STRESS3_DIVV ? 4 : ((LD == LD_STRESS3 || LD == LD_STRESS_STB3) ? 3 : 1);
}
// LZ >= 1: the lap grad psi components come from lz[0..2] (formed by the
// tile's own inverse z round) instead of a[6..8]; LZ = 2: grad rho as well,
// from lz[3..5] instead of a[13..15].
template <typename T, int LD, int LZ = 0>
__device__ __forceinline__ void multi_eval(const LoadOp& L, const Phys& P, Ix idx,
                                           T (&out)[multi_count<LD>()], const T* lz = nullptr) {
  if constexpr (LD == LD_STRESS3 || LD == LD_STRESS3_DIVV) {
    T in[5] = {ld<T>(L.a[0], idx), ld<T>(L.a[1], idx), ld<T>(L.a[2], idx), (T)0, (T)0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      in[3] = ld<T>(L.a[3 + c], idx);
      if constexpr (LZ >= 1) in[4] = lz[c];
      else in[4] = ld<T>(L.a[6 + c], idx);
      out[c] = op_apply<T, LD_STRESS>(in, P);
    }
    if constexpr (LD == LD_STRESS3_DIVV) {
      T d[8] = {in[0]};
#pragma unroll
      for (int f = 1; f < 8; ++f) {
        if constexpr (LZ == 2) d[f] = f >= 5 ? lz[f - 2] : ld<T>(L.a[8 + f], idx);
        else d[f] = ld<T>(L.a[8 + f], idx);
      }
      out[3] = op_apply<T, LD_DIVV>(d, P);
    }
  } else {
    static_assert(LD == LD_STRESS_STB3, "not a multi-output operator");
    T in[3] = {ld<T>(L.a[0], idx), ld<T>(L.a[1], idx), (T)0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      in[2] = ld<T>(L.a[2 + c], idx);
      out[c] = op_apply<T, LD_STRESS_STB>(in, P);
    }
  }
}

// Store operators that read fields besides the pass output (the update reads
// psihat, psihat_old, N_old) split into a load phase and an apply phase, so a
// caller can issue the loads of many elements before any store: the compiler
// cannot move a load above an earlier store to a possibly aliasing array.
struct StIn {
  double ph, pho, nold;
};

template <int STX>
constexpr bool store_reads() {
  constexpr int ST = st_base(STX);
  return ST == ST_UPDATE || ST == ST_UPDATE_STB || ST == ST_FILTER;
}

template <typename T, int STX>
__device__ __forceinline__ StIn store_inputs(const StoreOp& S, const Kv& K, Ix idx) {
  constexpr int ST = st_base(STX);
  StIn in{0.0, 0.0, 0.0};
  if constexpr (ST == ST_UPDATE || ST == ST_UPDATE_STB) {
    in.ph = ld<double>(S.s[0], idx);
    in.pho = ld<double>(S.d[0], idx);
    in.nold = (double)ld<T>(S.d[1], idx);
  } else if constexpr (ST == ST_FILTER) {
    in.ph = K.fx * K.fy * K.fz;
  }
  return in;
}

// Mixed precision: the FP64 (0,0,0) values a store operator substitutes at
// index 0, read once per thread (uniform loads) instead of per element.
struct DcIn {
  double dc, old;
  bool has_dc, has_old;
};
__device__ __forceinline__ DcIn dc_in(const StoreOp& S) {
  return DcIn{S.dc ? S.dc[0] : 0.0, S.dc_old ? S.dc_old[0] : 0.0, S.dc != nullptr,
              S.dc_old != nullptr};
}

// d0: psihat_old (double, in) -> psihat_new (out); d1: N_old (T, in) -> N (out);
// s0: psihat (double, in); mixed precision (T = float): d2 receives psihat_new
// rounded to FP32, the copy the next step's spectral prologues read.
// quasi.cpp:1949-1951, quasi-stb.cpp:1954-1956.
// Returns the stored spectral value in the field precision (the filtered
// density; psihat_new, rounded to FP32 in mixed precision), the input of the
// x inverse of the X codes.
template <typename T, int STX>
__device__ __forceinline__ T store_apply(const StoreOp& S, const Phys& P, const Kv& K, Ix idx,
                                         T v, const StIn& in, const DcIn& D) {
  constexpr int ST = st_base(STX);
  const bool z = idx == 0;
  if constexpr (ST == ST_FILTER) {
    // quasi.cpp:1260-1261 with exp(-filter K^2) = fx[i] fy[j] fz[k] (in.ph)
    const double raw = (D.has_dc & z) ? 8.0 * D.dc : (double)v;
    const double r = P.scale * raw * in.ph;
    st<T>(S.d[0], idx, (T)r);
    if (z && S.dc_out) S.dc_out[0] = r;
    return (T)r;
  }
  const double mq2 = sq(K.qx) + sq(K.qy) + sq(K.qz);
  const double alin = P.eps - P.alpha * sq(P.q02 - mq2);
  const double nh = (D.has_dc & z) ? 8.0 * D.dc : (double)v;
  const double nold = S.first ? nh : ((D.has_old & z) ? D.old : in.nold);
  double nw;
  if constexpr (ST == ST_UPDATE) {
    const double c1 = 2.0 - P.dt * P.rho_m * alin, c2 = 1.5 - P.dt * P.rho_m * alin;
    nw = (c1 * in.ph - 0.5 * in.pho + (P.dt / 2) * P.scale * (3.0 * nh - nold)) / c2;
  } else {
    const double c1 = 1.0 + (P.dt / 2) * alin, c2 = 1.0 - (P.dt / 2) * alin;
    nw = (c1 * in.ph + (P.dt / 2) * P.scale * (3.0 * nh - nold)) / c2;
  }
  st<double>(S.d[0], idx, nw);
  st<T>(S.d[1], idx, (T)nh);
  if constexpr (sizeof(T) < sizeof(double)) st<T>(S.d[2], idx, (T)nw);
  if (z && S.dc_out) S.dc_out[0] = nh;
  return (T)nw;
}

// Load operators whose forward transform keeps its (0,0,0) mode in FP64 in
// mixed precision (the density law and the nonlinear terms): their forward
// z pass accumulates an FP64 sum of the loaded values per workgroup.
__host__ __device__ constexpr bool dc_capable(int ld) {
  return ld == LD_RHOSTAR || ld == LD_NONLIN || ld == LD_NONLIN_P || ld == LD_NONLIN_PY ||
         ld == LD_NONLIN_STB || ld == LD_NONLIN_STB0;
}

// Forward z passes that first complete the inverse z transforms of some of
// their inputs (y-passed), holding the result in registers: they need the
// DCT-III pre-processing twiddles as well.
__host__ __device__ constexpr bool pre_rounds(int ld) { return ld == LD_NONLIN_PY; }

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
__device__ __forceinline__ void store_eval(const StoreOp& S, Ix idx, T v) {
  static_assert(ST == ST_PLAIN, "reading store operators use store_inputs/store_apply");
  st<T>(S.d[0], idx, v);
}

}  // namespace qsm
