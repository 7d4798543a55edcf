// qsm_fft.hpp -- device-side building blocks of the batched r2r passes.
//
// One pass transforms TK pencils of length N along one axis of a row-major
// [nx][ny][nz] field.  Each pencil lives in LDS as N reals.  The real
// DCT-II/III of length N is computed through an M = N/2 point complex FFT
// (Makhoul 1980, "A fast cosine transform in one and two dimensions"):
//
//   forward (REDFT10):  v[n] = x[2n], v[N-1-n] = x[2n+1];  z[m] = v[2m] + i v[2m+1]
//                       Z = FFT_M(z);  V_k = E_k + e^{-2 pi i k/N} O_k  with
//                       E_k = Z_k + conj Z_{M-k},  O_k = -i (Z_k - conj Z_{M-k})
//                       Y_k = Re(e^{-i pi k/2N} V_k),  Y_{N-k} = -Im(e^{-i pi k/2N} V_k)
//   inverse (REDFT01):  V_k = e^{i pi k/2N} (X_k - i X_{N-k})   (X_N = 0)
//                       Z_k = (V_k + V_{k+M}) + i e^{2 pi i k/N} (V_k - V_{k+M})
//                       z = IFFT_M(Z) (unnormalised);  y at the v positions above
//   DST-II / DST-III by the index identities
//                       RODFT10(x)[k] = REDFT10((-1)^n x)[N-1-k]
//                       RODFT01(X)[n] = (-1)^n REDFT01(rev X)[n]
//
// The interleaved complex view of the "v-ordered" real array is exactly z, so
// the Makhoul reorder costs nothing beyond an index map at load (forward) or
// store (inverse).  These are FFTW's unnormalised definitions: a forward and
// an inverse pass along one axis multiply by 2N.
//
// The M-point FFT is a mixed-radix Stockham autosort (radices 2,3,4,8,16)
// with E elements per thread and M/E threads per pencil; values move between
// stages through LDS, twiddles come from a quarter-turn table of
// e^{-2 pi i t/(4N)} built on the host in FP64 and staged in LDS.
#pragma once
#include <hip/hip_runtime.h>

namespace qsm {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// multiply by S*i (S = +1 or -1)
template <int S, typename T>
__device__ __forceinline__ cx<T> mul_si(cx<T> a) {
  return S > 0 ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x};
}

// cos/sin of 2*pi*m/16, m = 0..15 (exact decimal expansions of the roots)
__device__ constexpr double kC16[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
__device__ constexpr double kS16[16] = {
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

// a * e^{S 2 pi i m / 16}
template <int S, typename T>
__device__ __forceinline__ cx<T> rot16(cx<T> a, int m) {
  const T c = (T)kC16[m & 15], s = (T)(S * kS16[m & 15]);
  return {a.x * c - a.y * s, a.x * s + a.y * c};
}

// ---------------- small DFTs in registers: v[k] <- sum_r v[r] e^{S 2 pi i r k / R}

template <int S, typename T>
__device__ __forceinline__ void dft2(cx<T>& a, cx<T>& b) {
  cx<T> t = a;
  a = t + b;
  b = t - b;
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cx<T>& a0, cx<T>& a1, cx<T>& a2, cx<T>& a3) {
  cx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_si<S>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

template <int S, typename T>
__device__ __forceinline__ void dft3(cx<T>& a0, cx<T>& a1, cx<T>& a2) {
  const T h = (T)0.86602540378443864676;  // sqrt(3)/2
  cx<T> s = a1 + a2, d = a1 - a2;
  cx<T> m = {a0.x - (T)0.5 * s.x, a0.y - (T)0.5 * s.y};
  cx<T> r = mul_si<S>(cx<T>{h * d.x, h * d.y});
  a0 = a0 + s;
  a1 = m + r;
  a2 = m - r;
}

// R = 8 as 2 x 4: sub-DFT4 over even/odd (stride 2), twiddle, DFT2.
template <int S, typename T>
__device__ __forceinline__ void dft8(cx<T>* v) {
  cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<S>(e0, e1, e2, e3);
  dft4<S>(o0, o1, o2, o3);
  o1 = rot16<S>(o1, 2);
  o2 = mul_si<S>(o2);
  o3 = rot16<S>(o3, 6);
  v[0] = e0 + o0; v[4] = e0 - o0;
  v[1] = e1 + o1; v[5] = e1 - o1;
  v[2] = e2 + o2; v[6] = e2 - o2;
  v[3] = e3 + o3; v[7] = e3 - o3;
}

// R = 16 as 4 x 4 (decimation in time, stride 4 sub-DFTs).
template <int S, typename T>
__device__ __forceinline__ void dft16(cx<T>* v) {
  cx<T> y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    y[n2][0] = v[n2]; y[n2][1] = v[n2 + 4]; y[n2][2] = v[n2 + 8]; y[n2][3] = v[n2 + 12];
    dft4<S>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
  }
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) y[n2][k1] = rot16<S>(y[n2][k1], n2 * k1);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cx<T> a0 = y[0][k1], a1 = y[1][k1], a2 = y[2][k1], a3 = y[3][k1];
    dft4<S>(a0, a1, a2, a3);
    v[k1] = a0; v[k1 + 4] = a1; v[k1 + 8] = a2; v[k1 + 12] = a3;
  }
}

template <int R, int S, typename T>
__device__ __forceinline__ void dft(cx<T>* v) {
  if constexpr (R == 2) dft2<S>(v[0], v[1]);
  else if constexpr (R == 3) dft3<S>(v[0], v[1], v[2]);
  else if constexpr (R == 4) dft4<S>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 8) dft8<S>(v);
  else if constexpr (R == 16) dft16<S>(v);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------- twiddles ---------------------------------------------------
// Per length N (M = N/2) the host builds one table of complex values, all
// taken from the quarter turn e^{-2 pi i t/(4N)}, t < N, by exact symmetries
// (twq(u) = e^{-2 pi i u/(4N)} below):
//   [0, M)               stage twiddles    e^{-2 pi i v/M} = twq(8v)
//   [M, 3M)              DCT-III pre-processing, per m < M:
//                        conj twq(m) = e^{i pi m/2N},
//                        conj twq(5m+N) = i e^{2 pi i m/N} e^{i pi m/2N}
//   [3M, 3M + 3(M/2+1))  DCT-II post-processing, per k <= M/2:
//                        twq(4k), twq(k), twq(M-k)
// A pass stages the stage part and its direction's part in LDS; every
// twiddle is then one LDS read at a per-thread base plus a constant.
template <int M>
struct TwLayout {
  static constexpr int STAGE = 0, PRE = M, POST = 3 * M;
  static constexpr int NPRE = 2 * M, NPOST = 3 * (M / 2 + 1);
  static constexpr int TOTAL = 3 * M + NPOST;
};

// ---------------- Stockham stages ------------------------------------------
// The M-point FFT of TK pencils lives in LDS as buf[m * TKP + p] (complex
// slot m of pencil p): the lane-varying index in every access is the pencil
// p, so gathers and scatters of any stage hit consecutive addresses.
// A thread (p, t), t < TP = M/E, holds E values: at stage (radix R, stride NS)
// it owns butterflies j = t + b TP, b < E/R, whose inputs are slots
// j + r M/R and whose outputs go to slots (j/NS) NS R + j%NS + r NS.
template <int R, int... RS>
struct First {
  static constexpr int value = R;
};

template <int... Rs>
struct Radices {};

template <class RL>
struct Last;
template <int R>
struct Last<Radices<R>> {
  static constexpr int value = R;
};
template <int R, int R2, int... RS>
struct Last<Radices<R, R2, RS...>> {
  static constexpr int value = Last<Radices<R2, RS...>>::value;
};

template <typename T, int M, int E, int TKP, int R>
__device__ __forceinline__ void gather(cx<T>* v, const cx<T>* buf, int p, int t) {
  constexpr int TP = M / E, NB = E / R;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TP;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b * R + r] = buf[(j + r * (M / R)) * TKP + p];
  }
}

template <typename T, int M, int E, int TKP, int R, int NS>
__device__ __forceinline__ void scatter(const cx<T>* v, cx<T>* buf, int p, int t) {
  constexpr int TP = M / E, NB = E / R;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TP;
    const int d = (j / NS) * NS * R + (j % NS);
#pragma unroll
    for (int r = 0; r < R; ++r) buf[(d + r * NS) * TKP + p] = v[b * R + r];
  }
}

// twiddle + radix-R DFT on the thread's registers; S = -1 forward, +1 inverse
template <typename T, int N, int M, int E, int R, int NS, int S>
__device__ __forceinline__ void stage_compute(cx<T>* v, int t, const cx<T>* ts) {
  constexpr int TP = M / E, NB = E / R;
  static_assert(E % R == 0, "E must be a multiple of every radix");
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if constexpr (NS > 1) {
      const int jm = (t + b * TP) % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = ts[jm * r * (M / (NS * R))];  // e^{-2 pi i jm r/(NS R)}
        if (S > 0) w.y = -w.y;
        v[b * R + r] = cmul(v[b * R + r], w);
      }
    }
    dft<R, S>(v + b * R);
  }
}

// Runs the stages from the one whose inputs are already in v.  Between stages
// the values go through LDS.  LAST_TO_LDS: scatter the last stage's outputs to
// LDS (natural order); otherwise they stay in v, value b*R+r at slot
// t + b TP + r M/R.
template <typename T, int N, int M, int E, int TKP, int S, bool LAST_TO_LDS, int NS, int R,
          int... RS>
__device__ __forceinline__ void fft_from_regs(cx<T>* v, cx<T>* buf, int p, int t,
                                              const cx<T>* ts) {
  stage_compute<T, N, M, E, R, NS, S>(v, t, ts);
  if constexpr (sizeof...(RS) == 0) {
    if constexpr (LAST_TO_LDS) {
      __syncthreads();
      scatter<T, M, E, TKP, R, NS>(v, buf, p, t);
      __syncthreads();
    }
  } else {
    __syncthreads();  // every thread has read this stage's inputs
    scatter<T, M, E, TKP, R, NS>(v, buf, p, t);
    __syncthreads();
    gather<T, M, E, TKP, First<RS...>::value>(v, buf, p, t);
    fft_from_regs<T, N, M, E, TKP, S, LAST_TO_LDS, NS * R, RS...>(v, buf, p, t, ts);
  }
}

template <typename T, int N, int M, int E, int TKP, int S, bool LAST_TO_LDS, int... RS>
__device__ __forceinline__ void run_fft(Radices<RS...>, cx<T>* v, cx<T>* buf, int p, int t,
                                        const cx<T>* ts) {
  fft_from_regs<T, N, M, E, TKP, S, LAST_TO_LDS, 1, RS...>(v, buf, p, t, ts);
}

// Radix plan per half-length M: E elements per thread and the radix list.
template <int M>
struct Plan;
#define QSM_PLAN(M_, E_, ...)                        \
  template <>                                        \
  struct Plan<M_> {                                  \
    static constexpr int E = E_;                     \
    using R = Radices<__VA_ARGS__>;                  \
    static constexpr int R1 = First<__VA_ARGS__>::value; \
  };
QSM_PLAN(4, 4, 4)
QSM_PLAN(8, 8, 8)
QSM_PLAN(12, 12, 4, 3)
QSM_PLAN(16, 8, 8, 2)
QSM_PLAN(24, 12, 4, 3, 2)
QSM_PLAN(32, 8, 8, 4)
QSM_PLAN(48, 12, 4, 4, 3)
QSM_PLAN(64, 8, 8, 8)
QSM_PLAN(96, 12, 4, 4, 3, 2)
QSM_PLAN(128, 16, 16, 8)
QSM_PLAN(192, 12, 4, 4, 4, 3)
QSM_PLAN(256, 16, 16, 16)
QSM_PLAN(384, 12, 4, 4, 4, 3, 2)
QSM_PLAN(512, 16, 16, 16, 2)
#undef QSM_PLAN

// Plan of a pass in precision T.  FP64 at M = 256 (N = 512) takes 8
// elements per thread (radices 8, 8, 4): with 16 FP64 elements a strided
// tile is one 256-thread workgroup of 8 waves per CU, with 8 twice the waves
// (measured at 512^3: FP64 step 39.4 -> 38.1 ms; the FP32 step is faster
// with 16).
template <typename T, int M>
struct PlanFor : Plan<M> {};
template <>
struct PlanFor<double, 256> {
  static constexpr int E = 8;
  using R = Radices<8, 8, 4>;
  static constexpr int R1 = 8;
};

// Alternative radix plans A = 0, 1 per half-length M, for the autotuner to
// time against the default (tile variants 2 and 3 of the plain passes).
// Each trades stages against elements per thread: fewer, larger radices
// (more registers, fewer LDS round trips, fewer threads) or more, smaller
// ones.  Constraints: the radices multiply to M, E is a multiple of each,
// and M/E > 1 (at least two threads per pencil, so a strided tile of TK
// pencils fills a wavefront).  `ok` is false where M has no alternative.
template <typename T, int M, int A>
struct AltPlan {
  static constexpr bool ok = false;
};
#define QSM_ALT(T_, M_, A_, E_, ...)                      \
  template <>                                             \
  struct AltPlan<T_, M_, A_> {                            \
    static constexpr bool ok = true;                      \
    static constexpr int E = E_;                          \
    using R = Radices<__VA_ARGS__>;                       \
    static constexpr int R1 = First<__VA_ARGS__>::value;  \
  };
#define QSM_ALT2(M_, A_, E_, ...) QSM_ALT(float, M_, A_, E_, __VA_ARGS__) QSM_ALT(double, M_, A_, E_, __VA_ARGS__)
QSM_ALT2(16, 0, 4, 4, 4)
QSM_ALT2(16, 1, 8, 2, 8)
QSM_ALT2(24, 0, 6, 3, 2, 2, 2)
QSM_ALT2(24, 1, 12, 2, 3, 4)
QSM_ALT2(32, 0, 16, 16, 2)
QSM_ALT2(32, 1, 4, 4, 4, 2)
QSM_ALT2(48, 0, 6, 3, 2, 2, 2, 2)
QSM_ALT2(48, 1, 24, 4, 3, 4)
QSM_ALT2(64, 0, 16, 16, 4)
QSM_ALT2(64, 1, 4, 4, 4, 4)
QSM_ALT2(96, 0, 6, 3, 2, 2, 2, 2, 2)
QSM_ALT2(96, 1, 24, 4, 4, 3, 2)
QSM_ALT2(128, 0, 8, 8, 8, 2)
QSM_ALT2(128, 1, 4, 4, 4, 4, 2)
QSM_ALT2(192, 0, 24, 4, 4, 4, 3)
QSM_ALT2(192, 1, 12, 3, 4, 4, 4)
// M = 256: the other precision's default, and a reordered E = 16 plan
QSM_ALT(float, 256, 0, 8, 8, 8, 4)
QSM_ALT(double, 256, 0, 16, 16, 16)
QSM_ALT2(256, 1, 16, 4, 4, 16)
#undef QSM_ALT2
#undef QSM_ALT

// Tile variants of the plain passes: 0 default, 1 half-width tiles, 2 and 3
// the alternative radix plans, 4 and 5 the software-pipelined forms of 0
// and 2 (same tile and plan; qsm_pass_impl.hpp, k_pass_pipe).  plan_base
// gives the variant whose tile and plan a variant uses.
constexpr int plan_base(int V) { return V == 4 ? 0 : (V == 5 ? 2 : V); }

// Plan of tile variant V: the default plan, or for plan_base(V) = 2, 3 the
// alternatives 0 and 1 where they exist (else the default).
template <typename T, int M, int V,
          bool ALT = (plan_base(V) >= 2 && AltPlan<T, M, (plan_base(V) >= 2 ? plan_base(V) - 2 : 0)>::ok)>
struct PlanV : PlanFor<T, M> {};
template <typename T, int M, int V>
struct PlanV<T, M, V, true> : AltPlan<T, M, plan_base(V) - 2> {};

template <typename T, int M, int V>
constexpr bool plan_distinct() {
  return (V == 2 || V == 3) && AltPlan<T, M, (V >= 2 && V <= 3 ? V - 2 : 0)>::ok;
}

}  // namespace qsm
