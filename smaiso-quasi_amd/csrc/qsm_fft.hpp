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
// The M-point FFT is a mixed-radix Stockham autosort in LDS (radices 2,3,4,8,16),
// E elements per thread, M/E threads per pencil, with twiddles read from a
// table of e^{-2 pi i t/(4N)}, t < 4N, built on the host in FP64.
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

// ---------------- Stockham stages ------------------------------------------
// buf: this pencil's complex array (M entries).  t: thread index within the
// pencil, TP = M/E threads.  tw: e^{-2 pi i u/(4N)} table, TWM = 4N/M.
template <typename T, int M, int E, int R, int NS, int S, int TWM>
__device__ __forceinline__ void stockham_stage(cx<T>* buf, int t, const cx<T>* __restrict__ tw) {
  constexpr int TP = M / E;
  constexpr int NB = E / R;
  static_assert(E % R == 0, "E must be a multiple of every radix");
  cx<T> v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TP;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = buf[j + r * (M / R)];
    if constexpr (NS > 1) {
      const int jm = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int u = jm * r * (M / (NS * R));
        cx<T> w = tw[u * TWM];
        if (S > 0) w.y = -w.y;
        v[b][r] = cmul(v[b][r], w);
      }
    }
    dft<R, S>(v[b]);
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TP;
    const int d = (j / NS) * NS * R + (j % NS);
#pragma unroll
    for (int r = 0; r < R; ++r) buf[d + r * NS] = v[b][r];
  }
  __syncthreads();
}

template <typename T, int M, int E, int S, int TWM, int NS, int R, int... RS>
__device__ __forceinline__ void stockham_run(cx<T>* buf, int t, const cx<T>* __restrict__ tw) {
  stockham_stage<T, M, E, R, NS, S, TWM>(buf, t, tw);
  if constexpr (sizeof...(RS) > 0) stockham_run<T, M, E, S, TWM, NS * R, RS...>(buf, t, tw);
}

// Radix plan per half-length M: E elements per thread and the radix list.
template <int M>
struct Plan;
#define QSM_PLAN(M_, E_, ...)                                                            \
  template <>                                                                            \
  struct Plan<M_> {                                                                      \
    static constexpr int E = E_;                                                         \
    template <typename T, int S, int TWM>                                                \
    __device__ static __forceinline__ void run(cx<T>* b, int t, const cx<T>* tw) {       \
      stockham_run<T, M_, E_, S, TWM, 1, __VA_ARGS__>(b, t, tw);                         \
    }                                                                                    \
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

}  // namespace qsm
