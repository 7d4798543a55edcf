// qsm_pass.hip -- the batched r2r pass kernel (one axis of a 3D r2r transform)
// and the FP64 compensated reduction kernels.
//
// A pass reads TK whole pencils of one axis into LDS, transforms them there
// (Makhoul reorder + N/2-point Stockham FFT, see qsm_fft.hpp) and writes them
// back, so each pass moves exactly one read and one write of the field
// through HBM, plus whatever extra operands the fused load/store operator
// touches (qsm_ops.hpp).
//
// Tile shapes:
//   z pass (contiguous): TK consecutive rows of nz elements; a wavefront reads
//                        64 consecutive elements of one row.
//   y / x pass (stride nz / ny*nz): TK consecutive z-columns; a wavefront
//                        reads TK contiguous elements (128 B) of each of
//                        64/TK axis positions, so every HBM request is a full
//                        128-byte segment.
// Passes are in place when src == dst: every workgroup reads exactly the
// elements it writes.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "qsm_fft.hpp"
#include "qsm_kernels.h"
#include "qsm_ops.hpp"

namespace qsm {

constexpr int kAxX = 0, kAxY = 1, kAxZ = 2;

template <typename T, int N, int AX>
struct TileCfg {
  static constexpr int M = N / 2;
  static constexpr int E = Plan<M>::E;
  static constexpr int TP = M / E;  // threads per pencil
  static constexpr int LDP = N + 2; // LDS pencil stride (even: complex aligned)
  static constexpr int tk() {
    if (AX == kAxZ) {
      int t = 256 / TP;
      if (t < 1) t = 1;
      while (t > 1 && (long)t * LDP * (long)sizeof(T) > 66000) t /= 2;
      return t;
    }
    int t = 128 / (int)sizeof(T);
    while (t > 1 && (long)t * LDP * (long)sizeof(T) > 70000) t /= 2;
    return t;
  }
  static constexpr int TK = tk();
  static constexpr int BLOCK = TK * TP;
};

// position of sample n in the Makhoul-reordered array v
template <int N>
__device__ __forceinline__ int vpos(int n) {
  return (n & 1) ? (N - 1 - (n >> 1)) : (n >> 1);
}

template <typename T, int N, int AX, bool INV>
__global__ void __launch_bounds__((TileCfg<T, N, AX>::BLOCK))
    k_pass(const PassArgs A) {
  using C = TileCfg<T, N, AX>;
  constexpr int M = C::M, E = C::E, TP = C::TP, TK = C::TK, BLOCK = C::BLOCK, LDP = C::LDP;
  __shared__ T lds[TK * LDP];
  const cx<T>* __restrict__ tw = (const cx<T>*)A.tw;
  const Geo G = A.geo;
  const int tid = threadIdx.x;
  const long nyz = (long)G.ny * G.nz;
  const bool dst = A.dst != 0;

  // ---- tile decode
  long row0 = 0;  // z pass: first row of the tile
  int t_i = 0, t_j = 0, k0 = 0;
  if constexpr (AX == kAxZ) {
    row0 = (long)blockIdx.x * TK;
  } else {
    const int tiles_k = (G.nz + TK - 1) / TK;
    const int a = blockIdx.x / tiles_k;
    k0 = (blockIdx.x % tiles_k) * TK;
    if constexpr (AX == kAxY) t_i = a; else t_j = a;
  }
  const long nrows = (long)G.nx * G.ny;

  double dsum = 0.0;  // FP64 sum of loaded values (mixed-precision DC mode)

  // ---- load
  constexpr int PER = TK * N / BLOCK;  // = 2E
#pragma unroll 4
  for (int s = 0; s < PER; ++s) {
    const int e = tid + s * BLOCK;
    int p, n, i, j, k;
    long idx;
    bool ok;
    if constexpr (AX == kAxZ) {
      p = e / N; n = e % N;
      const long row = row0 + p;
      ok = row < nrows;
      i = (int)(row / G.ny); j = (int)(row % G.ny); k = n;
      idx = row * G.nz + n;
    } else {
      p = e % TK; n = e / TK;
      k = k0 + p;
      ok = k < G.nz;
      if constexpr (AX == kAxY) { i = t_i; j = n; } else { i = n; j = t_j; }
      idx = (long)i * nyz + (long)j * G.nz + k;
    }
    T v = (T)0;
    if (ok) {
      v = load_eval<T>(A.ld, A.phys, A.waves, G, i, j, k, idx);
      if (A.part) dsum += (double)v;
    }
    int pos;
    if constexpr (!INV) {
      if (dst && (n & 1)) v = -v;
      pos = vpos<N>(n);
    } else {
      pos = dst ? (N - 1 - n) : n;
    }
    lds[p * LDP + pos] = v;
  }
  __syncthreads();

  const int p = tid / TP, t = tid % TP;
  cx<T>* zb = (cx<T>*)(lds + p * LDP);
  T* xb = lds + p * LDP;
  constexpr int NPRE = M / 2 / TP;  // pair items per thread (k < M/2)
  static_assert(M / 2 >= TP || M < 4, "pair work must cover all threads");

  if constexpr (INV) {
    // ---- DCT-III pre-processing: X -> Z (two Z per pair item)
    cx<T> za[NPRE + 1], zc[NPRE + 1];
    const cx<T> r8 = {(T)0.70710678118654752440, (T)0.70710678118654752440};  // e^{i pi/4}
#pragma unroll
    for (int b = 0; b <= NPRE; ++b) {
      int kk = t + b * TP;
      if (b == NPRE) kk = (t == 0) ? M / 2 : -1;
      if (kk < 0) continue;
      const T Xk = xb[kk], XNk = kk ? xb[N - kk] : (T)0, XMk = xb[M + kk], XMm = xb[M - kk];
      const cx<T> w1 = conj(tw[kk]);          // e^{+i pi k/2N}
      const cx<T> vk = cmul(w1, cx<T>{Xk, -XNk});
      const cx<T> vm = cmul(cmul(w1, r8), cx<T>{XMk, -XMm});
      const cx<T> Ev = vk + vm;
      const cx<T> Ov = cmul(conj(tw[4 * kk]), vk - vm);  // e^{+2 pi i k/N}
      za[b] = Ev + mul_si<1>(Ov);
      zc[b] = conj(Ev) + mul_si<1>(conj(Ov));
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b <= NPRE; ++b) {
      int kk = t + b * TP;
      if (b == NPRE) kk = (t == 0) ? M / 2 : -1;
      if (kk < 0) continue;
      zb[kk] = za[b];
      if (kk > 0 && kk < M / 2) zb[M - kk] = zc[b];
    }
    __syncthreads();
    Plan<M>::template run<T, 1, 8>(zb, t, tw);
  } else {
    Plan<M>::template run<T, -1, 8>(zb, t, tw);
    // ---- DCT-II post-processing: Z -> Y (four outputs per pair item)
    T y0[NPRE + 1], y1[NPRE + 1], y2[NPRE + 1], y3[NPRE + 1];
#pragma unroll
    for (int b = 0; b <= NPRE; ++b) {
      int kk = t + b * TP;
      if (b == NPRE) kk = (t == 0) ? M / 2 : -1;
      if (kk < 0) continue;
      const cx<T> a = zb[kk], c = conj(zb[kk ? M - kk : 0]);
      const cx<T> Ev = a + c;
      const cx<T> Ov = mul_si<-1>(a - c);
      const cx<T> wO = cmul(tw[4 * kk], Ov);        // e^{-2 pi i k/N} O
      const cx<T> Wk = cmul(tw[kk], Ev + wO);        // e^{-i pi k/2N} V_k
      const cx<T> Wm = cmul(tw[M - kk], conj(Ev - wO));
      y0[b] = Wk.x; y1[b] = -Wk.y; y2[b] = Wm.x; y3[b] = -Wm.y;
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b <= NPRE; ++b) {
      int kk = t + b * TP;
      if (b == NPRE) kk = (t == 0) ? M / 2 : -1;
      if (kk < 0) continue;
      // outputs Y_k, Y_{N-k}, Y_{M-k}, Y_{M+k}; DST-II reverses the index
      const int o0 = kk, o1 = N - kk, o2 = M - kk, o3 = M + kk;
#define QSM_PUT(o, val) xb[dst ? (N - 1 - (o)) : (o)] = (val)
      QSM_PUT(o0, y0[b]);
      if (kk > 0) QSM_PUT(o1, y1[b]);
      if (kk < M / 2) {
        QSM_PUT(o2, y2[b]);
        if (kk > 0) QSM_PUT(o3, y3[b]);
      }
#undef QSM_PUT
    }
  }
  __syncthreads();

  // ---- store
#pragma unroll 4
  for (int s = 0; s < PER; ++s) {
    const int e = tid + s * BLOCK;
    int q, n, i, j, k;
    long idx;
    bool ok;
    if constexpr (AX == kAxZ) {
      q = e / N; n = e % N;
      const long row = row0 + q;
      ok = row < nrows;
      i = (int)(row / G.ny); j = (int)(row % G.ny); k = n;
      idx = row * G.nz + n;
    } else {
      q = e % TK; n = e / TK;
      k = k0 + q;
      ok = k < G.nz;
      if constexpr (AX == kAxY) { i = t_i; j = n; } else { i = n; j = t_j; }
      idx = (long)i * nyz + (long)j * G.nz + k;
    }
    if (!ok) continue;
    T v;
    if constexpr (INV) {
      v = lds[q * LDP + vpos<N>(n)];
      if (dst && (n & 1)) v = -v;
    } else {
      v = lds[q * LDP + n];
    }
    store_eval<T>(A.st, A.phys, A.waves, i, j, k, idx, v);
  }

  // ---- FP64 block sum of the loaded values (deterministic order)
  if (A.part) {
    __syncthreads();
    double* red = (double*)lds;
    for (int off = 32; off > 0; off >>= 1) dsum += __shfl_down(dsum, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = dsum;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < (BLOCK + 63) / 64; ++w) s += red[w];
      A.part[blockIdx.x] = s;
    }
  }
}

// ----------------------------------------------------------------------------
// host-side launch

template <typename T, int N, int AX, bool INV>
static hipError_t launch_one(const PassArgs& A, long* nblocks_out, hipStream_t s) {
  using C = TileCfg<T, N, AX>;
  long nb;
  if (AX == kAxZ) nb = ((long)A.geo.nx * A.geo.ny + C::TK - 1) / C::TK;
  else {
    const long tiles_k = (A.geo.nz + C::TK - 1) / C::TK;
    nb = (AX == kAxY ? (long)A.geo.nx : (long)A.geo.ny) * tiles_k;
  }
  if (nblocks_out) *nblocks_out = nb;
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_pass<T, N, AX, INV>), dim3((unsigned)nb), dim3(C::BLOCK), 0, s, A);
  return hipGetLastError();
}

template <typename T, int N>
static hipError_t launch_n(int axis, bool inv, const PassArgs& A, long* nb, hipStream_t s) {
  if (axis == kAxX) return inv ? launch_one<T, N, kAxX, true>(A, nb, s) : launch_one<T, N, kAxX, false>(A, nb, s);
  if (axis == kAxY) return inv ? launch_one<T, N, kAxY, true>(A, nb, s) : launch_one<T, N, kAxY, false>(A, nb, s);
  return inv ? launch_one<T, N, kAxZ, true>(A, nb, s) : launch_one<T, N, kAxZ, false>(A, nb, s);
}

template <typename T>
static hipError_t launch_pass_t(int n, int axis, bool inv, const PassArgs& A, long* nb,
                                hipStream_t s) {
  switch (n) {
#define QSM_CASE(NN) case NN: return launch_n<T, NN>(axis, inv, A, nb, s);
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_pass(int prec_fp32, int n, int axis, bool inv, const PassArgs& A, long* nb,
                       hipStream_t s) {
  return prec_fp32 ? launch_pass_t<float>(n, axis, inv, A, nb, s)
                   : launch_pass_t<double>(n, axis, inv, A, nb, s);
}

bool pass_supported(int n) {
  switch (n) {
#define QSM_CASE(NN) case NN: return true;
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return false;
  }
}

long pass_blocks(int prec_fp32, int n, int axis, const Geo& g) {
  // number of workgroups of a pass (size of the partial-sum array it fills)
  PassArgs A{};
  A.geo = g;
  long nb = 0;
  // compute without launching: mirror launch_one's arithmetic
  int tk = pass_tile_k(prec_fp32, n, axis);
  if (tk <= 0) return -1;
  if (axis == kAxZ) nb = ((long)g.nx * g.ny + tk - 1) / tk;
  else nb = (axis == kAxY ? (long)g.nx : (long)g.ny) * ((g.nz + tk - 1) / tk);
  (void)A;
  return nb;
}

template <typename T>
static int tile_k_t(int n, int axis) {
  switch (n) {
#define QSM_CASE(NN)                                            \
  case NN:                                                      \
    return axis == kAxZ ? TileCfg<T, NN, kAxZ>::TK              \
                        : (axis == kAxY ? TileCfg<T, NN, kAxY>::TK : TileCfg<T, NN, kAxX>::TK);
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return -1;
  }
}

int pass_tile_k(int prec_fp32, int n, int axis) {
  return prec_fp32 ? tile_k_t<float>(n, axis) : tile_k_t<double>(n, axis);
}

// ----------------------------------------------------------------------------
// FP64 compensated reductions

// TwoSum-based compensated accumulation (Kahan-Babuska / Neumaier).
struct KSum {
  double s, c;
  __device__ void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ void merge(const KSum& o) {
    add(o.s);
    add(o.c);
  }
};

// Diagnostics of quasi.cpp:1976-2000 over one step's fields.  Per block:
// out[4*b + {0,1,2,3}] = sum|psi-psi_old|, sum|psi|, sum rho, sum energy density.
template <typename T>
__global__ void __launch_bounds__(256)
    k_diag(const T* __restrict__ psi, const T* __restrict__ pold, const T* __restrict__ rho,
           const T* __restrict__ F, long n, Phys P, double* __restrict__ part) {
  KSum a{0, 0}, b{0, 0}, m{0, 0}, e{0, 0};
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const double ps = psi[t], r = rho[t], f = F[t];
    a.add(fabs(ps - (double)pold[t]));
    b.add(fabs(ps));
    m.add(r);
    const double p2 = ps * ps;
    e.add(r * (-0.5 * P.eps * p2 + 0.5 * P.alpha * f * f - 0.25 * P.beta * p2 * p2 +
               (P.gamma / 6) * p2 * p2 * p2));
  }
  __shared__ KSum red[4][256];
  red[0][threadIdx.x] = a; red[1][threadIdx.x] = b; red[2][threadIdx.x] = m; red[3][threadIdx.x] = e;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x].merge(red[q][threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0].s + red[threadIdx.x][0].c;
}

// out[q] = compensated sum over b of part[stride*b + q], q < nq.  One block.
__global__ void __launch_bounds__(256)
    k_finish(const double* __restrict__ part, long nparts, int stride, int nq, double* __restrict__ out) {
  __shared__ KSum red[256];
  for (int q = 0; q < nq; ++q) {
    KSum s{0, 0};
    for (long b = threadIdx.x; b < nparts; b += blockDim.x) s.add(part[stride * b + q]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x].merge(red[threadIdx.x + off]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0].s + red[0].c;
    __syncthreads();
  }
}

hipError_t launch_diag(int prec_fp32, const void* psi, const void* pold, const void* rho,
                       const void* F, long n, const Phys& P, double* part, int nblocks,
                       double* out4, hipStream_t s) {
  if (prec_fp32)
    hipLaunchKernelGGL(k_diag<float>, dim3(nblocks), dim3(256), 0, s, (const float*)psi,
                       (const float*)pold, (const float*)rho, (const float*)F, n, P, part);
  else
    hipLaunchKernelGGL(k_diag<double>, dim3(nblocks), dim3(256), 0, s, (const double*)psi,
                       (const double*)pold, (const double*)rho, (const double*)F, n, P, part);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, part, (long)nblocks, 4, 4, out4);
  return hipGetLastError();
}

hipError_t launch_finish(const double* part, long nparts, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, part, nparts, 1, 1, out);
  return hipGetLastError();
}

}  // namespace qsm
