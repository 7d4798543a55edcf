// qsm_pass_impl.hpp -- the batched r2r pass kernel (one axis of a 3D r2r
// transform).  Included by qsm_pass_inst.hip, which instantiates it for one
// field precision and one axis per translation unit.
//
// A pass transforms TK whole pencils of one axis per workgroup and moves one
// read and one write of the field through HBM, plus whatever extra operands
// the fused load/store operator touches (qsm_ops.hpp).
//
// Tile shapes:
//   z pass (contiguous): TK consecutive rows of nz elements, loaded and
//                        stored through LDS so a wavefront reads 64
//                        consecutive elements of one row.
//   y / x pass (stride nz / ny*nz): TK consecutive z-columns; a wavefront
//                        reads TK contiguous elements (128 B) of each row it
//                        touches, and the strided passes load their first
//                        FFT stage and store their last straight from/to
//                        registers.
// Passes are in place when src == dst: every workgroup reads exactly the
// elements it writes.
#pragma once
#include <hip/hip_runtime.h>

#include "qsm_fft.hpp"
#include "qsm_kernels.h"
#include "qsm_ops.hpp"

namespace qsm {

constexpr int kAxX = 0, kAxY = 1, kAxZ = 2;




// Tile configuration of a pass.  V selects the tile-width variant: V = 0 the
// default width below, V = 1 half of it (64-byte row segments on strided axes,
// half the LDS per workgroup, twice the resident workgroups per CU).  The
// autotuner picks a variant per (axis, direction) for the plain passes.
template <typename T, int N, int AX, int V = 0>
struct TileCfg {
  static constexpr int M = N / 2;
  static constexpr int E = Plan<M>::E;
  static constexpr int TP = M / E;  // threads per pencil
  static constexpr int tk() {
    if (AX == kAxZ) {
      // contiguous rows: enough pencils for a 256-thread block, LDS <= ~72 KB
      int t = 256 / TP;
      if (t < 1) t = 1;
      while (t > 1 && (long)M * (t + 1) * 2 * (long)sizeof(T) > 72000) t /= 2;
      return t;
    }
    // strided axes: 128 contiguous bytes of each row per tile
    int t = 128 / (int)sizeof(T);
    while (t > 1 && (long)M * (t + 1) * 2 * (long)sizeof(T) > 72000) t /= 2;
    return t;
  }
  static constexpr int TK = (V == 1 && tk() > 1) ? tk() / 2 : tk();
  static constexpr int TKP = TK + 1;  // LDS slot stride: odd, so row-wise z-pass access spreads banks
  static constexpr int BLOCK = TK * TP;
  static_assert(BLOCK % TK == 0, "a thread's column is fixed across its loads");
};

// position of sample n in the Makhoul-reordered array v, and its inverse
template <int N>
__device__ __forceinline__ int vpos(int n) {
  return (n & 1) ? (N - 1 - (n >> 1)) : (n >> 1);
}
template <int N>
__device__ __forceinline__ int xpos(int q) {
  return q < N / 2 ? 2 * q : 2 * (N - 1 - q) + 1;
}

// coordinates of element n of pencil p of the tile.  Fields have fewer than
// 2^31 points (checked on the host), so element indices are 32-bit.
// at() clamps an out-of-range pencil to the tile's first one, so every
// address it returns is inside the field; the caller discards such values.
template <int AX>
struct TileGeo {
  int row0;
  int t_i, t_j, k0;
  int nrows;
  int nx, ny, nz;
  __device__ __forceinline__ bool at(int p, int n, int& i, int& j, int& k, long& idx) const {
    if constexpr (AX == kAxZ) {
      const int row = row0 + p;
      const bool ok = row < nrows;
      const int rr = ok ? row : row0;
      i = rr / ny; j = rr - i * ny; k = n;
      idx = (long)(rr * nz + n);
      return ok;
    } else {
      const bool ok = k0 + p < nz;
      k = ok ? k0 + p : k0;
      if constexpr (AX == kAxY) { i = t_i; j = n; } else { i = n; j = t_j; }
      idx = (long)((i * ny + j) * nz + k);
      return ok;
    }
  }
};

// One workgroup's work: tile `tile` of the pass described by A.
template <typename T, int N, int AX, bool INV, int LD, int ST, int V>
__device__ __forceinline__ void pass_tile(const PassArgs& A, const int tile) {
  using C = TileCfg<T, N, AX, V>;
  using PL = Plan<N / 2>;
  constexpr int M = C::M, E = C::E, TP = C::TP, TK = C::TK, TKP = C::TKP, BLOCK = C::BLOCK;
  constexpr int R1 = PL::R1, NB1 = E / R1;
  __shared__ cx<T> buf[M * TKP];
  __shared__ cx<T> tq[N + 1];
  T* rbuf = (T*)buf;  // real view: sample n of pencil p at ((n>>1)*TKP + p)*2 + (n&1)
  const Geo G = A.geo;
  const int tid = threadIdx.x;
  const bool dst = A.dst != 0;

  // The x-pass spectral operators read one i-indexed wavenumber table per
  // launch (itab); it is staged in LDS where two workgroups per CU still fit
  // (160 KiB), else read from global memory per element.
  constexpr bool KV = uses_kv<LD, ST>();
  static_assert(!KV || AX == kAxX, "wavenumber operators run in x passes");
  constexpr bool XT_LDS =
      KV && sizeof(cx<T>) * ((long)M * TKP + N + 1) + sizeof(double) * N <= 81920;
  __shared__ double xt[XT_LDS ? N : 1];
  const double* xsrc = nullptr;
  if constexpr (KV) xsrc = itab<LD, ST>(A.ld, A.waves);
  {
    const cx<T>* __restrict__ tg = (const cx<T>*)A.tw;
    for (int u = tid; u <= N; u += BLOCK) tq[u] = tg[u];
    if constexpr (XT_LDS) {
      for (int u = tid; u < N; u += BLOCK) xt[u] = xsrc[u];
      __syncthreads();
    }
  }

  TileGeo<AX> tg{};
  tg.nx = G.nx; tg.ny = G.ny; tg.nz = G.nz;
  tg.nrows = G.nx * G.ny;
  if constexpr (AX == kAxZ) {
    tg.row0 = tile * TK;
  } else {
    const int tiles_k = (G.nz + TK - 1) / TK;
    const int a = tile / tiles_k;
    tg.k0 = (tile % tiles_k) * TK;
    if constexpr (AX == kAxY) tg.t_i = a; else tg.t_j = a;
  }

  // FFT-phase thread coordinates: pencil p, thread t within the pencil
  const int p = tid % TK, t = tid / TK;
  // wavenumbers of this thread's (j, k); an x-pass thread keeps its column
  // p = tid % TK in every phase
  Kv kv{};
  if constexpr (KV) {
    int i0, j0, k0;
    long idx0;
    tg.at(p, 0, i0, j0, k0, idx0);
    kv = kv_thread(A.waves, j0, k0);
  }
  auto kv_at = [&](int i) {
    Kv K = kv;
    if constexpr (XT_LDS) K.xi = xt[i];
    else if constexpr (KV) K.xi = xsrc[i];
    return K;
  };
  double dsum = 0.0;  // FP64 sum of the loaded values (mixed-precision DC modes)
  cx<T> v[E];

  if constexpr (!INV) {
    {
      // row-sweeping loads into LDS in Makhoul order (a wavefront reads 64
      // consecutive elements of a z row, or TK-column segments of
      // consecutive rows of a strided axis); the loads of a chunk are all
      // issued before any of its values is used
      constexpr int PER = TK * N / BLOCK;
      constexpr int CH = load_heavy<LD>() ? (PER < 8 ? PER : 8) : PER;
#pragma unroll 1
      for (int s0 = 0; s0 < PER; s0 += CH) {
        T xs[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int e = tid + (s0 + c) * BLOCK;
          int q, n;
          if constexpr (AX == kAxZ) { q = e / N; n = e % N; } else { q = tid % TK; n = e / TK; }
          int i, j, k;
          long idx;
          tg.at(q, n, i, j, k, idx);
          xs[c] = load_eval<T, LD>(A.ld, A.phys, kv_at(i), G, i, j, k, idx);
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int e = tid + (s0 + c) * BLOCK;
          int q, n;
          if constexpr (AX == kAxZ) { q = e / N; n = e % N; } else { q = tid % TK; n = e / TK; }
          int i, j, k;
          long idx;
          T x = tg.at(q, n, i, j, k, idx) ? xs[c] : (T)0;
          if (A.part) dsum += (double)x;
          if (dst && (n & 1)) x = -x;
          const int vq = vpos<N>(n);
          rbuf[((vq >> 1) * TKP + q) * 2 + (vq & 1)] = x;
        }
      }
      __syncthreads();
      gather<T, M, E, TKP, R1>(v, buf, p, t);
    }
    run_fft<T, N, M, E, TKP, -1, true>(typename PL::R{}, v, buf, p, t, tq);

    // ---- DCT-II post-processing: Z -> Y, four outputs per pair item
    constexpr int NPRE = E / 2;
    T y0[NPRE + 1], y1[NPRE + 1], y2[NPRE + 1], y3[NPRE + 1];
#pragma unroll
    for (int b = 0; b <= NPRE; ++b) {
      int kk = t + b * TP;
      if (b == NPRE) kk = (t == 0) ? M / 2 : -1;
      if (kk < 0) continue;
      const cx<T> a = buf[kk * TKP + p], c = conj(buf[(kk ? M - kk : 0) * TKP + p]);
      const cx<T> Ev = a + c;
      const cx<T> Ov = mul_si<-1>(a - c);
      const cx<T> wO = cmul(twq<T, N>(tq, 4 * kk), Ov);          // e^{-2 pi i k/N} O
      const cx<T> Wk = cmul(twq<T, N>(tq, kk), Ev + wO);         // e^{-i pi k/2N} V_k
      const cx<T> Wm = cmul(twq<T, N>(tq, M - kk), conj(Ev - wO));
      y0[b] = Wk.x; y1[b] = -Wk.y; y2[b] = Wm.x; y3[b] = -Wm.y;
    }
    // z passes and store operators that read other fields write Y through
    // LDS and store it row-coalesced in chunks; plain strided stores go
    // straight from registers.
    constexpr bool STAGED = (AX == kAxZ) || store_reads<ST>();
    auto put = [&](int o, T val) {
      const int n = dst ? (N - 1 - o) : o;
      if constexpr (STAGED) {
        rbuf[((n >> 1) * TKP + p) * 2 + (n & 1)] = val;
      } else {
        int i, j, k;
        long idx;
        if (tg.at(p, n, i, j, k, idx)) store_eval<T, ST>(A.st, idx, val);
      }
    };
    if constexpr (STAGED) __syncthreads();
#pragma unroll
    for (int b = 0; b <= NPRE; ++b) {
      int kk = t + b * TP;
      if (b == NPRE) kk = (t == 0) ? M / 2 : -1;
      if (kk < 0) continue;
      put(kk, y0[b]);
      if (kk > 0) put(N - kk, y1[b]);
      if (kk < M / 2) {
        put(M - kk, y2[b]);
        if (kk > 0) put(M + kk, y3[b]);
      }
    }
    if constexpr (STAGED) {
      __syncthreads();
      constexpr int PER = TK * N / BLOCK;
      constexpr int CH = store_reads<ST>() ? (PER < 8 ? PER : 8) : PER;
#pragma unroll 1
      for (int s0 = 0; s0 < PER; s0 += CH) {
        StIn in[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int e = tid + (s0 + c) * BLOCK;
          int q, n;
          if constexpr (AX == kAxZ) { q = e / N; n = e % N; } else { q = tid % TK; n = e / TK; }
          int i, j, k;
          long idx;
          tg.at(q, n, i, j, k, idx);
          in[c] = store_inputs<T, ST>(A.st, kv_at(i), idx);
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int e = tid + (s0 + c) * BLOCK;
          int q, n;
          if constexpr (AX == kAxZ) { q = e / N; n = e % N; } else { q = tid % TK; n = e / TK; }
          int i, j, k;
          long idx;
          if (!tg.at(q, n, i, j, k, idx)) continue;
          const T y = rbuf[((n >> 1) * TKP + q) * 2 + (n & 1)];
          if constexpr (store_reads<ST>()) store_apply<T, ST>(A.st, A.phys, kv_at(i), idx, y, in[c]);
          else store_eval<T, ST>(A.st, idx, y);
        }
      }
    }
  } else {
    // ---- load X in natural order (DST-III: reversed) into LDS
    {
      constexpr int PER = TK * N / BLOCK;
      constexpr int CH = load_heavy<LD>() ? (PER < 8 ? PER : 8) : PER;
#pragma unroll 1
      for (int s0 = 0; s0 < PER; s0 += CH) {
        T xs[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int e = tid + (s0 + c) * BLOCK;
          int q, n;
          if constexpr (AX == kAxZ) { q = e / N; n = e % N; } else { q = tid % TK; n = e / TK; }
          int i, j, k;
          long idx;
          tg.at(q, n, i, j, k, idx);
          xs[c] = load_eval<T, LD>(A.ld, A.phys, kv_at(i), G, i, j, k, idx);
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int e = tid + (s0 + c) * BLOCK;
          int q, n;
          if constexpr (AX == kAxZ) { q = e / N; n = e % N; } else { q = tid % TK; n = e / TK; }
          int i, j, k;
          long idx;
          const T x = tg.at(q, n, i, j, k, idx) ? xs[c] : (T)0;
          if (A.part) dsum += (double)x;
          const int pos = dst ? (N - 1 - n) : n;
          rbuf[((pos >> 1) * TKP + q) * 2 + (pos & 1)] = x;
        }
      }
    }
    __syncthreads();
    // ---- DCT-III pre-processing straight into the stage-1 inputs:
    // Z_m = (V_m + V_{m+M}) + i e^{2 pi i m/N} (V_m - V_{m+M}),
    // V_m = e^{i pi m/2N}(X_m - i X_{N-m}),  V_{m+M} = e^{i pi m/2N} e^{i pi/4}(X_{M+m} - i X_{M-m})
    const cx<T> r8 = {(T)0.70710678118654752440, (T)0.70710678118654752440};
    auto X = [&](int n) -> T { return rbuf[((n >> 1) * TKP + p) * 2 + (n & 1)]; };
#pragma unroll
    for (int b = 0; b < NB1; ++b) {
      const int jj = t + b * TP;
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        const int m = jj + r * (M / R1);
        const T Xm = X(m), XNm = m ? X(N - m) : (T)0, XMp = X(M + m), XMm = X(M - m);
        const cx<T> w1 = conj(twq<T, N>(tq, m));
        const cx<T> vk = cmul(w1, cx<T>{Xm, -XNm});
        const cx<T> vm = cmul(cmul(w1, r8), cx<T>{XMp, -XMm});
        const cx<T> Ev = vk + vm;
        const cx<T> Ov = cmul(conj(twq<T, N>(tq, 4 * m)), vk - vm);
        v[b * R1 + r] = Ev + mul_si<1>(Ov);
      }
    }
    if constexpr (AX == kAxZ) {
      run_fft<T, N, M, E, TKP, 1, true>(typename PL::R{}, v, buf, p, t, tq);
      constexpr int PER = TK * N / BLOCK;
#pragma unroll
      for (int s = 0; s < PER; ++s) {
        const int e = tid + s * BLOCK;
        const int q = e / N, n = e % N;
        int i, j, k;
        long idx;
        if (!tg.at(q, n, i, j, k, idx)) continue;
        const int vq = vpos<N>(n);
        T y = rbuf[((vq >> 1) * TKP + q) * 2 + (vq & 1)];
        if (dst && (n & 1)) y = -y;
        store_eval<T, ST>(A.st, idx, y);
      }
    } else {
      run_fft<T, N, M, E, TKP, 1, false>(typename PL::R{}, v, buf, p, t, tq);
      // last stage outputs stay in registers: value b*RL+r is slot m = t + b TP + r M/RL
      constexpr int RL = Last<typename PL::R>::value, NBL = E / RL;
      int i, j, k;
      long idx;
      if (tg.at(p, 0, i, j, k, idx)) {
#pragma unroll
        for (int b = 0; b < NBL; ++b) {
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int m = t + b * TP + r * (M / RL);
            const cx<T> z = v[b * RL + r];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
              const int n = xpos<N>(2 * m + c);
              T y = c ? z.y : z.x;
              if (dst && (n & 1)) y = -y;
              tg.at(p, n, i, j, k, idx);
              store_eval<T, ST>(A.st, idx, y);
            }
          }
        }
      }
    }
  }

  // ---- FP64 block sum of the loaded values (deterministic order)
  if (A.part) {
    __syncthreads();
    double* red = (double*)buf;
    for (int off = 32; off > 0; off >>= 1) dsum += __shfl_down(dsum, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = dsum;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < (BLOCK + 63) / 64; ++w) s += red[w];
      A.part[tile] = s;
    }
  }
}

template <typename T, int N, int AX, bool INV, int LD, int ST, int V>
__global__ void __launch_bounds__((TileCfg<T, N, AX, V>::BLOCK)) k_pass(const PassArgs A) {
  pass_tile<T, N, AX, INV, LD, ST, V>(A, (int)blockIdx.x);
}

// Batched pass over the instances of an ensemble that share this pass's
// length N: workgroup b runs tile b - prefix[q] of instance q, where
// prefix[q] <= b < prefix[q+1] (prefix[0] = 0, prefix[count] = grid size).
template <typename T, int N, int AX, bool INV, int LD, int ST, int V>
__global__ void __launch_bounds__((TileCfg<T, N, AX, V>::BLOCK))
    k_pass_batch(const PassArgs* __restrict__ descs, const int* __restrict__ prefix, int count) {
  const int b = (int)blockIdx.x;
  int lo = 0, hi = count;  // invariant: prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid;
  }
  pass_tile<T, N, AX, INV, LD, ST, V>(descs[lo], b - prefix[lo]);
}


template <typename T, int N, int AX, int V = 0>
long tiles_of(const Geo& g) {
  using C = TileCfg<T, N, AX, V>;
  if (AX == kAxZ) return ((long)g.nx * g.ny + C::TK - 1) / C::TK;
  const long tiles_k = (g.nz + C::TK - 1) / C::TK;
  return (AX == kAxY ? (long)g.nx : (long)g.ny) * tiles_k;
}

// Single pass (B == nullptr) or batched pass (B: device descriptors).
template <typename T, int N, int AX, bool INV, int LD, int ST, int V = 0>
hipError_t launch_one(const PassArgs& A, long* nblocks_out, hipStream_t s, const BatchLaunch* B) {
  using C = TileCfg<T, N, AX, V>;
  if (B) {
    if (B->total_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_pass_batch<T, N, AX, INV, LD, ST, V>), dim3((unsigned)B->total_tiles),
                       dim3(C::BLOCK), 0, s, B->descs, B->prefix, B->count);
    return hipGetLastError();
  }
  const long nb = tiles_of<T, N, AX, V>(A.geo);
  if (nblocks_out) *nblocks_out = nb;
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_pass<T, N, AX, INV, LD, ST, V>), dim3((unsigned)nb), dim3(C::BLOCK), 0, s, A);
  return hipGetLastError();
}

// The (direction, load operator, store operator) combinations compiled for
// one precision T, axis AX and length N.  Plain passes exist for every
// compiled length; the fused operators exist where the step uses them:
// forward z passes load the real-space operators, forward x passes store the
// spectral epilogues, inverse x passes load the spectral prologues.
template <typename T, int AX, int N, bool STEP>
hipError_t dispatch_ops(bool inv, int ld, int st, int variant, const PassArgs& A, long* nb,
                        hipStream_t s, const BatchLaunch* B) {
  if (ld == LD_PLAIN && st == ST_PLAIN) {
    if (variant == 1)
      return inv ? launch_one<T, N, AX, true, LD_PLAIN, ST_PLAIN, 1>(A, nb, s, B)
                 : launch_one<T, N, AX, false, LD_PLAIN, ST_PLAIN, 1>(A, nb, s, B);
    return inv ? launch_one<T, N, AX, true, LD_PLAIN, ST_PLAIN>(A, nb, s, B)
               : launch_one<T, N, AX, false, LD_PLAIN, ST_PLAIN>(A, nb, s, B);
  }
  if constexpr (STEP && AX == kAxZ) {
    if (!inv && st == ST_PLAIN) {
      switch (ld) {
#define QSM_L(OP) case OP: return launch_one<T, N, AX, false, OP, ST_PLAIN>(A, nb, s, B);
        QSM_L(LD_ZERO) QSM_L(LD_RHOSTAR) QSM_L(LD_DIVV) QSM_L(LD_STRESS) QSM_L(LD_NONLIN)
        QSM_L(LD_NONLIN_P) QSM_L(LD_STRESS_STB) QSM_L(LD_NONLIN_STB) QSM_L(LD_NONLIN_STB0)
#undef QSM_L
        default: break;
      }
    }
  }
  if constexpr (STEP && AX == kAxX) {
    if (!inv && ld == LD_PLAIN) {
      switch (st) {
#define QSM_S(OP) case OP: return launch_one<T, N, AX, false, LD_PLAIN, OP>(A, nb, s, B);
        QSM_S(ST_FILTER) QSM_S(ST_UPDATE) QSM_S(ST_UPDATE_STB)
#undef QSM_S
        default: break;
      }
    }
    if (inv && st == ST_PLAIN) {
      switch (ld) {
#define QSM_L(OP) case OP: return launch_one<T, N, AX, true, OP, ST_PLAIN>(A, nb, s, B);
        QSM_L(LD_PLAIN_D) QSM_L(LD_COEF_F) QSM_L(LD_COEF_G) QSM_L(LD_GRAD) QSM_L(LD_GRAD_D)
        QSM_L(LD_LAPGRAD) QSM_L(LD_SH) QSM_L(LD_PRESSURE) QSM_L(LD_VEL0) QSM_L(LD_VEL1)
        QSM_L(LD_VEL2) QSM_L(LD_SCALE)
#undef QSM_L
        default: break;
      }
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace qsm
