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

#include <algorithm>

#include "qsm_fft.hpp"
#include "qsm_kernels.h"
#include "qsm_ops.hpp"

namespace qsm {

constexpr int kAxX = 0, kAxY = 1, kAxZ = 2;
// Walk elements whose loads a fused group pass issues together: 4 keeps the
// FP32 projection groups at 119 VGPRs with the held values and no spills
// (8: 53 VGPRs spilled).
constexpr int kFuseChunk = 4;




// Tile configuration of a pass.  V selects the tile-width variant: V = 0 the
// default width below, V = 1 half of it (64-byte row segments on strided axes,
// half the LDS per workgroup, twice the resident workgroups per CU).  The
// autotuner picks a variant per (axis, direction) for the plain passes.
template <typename T, int N, int AX, int V = 0>
struct TileCfg {
  static constexpr int M = N / 2;
  static constexpr int E = PlanV<T, M, V>::E;
  static constexpr int TP = M / E;  // threads per pencil
  static constexpr int tk() {
    if (AX == kAxZ) {
      // contiguous rows: enough pencils for a 256-thread block, LDS <= ~72 KB
      int t = 256 / TP;
      if (t < 1) t = 1;
      while (t > 1 && (long)M * (t + 1) * 2 * (long)sizeof(T) > 72000) t /= 2;
      return t;
    }
    // strided axes: 128 contiguous bytes of each row per tile (FP64 x
    // passes: 256), narrowed only where the tile does not fit one workgroup
    // per CU (tile slots (t+1)*M plus the 3M twiddle table within the
    // CU's 160 KiB of LDS).  Wider row segments carry the strided
    // access pattern (copy microbenchmarks: x stride 4.2 / 5.1 / 5.2 TB/s
    // with 64 / 128 / 256-byte segments in FP64 and FP32); at 512^3 the
    // 128-byte tiles also fit two workgroups per CU.  FP64 x at 256 bytes:
    // FP64 step 38.2 -> 36.0 ms at 512^3 (FP64 y and FP32 at 256 are slower).
    const int want = (AX == kAxX && sizeof(T) == 8) ? 32 : 128 / (int)sizeof(T);
    int t = want;
    while (t > 1 && (long)M * (t + 4) * 2 * (long)sizeof(T) > 160 * 1024 - 256) t /= 2;
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

// Real sample n of pencil q sits in LDS at rbuf[lpos(n, q)] (complex slot
// n/2 of the interleaved buf[slot * TKP + q]).
template <int TKP>
__device__ __forceinline__ int lpos(int n, int q) {
  return ((n >> 1) * TKP + q) * 2 + (n & 1);
}

// LDS positions X(t + K) and X(K - t) of a thread's own pencil p for
// compile-time K, from four per-thread bases: the parity of n and the half
// index n/2 follow from the parities of t and K, so each access is a base
// plus a constant.
template <int TKP>
struct PencilPos {
  int ue, uo, de, dd;
  __device__ __forceinline__ void init(int p, int t) {
    ue = lpos<TKP>(t, p);                                   // t + K, K even
    uo = (((t + 1) >> 1) * TKP + p) * 2 + 1 - (t & 1);      // t + K, K odd
    de = (p - ((t + 1) >> 1) * TKP) * 2 + (t & 1);          // K - t, K even
    dd = (p - (t >> 1) * TKP) * 2 + 1 - (t & 1);            // K - t, K odd
  }
  __device__ __forceinline__ int up(int K) const { return (K & 1) ? uo + (K - 1) * TKP : ue + K * TKP; }
  __device__ __forceinline__ int dn(int K) const { return (K & 1) ? dd + (K - 1) * TKP : de + K * TKP; }
};

// LDS position of element s of a thread's row-sweeping walk (the order of
// TileGeo::walk) for sample n at: n itself (MODE 0), its Makhoul position
// vpos(n) (MODE 1) or the reversed position N-1-n (MODE 2), with the parity
// of n for the DST sign.  Where the tile shape allows, pos(s) = a + (s mod R) b
// + (s div R) c with per-thread a, b and constant c; otherwise it is computed
// per element.
template <int AX, int N, int TK, int TKP, int BLOCK, int MODE>
struct WalkPos {
  static constexpr int TP = BLOCK / TK;
  // 1: strided axis, n = t + s TP; 2: z axis, N a multiple of BLOCK;
  // 3: z axis, BLOCK a multiple of N; 0: general
  static constexpr int KIND =
      AX != 2 ? ((TP % (MODE == 1 ? 4 : 2) == 0) ? 1 : 0)
              : ((N % BLOCK == 0 && BLOCK % 4 == 0) ? 2 : (BLOCK % N == 0 ? 3 : 0));
  static constexpr int R = KIND == 2 ? N / BLOCK : 1;
  int a, b;
  bool odd;
  __device__ __forceinline__ static int place(int n) {
    return MODE == 1 ? vpos<N>(n) : (MODE == 2 ? N - 1 - n : n);
  }
  __device__ __forceinline__ void init(int tid) {
    // n0: the thread's sample at s = 0; dn: the change of n per step of s
    // (strided) or per BLOCK (z); q0: its pencil
    int n0, q0, dn;
    if constexpr (KIND == 1) { n0 = tid / TK; q0 = tid % TK; dn = TP; }
    else if constexpr (KIND == 2) { n0 = tid; q0 = 0; dn = BLOCK; }
    else { n0 = tid % N; q0 = tid / N; dn = 0; }
    odd = (n0 & 1) != 0;
    const int v0 = place(n0);
    a = lpos<TKP>(v0, q0);
    // place() moves by +dn (MODE 0), -dn (MODE 2) or +-dn/2 (MODE 1, by the
    // parity of n0); the moves are even, so the half index moves by half that
    const int mv = MODE == 0 ? dn : (MODE == 2 ? -dn : (odd ? -dn / 2 : dn / 2));
    b = (mv / 2) * TKP * 2;
  }
  __device__ __forceinline__ int pos(int tid, int s, bool& oddn) const {
    if constexpr (KIND == 1) {
      oddn = odd;
      return a + s * b;
    } else if constexpr (KIND == 2) {
      oddn = odd;
      return a + (s % R) * b + (s / R) * 2;
    } else if constexpr (KIND == 3) {
      oddn = odd;
      return a + s * (BLOCK / N) * 2;
    } else {
      int q, n;
      if constexpr (AX == 2) { q = (tid + s * BLOCK) / N; n = (tid + s * BLOCK) % N; }
      else { q = tid % TK; n = tid / TK + s * TP; }
      oddn = (n & 1) != 0;
      return lpos<TKP>(place(n), q);
    }
  }
};

// Element coordinates of one workgroup's tile.  Fields have at most 2^29
// elements (checked on the host), so indices are 32-bit and every address is
// a uniform base plus a 32-bit offset.  Out-of-range elements (a partial
// last tile) get a valid index inside the field and ok = false; the caller
// discards their values and skips their stores.
//
// Two enumerations:
//   walk(s): element s of the thread's row-sweeping walk e = tid + s BLOCK
//            (loads, staged stores): z passes read rows, strided passes read
//            TK-column segments of successive rows; the index advances by a
//            uniform step per s.
//   at(n):   element n of this thread's pencil (strided passes; the FFT
//            phase keeps pencil p = tid % TK).
template <int AX, int N, int TK, int BLOCK>
struct TileGeo {
  int ny, nrows;
  int row0;        // z: first row (i ny + j) of the tile
  int t_i, t_j;    // strided: the tile's fixed i (y pass) or j (x pass)
  int kc;          // strided: the thread's column, clamped into the field
  bool colok;      // strided: the thread's column is inside the field
  Ix cbase;        // strided: index of the thread's column at n = 0
  Ix axs;          // strided: index stride along the pass axis
  Ix wbase, wstep; // walk: index of element 0 and the uniform step per s

  __device__ __forceinline__ void init(const Geo& G, int tile, int tid) {
    ny = G.ny;
    nrows = G.nx * G.ny;
    if constexpr (AX == kAxZ) {
      row0 = tile * TK;
      wbase = (Ix)row0 * N + (Ix)tid;
      wstep = BLOCK;
    } else {
      const int tiles_k = (G.nz + TK - 1) / TK;
      const int a = tile / tiles_k;
      const int k0 = (tile - a * tiles_k) * TK;
      const int q = tid % TK;
      colok = k0 + q < G.nz;
      kc = colok ? k0 + q : k0;
      if constexpr (AX == kAxY) {
        t_i = a;
        cbase = (Ix)a * G.ny * G.nz + (Ix)kc;
        axs = G.nz;
      } else {
        t_j = a;
        cbase = (Ix)a * G.nz + (Ix)kc;
        axs = (Ix)G.ny * G.nz;
      }
      wbase = cbase + (Ix)(tid / TK) * axs;
      wstep = (Ix)(BLOCK / TK) * axs;
    }
  }

  __device__ __forceinline__ void walk(int tid, int s, int& q, int& n, int& i, int& j, int& k,
                                       Ix& idx, bool& ok) const {
    if constexpr (AX == kAxZ) {
      const int e = tid + s * BLOCK;
      q = e / N;
      n = e % N;
      const int row = row0 + q;
      ok = row < nrows;
      idx = ok ? wbase + (Ix)s * wstep : (Ix)n;
      const int rr = ok ? row : 0;
      i = rr / ny;
      j = rr - i * ny;
      k = n;
    } else {
      q = tid % TK;
      n = tid / TK + s * (BLOCK / TK);
      ok = colok;
      idx = wbase + (Ix)s * wstep;
      if constexpr (AX == kAxY) { i = t_i; j = n; } else { i = n; j = t_j; }
      k = kc;
    }
  }

  // strided passes only: element n of this thread's column
  __device__ __forceinline__ void at(int n, int& i, int& j, int& k, Ix& idx) const {
    static_assert(AX != kAxZ, "at() is for strided passes");
    idx = cbase + (Ix)n * axs;
    if constexpr (AX == kAxY) { i = t_i; j = n; } else { i = n; j = t_j; }
    k = kc;
  }
};

// LDS of a pass: the tile, the twiddles and the pass-axis wavenumber tables
// of mask TAB (bit 0: q, 1: s, 2: f; each N doubles, in that order) with, in
// z passes, the tile's per-row q_x, s_x, q_y, s_y.  It is declared once per
// kernel, so the passes of a group kernel share it.
template <typename T, int N, int AX, bool INV, int V, int TAB>
struct Smem {
  using C = TileCfg<T, N, AX, V>;
  static constexpr int M = C::M, TK = C::TK, TKP = C::TKP;
  static constexpr int VAR = V;
  static constexpr int NTP = INV ? TwLayout<M>::NPRE : TwLayout<M>::NPOST;
  static constexpr int tab = TAB, ntab = kv_count(TAB);
  static constexpr int nrowt = (AX == 2 && ntab > 0) ? 4 * TK : 0;
  static constexpr long bytes() {
    return (long)sizeof(cx<T>) * ((long)M * TKP + M + NTP) +
           (long)sizeof(double) * (ntab * N + nrowt);
  }
  cx<T> buf[M * TKP];  // the tile: complex slot m of pencil p at m * TKP + p
  cx<T> ts[M];         // stage twiddles
  cx<T> tp[NTP];       // pre- (inverse) or post-processing (forward) twiddles
  // the tables (ntab * N values), then the per-row values (nrowt); one array,
  // of length 0 in a pass without tables: a placeholder element would take
  // the FP64 inverse y tile at 512 (81920 bytes) past half the CU's LDS, to
  // one workgroup per CU
  double xt[ntab * N + nrowt];
};

// Tables staged in LDS: those of mask NT where they fit beside the tile while
// two workgroups share a CU (160 KiB); otherwise none, and the operator reads
// them from global memory.
template <typename T, int N, int AX, bool INV, int V, int NT>
constexpr int smem_tables() {
  return (NT != 0 && Smem<T, N, AX, INV, V, NT>::bytes() <= 81920) ? NT : 0;
}

// Passes with inverse pre-rounds (pre_rounds) stage the DCT-III
// pre-processing twiddles too: their LDS takes the inverse layout (2M
// entries, more than the forward post-processing's 3(M/2+1)).
template <typename T, int N, int AX, bool INV, int LD, int ST, int V>
using SmemFor = Smem<T, N, AX, INV || pre_rounds(LD) || st_inverts(ST), V,
                     smem_tables<T, N, AX, INV || pre_rounds(LD) || st_inverts(ST), V,
                                 kv_tables<AX, LD, ST>()>()>;

// DCT-III pre-processing of pencil p straight into the stage-1 inputs of
// thread t (the real samples X of the pencil in LDS, natural order):
// Z_m = (V_m + V_{m+M}) + i e^{2 pi i m/N} (V_m - V_{m+M}),
// V_m = e^{i pi m/2N}(X_m - i X_{N-m}),  V_{m+M} = e^{i pi m/2N} e^{i pi/4}(X_{M+m} - i X_{M-m}).
// With a = X_m - i X_{N-m}, c = e^{i pi/4}(X_{M+m} - i X_{M-m}):
// Z_m = A (a + c) + B (a - c),  A = e^{i pi m/2N},  B = i e^{2 pi i m/N} A = e^{i pi (5m+N)/2N}
// (tp[2m], tp[2m+1]).
template <typename T, int N, int M, int E, int TKP, int R1>
__device__ __forceinline__ void inv_preprocess(cx<T>* v, const T* rbuf, const cx<T>* tp, int p,
                                               int t) {
  constexpr int TP = M / E, NB1 = E / R1;
  const T h = (T)0.70710678118654752440;
  // samples m, N-m, M+m, M-m of item m = t + C from per-thread bases
  PencilPos<TKP> pp;
  pp.init(p, t);
#pragma unroll
  for (int b = 0; b < NB1; ++b) {
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int C = b * TP + r * (M / R1);
      const int m = t + C;
      const T Xm = rbuf[pp.up(C)], XMp = rbuf[pp.up(M + C)], XMm = rbuf[pp.dn(M - C)];
      T XNm;  // X_N = 0
      if (C == 0) XNm = t == 0 ? (T)0 : rbuf[t == 0 ? pp.up(0) : pp.dn(N)];
      else XNm = rbuf[pp.dn(N - C)];
      // c = h (d1 + i d2)
      const T d1 = XMp + XMm, d2 = XMp - XMm;
      const cx<T> P = {fma(h, d1, Xm), fma(h, d2, -XNm)};
      const cx<T> Q = {fma(-h, d1, Xm), fma(-h, d2, -XNm)};
      const cx<T> A = tp[2 * m], B = tp[2 * m + 1];
      v[b * R1 + r] = {fma(A.x, P.x, fma(-A.y, P.y, fma(B.x, Q.x, -B.y * Q.y))),
                       fma(A.x, P.y, fma(A.y, P.x, fma(B.x, Q.y, B.y * Q.x)))};
    }
  }
}

// Values per thread of a pass's load walk (the tile's elements over the block).
template <typename T, int N, int AX, int V>
constexpr int walk_count() {
  return TileCfg<T, N, AX, V>::TK * N / TileCfg<T, N, AX, V>::BLOCK;
}

// One workgroup's work: tile `tile` of the pass described by A, in the LDS
// sm.  stage = false skips the staging of twiddles and tables a previous
// pass of the same kernel (same axis and direction) left in sm.  NTS: the
// direct stores of inverse strided passes are non-temporal.
// Two further modes of inverse passes serve fused group outputs:
//   XLD >= 0  the load walk also evaluates load operator XLD (arguments
//             *xop) at every element it reads, into held[s] for walk element
//             s, so a second output that reads the same fields is formed
//             from the same reads;
//   FROMH     the tile is not loaded: held[s] (from an XLD pass of this
//             tile) is the input at walk element s.
//   PF        (with FROMH) called once the tile is in LDS, when held[] is
//             free again: the pipelined passes load the next tile into it
struct NoPrefetch {
  __device__ void operator()() const {}
};

template <typename T, int N, int AX, bool INV, int LD, int ST, int V, class SM, bool NTS = false,
          int XLD = -1, bool FROMH = false, class PF = NoPrefetch>
__device__ __forceinline__ void pass_tile(const Geo& G, const Waves& W, const Phys& Ph,
                                          const LoadOp& Lop, const StoreOp& Sop,
                                          const void* twp, double* part, const int dstf,
                                          const int tile, SM& sm, const bool stage = true,
                                          const int tid = (int)threadIdx.x,
                                          const LoadOp* xop = nullptr, T* held = nullptr,
                                          const PF& prefetch = PF{}) {
  static_assert(XLD < 0 || (INV && AX != 2), "fused outputs: strided inverse passes");
  static_assert(!FROMH || AX != 2, "held inputs: strided passes");
  static_assert(XLD < 0 || load_split<LD>() == load_split<XLD>(), "fused operators load alike");
  using C = TileCfg<T, N, AX, V>;
  using PL = PlanV<T, N / 2, V>;
  constexpr int M = C::M, E = C::E, TP = C::TP, TK = C::TK, TKP = C::TKP, BLOCK = C::BLOCK;
  constexpr int R1 = PL::R1;
  using TW = TwLayout<M>;
  constexpr int NTP = INV ? TW::NPRE : TW::NPOST;
  static_assert(SM::NTP == (INV || pre_rounds(LD) || st_inverts(ST) ? TW::NPRE : NTP) &&
                    SM::TKP == TKP,
                "LDS layout of another pass");
  static_assert(!st_inverts(ST) || (AX == kAxX && !INV), "x inverse after the store: forward x passes");
  static_assert(!pre_rounds(LD) || (AX == kAxZ && !INV), "pre-rounds: forward z passes");
  cx<T>* const buf = sm.buf;
  cx<T>* const ts = sm.ts;
  cx<T>* const tp = sm.tp;
  T* rbuf = (T*)buf;  // real view: sample n of pencil p at ((n>>1)*TKP + p)*2 + (n&1)
  const bool dst = dstf != 0;

  // Wavenumbers.  The pass-axis tables (q and s, or the filter factors f)
  // are staged in LDS where they fit (SM::ntab), else read per element; the
  // other axes' values are fixed per thread (strided passes: one i or j per
  // tile, one column k per thread) or per row (z passes: staged per row).
  constexpr int NT = kv_tables<AX, LD, ST>() | (XLD >= 0 ? kv_tables<AX, XLD, ST_PLAIN>() : 0);
  constexpr bool KV = uses_kv<LD, ST>() || (XLD >= 0 && uses_kv<XLD, ST_PLAIN>());
  constexpr bool XT_LDS = NT != 0 && (SM::tab & NT) == NT;
  constexpr bool ROWT = AX == kAxZ && SM::ntab > 0;  // per-row wavenumbers staged (z)
  // slot of table bit b in sm.xt
  constexpr int SQ = 0, SS = (SM::tab & 1) ? 1 : 0, SF = ((SM::tab & 1) ? 1 : 0) + ((SM::tab & 2) ? 1 : 0);
  const double* tq = AX == 0 ? W.qx : (AX == 1 ? W.qy : W.qz);
  const double* tsn = AX == 0 ? W.sx : (AX == 1 ? W.sy : W.sz);
  const double* tf = AX == 0 ? W.fx : (AX == 1 ? W.fy : W.fz);

  TileGeo<AX, N, TK, BLOCK> tg;
  tg.init(G, tile, tid);
  // Split addressing of the walk for plain loads and stores: element s at
  // the per-thread tg.wbase plus the uniform walk_uni(s).  In z tiles of
  // N % BLOCK == 0, element s lies in row row0 + s / R of the tile (uniform);
  // rows past the field read the tile's first row instead.
  constexpr bool SPLIT = AX != kAxZ || N % BLOCK == 0;
  constexpr int RZ = (AX == kAxZ && N % BLOCK == 0) ? N / BLOCK : 1;
  auto row_ok = [&](int s) { return AX != kAxZ || tg.row0 + s / RZ < tg.nrows; };
  auto walk_uni = [&](int s) -> int {
    if constexpr (AX != kAxZ) return s * (int)tg.wstep;
    else return row_ok(s) ? s * BLOCK : (s % RZ) * BLOCK;
  };

  if (stage) {
    const cx<T>* __restrict__ tw = (const cx<T>*)twp;
    for (int u = tid; u < M; u += BLOCK) ts[u] = tw[TW::STAGE + u];
    if constexpr (pre_rounds(LD))  // the pre-rounds first; the forward's tables after them
      for (int u = tid; u < TW::NPRE; u += BLOCK) tp[u] = tw[TW::PRE + u];
    else
      for (int u = tid; u < NTP; u += BLOCK) tp[u] = tw[(INV ? TW::PRE : TW::POST) + u];
    if constexpr (SM::ntab > 0) {
      for (int u = tid; u < N; u += BLOCK) {
        if constexpr ((SM::tab & 1) != 0) sm.xt[SQ * N + u] = tq[u];
        if constexpr ((SM::tab & 2) != 0) sm.xt[SS * N + u] = tsn[u];
        if constexpr ((SM::tab & 4) != 0) sm.xt[SF * N + u] = tf[u];
      }
      if constexpr (ROWT) {
        for (int r = tid; r < TK; r += BLOCK) {
          int row = tg.row0 + r;
          if (row >= tg.nrows) row = tg.row0;
          const int i = row / G.ny, j = row - i * G.ny;
          sm.xt[SM::ntab * N + 4 * r + 0] = W.qx[i];
          sm.xt[SM::ntab * N + 4 * r + 1] = W.sx[i];
          sm.xt[SM::ntab * N + 4 * r + 2] = W.qy[j];
          sm.xt[SM::ntab * N + 4 * r + 3] = W.sy[j];
        }
      }
      __syncthreads();
    }
  }

  // FFT-phase thread coordinates: pencil p, thread t within the pencil
  const int p = tid % TK, t = tid / TK;
  // wavenumbers fixed per thread in strided passes (a thread keeps its
  // column p = tid % TK in every phase)
  Kv kv{};
  if constexpr (KV && AX == kAxX) {
    const int j = tg.t_j, k = tg.kc;
    kv.qy = W.qy[j]; kv.sy = W.sy[j]; kv.fy = W.fy[j];
    kv.qz = W.qz[k]; kv.sz = W.sz[k]; kv.fz = W.fz[k];
  } else if constexpr (KV && AX == kAxY) {
    const int i = tg.t_i, k = tg.kc;
    kv.qx = W.qx[i]; kv.sx = W.sx[i]; kv.fx = W.fx[i];
    kv.qz = W.qz[k]; kv.sz = W.sz[k]; kv.fz = W.fz[k];
  }
  // element (i, j, k) of pencil q
  auto kv_at = [&](int i, int j, int k, int q) {
    Kv K = kv;
    if constexpr (KV) {
      const int m = AX == 0 ? i : (AX == 1 ? j : k);
      double a = 0.0, b = 0.0, f = 0.0;  // pass-axis q, s, f
      if constexpr (XT_LDS) {
        if constexpr ((NT & 1) != 0) a = sm.xt[SQ * N + m];
        if constexpr ((NT & 2) != 0) b = sm.xt[SS * N + m];
        if constexpr ((NT & 4) != 0) f = sm.xt[SF * N + m];
      } else {
        if constexpr ((NT & 1) != 0) a = tq[m];
        if constexpr ((NT & 2) != 0) b = tsn[m];
        if constexpr ((NT & 4) != 0) f = tf[m];
      }
      if constexpr (AX == kAxX) {
        K.qx = a; K.sx = b; K.fx = f;
      } else if constexpr (AX == kAxY) {
        K.qy = a; K.sy = b; K.fy = f;
      } else {
        K.qz = a; K.sz = b; K.fz = f;
        if constexpr (ROWT) {
          K.qx = sm.xt[SM::ntab * N + 4 * q + 0]; K.sx = sm.xt[SM::ntab * N + 4 * q + 1];
          K.qy = sm.xt[SM::ntab * N + 4 * q + 2]; K.sy = sm.xt[SM::ntab * N + 4 * q + 3];
        } else {
          K.qx = W.qx[i]; K.sx = W.sx[i]; K.qy = W.qy[j]; K.sy = W.sy[j];
        }
        K.fx = W.fx[i]; K.fy = W.fy[j];
      }
    }
    return K;
  };
  // FP64 sum of the loaded values (mixed-precision DC modes): the forward
  // z passes of the transforms whose (0,0,0) mode is kept in FP64
  constexpr bool DSUM = !INV && AX == kAxZ && dc_capable(LD);
  double dsum = 0.0;
  cx<T> v[E];

  // LD_NONLIN_PY: p (a10) arrives y-passed, and |grad rho|^2 (a11) as the
  // stress pass formed it from the grad rho values it transformed in
  // registers.  One inverse z round over the tile completes p's transform;
  // each thread folds the values of its walk elements into the pressure term
  // p kp |grad rho| / rho_s^2, the one value the operator then takes from
  // registers.  p's separate inverse z pass is not run.
  T pterm[pre_rounds(LD) ? TK * N / BLOCK : 1];
  if constexpr (pre_rounds(LD)) {
    constexpr int PER = TK * N / BLOCK;
    WalkPos<AX, N, TK, TKP, BLOCK, 0> wn;  // natural order (DCT-III)
    WalkPos<AX, N, TK, TKP, BLOCK, 1> wq;  // the transform's output positions
    wn.init(tid);
    wq.init(tid);
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      int q, n, i, j, k;
      Ix idx;
      bool ok, oddn;
      tg.walk(tid, s, q, n, i, j, k, idx, ok);
      T x;
      if constexpr (SPLIT) {
        x = ld_at<T>(Lop.a[10], tg.wbase, walk_uni(s));
        pterm[s] = ld_at<T>(Lop.a[11], tg.wbase, walk_uni(s));
      } else {
        x = ld<T>(Lop.a[10], idx);
        pterm[s] = ld<T>(Lop.a[11], idx);
      }
      rbuf[wn.pos(tid, s, oddn)] = x;
    }
    __syncthreads();
    cx<T> vr[E];
    inv_preprocess<T, N, M, E, TKP, R1>(vr, rbuf, tp, p, t);
    run_fft<T, N, M, E, TKP, 1, true>(typename PL::R{}, vr, buf, p, t, ts);
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      bool oddn;
      pterm[s] = pressure_term<T>(rbuf[wq.pos(tid, s, oddn)], pterm[s], Ph);
    }
    __syncthreads();  // the load phase rewrites the tile
    // the forward post-processing twiddles, read after the load phase's barrier
    const cx<T>* __restrict__ tw = (const cx<T>*)twp;
    for (int u = tid; u < TW::NPOST; u += BLOCK) tp[u] = tw[TW::POST + u];
  }

  if constexpr (!INV) {
    {
      // row-sweeping loads into LDS in Makhoul order (a wavefront reads 64
      // consecutive elements of a z row, or TK-column segments of
      // consecutive rows of a strided axis); the loads of a chunk are all
      // issued before any of its values is used
      constexpr int PER = TK * N / BLOCK;
      constexpr int CH = FROMH ? PER : (load_heavy<LD>() ? (PER < 8 ? PER : 8) : PER);
      static_assert(!FROMH || PER == walk_count<T, N, AX, V>(), "walk length");
      static_assert(!pre_rounds(LD) || CH == PER, "held values: one chunk");
      WalkPos<AX, N, TK, TKP, BLOCK, 1> wp;
      wp.init(tid);
#pragma unroll 1
      for (int s0 = 0; s0 < PER; s0 += CH) {
        T xs[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          int q, n, i, j, k;
          Ix idx;
          bool ok;
          tg.walk(tid, s0 + c, q, n, i, j, k, idx, ok);
          if constexpr (FROMH) xs[c] = held[c];  // one chunk (CH = PER): s0 = 0
          else if constexpr (LD == LD_PLAIN && SPLIT) xs[c] = ld_at<T>(Lop.a[0], tg.wbase, walk_uni(s0 + c));
          else if constexpr (pre_rounds(LD))  // one chunk (CH = PER): s0 = 0
            xs[c] = load_eval<T, LD>(Lop, Ph, kv_at(i, j, k, q), G, i, j, k, idx, pterm[c]);
          else xs[c] = load_eval<T, LD>(Lop, Ph, kv_at(i, j, k, q), G, i, j, k, idx);
        }
        // values of rows or columns outside the field (read from a valid
        // address) go through the transform of their own pencil only and
        // are never stored; only the DC sum must leave them out
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          int q, n, i, j, k;
          Ix idx;
          bool ok, oddn;
          tg.walk(tid, s0 + c, q, n, i, j, k, idx, ok);
          T x = xs[c];
          if constexpr (DSUM) {
            if (!ok) x = (T)0;
            dsum += (double)x;
          }
          const int lp = wp.pos(tid, s0 + c, oddn);
          if (dst & oddn) x = -x;
          rbuf[lp] = x;
        }
      }
      __syncthreads();
      if constexpr (FROMH) prefetch();
      gather<T, M, E, TKP, R1>(v, buf, p, t);
    }
    run_fft<T, N, M, E, TKP, -1, true>(typename PL::R{}, v, buf, p, t, ts);

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
      const cx<T> wO = cmul(tp[3 * kk], Ov);                 // e^{-2 pi i k/N} O
      const cx<T> Wk = cmul(tp[3 * kk + 1], Ev + wO);        // e^{-i pi k/2N} V_k
      const cx<T> Wm = cmul(tp[3 * kk + 2], conj(Ev - wO));  // e^{-i pi (M-k)/2N} V_{M-k}
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
        Ix idx;
        tg.at(n, i, j, k, idx);
        if (tg.colok) store_eval<T, ST>(Sop, idx, val);
      }
    };
    // outputs o = kk, N-kk, M-kk, M+kk of item kk = t + C are t + K or K - t
    // for constant K: positions from per-thread bases (the DST output n =
    // N-1-o swaps the two families)
    PencilPos<TKP> pp;
    pp.init(p, t);
    // split addressing of the direct stores: thread bases at samples t and
    // M - t (inside the field), uniform offsets along the axis
    Ix gu = 0, gd = 0;
    int ax = 0;
    if constexpr (!STAGED) {
      gu = tg.cbase + (Ix)t * tg.axs;
      gd = tg.cbase + (Ix)(M - t) * tg.axs;
      ax = (int)tg.axs;
    }
    auto put_k = [&](bool up, int K, T val) {
      if constexpr (STAGED) {
        const int pa = up ? pp.up(K) : pp.dn(K);
        const int pb = up ? pp.dn(N - 1 - K) : pp.up(N - 1 - K);
        rbuf[dst ? pb : pa] = val;
      } else {
        // sample t + K (up) or K - t (down); DST: N-1 minus that
        if (tg.colok) {
          if (dst) {
            if (up) st_at<T>(Sop.d[0], gd, (N - 1 - K - M) * ax, val);
            else st_at<T>(Sop.d[0], gu, (N - 1 - K) * ax, val);
          } else {
            if (up) st_at<T>(Sop.d[0], gu, K * ax, val);
            else st_at<T>(Sop.d[0], gd, (K - M) * ax, val);
          }
        }
      }
    };
    if constexpr (STAGED) __syncthreads();
#pragma unroll
    for (int b = 0; b < NPRE; ++b) {
      const int C = b * TP;  // kk = t + C < M/2
      put_k(true, C, y0[b]);
      if (C > 0 || t > 0) put_k(false, N - C, y1[b]);
      put_k(false, M - C, y2[b]);
      if (C > 0 || t > 0) put_k(true, M + C, y3[b]);
    }
    if (t == 0) {  // kk = M/2
      put(M / 2, y0[NPRE]);
      put(N - M / 2, y1[NPRE]);
    }
    if constexpr (STAGED) {
      __syncthreads();
      constexpr int PER = TK * N / BLOCK;
      constexpr int CH = store_reads<ST>() ? (PER < 8 ? PER : 8) : PER;
      DcIn dcv{};
      if constexpr (store_reads<ST>()) dcv = dc_in(Sop);
      WalkPos<AX, N, TK, TKP, BLOCK, 0> wp;
      wp.init(tid);
#pragma unroll 1
      for (int s0 = 0; s0 < PER; s0 += CH) {
        StIn in[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          int q, n, i, j, k;
          Ix idx;
          bool ok;
          tg.walk(tid, s0 + c, q, n, i, j, k, idx, ok);
          in[c] = store_inputs<T, ST>(Sop, kv_at(i, j, k, q), idx);
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          bool oddn;
          if constexpr (!store_reads<ST>() && SPLIT && AX == kAxZ) {
            if (!row_ok(s0 + c)) continue;
            const T y = rbuf[wp.pos(tid, s0 + c, oddn)];
            st_at<T>(Sop.d[0], tg.wbase, (s0 + c) * BLOCK, y);
          } else {
            int q, n, i, j, k;
            Ix idx;
            bool ok;
            tg.walk(tid, s0 + c, q, n, i, j, k, idx, ok);
            if (!ok) continue;
            const int lp = wp.pos(tid, s0 + c, oddn);
            const T y = rbuf[lp];
            if constexpr (store_reads<ST>()) {
              const T sv = store_apply<T, ST>(Sop, Ph, kv_at(i, j, k, q), idx, y, in[c], dcv);
              // the stored value at this element's natural-order slot (read
              // and rewritten by this thread only), the x inverse's input
              if constexpr (st_inverts(ST)) rbuf[lp] = sv;
            } else {
              store_eval<T, ST>(Sop, idx, y);
            }
          }
        }
      }
    }
    if constexpr (st_inverts(ST)) {
      // ---- the inverse x transform (DCT-III) of the stored values into d3:
      // the pre-processing twiddles replace the post-processing ones, then
      // the same steps and register stores as an inverse x pass
      __syncthreads();
      {
        const cx<T>* __restrict__ tw = (const cx<T>*)twp;
        for (int u = tid; u < TW::NPRE; u += BLOCK) tp[u] = tw[TW::PRE + u];
      }
      __syncthreads();
      cx<T> vi[E];
      inv_preprocess<T, N, M, E, TKP, R1>(vi, rbuf, tp, p, t);
      run_fft<T, N, M, E, TKP, 1, false>(typename PL::R{}, vi, buf, p, t, ts);
      constexpr int RL = Last<typename PL::R>::value, NBL = E / RL;
      if (tg.colok) {
        const Ix g4u = tg.cbase + (Ix)(4 * t) * tg.axs, g4d = tg.cbase - (Ix)(4 * t) * tg.axs;
        const Ix tu = g4u, td = tg.cbase + (Ix)(N - 1 - 4 * t) * tg.axs;
        const int ax = (int)tg.axs;
#pragma unroll
        for (int b = 0; b < NBL; ++b) {
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int C = b * TP + r * (M / RL);
            const cx<T> zv = vi[b * RL + r];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
              const int u0 = 2 * C + c;
              const T y = c ? zv.y : zv.x;
              if (2 * (TP - 1) + u0 < M) {
                st_at<T>(Sop.d[3], tu, 2 * u0 * ax, y);
              } else if (u0 >= M) {
                st_at<T>(Sop.d[3], td, (N - 2 * u0) * ax, y);
              } else {
                const Ix iu = g4u + (Ix)(2 * u0) * tg.axs;
                const Ix idn = g4d + (Ix)(2 * N - 1 - 2 * u0) * tg.axs;
                st<T>(Sop.d[3], 2 * t + u0 < M ? iu : idn, y);
              }
            }
          }
        }
      }
    }
  } else {
    // ---- load X in natural order (DST-III: reversed) into LDS
    {
      constexpr int PER = TK * N / BLOCK;
      static_assert(PER == walk_count<T, N, AX, V>(), "walk length");
      // chunk of the walk whose loads are issued together (fused: fewer, as
      // held[] already takes walk_count registers)
      constexpr int CHL = XLD >= 0 ? kFuseChunk : (load_heavy<LD>() ? 8 : PER);
      constexpr int CH = FROMH ? PER : (PER < CHL ? PER : CHL);
      static_assert(PER % CH == 0, "whole chunks");
      WalkPos<AX, N, TK, TKP, BLOCK, 0> wn;
      WalkPos<AX, N, TK, TKP, BLOCK, 2> wr;
      wn.init(tid);
      wr.init(tid);
      auto chunk = [&](const int s0) {
        T xs[CH];
        if constexpr (FROMH) {
#pragma unroll
          for (int c = 0; c < CH; ++c) xs[c] = held[s0 + c];
        } else if constexpr (load_split<LD>()) {
          Raw4<T> raw[CH];
          Raw4<T> xraw[XLD >= 0 ? CH : 1];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            int q, n, i, j, k;
            Ix idx;
            bool ok;
            tg.walk(tid, s0 + c, q, n, i, j, k, idx, ok);
            raw[c] = fetch<T, LD>(Lop, G, i, j, k, idx);
            if constexpr (XLD >= 0) xraw[c] = fetch<T, (XLD < 0 ? LD : XLD)>(*xop, G, i, j, k, idx);
          }
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            int q, n, i, j, k;
            Ix idx;
            bool ok;
            tg.walk(tid, s0 + c, q, n, i, j, k, idx, ok);
            const Kv K = kv_at(i, j, k, q);
            xs[c] = finish<T, LD>(raw[c], Lop, Ph, K, G, i, j, k);
            if constexpr (XLD >= 0) {
              // held[] as a shift register: after the last chunk, held[s]
              // is walk element s (constant indices in a rolled loop)
              const T xv = finish<T, (XLD < 0 ? LD : XLD)>(xraw[c], *xop, Ph, K, G, i, j, k);
#pragma unroll
              for (int u = 0; u + 1 < PER; ++u) held[u] = held[u + 1];
              held[PER - 1] = xv;
            }
          }
        } else if constexpr (LD == LD_GRAD_X + AX && SPLIT) {
          // the gradient along the pass axis (of a field transformed along
          // the axes before it): the shifted reads at the walk's split
          // address plus a uniform offset (inside the field or its padding),
          // then the operator's expression as in load_eval
          const int o = AX == kAxX ? G.ny * G.nz : (AX == kAxY ? G.nz : 1);
          const int nc = AX == kAxX ? G.nx : (AX == kAxY ? G.ny : G.nz);
          T raw[CH];
#pragma unroll
          for (int c = 0; c < CH; ++c) raw[c] = ld_at<T>(Lop.a[0], tg.wbase, walk_uni(s0 + c) + o);
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            int q, n, i, j, k;
            Ix idx;
            bool ok;
            tg.walk(tid, s0 + c, q, n, i, j, k, idx, ok);
            const Kv K = kv_at(i, j, k, q);
            const int m = AX == kAxX ? i : (AX == kAxY ? j : k);
            const double st = AX == kAxX ? K.sx : (AX == kAxY ? K.sy : K.sz);
            if constexpr (sizeof(T) == sizeof(float)) xs[c] = m < nc - 1 ? (T)(-st) * raw[c] : (T)0;
            else xs[c] = (T)(-st * (m < nc - 1 ? (double)raw[c] : 0.0));
          }
        } else {
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            int q, n, i, j, k;
            Ix idx;
            bool ok;
            tg.walk(tid, s0 + c, q, n, i, j, k, idx, ok);
            if constexpr (LD == LD_PLAIN && SPLIT) xs[c] = ld_at<T>(Lop.a[0], tg.wbase, walk_uni(s0 + c));
            else xs[c] = load_eval<T, LD>(Lop, Ph, kv_at(i, j, k, q), G, i, j, k, idx);
            if constexpr (XLD >= 0)
              held[s0 + c] = load_eval<T, (XLD < 0 ? LD : XLD)>(*xop, Ph, kv_at(i, j, k, q), G, i, j, k, idx);
          }
        }
        // values outside the field stay in their own pencil and are never
        // stored; dst (uniform) picks the reversed positions
        bool oddn;
        if (dst) {
#pragma unroll
          for (int c = 0; c < CH; ++c) rbuf[wr.pos(tid, s0 + c, oddn)] = xs[c];
        } else {
#pragma unroll
          for (int c = 0; c < CH; ++c) rbuf[wn.pos(tid, s0 + c, oddn)] = xs[c];
        }
      };
      if constexpr (FROMH) {
        // held[] is indexed by walk element: every index a constant
#pragma unroll
        for (int s0 = 0; s0 < PER; s0 += CH) chunk(s0);
      } else {
#pragma unroll 1
        for (int s0 = 0; s0 < PER; s0 += CH) chunk(s0);
      }
    }
    __syncthreads();
    if constexpr (FROMH) prefetch();
    inv_preprocess<T, N, M, E, TKP, R1>(v, rbuf, tp, p, t);
    if constexpr (AX == kAxZ) {
      run_fft<T, N, M, E, TKP, 1, true>(typename PL::R{}, v, buf, p, t, ts);
      constexpr int PER = TK * N / BLOCK;
      WalkPos<AX, N, TK, TKP, BLOCK, 1> wp;
      wp.init(tid);
#pragma unroll
      for (int s = 0; s < PER; ++s) {
        bool oddn;
        if constexpr (SPLIT) {
          if (!row_ok(s)) continue;
          T y = rbuf[wp.pos(tid, s, oddn)];
          if (dst & oddn) y = -y;
          st_at<T>(Sop.d[0], tg.wbase, s * BLOCK, y);
        } else {
          int q, n, i, j, k;
          Ix idx;
          bool ok;
          tg.walk(tid, s, q, n, i, j, k, idx, ok);
          if (!ok) continue;
          T y = rbuf[wp.pos(tid, s, oddn)];
          if (dst & oddn) y = -y;
          store_eval<T, ST>(Sop, idx, y);
        }
      }
    } else {
      run_fft<T, N, M, E, TKP, 1, false>(typename PL::R{}, v, buf, p, t, ts);
      // last stage outputs stay in registers: value b*RL+r is slot m = t + C,
      // C = b TP + r M/RL, holding samples u = 2m + c at n = xpos(u):
      // n = 4t + 2u0 (even) while u < M, else 2N-1-2u0-4t (odd), u0 = 2C + c
      constexpr int RL = Last<typename PL::R>::value, NBL = E / RL;
      if (tg.colok) {
        const Ix g4u = tg.cbase + (Ix)(4 * t) * tg.axs, g4d = tg.cbase - (Ix)(4 * t) * tg.axs;
        // split addressing: samples 4t (+ 2 u0) and N-1-4t (+ N - 2 u0), both
        // thread bases inside the field (4t < N)
        const Ix tu = g4u, td = tg.cbase + (Ix)(N - 1 - 4 * t) * tg.axs;
        const int ax = (int)tg.axs;
#pragma unroll
        for (int b = 0; b < NBL; ++b) {
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int C = b * TP + r * (M / RL);
            const cx<T> z = v[b * RL + r];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
              const int u0 = 2 * C + c;
              const T y = c ? z.y : z.x;
              if (2 * (TP - 1) + u0 < M) {  // every thread: n = 4t + 2 u0
                if constexpr (NTS) st_at_nt<T>(Sop.d[0], tu, 2 * u0 * ax, y);
                else st_at<T>(Sop.d[0], tu, 2 * u0 * ax, y);
              } else if (u0 >= M) {  // every thread: n = 2N-1-2u0-4t
                if constexpr (NTS) st_at_nt<T>(Sop.d[0], td, (N - 2 * u0) * ax, dst ? -y : y);
                else st_at<T>(Sop.d[0], td, (N - 2 * u0) * ax, dst ? -y : y);
              } else {  // the thread range straddles M
                const Ix iu = g4u + (Ix)(2 * u0) * tg.axs;
                const Ix idn = g4d + (Ix)(2 * N - 1 - 2 * u0) * tg.axs;
                const bool low = 2 * t + u0 < M;
                store_eval<T, ST>(Sop, low ? iu : idn, (!low && dst) ? -y : y);
              }
            }
          }
        }
      }
    }
  }

  // ---- FP64 block sum of the loaded values (deterministic order)
  if constexpr (DSUM) {
    if (!part) return;
    __syncthreads();
    double* red = (double*)buf;
    for (int off = 32; off > 0; off >>= 1) dsum += __shfl_down(dsum, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = dsum;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < (BLOCK + 63) / 64; ++w) s += red[w];
      part[tile] = s;
    }
  }
}

// ---- multi-output forward z pass.  The NO = 3 or 4 values of an operator
// that shares its inputs (the stress components, and the divergence when it
// is live) are transformed in one tile: its TK pencils form MO = 4 groups of
// R = TK/4 rows, group o holding value o of rows row0 .. row0+R-1 (with three
// values, group 3 idles).  The load phase reads every input element once and
// writes the NO values; the transform runs on all TK pencils; the store
// phase writes group o to A.out[o].  dmask bit o: DST along z for value o.
constexpr int kMultiGroups = 4;

template <typename T, int N>
constexpr bool multi_ok() {
  return TileCfg<T, N, kAxZ, 0>::TK % kMultiGroups == 0;
}

template <typename T, int N>
__host__ __device__ inline long multi_tiles(const Geo& g) {
  constexpr int R = TileCfg<T, N, kAxZ, 0>::TK / kMultiGroups;
  return ((long)g.nx * g.ny + R - 1) / R;
}

// One inverse z round of a multi-output tile: field a[F_g] (F_g < 0: zero)
// in pencil group g, DST-III for group DG; each thread keeps the outputs of
// its walk elements, group g at lz[s][BASE + g].  Group GZ (>= 0) takes the
// z gradient of its field instead, -s_z[k] a[k+1] (0 at k = nz - 1; the
// LD_GRAD_Z operator): the field then holds a y-passed scalar whose z
// derivative the round transforms.  Every load of the round is issued
// before the first LDS write, so they are in flight together.
template <typename T, int N, int R, int TKP, int BLOCK, int M, int E, int R1, class RADICES>
struct MultiRound {
  static constexpr int PER = R * N / BLOCK;
  template <int F0, int F1, int F2, int F3, int DG, int BASE, int GZ, class WI, class WK>
  __device__ __forceinline__ static void run(const LoadOp& L, const WI& walk_idx, const WK& gz_of,
                                             T* rbuf, cx<T>* buf, const cx<T>* tp,
                                             const cx<T>* ts, int p, int t, int tid, T (*lz)[6]) {
    constexpr int F[4] = {F0, F1, F2, F3};
    {
      T x[PER][4];
#pragma unroll
      for (int s = 0; s < PER; ++s) {
        const Ix idx = walk_idx(s);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          if constexpr (GZ >= 0) {
            if (g == GZ) {
              // inside the field or its padding (one y-z slab past the end)
              x[s][g] = gz_of(s, ld<T>(L.a[F[GZ]], idx + 1));
              continue;
            }
          }
          x[s][g] = F[g] >= 0 ? ld<T>(L.a[F[g] >= 0 ? F[g] : 0], idx) : (T)0;
        }
      }
      WalkPos<kAxZ, N, R, TKP, BLOCK, 0> wn;  // natural order (DCT-III)
      WalkPos<kAxZ, N, R, TKP, BLOCK, 2> wr;  // reversed (DST-III)
      wn.init(tid);
      wr.init(tid);
#pragma unroll
      for (int s = 0; s < PER; ++s) {
        bool oddn;
        const int pn = wn.pos(tid, s, oddn), pr = wr.pos(tid, s, oddn);
#pragma unroll
        for (int g = 0; g < 4; ++g) rbuf[(g == DG ? pr : pn) + 2 * g * R] = x[s][g];
      }
    }
    __syncthreads();
    cx<T> v[E];
    inv_preprocess<T, N, M, E, TKP, R1>(v, rbuf, tp, p, t);
    run_fft<T, N, M, E, TKP, 1, true>(RADICES{}, v, buf, p, t, ts);
    WalkPos<kAxZ, N, R, TKP, BLOCK, 1> wq;
    wq.init(tid);
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      bool oddn;
      const int lp = wq.pos(tid, s, oddn);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        if (F[g] >= 0) {  // constant after unrolling; the index guard keeps unused groups in bounds
          const T y = rbuf[lp + 2 * g * R];
          lz[s][BASE + g < 6 ? BASE + g : 5] = (g == DG && oddn) ? -y : y;
        }
      }
    }
    __syncthreads();  // the next round or the forward transform rewrites the tile
  }
};

// LZ (stress operators) >= 1: the lap grad psi inputs a[6..8] hold their
// y-passed values (the inverse transforms of lap grad psi stop before z).
// The tile first runs their inverse z pass itself: l_x, l_y, l_z go to
// pencil groups 0-2, one inverse transform runs on all TK pencils (DST-III
// for l_z, by reversed positions and signs as in pass_tile), and each thread
// keeps the real-space values of its walk elements in registers for the
// operator.  LZ = 2 (with the divergence): grad rho as well, r_x (a[13],
// y-passed) in group 3 of that round and r_y (a[14], y-passed) and r_z
// (DST-III) in groups 0-1 of a second.  r_z is formed in the round from
// rho's y-passed field a[15] (the z gradient operator), so it has no
// transform of its own.  Unused groups are zero.  The separate inverse z passes, a read and a write
// of each field, are then not run.
template <typename T, int N, int LD, int LZ, class SM>
__device__ __forceinline__ void pass_tile_multi(const MultiArgs& A, const int tile, SM& sm,
                                                const int tid) {
  using C = TileCfg<T, N, kAxZ, 0>;
  using PL = PlanFor<T, N / 2>;
  constexpr int M = C::M, E = C::E, TP = C::TP, TK = C::TK, TKP = C::TKP, BLOCK = C::BLOCK;
  constexpr int R1 = PL::R1;
  constexpr int MO = kMultiGroups, R = TK / MO, NO = multi_count<LD>();
  static_assert(TK % MO == 0 && NO <= MO, "multi-output tile");
  using TW = TwLayout<M>;
  cx<T>* const buf = sm.buf;
  cx<T>* const ts = sm.ts;
  cx<T>* const tp = sm.tp;
  T* rbuf = (T*)buf;

  static_assert(LZ == 0 || LD == LD_STRESS3 || LD == LD_STRESS3_DIVV, "lap grad psi inputs");
  static_assert(LZ < 2 || LD == LD_STRESS3_DIVV, "grad rho inputs");
  static_assert(SM::NTP == (LZ ? TW::NPRE : TW::NPOST), "LDS layout");
  const cx<T>* __restrict__ tw = (const cx<T>*)A.tw;
  for (int u = tid; u < M; u += BLOCK) ts[u] = tw[TW::STAGE + u];
  if constexpr (LZ != 0)
    for (int u = tid; u < TW::NPRE; u += BLOCK) tp[u] = tw[TW::PRE + u];
  else
    for (int u = tid; u < TW::NPOST; u += BLOCK) tp[u] = tw[TW::POST + u];
  // the walk over the tile's R rows (element s of a thread: row q, sample n)
  TileGeo<kAxZ, N, R, BLOCK> tg;
  tg.init(A.geo, tile, tid);
  constexpr int PER = R * N / BLOCK;
  constexpr bool SPLIT = N % BLOCK == 0;
  constexpr int RZ = SPLIT ? N / BLOCK : 1;
  auto row_ok = [&](int s) { return tg.row0 + s / RZ < tg.nrows; };
  auto walk_idx = [&](int s) {
    int q, n, i, j, k;
    Ix idx;
    bool ok;
    tg.walk(tid, s, q, n, i, j, k, idx, ok);
    if constexpr (SPLIT) idx = tg.wbase + (Ix)(row_ok(s) ? s * BLOCK : (s % RZ) * BLOCK);
    return idx;
  };
  const int p = tid % TK, t = tid / TK;
  // real-space inputs at the walk elements: l_x, l_y, l_z, r_x, r_y, r_z
  T lz[LZ ? PER : 1][6];
  if constexpr (LZ != 0) {
    using IR = MultiRound<T, N, R, TKP, BLOCK, M, E, R1, typename PL::R>;
    // the z gradient operator at walk element s (the expression of
    // LD_GRAD_Z in load_eval) of the value v read at k + 1.  Split walks
    // repeat the same RZ columns k, so their -s_z[k] are read once.
    constexpr int KZ = SPLIT ? RZ : PER;
    double nsz[LZ == 2 ? KZ : 1];
    bool okz[LZ == 2 ? KZ : 1];
    if constexpr (LZ == 2) {
#pragma unroll
      for (int u = 0; u < KZ; ++u) {
        int q, n, i, j, k;
        Ix idx;
        bool ok;
        tg.walk(tid, u, q, n, i, j, k, idx, ok);
        nsz[u] = -A.sz[k];
        okz[u] = k < A.geo.nz - 1;
      }
    }
    auto gz_of = [&](int s, T v) -> T {
      const int u = LZ == 2 ? s % KZ : 0;
      if constexpr (sizeof(T) == sizeof(float)) return okz[u] ? (T)nsz[u] * v : (T)0;
      else return (T)(nsz[u] * (okz[u] ? (double)v : 0.0));
    };
    IR::template run<6, 7, 8, (LZ == 2 ? 13 : -1), 2, 0, -1>(A.ld, walk_idx, gz_of, rbuf, buf, tp, ts,
                                                            p, t, tid, lz);
    if constexpr (LZ == 2) {
      IR::template run<14, 15, -1, -1, 1, 4, 1>(A.ld, walk_idx, gz_of, rbuf, buf, tp, ts, p, t, tid,
                                               lz);
      // |grad rho|^2 for the nonlinear pass's pressure term (the expression
      // of LD_NONLIN_P), so that pass need not transform grad rho again
      if (A.aux) {
#pragma unroll
        for (int s = 0; s < PER; ++s) {
          if (!row_ok(s)) continue;
          const T rx = lz[s][3], ry = lz[s][4], rz = lz[s][5];
          const T r2 = rx * rx + ry * ry + rz * rz;
          if constexpr (SPLIT) st_at<T>(A.aux, tg.wbase, s * BLOCK, r2);
          else st<T>(A.aux, walk_idx(s), r2);
        }
      }
    }
    for (int u = tid; u < TW::NPOST; u += BLOCK) tp[u] = tw[TW::POST + u];
  }
  {
    // each element's NO values go to LDS as soon as they are formed (Makhoul
    // order; component o of row q in pencil o R + q).  Holding every
    // element's values first (with the rounds' held inputs) filled the
    // registers, and the allocator issued the loads in waves: the stress and
    // divergence pass at 512^3 took 2747 us against 2169 this way.
    // The stb stress pass and the stress-only pass are unchanged by it.
    WalkPos<kAxZ, N, R, TKP, BLOCK, 1> wp;
    wp.init(tid);
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      T xs[NO];
      multi_eval<T, LD, LZ>(A.ld, A.phys, walk_idx(s), xs, lz[LZ ? s : 0]);
      bool oddn;
      const int lp = wp.pos(tid, s, oddn);
#pragma unroll
      for (int o = 0; o < NO; ++o) {
        T x = xs[o];
        if (((A.dmask >> o) & 1) & oddn) x = -x;
        rbuf[lp + 2 * o * R] = x;
      }
    }
  }
  __syncthreads();
  const bool dst = (A.dmask >> (p / R)) & 1;  // this thread's pencil group
  cx<T> v[E];
  gather<T, M, E, TKP, R1>(v, buf, p, t);
  run_fft<T, N, M, E, TKP, -1, true>(typename PL::R{}, v, buf, p, t, ts);

  // DCT-II post-processing (as in pass_tile), staged in LDS
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
    const cx<T> wO = cmul(tp[3 * kk], Ov);
    const cx<T> Wk = cmul(tp[3 * kk + 1], Ev + wO);
    const cx<T> Wm = cmul(tp[3 * kk + 2], conj(Ev - wO));
    y0[b] = Wk.x; y1[b] = -Wk.y; y2[b] = Wm.x; y3[b] = -Wm.y;
  }
  PencilPos<TKP> pp;
  pp.init(p, t);
  auto put_k = [&](bool up, int K, T val) {
    const int pa = up ? pp.up(K) : pp.dn(K);
    const int pb = up ? pp.dn(N - 1 - K) : pp.up(N - 1 - K);
    rbuf[dst ? pb : pa] = val;
  };
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NPRE; ++b) {
    const int C0 = b * TP;
    put_k(true, C0, y0[b]);
    if (C0 > 0 || t > 0) put_k(false, N - C0, y1[b]);
    put_k(false, M - C0, y2[b]);
    if (C0 > 0 || t > 0) put_k(true, M + C0, y3[b]);
  }
  if (t == 0) {
    const int n0 = dst ? N - 1 - M / 2 : M / 2, n1 = dst ? M / 2 - 1 : N - M / 2;
    rbuf[lpos<TKP>(n0, p)] = y0[NPRE];
    rbuf[lpos<TKP>(n1, p)] = y1[NPRE];
  }
  __syncthreads();
  WalkPos<kAxZ, N, R, TKP, BLOCK, 0> wq;
  wq.init(tid);
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    bool oddn;
    int q, n, i, j, k;
    Ix idx;
    bool ok;
    tg.walk(tid, s, q, n, i, j, k, idx, ok);
    if constexpr (SPLIT) ok = row_ok(s);
    if (!ok) continue;
    const int lp = wq.pos(tid, s, oddn);
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      if constexpr (SPLIT) st_at<T>(A.out[o], tg.wbase, s * BLOCK, rbuf[lp + 2 * o * R]);
      else st<T>(A.out[o], idx, rbuf[lp + 2 * o * R]);
    }
  }
}

template <typename T, int N, int LD, int LZ>
__global__ void __launch_bounds__((TileCfg<T, N, kAxZ, 0>::BLOCK)) k_multi(const MultiArgs A) {
  // LZ: the tile's inverse rounds need the DCT-III pre-processing twiddles
  // (2M), more than the forward post-processing's 3(M/2+1)
  __shared__ Smem<T, N, kAxZ, (LZ != 0), 0, 0> sm;
  const int nb = (int)gridDim.x;
  pass_tile_multi<T, N, LD, LZ>(A, nb - 1 - (int)blockIdx.x, sm, (int)threadIdx.x);
}

template <typename T, int N>
hipError_t launch_multi_n(const MultiArgs& A, long* nblocks, hipStream_t s) {
  if constexpr (!multi_ok<T, N>()) {
    return hipErrorInvalidValue;
  } else {
    const long nb = multi_tiles<T, N>(A.geo);
    if (nblocks) *nblocks = nb;
    if (nb <= 0) return hipSuccess;
    constexpr int BLOCK = TileCfg<T, N, kAxZ, 0>::BLOCK;
    // the quasi stress takes lap grad psi y-passed (A.lz = 1), and with the
    // divergence grad rho as well (A.lz = 2)
    if (A.ld.code == LD_STRESS3 && A.lz == 1)
      hipLaunchKernelGGL((k_multi<T, N, LD_STRESS3, 1>), dim3((unsigned)nb), dim3(BLOCK), 0, s, A);
    else if (A.ld.code == LD_STRESS_STB3 && A.lz == 0)
      hipLaunchKernelGGL((k_multi<T, N, LD_STRESS_STB3, 0>), dim3((unsigned)nb), dim3(BLOCK), 0, s, A);
    else if (A.ld.code == LD_STRESS3_DIVV && A.lz == 2)
      hipLaunchKernelGGL((k_multi<T, N, LD_STRESS3_DIVV, 2>), dim3((unsigned)nb), dim3(BLOCK), 0, s, A);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2: workgroup b runs on XCD b mod 8.  xcd_tile gives XCD x one contiguous
// range of tiles (any tile count).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int NX = 8;
  const int x = b % NX, s = b / NX;
  const int per = nb / NX, rem = nb % NX;
  return x * per + (x < rem ? x : rem) + s;
}

// Tile order of a single pass, measured at 512^3:
//  * strided passes: the G column blocks of one row (x pass) or slab (y pass)
//    form a unit that runs on one XCD, units going round-robin over the XCDs,
//    so neighbouring columns (shifted-neighbour reads) meet in one L2 while
//    the chip as a whole sweeps rows/slabs in increasing order;
//  * z passes: rows in decreasing order.  The y pass before a z pass (inverse
//    transforms) ends on the high slabs, and the y pass after one (forward
//    transforms) starts on the low slabs, so the z pass begins, and the next
//    y pass begins, on the data written last, still in the 256 MB
//    memory-side cache.
template <typename T, int N, int AX, int V>
__device__ __forceinline__ int tile_order(const Geo& g, int b, int nb) {
  if constexpr (AX == kAxZ) {
    return nb - 1 - b;
  } else {
    constexpr int NX = 8;
    const int G = (g.nz + TileCfg<T, N, AX, V>::TK - 1) / TileCfg<T, N, AX, V>::TK;
    if (nb % (NX * G) != 0) return xcd_tile(b, nb);
    const int x = b % NX, s = b / NX;
    return ((s / G) * NX + x) * G + s % G;
  }
}

template <typename T, int N, int AX, bool INV, int LD, int ST, int V>
__global__ void __launch_bounds__((TileCfg<T, N, AX, V>::BLOCK)) k_pass(const PassArgs A) {
  __shared__ SmemFor<T, N, AX, INV, LD, ST, V> sm;
  pass_tile<T, N, AX, INV, LD, ST, V>(
      A.geo, A.waves, A.phys, A.ld, A.st, A.tw, A.part, A.dst,
      tile_order<T, N, AX, V>(A.geo, (int)blockIdx.x, (int)gridDim.x), sm);
}

// Forward x passes that also run the x inverse of their stored values
// (ST_FILTER_X, ST_UPDATE_X, ST_UPDATE_STB_X): the register allocator is held
// to two workgroups per CU.  Left alone it takes 132 VGPRs in FP32 at 512, one
// 512-thread workgroup per CU instead of the two the LDS allows.  Smaller
// workgroups are left alone (held to 128 VGPRs, the 96-point tile spills).
template <typename T, int N, int V>
constexpr int inv_store_waves() {
  constexpr int w = (TileCfg<T, N, kAxX, V>::BLOCK + 255) / 256;  // waves per SIMD per workgroup
  return 2 * w > 8 ? 8 : 2 * w;
}
template <typename T, int N, int AX, bool INV, int LD, int ST, int V>
__global__ void __launch_bounds__((TileCfg<T, N, AX, V>::BLOCK))
    __attribute__((amdgpu_waves_per_eu(inv_store_waves<T, N, V>()))) k_pass_sx(const PassArgs A) {
  __shared__ SmemFor<T, N, AX, INV, LD, ST, V> sm;
  pass_tile<T, N, AX, INV, LD, ST, V>(
      A.geo, A.waves, A.phys, A.ld, A.st, A.tw, A.part, A.dst,
      tile_order<T, N, AX, V>(A.geo, (int)blockIdx.x, (int)gridDim.x), sm);
}

// ---- software-pipelined plain inverse passes (variants 4 and 5).  Where a tile
// fills a CU's LDS alone (FP64 x passes at N = 512: 256-byte rows of 32
// columns, 135 KB), one workgroup per CU runs its load, transform and store
// phases one after another and the CU's memory traffic stops during the
// transform.  A persistent workgroup instead loads the walk of its next tile
// into registers before transforming the current one (from those registers
// into LDS, pass_tile's FROMH mode), so the next tile's loads are in flight
// while this one is transformed and stored.  Tiles are disjoint, so passes
// may run in place.
// Inverse passes only: a forward pass keeps its DCT-II post-processing
// outputs in registers for the direct strided stores, and with the next
// tile's walk held beside them it spills (FP64 x at 512: 66-94 VGPRs
// spilled, 907-998 us per pass against 571 unpipelined).
template <typename T, int N, int AX, bool INV, int V>
constexpr bool pipe_ok() {
  return INV && AX != kAxZ && (V == 4 || (V == 5 && plan_distinct<T, N / 2, 2>())) &&
         sizeof(SmemFor<T, N, AX, INV, LD_PLAIN, ST_PLAIN, V>) > 81920;
}

// the plain load walk of one tile into held[] (pass_tile's load order)
template <typename T, int N, int AX, int V>
__device__ __forceinline__ void walk_prefetch(const PassArgs& A, int tile, int tid, T* held) {
  using C = TileCfg<T, N, AX, V>;
  TileGeo<AX, N, C::TK, C::BLOCK> tg;
  tg.init(A.geo, tile, tid);
#pragma unroll
  for (int s = 0; s < walk_count<T, N, AX, V>(); ++s)
    held[s] = ld_at<T>(A.ld.a[0], tg.wbase, s * (int)tg.wstep);
}

template <typename T, int N, int AX, bool INV, int V>
__global__ void __launch_bounds__((TileCfg<T, N, AX, V>::BLOCK)) k_pass_pipe(const PassArgs A,
                                                                            const int ntiles) {
  static_assert(pipe_ok<T, N, AX, INV, V>(), "pipelined pass");
  __shared__ SmemFor<T, N, AX, INV, LD_PLAIN, ST_PLAIN, V> sm;
  constexpr int PER = walk_count<T, N, AX, V>();
  const int tid = (int)threadIdx.x, nb = (int)gridDim.x;
  T held[PER];  // the walk of the tile being transformed, then of the next
  int vb = (int)blockIdx.x;  // virtual block: the tile order of a one-shot launch
  if (vb < ntiles) walk_prefetch<T, N, AX, V>(A, tile_order<T, N, AX, V>(A.geo, vb, ntiles), tid, held);
  bool first = true;
  for (; vb < ntiles; vb += nb) {
    const int nvb = vb + nb;
    // issued once this tile's values are in LDS; in flight during its
    // transform and stores (barriers wait on LDS operations only)
    auto next = [&]() {
      if (nvb < ntiles)
        walk_prefetch<T, N, AX, V>(A, tile_order<T, N, AX, V>(A.geo, nvb, ntiles), tid, held);
    };
    pass_tile<T, N, AX, INV, LD_PLAIN, ST_PLAIN, V, decltype(sm), false, -1, true, decltype(next)>(
        A.geo, A.waves, A.phys, A.ld, A.st, A.tw, nullptr, A.dst,
        tile_order<T, N, AX, V>(A.geo, vb, ntiles), sm, first, tid, nullptr, held, next);
    first = false;
    __syncthreads();  // the next tile rewrites LDS
  }
}

template <typename T, int N, int AX, bool INV, int V>
hipError_t launch_pipe(const PassArgs& A, long* nblocks_out, hipStream_t s);

// ---- group passes (the x pass of inverse transforms): the operators of
// each group in output order, the workgroups that share each tile in FP32
// (PARTS) and the FP64 tile variant (VAR64); times per launch at 512^3.
// FP64 psi and rho groups at half width: 3765 -> 3370 and 1630 -> 1530 us
// (the projection group: no gain; FP32 groups at half width: slower)
template <int... LDs>
struct OpList {};
template <int G>
struct GroupOps;
template <>
struct GroupOps<GRP_PSI> {
  // grad psi y and z are not here: their transforms start from psi's x- and
  // y-passed fields (the y or z pass with the gradient operator)
  using L = OpList<LD_COEF_F, LD_COEF_G, LD_GRAD_X, LD_LAPGRAD_X, LD_LAPGRAD_Y, LD_LAPGRAD_Z>;
  static constexpr bool INV = true;
  static constexpr int AX = kAxX;
  static constexpr bool XORD = false;
  static constexpr int PARTS = 2;
  static constexpr int VAR64 = 1;
  static constexpr bool FUSED = false;
};
template <>
struct GroupOps<GRP_PROJ> {
  using L = OpList<LD_PRESSURE, LD_VEL0, LD_VEL1, LD_VEL2>;
  static constexpr bool INV = true;
  static constexpr int AX = kAxX;
  static constexpr bool XORD = true;
  static constexpr int PARTS = 1;  // 2 parts: 1478 -> 1628 us
  static constexpr int VAR64 = 0;
  // v_x and v_z read the same rows of S-hat and D-hat (slab i and i+1 of
  // row j, S_y of row j-1): one load walk forms both, then p, then v_y
  // (FP32 at 512^3: 1691 -> 1542 us)
  static constexpr bool FUSED = true;
  using SEQ = OpList<1, 3, 0, 2>;
};
template <>
struct GroupOps<GRP_STB_PSI> {
  // grad psi y and z start from psi's passed fields, as in the quasi step
  using L = OpList<LD_SH, LD_GRAD_X>;
  static constexpr bool INV = true;
  static constexpr int AX = kAxX;
  static constexpr bool XORD = false;
  static constexpr int PARTS = 1;
  static constexpr int VAR64 = 0;
  static constexpr bool FUSED = false;
};
template <>
struct GroupOps<GRP_VEL> {
  using L = OpList<LD_VEL0, LD_VEL1, LD_VEL2>;
  static constexpr bool INV = true;
  static constexpr int AX = kAxX;
  static constexpr bool XORD = true;
  static constexpr int PARTS = 1;
  static constexpr int VAR64 = 0;
  // v_x and v_z from one load walk, then v_y (FP32 at 512^3: 1186 -> 1125 us)
  static constexpr bool FUSED = true;
  using SEQ = OpList<0, 2, 1>;
};

template <int AX, int... LDs>
constexpr int group_tables(OpList<LDs...>) {
  return (kv_tables<AX, LDs, ST_PLAIN>() | ... | 0);
}
template <int... LDs>
constexpr int group_count(OpList<LDs...>) {
  return (int)sizeof...(LDs);
}

// Outputs LO..HI-1 of a group: output O's pass on this tile, then the next
// one's.  The passes index the same tile, so the compiler would keep every
// per-element address of the first pass live for reuse in the next ones; an
// opaque copy of the thread index per pass makes each pass derive its own,
// so each keeps the register budget of a single pass.
template <typename T, int N, int AX, bool INV, class SM, int O, int LO, int HI, int LD,
          int... REST>
__device__ __forceinline__ void run_group(OpList<LD, REST...>, const GroupArgs& G, int tile,
                                          SM& sm) {
  if constexpr (O >= LO && O < HI) {
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    // non-temporal stores: the outputs are read again only by the next
    // passes, after the whole group, so they should not evict the group's
    // inputs, which every output re-reads
    pass_tile<T, N, AX, INV, LD, ST_PLAIN, SM::VAR, SM, true>(G.geo, G.waves, G.phys, G.out[O].ld,
                                                        G.out[O].st, G.tw, nullptr, G.out[O].dst,
                                                        tile, sm, O == LO, tid);
  }
  if constexpr (sizeof...(REST) > 0 && O + 1 < HI) {
    if constexpr (O >= LO) __syncthreads();  // the next pass rewrites the tile in LDS
    run_group<T, N, AX, INV, SM, O + 1, LO, HI>(OpList<REST...>{}, G, tile, sm);
  }
}

// Load operator of output I of a group.
template <int I, int... LDs>
constexpr int op_at(OpList<LDs...>) {
  constexpr int a[] = {LDs...};
  return a[I];
}

// Outputs of a group in the order SEQ (output indices).  The first two are
// fused: the first output's load walk also evaluates the second's operator at
// every element, keeping those values in registers (walk_count of them per
// thread), so the fields both read are read once; the second output's pass
// then starts from the registers.  The rest run one after another as in
// run_group.
template <typename T, int N, int AX, bool INV, class SM, class L, int... REST>
__device__ __forceinline__ void run_rest(OpList<REST...>, const GroupArgs& G, int tile, SM& sm) {
  (
      [&] {
        __syncthreads();  // the next pass rewrites the tile in LDS
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        pass_tile<T, N, AX, INV, op_at<REST>(L{}), ST_PLAIN, SM::VAR, SM, true>(
            G.geo, G.waves, G.phys, G.out[REST].ld, G.out[REST].st, G.tw, nullptr,
            G.out[REST].dst, tile, sm, false, tid);
      }(),
      ...);
}

template <typename T, int N, int AX, bool INV, class SM, class L, int OA, int OB, int... REST>
__device__ __forceinline__ void run_fused(OpList<OA, OB, REST...>, const GroupArgs& G, int tile,
                                          SM& sm) {
  constexpr int LDA = op_at<OA>(L{}), LDB = op_at<OB>(L{});
  T held[walk_count<T, N, AX, SM::VAR>()];
  {
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    pass_tile<T, N, AX, INV, LDA, ST_PLAIN, SM::VAR, SM, true, LDB, false>(
        G.geo, G.waves, G.phys, G.out[OA].ld, G.out[OA].st, G.tw, nullptr, G.out[OA].dst, tile,
        sm, true, tid, &G.out[OB].ld, held);
  }
  __syncthreads();
  {
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    pass_tile<T, N, AX, INV, LDB, ST_PLAIN, SM::VAR, SM, true, -1, true>(
        G.geo, G.waves, G.phys, G.out[OB].ld, G.out[OB].st, G.tw, nullptr, G.out[OB].dst, tile,
        sm, false, tid, nullptr, held);
  }
  run_rest<T, N, AX, INV, SM, L>(OpList<REST...>{}, G, tile, sm);
}

// FP32 only: in FP64 (1024-thread tiles, at most 128 VGPRs) the held values
// spill
template <typename T, class GO>
constexpr bool group_fused() {
  return sizeof(T) == 4 && GO::FUSED;
}
template <class GO, bool F = GO::FUSED>
struct GroupSeq {
  using S = OpList<>;
};
template <class GO>
struct GroupSeq<GO, true> {
  using S = typename GO::SEQ;
};

// Tile variant of group GRP in precision T: FP64 groups with VAR64 = 1 take
// half-width tiles (128-byte row segments, two workgroups per CU instead of
// one); FP32 groups keep full width.
template <typename T, int GRP>
constexpr int group_variant() {
  return sizeof(T) == 8 ? GroupOps<GRP>::VAR64 : 0;
}

template <typename T, int N, int GRP>
using GroupSmem =
    Smem<T, N, GroupOps<GRP>::AX, GroupOps<GRP>::INV, group_variant<T, GRP>(),
         smem_tables<T, N, GroupOps<GRP>::AX, GroupOps<GRP>::INV, group_variant<T, GRP>(),
                     group_tables<GroupOps<GRP>::AX>(typename GroupOps<GRP>::L{})>()>;

// Waves per SIMD a group kernel asks the register allocator for: those of
// two workgroups (as the single x passes run), capped at what the LDS
// (160 KB per CU) allows, and at least one workgroup's.
template <typename T, int N, int GRP>
constexpr int group_waves() {
  constexpr int AX = GroupOps<GRP>::AX;
  constexpr int w = (TileCfg<T, N, AX, group_variant<T, GRP>()>::BLOCK + 255) / 256;
  constexpr int lds = (int)(163840 / sizeof(GroupSmem<T, N, GRP>)) * w;
  constexpr int r = 2 * w < lds ? 2 * w : lds;
  return r < w ? w : (r > 8 ? 8 : r);
}

// Parts per tile of group GRP in precision T (FP32 only: splitting the FP64
// psi group, at either tile width, measured no gain).
template <typename T, int GRP>
constexpr int group_parts() {
  return sizeof(T) == 4 ? GroupOps<GRP>::PARTS : 1;
}

// Part P of PARTS of a group's outputs on this tile: outputs
// [P CNT / PARTS, (P + 1) CNT / PARTS).
template <typename T, int N, int AX, bool INV, class SM, class L, int CNT, int PARTS, int P>
__device__ __forceinline__ void run_part(int part, const GroupArgs& G, int tile, SM& sm) {
  if constexpr (P < PARTS) {
    if (part == P)
      run_group<T, N, AX, INV, SM, 0, P * CNT / PARTS, (P + 1) * CNT / PARTS>(L{}, G, tile, sm);
    else
      run_part<T, N, AX, INV, SM, L, CNT, PARTS, P + 1>(part, G, tile, sm);
  }
}

// The group runs the first (x) pass of its inverse (x, y, z) transforms.
// A group with PARTS > 1 runs each tile as PARTS workgroups, each on its
// share of the outputs: workgroups b, b + 8, b + 16, ... (one XCD,
// dispatched one after the other) take the parts of one tile, so they read
// its inputs at about the same time through one L2, and each re-reads them
// for fewer outputs.
template <typename T, int N, int GRP>
__global__ void __launch_bounds__((TileCfg<T, N, GroupOps<GRP>::AX, group_variant<T, GRP>()>::BLOCK))
    __attribute__((amdgpu_waves_per_eu(group_waves<T, N, GRP>())))
    k_group(const GroupArgs G) {
  using GO = GroupOps<GRP>;
  constexpr int AX = GO::AX;
  constexpr int CNT = group_count(typename GO::L{});
  constexpr int PARTS = group_parts<T, GRP>();
  static_assert(PARTS >= 1 && PARTS <= CNT, "every part runs at least one output");
  __shared__ GroupSmem<T, N, GRP> sm;
  int b = (int)blockIdx.x, nb = (int)gridDim.x, part = 0;
  if constexpr (PARTS > 1) {
    nb /= PARTS;
    if (nb % 8 == 0) {
      const int x = b % 8, s = b / 8;
      part = s % PARTS;
      b = (s / PARTS) * 8 + x;
    } else {
      part = b % PARTS;
      b /= PARTS;
    }
  }
  // projection groups read the j-1 and j+1 neighbours of their rows: one
  // contiguous range of rows per XCD keeps those in the XCD's L2
  const int tile = GO::XORD ? xcd_tile(b, nb)
                            : tile_order<T, N, AX, group_variant<T, GRP>()>(G.geo, b, nb);
  if constexpr (group_fused<T, GO>()) {
    static_assert(PARTS == 1, "a fused group runs each tile in one workgroup");
    run_fused<T, N, AX, GO::INV, decltype(sm), typename GO::L>(typename GroupSeq<GO>::S{}, G, tile,
                                                                sm);
  } else {
    run_part<T, N, AX, GO::INV, decltype(sm), typename GO::L, CNT, PARTS, 0>(part, G, tile, sm);
  }
}

// Batched pass over the instances of an ensemble that share this pass's
// length N: workgroup b runs tile b - prefix[q] of instance q, where
// prefix[q] <= b < prefix[q+1] (prefix[0] = 0, prefix[count] = grid size).
template <typename T, int N, int AX, bool INV, int LD, int ST, int V>
__global__ void __launch_bounds__((TileCfg<T, N, AX, V>::BLOCK))
    k_pass_batch(const PassArgs* __restrict__ descs, const int* __restrict__ prefix, int count) {
  const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  int lo = 0, hi = count;  // invariant: prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid;
  }
  __shared__ SmemFor<T, N, AX, INV, LD, ST, V> sm;
  const PassArgs& A = descs[lo];
  pass_tile<T, N, AX, INV, LD, ST, V>(A.geo, A.waves, A.phys, A.ld, A.st, A.tw, A.part, A.dst,
                                      b - prefix[lo], sm);
}

// the batched form of k_pass_sx (the same register bound)
template <typename T, int N, int AX, bool INV, int LD, int ST, int V>
__global__ void __launch_bounds__((TileCfg<T, N, AX, V>::BLOCK))
    __attribute__((amdgpu_waves_per_eu(inv_store_waves<T, N, V>()))) k_pass_batch_sx(const PassArgs* __restrict__ descs, const int* __restrict__ prefix, int count) {
  const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  int lo = 0, hi = count;  // invariant: prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid;
  }
  __shared__ SmemFor<T, N, AX, INV, LD, ST, V> sm;
  const PassArgs& A = descs[lo];
  pass_tile<T, N, AX, INV, LD, ST, V>(A.geo, A.waves, A.phys, A.ld, A.st, A.tw, A.part, A.dst,
                                      b - prefix[lo], sm);
}

template <typename T, int N, int AX, int V = 0>
long tiles_of(const Geo& g) {
  using C = TileCfg<T, N, AX, V>;
  if (AX == kAxZ) return ((long)g.nx * g.ny + C::TK - 1) / C::TK;
  const long tiles_k = (g.nz + C::TK - 1) / C::TK;
  return (AX == kAxY ? (long)g.nx : (long)g.ny) * tiles_k;
}

// group kernels of the groups that run on axis AX
template <typename T, int N, int AX, int GG>
hipError_t launch_group_g(const GroupArgs& G, long* nblocks, hipStream_t s) {
  if constexpr (GroupOps<GG>::AX != AX) {
    return hipErrorInvalidValue;
  } else {
    constexpr int GV = group_variant<T, GG>();
    using C = TileCfg<T, N, AX, GV>;
    const long nb = tiles_of<T, N, AX, GV>(G.geo) * group_parts<T, GG>();
    if (nblocks) *nblocks = nb;
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_group<T, N, GG>), dim3((unsigned)nb), dim3(C::BLOCK), 0, s, G);
    return hipGetLastError();
  }
}

template <typename T, int N, int AX>
hipError_t launch_group_n(int group, const GroupArgs& G, long* nblocks, hipStream_t s) {
  switch (group) {
    case GRP_PSI: return launch_group_g<T, N, AX, GRP_PSI>(G, nblocks, s);
    case GRP_PROJ: return launch_group_g<T, N, AX, GRP_PROJ>(G, nblocks, s);
    case GRP_STB_PSI: return launch_group_g<T, N, AX, GRP_STB_PSI>(G, nblocks, s);
    case GRP_VEL: return launch_group_g<T, N, AX, GRP_VEL>(G, nblocks, s);
    default: return hipErrorInvalidValue;
  }
}

// Single pass (B == nullptr) or batched pass (B: device descriptors).
template <typename T, int N, int AX, bool INV, int LD, int ST, int V = 0>
hipError_t launch_one(const PassArgs& A, long* nblocks_out, hipStream_t s, const BatchLaunch* B) {
  using C = TileCfg<T, N, AX, V>;
  if (B) {
    if (B->total_tiles <= 0) return hipSuccess;
    if constexpr (st_inverts(ST))
      hipLaunchKernelGGL((k_pass_batch_sx<T, N, AX, INV, LD, ST, V>), dim3((unsigned)B->total_tiles),
                         dim3(C::BLOCK), 0, s, B->descs, B->prefix, B->count);
    else
      hipLaunchKernelGGL((k_pass_batch<T, N, AX, INV, LD, ST, V>), dim3((unsigned)B->total_tiles),
                         dim3(C::BLOCK), 0, s, B->descs, B->prefix, B->count);
    return hipGetLastError();
  }
  const long nb = tiles_of<T, N, AX, V>(A.geo);
  if (nblocks_out) *nblocks_out = nb;
  if (nb <= 0) return hipSuccess;
  if constexpr (st_inverts(ST))
    hipLaunchKernelGGL((k_pass_sx<T, N, AX, INV, LD, ST, V>), dim3((unsigned)nb), dim3(C::BLOCK), 0, s, A);
  else
    hipLaunchKernelGGL((k_pass<T, N, AX, INV, LD, ST, V>), dim3((unsigned)nb), dim3(C::BLOCK), 0, s, A);
  return hipGetLastError();
}

// Persistent grid: as many workgroups as the GPU holds at once (at most the
// tile count), a multiple of the 8 XCDs so each keeps its tile order's XCD.
template <typename T, int N, int AX, bool INV, int V>
hipError_t launch_pipe(const PassArgs& A, long* nblocks_out, hipStream_t s) {
  using C = TileCfg<T, N, AX, V>;
  const long nt = tiles_of<T, N, AX, V>(A.geo);
  if (nblocks_out) *nblocks_out = nt;
  if (nt <= 0) return hipSuccess;
  static const int resident = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pass_pipe<T, N, AX, INV, V>, C::BLOCK, 0) !=
            hipSuccess)
      return 0;
    return cus * (per > 0 ? per : 1);
  }();
  if (resident <= 0) return hipErrorInvalidValue;
  long nb = std::min<long>(nt, resident);
  if (nb > 8) nb -= nb % 8;
  hipLaunchKernelGGL((k_pass_pipe<T, N, AX, INV, V>), dim3((unsigned)nb), dim3(C::BLOCK), 0, s, A,
                     (int)nt);
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
    if (variant == 2 || variant == 3) {
      // alternative radix plans, where the length has them
      if constexpr (plan_distinct<T, N / 2, 2>()) {
        if (variant == 2)
          return inv ? launch_one<T, N, AX, true, LD_PLAIN, ST_PLAIN, 2>(A, nb, s, B)
                     : launch_one<T, N, AX, false, LD_PLAIN, ST_PLAIN, 2>(A, nb, s, B);
      }
      if constexpr (plan_distinct<T, N / 2, 3>()) {
        if (variant == 3)
          return inv ? launch_one<T, N, AX, true, LD_PLAIN, ST_PLAIN, 3>(A, nb, s, B)
                     : launch_one<T, N, AX, false, LD_PLAIN, ST_PLAIN, 3>(A, nb, s, B);
      }
      return hipErrorInvalidValue;
    }
    if (variant == 4 || variant == 5) {
      // pipelined forms of variants 0 and 2, single passes only
      if (B || !inv) return hipErrorInvalidValue;
      if constexpr (pipe_ok<T, N, AX, true, 4>()) {
        if (variant == 4) return launch_pipe<T, N, AX, true, 4>(A, nb, s);
      }
      if constexpr (pipe_ok<T, N, AX, true, 5>()) {
        if (variant == 5) return launch_pipe<T, N, AX, true, 5>(A, nb, s);
      }
      return hipErrorInvalidValue;
    }
    return inv ? launch_one<T, N, AX, true, LD_PLAIN, ST_PLAIN>(A, nb, s, B)
               : launch_one<T, N, AX, false, LD_PLAIN, ST_PLAIN>(A, nb, s, B);
  }
  if constexpr (STEP && AX == kAxZ) {
    if (!inv && st == ST_PLAIN) {
      switch (ld) {
#define QSM_L(OP) case OP: return launch_one<T, N, AX, false, OP, ST_PLAIN>(A, nb, s, B);
        QSM_L(LD_ZERO) QSM_L(LD_RHOSTAR) QSM_L(LD_DIVV) QSM_L(LD_STRESS) QSM_L(LD_NONLIN)
        QSM_L(LD_NONLIN_P) QSM_L(LD_NONLIN_PY) QSM_L(LD_STRESS_STB) QSM_L(LD_NONLIN_STB)
        QSM_L(LD_NONLIN_STB0)
#undef QSM_L
        default: break;
      }
    }
  }
  if constexpr (STEP && AX != kAxX) {
    // the gradient along the pass axis of a field whose earlier axes are
    // already transformed (grad psi y from psi's x-passed field, z from its
    // y-passed one)
    constexpr int GLD = AX == kAxY ? LD_GRAD_Y : LD_GRAD_Z;
    if (inv && st == ST_PLAIN && ld == GLD) return launch_one<T, N, AX, true, GLD, ST_PLAIN>(A, nb, s, B);
  }
  if constexpr (STEP && AX == kAxX) {
    if (!inv && ld == LD_PLAIN) {
      switch (st) {
#define QSM_S(OP) case OP: return launch_one<T, N, AX, false, LD_PLAIN, OP>(A, nb, s, B);
        QSM_S(ST_FILTER) QSM_S(ST_UPDATE) QSM_S(ST_UPDATE_STB) QSM_S(ST_FILTER_X) QSM_S(ST_UPDATE_X)
        QSM_S(ST_UPDATE_STB_X)
#undef QSM_S
        default: break;
      }
    }
    if (inv && st == ST_PLAIN) {
      switch (ld) {
#define QSM_L(OP) case OP: return launch_one<T, N, AX, true, OP, ST_PLAIN>(A, nb, s, B);
        QSM_L(LD_COEF_F) QSM_L(LD_COEF_G) QSM_L(LD_GRAD_X) QSM_L(LD_GRAD_Y) QSM_L(LD_GRAD_Z)
        QSM_L(LD_LAPGRAD_X) QSM_L(LD_LAPGRAD_Y) QSM_L(LD_LAPGRAD_Z) QSM_L(LD_SH) QSM_L(LD_PRESSURE) QSM_L(LD_VEL0) QSM_L(LD_VEL1)
        QSM_L(LD_VEL2) QSM_L(LD_SCALE)
#undef QSM_L
        default: break;
      }
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace qsm
