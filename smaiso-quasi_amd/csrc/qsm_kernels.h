// qsm_kernels.h -- host-visible interface of the device kernels (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "qsm_ops.hpp"

// Pencil lengths with a compiled pass kernel.  Each is N = 2M with M a
// product of the radices 2, 3, 4, 8, 16 (qsm_fft.hpp).
#define QSM_FOR_EACH_N(X) \
  X(8) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(1024)
// Lengths that also carry the step's fused operators.
#define QSM_IS_STEP_N(n)                                                                      \
  ((n) == 32 || (n) == 48 || (n) == 64 || (n) == 96 || (n) == 128 || (n) == 256 || (n) == 512 || \
   (n) == 1024)

namespace qsm {

struct PassArgs {
  Geo geo;
  Waves waves;
  Phys phys;
  LoadOp ld;
  StoreOp st;
  const void* tw;   // e^{-2 pi i t/(4N)} table, t < 4N, for this axis length (field precision)
  double* part;     // optional per-block FP64 sums of the loaded values
  int dst;          // 1: RODFT10/RODFT01 along this axis, 0: REDFT10/REDFT01
  int pad_;
};

// A batched pass: `count` instances with device descriptors and the
// exclusive prefix sum of their tile counts (count + 1 entries).
struct BatchLaunch {
  const PassArgs* descs;
  const int* prefix;
  int count;
  long total_tiles;
};

// One pass of a 3D r2r transform along `axis` (0 = x, 1 = y, 2 = z) of length n.
// inv = REDFT01/RODFT01 (DCT-III/DST-III); otherwise REDFT10/RODFT10.
// ld / st: the LoadCode / StoreCode of the fused operators (A.ld.code and
// A.st.code are ignored); combinations that are not compiled return
// hipErrorInvalidValue.
// B != nullptr launches the batched kernel over B's instances instead; A is
// then ignored (every instance's arguments come from B->descs).
// variant: tile variant of plain passes (0 default, 1 half width, 2 and 3
// alternative radix plans, 4 and 5 pipelined); ignored for fused operators.
hipError_t launch_pass(int prec_fp32, int n, int axis, bool inv, int ld, int st, const PassArgs& A,
                       long* nblocks, hipStream_t s, const BatchLaunch* B = nullptr,
                       int variant = 0);
long pass_tiles(int prec_fp32, int n, int axis, const Geo& g, int variant = 0);
bool pass_supported(int n);
bool step_supported(int n);
int pass_tile_k(int prec_fp32, int n, int axis, int variant = 0);
// variant of the plain passes of length n along axis in direction inv: 0
// default, 1 half-width tiles, 2 and 3 the alternative radix plans (only
// where the length has them), 4 and 5 the software-pipelined forms of 0 and
// 2 (only single, not batched, inverse passes whose tile fills a CU's LDS
// alone)
bool pass_variant_ok(int prec_fp32, int n, int variant, int axis, bool inv);
constexpr int kVariants = 6;
constexpr int kBatchVariants = 4;  // variants a batched (ensemble) pass has
long pass_blocks(int prec_fp32, int n, int axis, const Geo& g);

// Group passes.  One kernel runs the first (x) pass of several inverse
// transforms that read the same fields: each workgroup runs them one after
// another on its tile, so later outputs re-read the shared inputs from the
// caches.
enum GroupCode : int {
  GRP_PSI = 0,     // F, G, grad psi x, lap grad psi (x, y, z) from psihat
  GRP_PROJ,        // p, v_x, v_y, v_z from S-hat (x, y, z) and D-hat
  GRP_STB_PSI,     // sh, grad psi x
  GRP_VEL,         // v_x, v_y, v_z (stb, and quasi steps whose p is read by nothing)
  GRP_COUNT
};
constexpr int kMaxGroup = 8;
struct GroupOut {
  LoadOp ld;   // the output's load operator (its code must match the group's list)
  StoreOp st;  // plain store into the output field
  int dst;     // 1: DST along the group's axis for this output
  int pad_;
};
struct GroupArgs {
  Geo geo;
  Waves waves;
  Phys phys;
  const void* tw;  // twiddle table of the group's axis
  GroupOut out[kMaxGroup];
};
// Multi-output forward z pass (the three stress components, and the
// divergence, from one read of their shared inputs): ld is a multi-output
// operator (LD_STRESS3, LD_STRESS_STB3, LD_STRESS3_DIVV), out[c] receives
// value c, dmask bit c selects DST along z for value c.  The quasi stress
// operators take lap grad psi y-passed (lz = 1), and with the divergence
// grad rho too (lz = 2: r_x, r_y y-passed in a[13], a[14], and rho's
// y-passed field in a[15], from which the pass forms r_z): the pass
// completes their inverse transforms itself.
struct MultiArgs {
  Geo geo;
  Phys phys;
  LoadOp ld;
  void* out[4];
  const void* tw;  // z-axis twiddle table
  int dmask;
  int lz;          // 1: a[6..8] (lap grad psi) are y-passed, 2: a[13..15] (grad rho) too;
                   // the pass runs their inverse z rounds
  void* aux;       // lz = 2: |grad rho|^2 of each element (null: not formed)
  const double* sz;  // lz = 2: the s_z wavenumbers (device)
};
hipError_t launch_multi(int prec_fp32, int n, const MultiArgs& A, long* nblocks, hipStream_t s);
bool multi_supported(int prec_fp32, int n);

// operators of a group in output order, their count, and the direction
int group_ops(int group, int* codes);
bool group_inverse(int group);
int group_axis(int group);  // 0 (x)
hipError_t launch_group(int prec_fp32, int n, int group, const GroupArgs& G, long* nblocks,
                        hipStream_t s);

// FP64 compensated reductions
hipError_t launch_diag(int prec_fp32, const void* psi, const void* pold, const void* rho,
                       const void* F, long n, const Phys& P, double* part, int nblocks,
                       double* out4, hipStream_t s);
hipError_t launch_finish(const double* part, long nparts, double* out, hipStream_t s);
// dst[i] = (float)src[i], i < n (the FP32 copy of psihat in mixed precision)
hipError_t launch_round32(const double* src, float* dst, long n, hipStream_t s);
// stb dtpsi = -(sh - beta psi^3 + gamma psi^5) [- v.g if hydro == 1]
// (quasi-stb.cpp:1904-1913) over n elements; in: sh, psi, vx, vy, vz, gx, gy, gz
hipError_t launch_dtpsi(int prec_fp32, const void* const* in, long n, const Phys& P, void* out,
                        hipStream_t s);
// count independent compensated sums: out[q][0] = sum of parts[q][0 .. n[q])
hipError_t launch_finish_batch(const double* const* parts, const long* n, int count,
                               double* const* out, hipStream_t s);

}  // namespace qsm
