// qsm_pass.hip -- dispatch of the pass kernels (compiled per precision and
// axis in qsm_pass_inst.hip) and the FP64 compensated reduction kernels.
#include <hip/hip_runtime.h>

#include "qsm_pass_impl.hpp"

namespace qsm {

#define QSM_DECL(T, AX) \
  hipError_t launch_pass_##T##_##AX(int n, bool inv, int ld, int st, int variant,          \
                                    const PassArgs& A, long* nb, hipStream_t s,                 \
                                    const BatchLaunch* B);
QSM_DECL(float, 0)
QSM_DECL(float, 1)
QSM_DECL(float, 2)
QSM_DECL(double, 0)
QSM_DECL(double, 1)
QSM_DECL(double, 2)
#undef QSM_DECL

hipError_t launch_group_float_0(int n, int group, const GroupArgs& G, long* nb, hipStream_t s);
hipError_t launch_group_double_0(int n, int group, const GroupArgs& G, long* nb, hipStream_t s);

hipError_t launch_group(int prec_fp32, int n, int group, const GroupArgs& G, long* nblocks,
                        hipStream_t s) {
  if (group_axis(group) != kAxX) return hipErrorInvalidValue;
  return prec_fp32 ? launch_group_float_0(n, group, G, nblocks, s)
                   : launch_group_double_0(n, group, G, nblocks, s);
}

template <int... LDs>
static int list_codes(OpList<LDs...>, int* codes) {
  int k = 0;
  ((codes ? (void)(codes[k] = LDs) : (void)0, ++k), ...);
  return k;
}

int group_ops(int group, int* codes) {
  switch (group) {
    case GRP_PSI: return list_codes(GroupOps<GRP_PSI>::L{}, codes);
    case GRP_PROJ: return list_codes(GroupOps<GRP_PROJ>::L{}, codes);
    case GRP_STB_PSI: return list_codes(GroupOps<GRP_STB_PSI>::L{}, codes);
    case GRP_VEL: return list_codes(GroupOps<GRP_VEL>::L{}, codes);
    default: return 0;
  }
}

bool group_inverse(int group) {
  switch (group) {
#define QSM_G(GG) case GG: return GroupOps<GG>::INV;
    QSM_G(GRP_PSI) QSM_G(GRP_PROJ) QSM_G(GRP_STB_PSI) QSM_G(GRP_VEL)
#undef QSM_G
    default: return false;
  }
}

hipError_t launch_multi_float_n(int n, const MultiArgs& A, long* nb, hipStream_t s);
hipError_t launch_multi_double_n(int n, const MultiArgs& A, long* nb, hipStream_t s);
bool multi_ok_float_n(int n);
bool multi_ok_double_n(int n);

hipError_t launch_multi(int prec_fp32, int n, const MultiArgs& A, long* nblocks, hipStream_t s) {
  return prec_fp32 ? launch_multi_float_n(n, A, nblocks, s) : launch_multi_double_n(n, A, nblocks, s);
}

bool multi_supported(int prec_fp32, int n) {
  return prec_fp32 ? multi_ok_float_n(n) : multi_ok_double_n(n);
}

int group_axis(int group) {
  switch (group) {
#define QSM_G(GG) case GG: return GroupOps<GG>::AX;
    QSM_G(GRP_PSI) QSM_G(GRP_PROJ) QSM_G(GRP_STB_PSI) QSM_G(GRP_VEL)
#undef QSM_G
    default: return -1;
  }
}

hipError_t launch_pass(int prec_fp32, int n, int axis, bool inv, int ld, int st, const PassArgs& A,
                       long* nb, hipStream_t s, const BatchLaunch* B, int variant) {
  const int v = (ld == LD_PLAIN && st == ST_PLAIN) ? variant : 0;
  if (prec_fp32) {
    if (axis == 0) return launch_pass_float_0(n, inv, ld, st, v, A, nb, s, B);
    if (axis == 1) return launch_pass_float_1(n, inv, ld, st, v, A, nb, s, B);
    return launch_pass_float_2(n, inv, ld, st, v, A, nb, s, B);
  }
  if (axis == 0) return launch_pass_double_0(n, inv, ld, st, v, A, nb, s, B);
  if (axis == 1) return launch_pass_double_1(n, inv, ld, st, v, A, nb, s, B);
  return launch_pass_double_2(n, inv, ld, st, v, A, nb, s, B);
}

long pass_tiles(int prec_fp32, int n, int axis, const Geo& g, int variant) {
  const int tk = pass_tile_k(prec_fp32, n, axis, variant);
  if (tk <= 0) return -1;
  if (axis == kAxZ) return ((long)g.nx * g.ny + tk - 1) / tk;
  return (axis == kAxY ? (long)g.nx : (long)g.ny) * ((g.nz + tk - 1) / tk);
}

bool step_supported(int n) { return QSM_IS_STEP_N(n); }

bool pass_supported(int n) {
  switch (n) {
#define QSM_CASE(NN) case NN: return true;
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return false;
  }
}

template <typename T, int V>
static int tile_k_t(int n, int axis) {
  switch (n) {
#define QSM_CASE(NN)                                                  \
  case NN:                                                            \
    return axis == kAxZ ? TileCfg<T, NN, kAxZ, V>::TK                 \
                        : (axis == kAxY ? TileCfg<T, NN, kAxY, V>::TK : TileCfg<T, NN, kAxX, V>::TK);
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return -1;
  }
}

int pass_tile_k(int prec_fp32, int n, int axis, int variant) {
  // variants 2 and 3 change the radix plan, not the tile width
  if (variant == 1) return prec_fp32 ? tile_k_t<float, 1>(n, axis) : tile_k_t<double, 1>(n, axis);
  return prec_fp32 ? tile_k_t<float, 0>(n, axis) : tile_k_t<double, 0>(n, axis);
}

template <typename T, int NN, int AX>
static bool variant_ok_n(int variant, bool inv) {
  switch (variant) {
    case 0: case 1: return true;
    case 2: return plan_distinct<T, NN / 2, 2>();
    case 3: return plan_distinct<T, NN / 2, 3>();
    case 4: return inv && pipe_ok<T, NN, AX, true, 4>();
    case 5: return inv && pipe_ok<T, NN, AX, true, 5>();
    default: return false;
  }
}

template <typename T>
static bool variant_ok_t(int n, int variant, int axis, bool inv) {
  switch (n) {
#define QSM_CASE(NN)                                                                  \
  case NN:                                                                            \
    return axis == 0 ? variant_ok_n<T, NN, 0>(variant, inv)                           \
                     : (axis == 1 ? variant_ok_n<T, NN, 1>(variant, inv)              \
                                  : variant_ok_n<T, NN, 2>(variant, inv));
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return false;
  }
}

bool pass_variant_ok(int prec_fp32, int n, int variant, int axis, bool inv) {
  if (axis < 0 || axis > 2) return false;
  return prec_fp32 ? variant_ok_t<float>(n, variant, axis, inv)
                   : variant_ok_t<double>(n, variant, axis, inv);
}

long pass_blocks(int prec_fp32, int n, int axis, const Geo& g) {
  const int tk = pass_tile_k(prec_fp32, n, axis);
  if (tk <= 0) return -1;
  if (axis == kAxZ) return ((long)g.nx * g.ny + tk - 1) / tk;
  return (axis == kAxY ? (long)g.nx : (long)g.ny) * ((g.nz + tk - 1) / tk);
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

// One block per sum: out[q][0] = compensated sum of parts[q][0 .. n[q]).
__global__ void __launch_bounds__(256)
    k_finish_batch(const double* const* __restrict__ parts, const long* __restrict__ n,
                   double* const* __restrict__ out) {
  __shared__ KSum red[256];
  const int q = blockIdx.x;
  const double* p = parts[q];
  KSum s{0, 0};
  for (long b = threadIdx.x; b < n[q]; b += blockDim.x) s.add(p[b]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x].merge(red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q][0] = red[0].s + red[0].c;
}

hipError_t launch_finish_batch(const double* const* parts, const long* n, int count,
                               double* const* out, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finish_batch, dim3(count), dim3(256), 0, s, parts, n, out);
  return hipGetLastError();
}

hipError_t launch_finish(const double* part, long nparts, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, part, nparts, 1, 1, out);
  return hipGetLastError();
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_dtpsi(const T* __restrict__ sh, const T* __restrict__ psi, const T* __restrict__ vx,
            const T* __restrict__ vy, const T* __restrict__ vz, const T* __restrict__ gx,
            const T* __restrict__ gy, const T* __restrict__ gz, long n, Phys P, T* __restrict__ out) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ps = psi[i];
    const double mu = (double)sh[i] - P.beta * (ps * ps * ps) + P.gamma * (ps * ps * ps * ps * ps);
    double d = -1.0 * mu;  // rho = 1 in the pure smectic
    if (P.hydro == 1)
      d -= (double)vx[i] * gx[i] + (double)vy[i] * gy[i] + (double)vz[i] * gz[i];
    out[i] = (T)d;
  }
}

hipError_t launch_dtpsi(int prec_fp32, const void* const* in, long n, const Phys& P, void* out,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long nb = (n + 255) / 256;
  if (nb > 8192) nb = 8192;
  if (prec_fp32)
    hipLaunchKernelGGL(k_dtpsi<float>, dim3((unsigned)nb), dim3(256), 0, s, (const float*)in[0],
                       (const float*)in[1], (const float*)in[2], (const float*)in[3],
                       (const float*)in[4], (const float*)in[5], (const float*)in[6],
                       (const float*)in[7], n, P, (float*)out);
  else
    hipLaunchKernelGGL(k_dtpsi<double>, dim3((unsigned)nb), dim3(256), 0, s, (const double*)in[0],
                       (const double*)in[1], (const double*)in[2], (const double*)in[3],
                       (const double*)in[4], (const double*)in[5], (const double*)in[6],
                       (const double*)in[7], n, P, (double*)out);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_round32(const double* __restrict__ src,
                                                 float* __restrict__ dst, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = (float)src[i];
}

hipError_t launch_round32(const double* src, float* dst, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long nb = (n + 255) / 256;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(k_round32, dim3((unsigned)nb), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

}  // namespace qsm
