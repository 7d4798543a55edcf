// qsm_api.hip -- host side of libqsm: the C ABI of include/qsm.h.
//
// Owns device memory, wavenumber and twiddle tables, and sequences the pass
// kernels of one time step (quasi/quasi.cpp:1017-1965, quasi-stb.cpp:986-1968).
// Every reference loop between two fftw_execute calls is fused into the
// first or last pass of the adjacent transform (qsm_ops.hpp), so a step is
// 3 passes per 3D transform and nothing else (plus FP64 sums in mixed mode).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/qsm.h"
#include "qsm_kernels.h"

using namespace qsm;

namespace {

thread_local std::string g_err;

// n elements of esz bytes with pad elements of (unused) padding on either
// side; returns the first element (null on failure)
void* alloc_padded(size_t n, size_t esz, size_t pad) {
  void* d = nullptr;
  if (hipMalloc(&d, (n + 2 * pad) * esz) != hipSuccess) return nullptr;
  return (char*)d + pad * esz;
}
void free_padded(void* p, size_t esz, size_t pad) {
  if (p) hipFree((char*)p - pad * esz);
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(QSM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                                  \
  } while (0)

// ---- twiddle tables per (device, N, precision), in the layout the pass
// kernel stages (stage twiddles, DCT-III pre- and DCT-II post-processing
// twiddles; qsm_fft.hpp).  Every entry is a quarter-turn value
// e^{-2 pi i t/(4N)}, t < N, computed in long double and rounded once, moved
// to its quadrant by exact sign flips and swaps.
std::mutex g_tw_mu;
std::map<std::tuple<int, int, int>, void*> g_tw;

const void* twiddles(int n, int fp32) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_tuple(dev, n, fp32);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  std::vector<double> qre(n), qim(n);
  for (int t = 0; t < n; ++t) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * t / (4.0L * n);
    qre[t] = (double)cosl(a);
    qim[t] = (double)sinl(a);
  }
  // e^{-2 pi i u/(4N)} = (-i)^(u div N) e^{-2 pi i (u mod N)/(4N)}, u < 4N
  auto twq = [&](int u, double& re, double& im) {
    const int q = u / n, r = u - q * n;
    const double x = qre[r], y = qim[r];
    const double a = (q & 1) ? y : x, b = (q & 1) ? x : y;
    re = (q & 2) ? -a : a;
    im = ((q ^ (q >> 1)) & 1) ? -b : b;
  };
  const int m = n / 2;
  std::vector<double> h;  // interleaved re, im
  auto put = [&](double re, double im) { h.push_back(re); h.push_back(im); };
  double re, im;
  for (int v = 0; v < m; ++v) { twq(8 * v, re, im); put(re, im); }  // e^{-2 pi i v/M}
  for (int k = 0; k < m; ++k) {  // conj twq(k) = e^{i pi k/2N}, conj twq(5k+N) = i e^{2 pi i k/N} e^{i pi k/2N}
    twq(k, re, im); put(re, -im);
    twq(5 * k + n, re, im); put(re, -im);
  }
  for (int k = 0; k <= m / 2; ++k) {                                 // twq(4k), twq(k), twq(M-k)
    twq(4 * k, re, im); put(re, im);
    twq(k, re, im); put(re, im);
    twq(m - k, re, im); put(re, im);
  }
  void* d = nullptr;
  const size_t el = fp32 ? sizeof(float) : sizeof(double);
  if (hipMalloc(&d, el * h.size()) != hipSuccess) return nullptr;
  hipError_t e;
  if (fp32) {
    std::vector<float> f(h.begin(), h.end());
    e = hipMemcpy(d, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice);
  } else {
    e = hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) { hipFree(d); return nullptr; }
  g_tw[key] = d;
  return d;
}

bool is_kind(int k) {
  return k == QSM_REDFT01 || k == QSM_REDFT10 || k == QSM_RODFT01 || k == QSM_RODFT10;
}
bool kind_inv(int k) { return k == QSM_REDFT01 || k == QSM_RODFT01; }
bool kind_dst(int k) { return k == QSM_RODFT01 || k == QSM_RODFT10; }

// ---- per-launch timing with HIP events.  Categories: one per kernel symbol
// k_pass<T, N, axis, inv, LD, ST, V> (key = (((axis*2 + inv)*64 + LD)*16 + ST)*2 + V).
struct Timer {
  bool on = false;
  int only = -1;  // >= 0: record only the launches of this kernel key
  bool want(int key) const { return on && (only < 0 || only == key); }
  struct Rec { hipEvent_t a, b; int cat; double bytes; };
  struct Acc { double ms = 0, bytes = 0; long long count = 0; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  std::map<int, Acc> acc;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  void drain() {
    for (auto& r : recs) {
      float t = 0;
      (void)hipEventSynchronize(r.b);
      (void)hipEventElapsedTime(&t, r.a, r.b);
      Acc& a = acc[r.cat];
      a.ms += t; a.bytes += r.bytes; a.count += 1;
      pool.push_back(r.a); pool.push_back(r.b);
    }
    recs.clear();
  }
  void reset() {
    drain();
    acc.clear();
  }
  ~Timer() {
    for (auto& r : recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// timer categories of the group kernels: load code kGroupCode0 + group
constexpr int kGroupCode0 = 40;

int timer_key(int axis, bool inv, int ld, int st, int v) {
  return (((axis * 2 + (inv ? 1 : 0)) * 64 + ld) * 16 + st) * kVariants + v;
}

}  // namespace

// ============================================================================
// transform level

struct qsm_plan_s {
  int n[3];
  int kind[3];
  int fp32;
  const void* tw[3];
  int variant[3] = {0, 0, 0};  // tile / radix variant of each axis's pass
};

namespace {
int plan_measure(qsm_plan_s* p);
}

extern "C" const char* qsm_version(void) { return "qsm 0.1 (gfx950)"; }
extern "C" const char* qsm_last_error(void) { return g_err.c_str(); }

extern "C" int qsm_r2r3d_plan(int nx, int ny, int nz, int kx, int ky, int kz, int prec,
                              unsigned flags, qsm_plan* out) {
  if (!out) return fail(QSM_ERR_ARG, "qsm_r2r3d_plan: null output");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return fail(QSM_ERR_ARG, "qsm_r2r3d_plan: bad size");
  if (!is_kind(kx) || !is_kind(ky) || !is_kind(kz))
    return fail(QSM_ERR_ARG, "qsm_r2r3d_plan: unknown r2r kind");
  if (prec != QSM_FP64 && prec != QSM_MIXED) return fail(QSM_ERR_ARG, "qsm_r2r3d_plan: bad precision");
  if ((flags & ~QSM_PLAN_MEASURE) != 0) return fail(QSM_ERR_ARG, "qsm_r2r3d_plan: unknown flags");
  // kernels address a field as a base pointer plus a 32-bit byte offset
  if ((long)nx * ny * nz * (prec == QSM_FP64 ? 8 : 4) > (1L << 32))
    return fail(QSM_ERR_UNSUPPORTED, "qsm_r2r3d_plan: field larger than 4 GiB");
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a)
    if (!pass_supported(n[a]))
      return fail(QSM_ERR_UNSUPPORTED, "qsm_r2r3d_plan: no pass kernel for length %d", n[a]);
  auto* p = new qsm_plan_s;
  const int k[3] = {kx, ky, kz};
  p->fp32 = prec == QSM_MIXED;
  for (int a = 0; a < 3; ++a) {
    p->n[a] = n[a];
    p->kind[a] = k[a];
    p->tw[a] = twiddles(n[a], p->fp32);
    if (!p->tw[a]) { delete p; return fail(QSM_ERR_NOMEM, "qsm_r2r3d_plan: twiddle table"); }
  }
  if (flags & QSM_PLAN_MEASURE) {
    const int rc = plan_measure(p);
    if (rc) { delete p; return rc; }
  }
  *out = p;
  return QSM_OK;
}

extern "C" int qsm_r2r3d_set_variant(qsm_plan p, int axis, int variant) {
  if (!p || axis < 0 || axis > 2 || variant < 0 || variant >= kVariants)
    return fail(QSM_ERR_ARG, "qsm_r2r3d_set_variant: bad argument");
  if (!pass_variant_ok(p->fp32, p->n[axis], variant, axis, kind_inv(p->kind[axis])))
    return fail(QSM_ERR_UNSUPPORTED, "qsm_r2r3d_set_variant: length %d has no variant %d",
                p->n[axis], variant);
  p->variant[axis] = variant;
  return QSM_OK;
}

extern "C" int qsm_r2r3d_execute(qsm_plan p, const void* d_in, void* d_out, void* stream) {
  if (!p || !d_in || !d_out) return fail(QSM_ERR_ARG, "qsm_r2r3d_execute: null argument");
  hipStream_t s = (hipStream_t)stream;
  PassArgs A{};
  A.geo = Geo{p->n[0], p->n[1], p->n[2], 0};
  // x, y, z passes; the first reads d_in, the rest run in place on d_out
  const int order[3] = {0, 1, 2};
  for (int q = 0; q < 3; ++q) {
    const int ax = order[q];
    A.ld = LoadOp{};
    A.ld.code = LD_PLAIN;
    A.ld.a[0] = q == 0 ? d_in : d_out;
    A.st = StoreOp{};
    A.st.code = ST_PLAIN;
    A.st.d[0] = d_out;
    A.tw = p->tw[ax];
    A.dst = kind_dst(p->kind[ax]);
    hipError_t e = launch_pass(p->fp32, p->n[ax], ax, kind_inv(p->kind[ax]), LD_PLAIN, ST_PLAIN, A,
                               nullptr, s, nullptr, p->variant[ax]);
    if (e != hipSuccess) return fail(QSM_ERR_HIP, "pass launch failed: %s", hipGetErrorString(e));
  }
  return QSM_OK;
}

extern "C" int qsm_r2r3d_execute_host(qsm_plan p, const void* h_in, void* h_out) {
  if (!p || !h_in || !h_out) return fail(QSM_ERR_ARG, "qsm_r2r3d_execute_host: null argument");
  const size_t bytes = (size_t)p->n[0] * p->n[1] * p->n[2] * (p->fp32 ? 4 : 8);
  void* d = nullptr;
  HIPCHK(hipMalloc(&d, bytes));
  int rc = QSM_OK;
  if (hipMemcpy(d, h_in, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = fail(QSM_ERR_HIP, "copy in");
  if (rc == QSM_OK) rc = qsm_r2r3d_execute(p, d, d, nullptr);
  if (rc == QSM_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(QSM_ERR_HIP, "execute");
  if (rc == QSM_OK && hipMemcpy(h_out, d, bytes, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(QSM_ERR_HIP, "copy out");
  hipFree(d);
  return rc;
}

// ---- the allocation-free form of qsm_r2r3d_execute_host.  Its staging
// buffer is this file-scope device array (kStageBytes: one 256^3 FP64 field,
// the reference's default grid), held by every GPU the library runs on until
// the process ends.  Host code never passes `scratch` itself to a kernel or a
// copy: the array's address on the current device comes from
// hipGetSymbolAddress, once per device, and is kept in g_stage.
constexpr size_t kStageBytes = (size_t)1 << 27;
__device__ __align__(256) unsigned char scratch[kStageBytes];

namespace {
std::mutex g_stage_mu;
std::map<int, void*> g_stage;  // device ordinal -> address of `scratch` there

int stage_address(void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_stage_mu);
  auto it = g_stage.find(dev);
  if (it == g_stage.end()) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(QSM_ERR_HIP, "staging buffer has no address on device %d", dev);
    it = g_stage.emplace(dev, ptr).first;
  }
  *out = it->second;
  return QSM_OK;
}
}  // namespace

extern "C" int qsm_r2r3d_execute_host_noalloc(qsm_plan p, const void* h_in, void* h_out) {
  if (!p || !h_in || !h_out) return fail(QSM_ERR_ARG, "qsm_r2r3d_execute_host_noalloc: null argument");
  const size_t bytes = (size_t)p->n[0] * p->n[1] * p->n[2] * (p->fp32 ? 4 : 8);
  if (bytes > kStageBytes)
    return fail(QSM_ERR_UNSUPPORTED,
                "qsm_r2r3d_execute_host_noalloc: %zu-byte field exceeds the %zu-byte staging buffer",
                bytes, kStageBytes);
  void* d = nullptr;
  int rc = stage_address(&d);
  if (rc) return rc;
  HIPCHK(hipMemcpy(d, h_in, bytes, hipMemcpyHostToDevice));
  rc = qsm_r2r3d_execute(p, d, d, nullptr);
  if (rc) return rc;
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(h_out, d, bytes, hipMemcpyDeviceToHost));
  return QSM_OK;
}

extern "C" int qsm_r2r3d_destroy(qsm_plan p) {
  delete p;
  return QSM_OK;
}

// ============================================================================
// step level

namespace {

enum Fld {
  F_NHAT_OLD, F_NTMP, F_RHOHAT, F_RHO_A, F_RHO_B, F_VX, F_VY, F_VZ, F_PSI_A, F_PSI_B,
  F_F, F_G, F_GX, F_GY, F_GZ, F_RX, F_RY, F_RZ, F_LX, F_LY, F_LZ, F_DHAT, F_DIVV,
  F_SX, F_SY, F_SZ, F_P,
  // psi's x-passed and y-passed fields: the transforms of grad psi y and z
  // start from them.  rho's x-passed field and its y-passed fields
  // (ping-pong: the latest rho's and the one before, which the lagged grad
  // rho reads): the same for grad rho y and z (quasi).  Derived from psihat
  // and rhohat; not in checkpoints.
  F_PSIX, F_PSIY, F_RHOX, F_RHOY0, F_RHOY1, F_COUNT
};
constexpr int F_CKPT = F_PSIX;  // fields a checkpoint holds: those before F_PSIX

// device FP64 scalars
enum Dsc {
  D_RHOSUM = 0,   // sum of rho* (mixed)
  D_NSUM = 1,     // sum of N (mixed)
  D_RHOHAT0 = 2,  // FP64 rhohat(0,0,0) (mixed)
  D_NHAT0 = 3,    // FP64 N(0,0,0) of the last step (mixed)
  D_DIAG = 8,     // 4 diagnostic sums
  D_COUNT = 16
};

}  // namespace

// One launch of a step, recorded instead of launched (ensembles batch the
// same launch of many instances into one kernel).
struct PassRec {
  int kind;  // 0: pass, 1: FP64 finish of per-block partial sums
  int axis, n, inv, ld, st, variant;
  double bytes;  // algorithmic HBM bytes of the pass (op_bytes)
  PassArgs A;
  const double* part;
  long nparts;
  double* out;
};

struct qsm_handle_s {
  qsm_params p{};
  int prec = 0, fp32 = 0;
  size_t n = 0, esz = 8;
  hipStream_t stream = nullptr;
  Geo geo{};
  Waves waves{};
  Phys phys{};
  double* wtab = nullptr;     // 6 wavenumber tables + 3 filter factors
  double* ones = nullptr;     // unit filter factors (setup scaling)
  const void* tw[3] = {nullptr, nullptr, nullptr};
  void* f[F_COUNT] = {};
  int psi_cur = F_PSI_A, psi_old = F_PSI_B, rho_cur = F_RHO_A, rho_old = F_RHO_B;
  double* psihat = nullptr;   // FP64 state
  double* psihat_old = nullptr;
  float* psihat32 = nullptr;  // mixed precision: psihat rounded to FP32 (the spectral prologues' source)
  size_t pad = 0;             // elements of padding before and after each field (alloc_padded)
  bool divv_valid = false;    // the D field holds IDCT(scale D-hat) of the current step
  bool p_valid = false;       // the p field holds the current step's pressure
  bool r_valid = false;       // the grad rho fields hold the current step's (lagged) gradient
  bool l_yspec = false;       // the lap grad psi fields hold their y-passed values (the stress
                              // pass completed their inverse z passes in registers); the
                              // real-space fields are formed on request (lapgrad_output)
  bool r_yspec = false;       // the same for grad rho (stress pass)
  bool p_yspec = false;       // the same for p (nonlinear pass)
  int rhoy = 0;               // F_RHOY0 + rhoy: y-passed field of the latest rho
  double* part = nullptr;     // per-block partial sums
  long part_n = 0;
  double* dsc = nullptr;
  long long step = 0;
  bool ready = false;
  double growth = 0;
  Timer timer;
  // hipGraph replay of whole steps, keyed by everything the step's kernel
  // arguments depend on (ping-pong buffer state and phase flags)
  bool use_graphs = true;
  bool dry = false;  // sequence the step without launching (state bookkeeping only)
  std::vector<PassRec>* rec = nullptr;  // when set, launches are recorded here
  int variant[3][2] = {{0, 0}, {0, 0}, {0, 0}};  // tile-width variant of plain passes [axis][inv]
  std::map<std::vector<long long>, hipGraphExec_t> graphs;
};

namespace {

void fill_phys(qsm_handle_s* h) {
  const qsm_params& p = h->p;
  Phys& P = h->phys;
  P.q02 = p.q0 * p.q0;
  P.alpha = p.alpha; P.beta = p.beta; P.gamma = p.gamma; P.eps = p.eps; P.nu = p.nu;
  P.kp = p.kp; P.rho0 = p.rho0; P.dt = p.dt; P.filter = p.filter; P.rho_m = p.rho_m;
  P.scale = 0.125 / ((double)p.nx * p.ny * p.nz);
  const double rs = p.kp * p.amp + p.rho0;
  P.rho_s2 = rs * rs;
  P.hydro = p.hydro;
}

// algorithmic HBM bytes of one pass: (fields read + fields written) x points,
// FP64 state fields counted at 8 bytes
double op_bytes(const qsm_handle_s* h, const LoadOp& L, const StoreOp& S, bool first_pass) {
  const double n = (double)h->n, e = (double)h->esz;
  double rd = 0, wr = 0;
  if (!first_pass) rd = e;
  else switch (L.code) {
    case LD_PLAIN: rd = e; break;
    case LD_COEF_F: case LD_COEF_G: case LD_SH: rd = e; break;
    case LD_ZERO: rd = 0; break;
    case LD_GRAD_X: case LD_GRAD_Y: case LD_GRAD_Z:
    case LD_LAPGRAD_X: case LD_LAPGRAD_Y: case LD_LAPGRAD_Z: rd = e; break;
    case LD_RHOSTAR: rd = 4 * e; break;
    case LD_DIVV: rd = 8 * e; break;
    case LD_STRESS: rd = 5 * e; break;
    case LD_PRESSURE: rd = 3 * e; break;
    case LD_VEL0: case LD_VEL1: case LD_VEL2: rd = (L.a[3] ? 4 : 3) * e; break;
    case LD_SCALE: rd = e; break;
    case LD_NONLIN: rd = 10 * e; break;
    case LD_NONLIN_P: rd = 14 * e; break;
    case LD_NONLIN_PY: rd = 12 * e; break;
    case LD_STRESS_STB: rd = 3 * e; break;
    case LD_NONLIN_STB: rd = 7 * e; break;
    case LD_NONLIN_STB0: rd = e; break;
    default: rd = e;
  }
  switch (S.code) {
    case ST_PLAIN: case ST_FILTER: wr = e; break;
    case ST_FILTER_X: wr = 2 * e; break;
    case ST_UPDATE: case ST_UPDATE_STB: rd += 16 + e; wr = 8 + e + (h->fp32 ? e : 0); break;
    case ST_UPDATE_X: case ST_UPDATE_STB_X: rd += 16 + e; wr = 8 + 2 * e + (h->fp32 ? e : 0); break;
    default: wr = e;
  }
  return (rd + wr) * n;
}

// var >= 0: the tile variant of a plain pass instead of the tuned one
int run_pass(qsm_handle_s* h, int axis, bool inv, bool dst, const LoadOp& L, const StoreOp& S,
             bool first_pass, double* part, int var_force = -1) {
  PassArgs A{};
  A.geo = h->geo;
  A.waves = h->waves;
  A.phys = h->phys;
  A.ld = L;
  A.st = S;
  A.tw = h->tw[axis];
  A.part = part;
  A.dst = dst ? 1 : 0;
  const int n = axis == 0 ? h->p.nx : (axis == 1 ? h->p.ny : h->p.nz);
  if (part && (inv || axis != 2 || !dc_capable(L.code)))
    return fail(QSM_ERR_STATE, "pass: FP64 mean-mode sum requested for a pass that has none");
  if (h->dry) return QSM_OK;
  const int var = (L.code == LD_PLAIN && S.code == ST_PLAIN)
                      ? (var_force >= 0 ? var_force : h->variant[axis][inv ? 1 : 0])
                      : 0;
  if (h->rec) {
    PassRec r{};
    r.kind = 0; r.axis = axis; r.n = n; r.inv = inv; r.ld = L.code; r.st = S.code; r.A = A;
    r.variant = var;
    r.bytes = op_bytes(h, L, S, first_pass);
    h->rec->push_back(r);
    return QSM_OK;
  }
  Timer& T = h->timer;
  const int key = timer_key(axis, inv, L.code, S.code, var);
  const bool timed = T.want(key);
  hipEvent_t ea = nullptr, eb = nullptr;
  if (timed) { ea = T.get(); eb = T.get(); hipEventRecord(ea, h->stream); }
  hipError_t e = launch_pass(h->fp32, n, axis, inv, L.code, S.code, A, nullptr, h->stream, nullptr, var);
  if (e != hipSuccess) return fail(QSM_ERR_HIP, "pass launch (axis %d, n %d): %s", axis, n, hipGetErrorString(e));
  if (timed) {
    hipEventRecord(eb, h->stream);
    T.recs.push_back({ea, eb, key, op_bytes(h, L, S, first_pass)});
  }
  return QSM_OK;
}

LoadOp plain_load(const void* src) {
  LoadOp L{};
  L.code = LD_PLAIN;
  L.a[0] = src;
  return L;
}
StoreOp plain_store(void* dst) {
  StoreOp S{};
  S.code = ST_PLAIN;
  S.d[0] = dst;
  return S;
}

// Inverse 3D transform (DCT-III on every axis except DST-III along dst_axis):
// x pass with the fused load operator L into `out`, then y and z in place.
int inv3(qsm_handle_s* h, int dst_axis, const LoadOp& L, void* out) {
  int rc = run_pass(h, 0, true, dst_axis == 0, L, plain_store(out), true, nullptr);
  if (rc) return rc;
  rc = run_pass(h, 1, true, dst_axis == 1, plain_load(out), plain_store(out), false, nullptr);
  if (rc) return rc;
  return run_pass(h, 2, true, dst_axis == 2, plain_load(out), plain_store(out), false, nullptr);
}

// Forward 3D transform (DCT-II, DST-II along dst_axis): z pass with the fused
// load operator L into `tmp`, y in place, x pass from `tmp` with store operator S.
// dc_slot >= 0 (mixed precision): FP64 sum of the loaded values into dsc[dc_slot].
int fwd3(qsm_handle_s* h, int dst_axis, const LoadOp& L, void* tmp, const StoreOp& S, int dc_slot) {
  double* part = dc_slot >= 0 ? h->part : nullptr;
  int rc = run_pass(h, 2, false, dst_axis == 2, L, plain_store(tmp), true, part);
  if (rc) return rc;
  if (part && !h->dry) {
    const long nb = pass_blocks(h->fp32, h->p.nz, 2, h->geo);
    if (h->rec) {
      PassRec r{};
      r.kind = 1; r.part = part; r.nparts = nb; r.out = h->dsc + dc_slot;
      h->rec->push_back(r);
    } else {
      hipError_t e = launch_finish(part, nb, h->dsc + dc_slot, h->stream);
      if (e != hipSuccess) return fail(QSM_ERR_HIP, "dc sum: %s", hipGetErrorString(e));
    }
  }
  rc = run_pass(h, 1, false, dst_axis == 1, plain_load(tmp), plain_store(tmp), false, nullptr);
  if (rc) return rc;
  return run_pass(h, 0, false, dst_axis == 0, plain_load(tmp), S, false, nullptr);
}

// ---- group transforms: several inverse 3D transforms whose first (x) pass
// reads the same fields run that pass as one kernel (launch_group), then
// their y and z passes one by one.  Used by single handles, as are the
// multi-output forward transforms below; an ensemble records its passes one
// transform at a time and keeps the per-transform path.
struct GroupItem {
  LoadOp L;
  void* out;     // output field; the group pass writes it, the others run in place
  int dst_axis;  // axis of the DST (-1: DCT on every axis)
};

// launch of the first pass of a group, with the timing of one pass; `reads`
// is the number of distinct fields the group reads (algorithmic bytes)
int run_group_first(qsm_handle_s* h, int group, const GroupItem* it, int count, int reads) {
  const int axis = group_axis(group);
  GroupArgs G{};
  G.geo = h->geo;
  G.waves = h->waves;
  G.phys = h->phys;
  G.tw = h->tw[axis];
  for (int o = 0; o < count; ++o) {
    G.out[o].ld = it[o].L;
    G.out[o].st = plain_store(it[o].out);
    G.out[o].dst = it[o].dst_axis == axis ? 1 : 0;
  }
  if (h->dry) return QSM_OK;
  if (h->rec) return fail(QSM_ERR_STATE, "group passes cannot be recorded");
  const int n = axis == 0 ? h->p.nx : (axis == 1 ? h->p.ny : h->p.nz);
  Timer& T = h->timer;
  const int key = timer_key(axis, group_inverse(group), kGroupCode0 + group, ST_PLAIN, 0);
  const bool timed = T.want(key);
  hipEvent_t ea = nullptr, eb = nullptr;
  if (timed) { ea = T.get(); eb = T.get(); hipEventRecord(ea, h->stream); }
  hipError_t e = launch_group(h->fp32, n, group, G, nullptr, h->stream);
  if (e != hipSuccess) return fail(QSM_ERR_HIP, "group %d launch: %s", group, hipGetErrorString(e));
  if (timed) {
    hipEventRecord(eb, h->stream);
    T.recs.push_back({ea, eb, key, (double)(reads + count) * (double)h->n * (double)h->esz});
  }
  return QSM_OK;
}

// inverse transforms x, y, z (the x pass as one group kernel).  The y and z
// passes take the outputs last-written first: the group kernel's last tiles
// are still in the memory-side cache when the next y pass starts on them.
// no_z bit o: output o stops after its y pass (its consumer runs the z pass)
int inv3_group(qsm_handle_s* h, int group, const GroupItem* it, int count, int reads,
               unsigned no_z = 0) {
  int rc = run_group_first(h, group, it, count, reads);
  for (int o = count - 1; o >= 0 && !rc; --o) {
    rc = run_pass(h, 1, true, it[o].dst_axis == 1, plain_load(it[o].out), plain_store(it[o].out), false, nullptr);
    if (!rc && !((no_z >> o) & 1))
      rc = run_pass(h, 2, true, it[o].dst_axis == 2, plain_load(it[o].out), plain_store(it[o].out), false, nullptr);
  }
  return rc;
}

// Forward transforms z, y, x of the nout values of a multi-output operator
// (L: LD_STRESS3 / LD_STRESS_STB3 / LD_STRESS3_DIVV; value c < 3 is the
// stress component c, with its DST along c; value 3 the divergence, DCT on
// every axis): the z pass of all of them as one kernel reading the shared
// inputs once (launch_multi), then y and x in place per value.  `reads`:
// distinct fields the operator reads (algorithmic bytes).  Returns 1 (no
// error set) where the pencil length has no multi-output kernel.
// aux (lz = 2): the pass also writes |grad rho|^2 of each element there.
int fwd3_multi(qsm_handle_s* h, const LoadOp& L, void* const* out, int nout, int reads,
               int lz = 0, void* aux = nullptr) {
  if (!multi_supported(h->fp32, h->p.nz)) return 1;
  if (!h->dry) {
    if (h->rec) return fail(QSM_ERR_STATE, "multi-output passes cannot be recorded");
    MultiArgs A{};
    A.geo = h->geo;
    A.phys = h->phys;
    A.ld = L;
    A.tw = h->tw[2];
    A.dmask = 1 << 2;  // component z: DST along z
    A.lz = lz;
    A.aux = lz == 2 ? aux : nullptr;
    A.sz = h->waves.sz;
    for (int c = 0; c < nout; ++c) A.out[c] = out[c];
    Timer& T = h->timer;
    const int key = timer_key(2, false, L.code, ST_PLAIN, 0);
    const bool timed = T.want(key);
    hipEvent_t ea = nullptr, eb = nullptr;
    if (timed) { ea = T.get(); eb = T.get(); hipEventRecord(ea, h->stream); }
    hipError_t e = launch_multi(h->fp32, h->p.nz, A, nullptr, h->stream);
    if (e != hipSuccess) return fail(QSM_ERR_HIP, "multi-output launch: %s", hipGetErrorString(e));
    if (timed) {
      hipEventRecord(eb, h->stream);
      const int writes = nout + (A.aux ? 1 : 0);
      T.recs.push_back({ea, eb, key, (double)(reads + writes) * (double)h->n * (double)h->esz});
    }
  }
  for (int c = 0; c < nout; ++c) {
    int rc = run_pass(h, 1, false, c == 1, plain_load(out[c]), plain_store(out[c]), false, nullptr);
    if (!rc) rc = run_pass(h, 0, false, c == 0, plain_load(out[c]), plain_store(out[c]), false, nullptr);
    if (rc) return rc;
  }
  return QSM_OK;
}

void* F(qsm_handle_s* h, int f) { return h->f[f]; }

// psihat in the field precision: the FP64 state, or its FP32 copy in mixed
// precision
const void* psihat_src(const qsm_handle_s* h) {
  return h->fp32 ? (const void*)h->psihat32 : (const void*)h->psihat;
}

// psi's x- and y-passed fields from psihat: the first two passes of psi =
// IDCT(psihat), kept for the transforms of grad psi y and z.  The step
// forms the x-passed field in its update pass (ST_UPDATE_X), by the default
// radix plan; the x pass here (set-up, checkpoint read) takes the same plan
// (variant 0), so either way the field has the same bits.  with_x = false:
// the x-passed field is already there (the update pass wrote it).
int psi_xy(qsm_handle_s* h, bool with_x = true) {
  int rc = with_x ? run_pass(h, 0, true, false, plain_load(psihat_src(h)), plain_store(F(h, F_PSIX)),
                             true, nullptr, 0)
                  : QSM_OK;
  if (!rc) rc = run_pass(h, 1, true, false, plain_load(F(h, F_PSIX)), plain_store(F(h, F_PSIY)), false, nullptr);
  return rc;
}

// psi = IDCT(psihat) into out (quasi.cpp:924-927, 1961; quasi-stb.cpp:1961)
// through F_PSIX and F_PSIY, which the next step reads.
int psi_inverse(qsm_handle_s* h, void* out, bool with_x = true) {
  const int rc = psi_xy(h, with_x);
  return rc ? rc : run_pass(h, 2, true, false, plain_load(F(h, F_PSIY)), plain_store(out), false, nullptr);
}

// grad psi y and z (quasi.cpp:1060-1118) from psi's x- and y-passed fields:
// -s_c psihat[m+1] (0 at m = N_c - 1) commutes with the transforms along the
// other axes, so grad psi y is the y pass of the gradient operator on
// F_PSIX, then the z pass; grad psi z the z pass of the operator on F_PSIY.
int grad_psi_yz(qsm_handle_s* h) {
  LoadOp L{};
  L.code = LD_GRAD_Y; L.axis = 1; L.a[0] = F(h, F_PSIX);
  int rc = run_pass(h, 1, true, true, L, plain_store(F(h, F_GY)), true, nullptr);
  if (!rc) rc = run_pass(h, 2, true, false, plain_load(F(h, F_GY)), plain_store(F(h, F_GY)), false, nullptr);
  if (rc) return rc;
  L.code = LD_GRAD_Z; L.axis = 2; L.a[0] = F(h, F_PSIY);
  return run_pass(h, 2, true, true, L, plain_store(F(h, F_GZ)), true, nullptr);
}

// rho = IDCT(rhohat) into out (quasi.cpp:1263-1268) through F_RHOX and the
// free one of the two y-passed fields, which then holds the latest rho's.
// The step's density filter pass forms F_RHOX itself (ST_FILTER_X, with_x =
// false); the x pass here takes the same radix plan (variant 0).
int rho_inverse(qsm_handle_s* h, void* out, bool with_x = true) {
  const int cur = h->rhoy ^ 1;
  void* y = F(h, F_RHOY0 + cur);
  int rc = with_x ? run_pass(h, 0, true, false, plain_load(F(h, F_RHOHAT)), plain_store(F(h, F_RHOX)),
                             true, nullptr, 0)
                  : QSM_OK;
  if (!rc) rc = run_pass(h, 1, true, false, plain_load(F(h, F_RHOX)), plain_store(y), false, nullptr);
  if (!rc && out) rc = run_pass(h, 2, true, false, plain_load(y), plain_store(out), false, nullptr);
  if (!rc) h->rhoy = cur;
  return rc;
}

// The lagged grad rho (quasi.cpp:1062-1118: that of the rhohat the step
// starts from), before the step's density pass.  r_x from rhohat; r_y, r_z
// from rho's x- and y-passed fields, which the previous step's rho_inverse
// left (the gradient operator commutes with the transforms along the other
// axes).  yspec: r_x and r_y stop after their y passes and r_z is not
// formed; the stress pass completes them (r_z from F_RHOY0 + rhoy).
int grad_rho(qsm_handle_s* h, bool yspec) {
  LoadOp L{};
  L.code = LD_GRAD_X; L.axis = 0; L.a[0] = F(h, F_RHOHAT);
  void* rx = F(h, F_RX);
  int rc = run_pass(h, 0, true, true, L, plain_store(rx), true, nullptr);
  if (!rc) rc = run_pass(h, 1, true, false, plain_load(rx), plain_store(rx), false, nullptr);
  if (!rc && !yspec) rc = run_pass(h, 2, true, false, plain_load(rx), plain_store(rx), false, nullptr);
  if (rc) return rc;
  L.code = LD_GRAD_Y; L.axis = 1; L.a[0] = F(h, F_RHOX);
  void* ry = F(h, F_RY);
  rc = run_pass(h, 1, true, true, L, plain_store(ry), true, nullptr);
  if (!rc && !yspec) rc = run_pass(h, 2, true, false, plain_load(ry), plain_store(ry), false, nullptr);
  if (rc || yspec) return rc;
  L.code = LD_GRAD_Z; L.axis = 2; L.a[0] = F(h, F_RHOY0 + h->rhoy);
  return run_pass(h, 2, true, true, L, plain_store(F(h, F_RZ)), true, nullptr);
}

// Whether the grad rho arrays hold the gradient that step n reads: steps
// after divv_switch form it; at step 1 it is that of rhohat = 0, the zeros
// the arrays hold from set-up; other steps do not form it (quasi_step).
bool grad_rho_formed(const qsm_params& p, long long n) { return n > p.divv_switch || n <= 1; }

// p-hat from the current S-hat (quasi.cpp:1385-1415), as the load of p's x pass
LoadOp pressure_load(qsm_handle_s* h) {
  LoadOp L{};
  L.code = LD_PRESSURE;
  L.a[0] = F(h, F_SX); L.a[1] = F(h, F_SY); L.a[2] = F(h, F_SZ);
  return L;
}

int quasi_step(qsm_handle_s* h) {
  const long long n = ++h->step;
  const qsm_params& p = h->p;
  const bool mixed = h->fp32;
  int rc;
  if (n >= 2) std::swap(h->rho_cur, h->rho_old);  // rho_old <- rho (quasi.cpp:1289)
  void *rho = F(h, h->rho_cur), *rho_old = F(h, h->rho_old), *psi = F(h, h->psi_cur);
  const void* psih = psihat_src(h);
  const bool dlive = n > p.divv_switch;  // D is computed from this step on
  // grad rho is read by D (n > divv_switch) and by N's pressure term
  // (n > divv_switch + 1); p only by that term (quasi.cpp:1273-1293,
  // 1905-1939).  Before that they are read by no stage of the step, so they
  // are not formed: grad rho at step 1 is that of rhohat = 0, which its
  // arrays hold from set-up, and p is formed when the field is asked for
  // (pressure_output).
  const bool plive = n > p.divv_switch + 1;
  const bool grp = !h->rec;  // group kernels (single handles)
  // Single handles form the stress components (and D) in one multi-output z
  // pass.  That pass completes the inverse transforms of lap grad psi and,
  // once D is live, of grad rho, whose |grad rho|^2 it leaves in the D field
  // for the nonlinear pass; that pass completes the transform of p.  Their z
  // passes are not run separately.
  const bool multi = grp && multi_supported(h->fp32, p.nz);
  h->divv_valid = false;
  h->p_valid = plive && !multi;
  h->p_yspec = plive && multi;
  h->r_valid = grad_rho_formed(p, n);
  h->r_yspec = dlive && multi;
  h->l_yspec = multi;
  LoadOp L{};
  if (grp) {
    // S1 + S2 (quasi.cpp:1034-1232): F, G, grad psi x, laplacian of grad
    // psi from psihat in one group, grad psi y and z from psi's passed
    // fields; grad rho (lagged) from rhohat in another group
    GroupItem it[6];
    it[0] = {LoadOp{}, F(h, F_F), -1};
    it[0].L.code = LD_COEF_F;
    it[1] = {LoadOp{}, F(h, F_G), -1};
    it[1].L.code = LD_COEF_G;
    it[2] = {LoadOp{}, F(h, F_GX), 0};
    it[2].L.code = LD_GRAD_X;
    for (int c = 0; c < 3; ++c) {
      it[3 + c] = {LoadOp{}, F(h, F_LX + c), c};
      it[3 + c].L.code = LD_LAPGRAD_X + c;
    }
    for (auto& x : it) { x.L.a[0] = psih; x.L.axis = x.dst_axis; }
    if ((rc = inv3_group(h, GRP_PSI, it, 6, 1, multi ? 0x38u : 0u))) return rc;
    if ((rc = grad_psi_yz(h))) return rc;
    if (dlive && (rc = grad_rho(h, multi))) return rc;
  } else {
    // S1: F, G (quasi.cpp:1034-1057)
    L.code = LD_COEF_F; L.a[0] = psih;
    if ((rc = inv3(h, -1, L, F(h, F_F)))) return rc;
    L.code = LD_COEF_G;
    if ((rc = inv3(h, -1, L, F(h, F_G)))) return rc;

    // S2: grad psi, grad rho (lagged), laplacian of grad psi (quasi.cpp:1060-1232)
    if ((rc = grad_psi_yz(h))) return rc;
    L = LoadOp{}; L.code = LD_GRAD_X; L.axis = 0; L.a[0] = psih;
    if ((rc = inv3(h, 0, L, F(h, F_GX)))) return rc;
    if (dlive && (rc = grad_rho(h, false))) return rc;
    for (int c = 0; c < 3; ++c) {
      L = LoadOp{}; L.code = LD_LAPGRAD_X + c; L.axis = c; L.a[0] = psih;
      if ((rc = inv3(h, c, L, F(h, F_LX + c)))) return rc;
    }
  }

  // S3: density law + filter; rho (quasi.cpp:1237-1268)
  L = LoadOp{}; L.code = LD_RHOSTAR;
  L.a[0] = psi; L.a[1] = F(h, F_GX); L.a[2] = F(h, F_GY); L.a[3] = F(h, F_GZ);
  // the filter pass also runs rho's inverse x pass (into F_RHOX)
  StoreOp S{}; S.code = ST_FILTER_X; S.d[0] = F(h, F_RHOHAT); S.d[3] = F(h, F_RHOX);
  if (mixed) { S.dc = h->dsc + D_RHOSUM; S.dc_out = h->dsc + D_RHOHAT0; }
  if ((rc = fwd3(h, -1, L, F(h, F_RHOHAT), S, mixed ? D_RHOSUM : -1))) return rc;
  if ((rc = rho_inverse(h, rho, false))) return rc;

  // S4: divergence (quasi.cpp:1273-1293).  Up to divv_switch the reference
  // transforms an all-zero D; D-hat = 0 then, which the D-hat array already
  // holds (it is zeroed at set-up and first written at step divv_switch+1),
  // so the transform of zeros is not run.  S4 and S5 read only fields of
  // earlier stages: single handles run D (when live) as the fourth value of
  // the stress components' multi-output z pass.
  if (dlive && !multi) {
    L = LoadOp{}; L.code = LD_DIVV;
    L.a[0] = rho; L.a[1] = rho_old; L.a[2] = F(h, F_VX); L.a[3] = F(h, F_VY); L.a[4] = F(h, F_VZ);
    L.a[5] = F(h, F_RX); L.a[6] = F(h, F_RY); L.a[7] = F(h, F_RZ);
    if ((rc = fwd3(h, -1, L, F(h, F_DHAT), plain_store(F(h, F_DHAT)), -1))) return rc;
  }

  // S5: stress (quasi.cpp:1298-1330)
  if (multi) {
    L = LoadOp{}; L.code = dlive ? LD_STRESS3_DIVV : LD_STRESS3;
    L.a[0] = rho; L.a[1] = F(h, F_G); L.a[2] = F(h, F_F);
    for (int c = 0; c < 3; ++c) { L.a[3 + c] = F(h, F_GX + c); L.a[6 + c] = F(h, F_LX + c); }
    L.a[9] = rho_old;
    for (int c = 0; c < 3; ++c) { L.a[10 + c] = F(h, F_VX + c); L.a[13 + c] = F(h, F_RX + c); }
    // with D live the pass forms r_z from rho's y-passed field of the step
    // before (the lagged gradient); the density pass has since written the
    // other one
    if (dlive) L.a[15] = F(h, F_RHOY0 + (h->rhoy ^ 1));
    void* const out[4] = {F(h, F_SX), F(h, F_SY), F(h, F_SZ), F(h, F_DHAT)};
    // with D live the pass also forms |grad rho|^2 for the nonlinear pass's
    // pressure term, in the D field (read by no stage; formed on request)
    if ((rc = fwd3_multi(h, L, out, dlive ? 4 : 3, dlive ? 16 : 9, dlive ? 2 : 1,
                         dlive ? F(h, F_DIVV) : nullptr)))
      return rc;
  } else {
    for (int c = 0; c < 3; ++c) {
      L = LoadOp{}; L.code = LD_STRESS; L.axis = c;
      L.a[0] = rho; L.a[1] = F(h, F_G); L.a[2] = F(h, F_F); L.a[3] = F(h, F_GX + c); L.a[4] = F(h, F_LX + c);
      if ((rc = fwd3(h, c, L, F(h, F_SX + c), plain_store(F(h, F_SX + c)), -1))) return rc;
    }
  }

  // S6 + S7: pressure and velocity projection (quasi.cpp:1336-1884)
  const int n0 = p.compat_ranks > 0 ? p.nx / p.compat_ranks : 0;
  if (grp) {
    GroupItem it[4];
    it[0] = {LoadOp{}, F(h, F_P), -1};
    it[0].L.code = LD_PRESSURE;
    for (int c = 0; c < 3; ++c) {
      it[1 + c] = {LoadOp{}, F(h, F_VX + c), c};
      it[1 + c].L.code = LD_VEL0 + c;
      it[1 + c].L.n0 = n0;
      it[1 + c].L.a[3] = dlive ? F(h, F_DHAT) : nullptr;  // D-hat = 0 before: its term vanishes
    }
    for (auto& x : it) { x.L.a[0] = F(h, F_SX); x.L.a[1] = F(h, F_SY); x.L.a[2] = F(h, F_SZ); }
    if (plive) rc = inv3_group(h, GRP_PROJ, it, 4, 4, multi ? 0x1u : 0u);
    else rc = inv3_group(h, GRP_VEL, it + 1, 3, dlive ? 4 : 3);
    if (rc) return rc;
  } else {
    if (plive && (rc = inv3(h, -1, pressure_load(h), F(h, F_P)))) return rc;
    for (int c = 0; c < 3; ++c) {
      L = LoadOp{}; L.code = LD_VEL0 + c; L.n0 = n0;
      L.a[0] = F(h, F_SX); L.a[1] = F(h, F_SY); L.a[2] = F(h, F_SZ);
      L.a[3] = dlive ? F(h, F_DHAT) : nullptr;  // D-hat = 0 before: its term vanishes
      if ((rc = inv3(h, c, L, F(h, F_VX + c)))) return rc;
    }
  }

  // S8: the D output, IDCT(scale D-hat) (quasi.cpp:1888-1900), is read by no
  // later stage of the step; it is formed when the field is asked for
  // (divv_output).

  // S9 + S10: N, its transform, and the update fused into the last pass
  L = LoadOp{}; L.code = plive ? (multi ? LD_NONLIN_PY : LD_NONLIN_P) : LD_NONLIN;
  L.a[0] = rho; L.a[1] = psi; L.a[2] = F(h, F_F); L.a[3] = F(h, F_G);
  L.a[4] = F(h, F_VX); L.a[5] = F(h, F_VY); L.a[6] = F(h, F_VZ);
  L.a[7] = F(h, F_GX); L.a[8] = F(h, F_GY); L.a[9] = F(h, F_GZ);
  L.a[10] = F(h, F_P); L.a[11] = F(h, F_RX); L.a[12] = F(h, F_RY); L.a[13] = F(h, F_RZ);
  if (L.code == LD_NONLIN_PY) L.a[11] = F(h, F_DIVV);  // |grad rho|^2 from the stress pass
  // the update pass also runs psi's inverse x pass (into F_PSIX)
  S = StoreOp{}; S.code = ST_UPDATE_X; S.first = n == 1;
  S.d[0] = h->psihat_old; S.d[1] = F(h, F_NHAT_OLD); S.d[2] = h->psihat32; S.s[0] = h->psihat;
  S.d[3] = F(h, F_PSIX);
  if (mixed) { S.dc = h->dsc + D_NSUM; S.dc_old = h->dsc + D_NHAT0; S.dc_out = h->dsc + D_NHAT0; }
  if ((rc = fwd3(h, -1, L, F(h, F_NTMP), S, mixed ? D_NSUM : -1))) return rc;
  std::swap(h->psihat, h->psihat_old);

  // psi = IDCT(psihat) into the psi_old buffer, then swap (psi_old <- psi)
  if ((rc = psi_inverse(h, F(h, h->psi_old), false))) return rc;
  std::swap(h->psi_cur, h->psi_old);
  return QSM_OK;
}

// The D field: IDCT(scale D-hat) of the current step (quasi.cpp:1888-1900),
// formed on request.  Zero while D-hat is (steps <= divv_switch).
int divv_output(qsm_handle_s* h) {
  if (h->divv_valid) return QSM_OK;
  LoadOp L{};
  L.code = LD_SCALE; L.a[0] = F(h, F_DHAT);
  const int rc = inv3(h, -1, L, F(h, F_DIVV));
  if (!rc) h->divv_valid = true;
  return rc;
}

// The p field of the current step, formed on request when the step did not
// need it (steps <= divv_switch + 1): IDCT of p-hat from the S-hat arrays,
// which hold this step's stress transforms until the next step's S5.  The
// same x pass as the projection group's first output, so the same bits.
int pressure_output(qsm_handle_s* h) {
  if (h->p_valid) return QSM_OK;
  // y-passed by the step (its nonlinear pass completed the transform in
  // registers): the z pass in place; else the whole transform from S-hat
  const int rc = h->p_yspec
                     ? run_pass(h, 2, true, false, plain_load(F(h, F_P)), plain_store(F(h, F_P)), false, nullptr)
                     : inv3(h, -1, pressure_load(h), F(h, F_P));
  if (!rc) {
    h->p_valid = true;
    h->p_yspec = false;
  }
  return rc;
}

// The lap grad psi fields of the current step in real space, formed on
// request where the step left them y-passed (the stress pass ran their z
// passes in registers): the same inverse z passes, in place.
int lapgrad_output(qsm_handle_s* h) {
  if (!h->l_yspec) return QSM_OK;
  for (int c = 0; c < 3; ++c) {
    void* f = F(h, F_LX + c);
    const int rc = run_pass(h, 2, true, c == 2, plain_load(f), plain_store(f), false, nullptr);
    if (rc) return rc;
  }
  h->l_yspec = false;
  return QSM_OK;
}

// The same for grad rho: r_x, r_y by the z passes in place; r_z (which the
// stress pass formed in registers only) by the z pass of the z gradient
// operator on the y-passed field of the rho before the latest, as the stress
// pass read it.
int gradrho_output(qsm_handle_s* h) {
  if (!h->r_yspec) return QSM_OK;
  for (int c = 0; c < 2; ++c) {
    void* f = F(h, F_RX + c);
    const int rc = run_pass(h, 2, true, false, plain_load(f), plain_store(f), false, nullptr);
    if (rc) return rc;
  }
  LoadOp L{};
  L.code = LD_GRAD_Z; L.axis = 2; L.a[0] = F(h, F_RHOY0 + (h->rhoy ^ 1));
  const int rc = run_pass(h, 2, true, true, L, plain_store(F(h, F_RZ)), true, nullptr);
  if (rc) return rc;
  h->r_yspec = false;
  return QSM_OK;
}

// stb: dtpsi = -rho mu - v.g (hydro = 1) of the last step, rho = 1, mu =
// sh - beta psi^3 + gamma psi^5 (quasi-stb.cpp:1904-1913), formed on request
// from the fields the step leaves: sh in F, the step's psi in the psi_old
// buffer, v and grad psi.  The result goes to the transform scratch.
int dtpsi_output(qsm_handle_s* h) {
  const void* in[8] = {F(h, F_F), F(h, h->psi_old), F(h, F_VX), F(h, F_VY), F(h, F_VZ),
                       F(h, F_GX), F(h, F_GY), F(h, F_GZ)};
  hipError_t e = launch_dtpsi(h->fp32, in, (long)h->n, h->phys, F(h, F_NTMP), h->stream);
  if (e != hipSuccess) return fail(QSM_ERR_HIP, "dtpsi launch: %s", hipGetErrorString(e));
  return QSM_OK;
}

int stb_step(qsm_handle_s* h) {
  const long long n = ++h->step;
  const bool mixed = h->fp32;
  int rc;
  void* psi = F(h, h->psi_cur);
  // sh = IDCT(-aLin psihat) (quasi-stb.cpp:1015-1038), kept in F
  LoadOp L{};
  const void* psih = psihat_src(h);
  const int n0 = h->p.compat_ranks > 0 ? h->p.nx / h->p.compat_ranks : 0;
  if (!h->rec) {  // group kernels (single handles)
    // sh and grad psi x in one group; grad psi y and z from psi's passed
    // fields (grad_psi_yz)
    GroupItem it[3];
    it[0] = {LoadOp{}, F(h, F_F), -1};
    it[0].L.code = LD_SH;
    it[1] = {LoadOp{}, F(h, F_GX), 0};
    it[1].L.code = LD_GRAD_X;
    it[1].L.axis = 0;
    for (int o = 0; o < 2; ++o) it[o].L.a[0] = psih;
    if ((rc = inv3_group(h, GRP_STB_PSI, it, 2, 1))) return rc;
    if ((rc = grad_psi_yz(h))) return rc;
    L = LoadOp{}; L.code = LD_STRESS_STB3;
    L.a[0] = F(h, F_F); L.a[1] = psi;
    for (int c = 0; c < 3; ++c) L.a[2 + c] = F(h, F_GX + c);
    void* const out[3] = {F(h, F_SX), F(h, F_SY), F(h, F_SZ)};
    if ((rc = fwd3_multi(h, L, out, 3, 5)) == 1) {
      for (int c = 0; c < 3; ++c) {
        L = LoadOp{}; L.code = LD_STRESS_STB; L.axis = c;
        L.a[0] = F(h, F_F); L.a[1] = psi; L.a[2] = F(h, F_GX + c);
        if ((rc = fwd3(h, c, L, F(h, F_SX + c), plain_store(F(h, F_SX + c)), -1))) return rc;
      }
    } else if (rc) {
      return rc;
    }
    for (int c = 0; c < 3; ++c) {
      it[c] = {LoadOp{}, F(h, F_VX + c), c};
      it[c].L.code = LD_VEL0 + c; it[c].L.n0 = n0;
      it[c].L.a[0] = F(h, F_SX); it[c].L.a[1] = F(h, F_SY); it[c].L.a[2] = F(h, F_SZ);
      it[c].L.a[3] = nullptr;
    }
    if ((rc = inv3_group(h, GRP_VEL, it, 3, 3))) return rc;
  } else {
    L.code = LD_SH; L.a[0] = psih;
    if ((rc = inv3(h, -1, L, F(h, F_F)))) return rc;
    L = LoadOp{}; L.code = LD_GRAD_X; L.axis = 0; L.a[0] = psih;
    if ((rc = inv3(h, 0, L, F(h, F_GX)))) return rc;
    if ((rc = grad_psi_yz(h))) return rc;
    for (int c = 0; c < 3; ++c) {
      L = LoadOp{}; L.code = LD_STRESS_STB; L.axis = c;
      L.a[0] = F(h, F_F); L.a[1] = psi; L.a[2] = F(h, F_GX + c);
      if ((rc = fwd3(h, c, L, F(h, F_SX + c), plain_store(F(h, F_SX + c)), -1))) return rc;
    }
    for (int c = 0; c < 3; ++c) {
      L = LoadOp{}; L.code = LD_VEL0 + c; L.n0 = n0;
      L.a[0] = F(h, F_SX); L.a[1] = F(h, F_SY); L.a[2] = F(h, F_SZ); L.a[3] = nullptr;
      if ((rc = inv3(h, c, L, F(h, F_VX + c)))) return rc;
    }
  }
  L = LoadOp{}; L.code = h->p.hydro == 1 ? LD_NONLIN_STB : LD_NONLIN_STB0;
  L.a[0] = psi; L.a[1] = F(h, F_VX); L.a[2] = F(h, F_VY); L.a[3] = F(h, F_VZ);
  L.a[4] = F(h, F_GX); L.a[5] = F(h, F_GY); L.a[6] = F(h, F_GZ);
  // the update pass also runs psi's inverse x pass (into F_PSIX)
  StoreOp S{}; S.code = ST_UPDATE_STB_X; S.first = n == 1;
  S.d[0] = h->psihat_old; S.d[1] = F(h, F_NHAT_OLD); S.d[2] = h->psihat32; S.s[0] = h->psihat;
  S.d[3] = F(h, F_PSIX);
  if (mixed) { S.dc = h->dsc + D_NSUM; S.dc_old = h->dsc + D_NHAT0; S.dc_out = h->dsc + D_NHAT0; }
  if ((rc = fwd3(h, -1, L, F(h, F_NTMP), S, mixed ? D_NSUM : -1))) return rc;
  std::swap(h->psihat, h->psihat_old);
  if ((rc = psi_inverse(h, F(h, h->psi_old), false))) return rc;
  std::swap(h->psi_cur, h->psi_old);
  return QSM_OK;
}

size_t idx3(const qsm_params& p, int i, int j, int k) {
  return ((size_t)i * p.ny + j) * p.nz + k;
}

int stb_growth(qsm_handle_s* h) {
  // quasi-stb.cpp:2147-2158 at mode (i*, 0, k*); i* from q_x[1], k* from q_z[1]
  const qsm_params& p = h->p;
  if (h->step < 2) { h->growth = 0; return QSM_OK; }
  const double Lx = p.nx * 2.0 * M_PI / (p.q0 * p.nw), Lz = p.nz * 2.0 * M_PI / (p.q0 * p.nw);
  const double qx1 = M_PI * 1 / ((Lx / p.nx) * p.nx), qz1 = M_PI * 1 / ((Lz / p.nz) * p.nz);
  const long is = lround(p.growth_q / qx1), ks = lround(p.q0 / qz1);
  if (is >= p.nx || ks >= p.nz) { h->growth = 0; return QSM_OK; }
  const size_t id = idx3(p, (int)is, 0, (int)ks);
  double a = 0, b = 0;
  HIPCHK(hipMemcpyAsync(&a, h->psihat + id, 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(&b, h->psihat_old + id, 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->growth = (a - b) / (p.dt * b);
  return QSM_OK;
}

void free_handle(qsm_handle_s* h) {
  if (!h) return;
  for (int f = 0; f < F_COUNT; ++f) free_padded(h->f[f], h->esz, h->pad);
  free_padded(h->psihat, 8, h->pad);
  free_padded(h->psihat_old, 8, h->pad);
  free_padded(h->psihat32, 4, h->pad);
  if (h->part) hipFree(h->part);
  if (h->dsc) hipFree(h->dsc);
  for (auto& g : h->graphs) hipGraphExecDestroy(g.second);
  h->graphs.clear();
  if (h->wtab) hipFree(h->wtab);
  if (h->ones) hipFree(h->ones);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

}  // namespace

extern "C" int qsm_default_params(int variant, qsm_params* p) {
  if (!p) return fail(QSM_ERR_ARG, "null params");
  std::memset(p, 0, sizeof *p);
  p->variant = variant;
  p->nw = 8; p->q0 = 1.0; p->alpha = 1.0; p->hydro = 1;
  p->filter = 1.57 * 1.57 * 8;
  p->compat_ranks = 0;
  if (variant == QSM_QUASI) {
    // quasi.cpp:159-219
    p->nx = p->ny = p->nz = 256;
    p->divv_switch = 200;
    p->gamma = 1.0; p->beta = 2.0; p->eps = 0.0; p->nu = 1.0;
    p->amp = 1.34164; p->rho0 = 0.005; p->kp = 0.5;
    p->rho_m = (p->kp * p->amp + p->rho0) / 2;
    p->dt = 0.001;
  } else if (variant == QSM_STB) {
    // quasi-stb.cpp:149-220
    p->nx = p->ny = p->nz = 512;
    p->divv_switch = 200;
    p->gamma = 3.0; p->beta = 0.4; p->eps = 0.0; p->nu = 0.1;
    p->amp = 0.3464101615; p->rho0 = 0.1; p->kp = 0.5; p->rho_m = 0.5;
    p->dt = 0.0005;
    p->growth_q = 0.5;
  } else {
    return fail(QSM_ERR_ARG, "unknown variant %d", variant);
  }
  return QSM_OK;
}

extern "C" int qsm_create(const qsm_params* pp, int prec, qsm_handle* out) {
  if (!pp || !out) return fail(QSM_ERR_ARG, "qsm_create: null argument");
  *out = nullptr;
  const qsm_params& p = *pp;
  if (p.nx < 2 || p.ny < 2 || p.nz < 2) return fail(QSM_ERR_ARG, "qsm_create: grid too small");
  if (p.variant != QSM_QUASI && p.variant != QSM_STB) return fail(QSM_ERR_ARG, "qsm_create: bad variant");
  if (prec != QSM_FP64 && prec != QSM_MIXED) return fail(QSM_ERR_ARG, "qsm_create: bad precision");
  if (p.compat_ranks < 0 || (p.compat_ranks > 0 && p.nx % p.compat_ranks != 0))
    return fail(QSM_ERR_ARG, "qsm_create: compat_ranks must divide nx");
  if (!(p.dt > 0) || !(p.nu > 0) || p.nw <= 0 || !(p.q0 > 0))
    return fail(QSM_ERR_ARG, "qsm_create: dt, nu, q0 and nw must be positive");
  // psihat is FP64 in both precisions and kernels address a field as a base
  // pointer plus a 32-bit byte offset: at most 2^29 points
  if ((long)p.nx * p.ny * p.nz > (1L << 29))
    return fail(QSM_ERR_UNSUPPORTED, "qsm_create: more than 2^29 grid points");
  const int nn[3] = {p.nx, p.ny, p.nz};
  for (int a = 0; a < 3; ++a)
    if (!step_supported(nn[a]))
      return fail(QSM_ERR_UNSUPPORTED, "qsm_create: no step kernels for length %d (compiled: 32, 48, 64, 96, 128, 256, 512, 1024)", nn[a]);

  auto* h = new qsm_handle_s;
  h->p = p;
  h->prec = prec;
  h->fp32 = prec == QSM_MIXED;
  h->esz = h->fp32 ? 4 : 8;
  h->n = (size_t)p.nx * p.ny * p.nz;
  h->geo = Geo{p.nx, p.ny, p.nz, 0};
  fill_phys(h);
  auto bail = [&](int code, const char* what) {
    free_handle(h);
    return fail(code, "qsm_create: %s", what);
  };
  if (hipStreamCreate(&h->stream) != hipSuccess) return bail(QSM_ERR_HIP, "stream");
  // wavenumbers, quasi.cpp:397-431 (same expressions: pi m/(dx N), pi (m+1)/(dx N)),
  // then the three 1-D factors of the density filter exp(-filter q^2)
  std::vector<double> w;
  for (int a = 0; a < 3; ++a) {
    const int N = nn[a];
    const double L = N * 2.0 * M_PI / (p.q0 * p.nw), d = L / N;
    for (int m = 0; m < N; ++m) w.push_back(1.0 * M_PI * m / (d * N));
    for (int m = 0; m < N; ++m) w.push_back(1.0 * M_PI * (m + 1) / (d * N));
  }
  const size_t nwav = w.size();
  for (int a = 0; a < 3; ++a) {
    const int N = nn[a];
    const double L = N * 2.0 * M_PI / (p.q0 * p.nw), d = L / N;
    for (int m = 0; m < N; ++m) {
      const long double q = 1.0 * M_PI * m / (d * N);
      w.push_back((double)expl(-(long double)p.filter * q * q));
    }
  }
  if (hipMalloc(&h->wtab, w.size() * 8) != hipSuccess) return bail(QSM_ERR_NOMEM, "wavenumber tables");
  if (hipMemcpy(h->wtab, w.data(), w.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
    return bail(QSM_ERR_HIP, "wavenumber upload");
  h->waves.qx = h->wtab; h->waves.sx = h->wtab + p.nx;
  h->waves.qy = h->wtab + 2 * p.nx; h->waves.sy = h->waves.qy + p.ny;
  h->waves.qz = h->wtab + 2 * p.nx + 2 * p.ny; h->waves.sz = h->waves.qz + p.nz;
  h->waves.fx = h->wtab + nwav; h->waves.fy = h->waves.fx + p.nx; h->waves.fz = h->waves.fy + p.ny;
  // the filter store also scales psihat at setup with a unit filter
  if (hipMalloc(&h->ones, (size_t)std::max(p.nx, std::max(p.ny, p.nz)) * 8) != hipSuccess)
    return bail(QSM_ERR_NOMEM, "unit filter");
  {
    std::vector<double> one((size_t)std::max(p.nx, std::max(p.ny, p.nz)), 1.0);
    if (hipMemcpy(h->ones, one.data(), one.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return bail(QSM_ERR_HIP, "unit filter upload");
  }
  for (int a = 0; a < 3; ++a) {
    h->tw[a] = twiddles(nn[a], h->fp32);
    if (!h->tw[a]) return bail(QSM_ERR_NOMEM, "twiddle table");
  }
  // Every field and psihat copy carries one y-z slab (ny nz elements) of
  // padding on either side: the spectral operators' shifted-neighbour reads
  // (+-1 along z, +-nz along y, +-ny nz along x, and sums of an x and a y or
  // z shift, all within one slab) stay inside the allocation at the field's
  // ends, so they read at a uniform offset with no address clamp and mask
  // the values instead (qsm_ops.hpp).
  h->pad = (size_t)p.ny * p.nz;
  for (int f = 0; f < F_COUNT; ++f)
    if (!(h->f[f] = alloc_padded(h->n, h->esz, h->pad))) return bail(QSM_ERR_NOMEM, "field");
  if (!(h->psihat = (double*)alloc_padded(h->n, 8, h->pad)) ||
      !(h->psihat_old = (double*)alloc_padded(h->n, 8, h->pad)))
    return bail(QSM_ERR_NOMEM, "psihat");
  if (h->fp32 && !(h->psihat32 = (float*)alloc_padded(h->n, 4, h->pad)))
    return bail(QSM_ERR_NOMEM, "psihat32");
  long nb = pass_blocks(h->fp32, p.nz, 2, h->geo);
  h->part_n = nb > 4096 ? nb : 4096;
  if (hipMalloc(&h->part, h->part_n * 4 * 8) != hipSuccess) return bail(QSM_ERR_NOMEM, "partials");
  if (hipMalloc(&h->dsc, D_COUNT * 8) != hipSuccess) return bail(QSM_ERR_NOMEM, "scalars");
  if (hipMemset(h->dsc, 0, D_COUNT * 8) != hipSuccess) return bail(QSM_ERR_HIP, "memset");
  *out = h;
  return QSM_OK;
}

extern "C" int qsm_set_psi(qsm_handle h, const double* psi) {
  if (!h || !psi) return fail(QSM_ERR_ARG, "qsm_set_psi: null argument");
  const size_t n = h->n;
  for (int f = 0; f < F_COUNT; ++f) HIPCHK(hipMemsetAsync(h->f[f], 0, n * h->esz, h->stream));
  HIPCHK(hipMemsetAsync(h->dsc, 0, D_COUNT * 8, h->stream));
  h->psi_cur = F_PSI_A; h->psi_old = F_PSI_B; h->rho_cur = F_RHO_A; h->rho_old = F_RHO_B;
  h->rhoy = 0;  // rhohat = 0: its passed fields are the zeros set above
  // psi (FP64 host) -> psihat = scale DCT(psi) in FP64, whatever the precision
  HIPCHK(hipMemcpyAsync(h->psihat_old, psi, n * 8, hipMemcpyHostToDevice, h->stream));
  // forward FP64 DCT of psi: z, y, x passes on psihat; scale in the last store
  const int saved_fp32 = h->fp32;
  h->fp32 = 0;
  const void* tw_saved[3] = {h->tw[0], h->tw[1], h->tw[2]};
  const int nn[3] = {h->p.nx, h->p.ny, h->p.nz};
  for (int a = 0; a < 3; ++a) h->tw[a] = twiddles(nn[a], 0);
  int rc = run_pass(h, 2, false, false, plain_load(h->psihat_old), plain_store(h->psihat), true, nullptr);
  if (!rc) rc = run_pass(h, 1, false, false, plain_load(h->psihat), plain_store(h->psihat), false, nullptr);
  // last pass: scale via the filter store with unit filter factors
  StoreOp S{}; S.code = ST_FILTER; S.d[0] = h->psihat;
  const Waves wsaved = h->waves;
  h->waves.fx = h->waves.fy = h->waves.fz = h->ones;
  if (!rc) rc = run_pass(h, 0, false, false, plain_load(h->psihat), S, false, nullptr);
  h->waves = wsaved;
  for (int a = 0; a < 3; ++a) h->tw[a] = tw_saved[a];
  h->fp32 = saved_fp32;
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h->psihat_old, h->psihat, n * 8, hipMemcpyDeviceToDevice, h->stream));
  if (h->fp32) HIPCHK(launch_round32(h->psihat, h->psihat32, (long)n, h->stream));
  // psi = IDCT(psihat) (quasi.cpp:924-927), psi_old = psi
  if ((rc = psi_inverse(h, F(h, h->psi_cur)))) return rc;
  HIPCHK(hipMemcpyAsync(F(h, h->psi_old), F(h, h->psi_cur), n * h->esz, hipMemcpyDeviceToDevice, h->stream));
  if (h->p.variant == QSM_STB) {
    // rho = 1 in the pure smectic (quasi-stb.cpp:1232): kept for the mass diagnostic
    if (h->fp32) {
      std::vector<float> ones(n, 1.0f);
      HIPCHK(hipMemcpy(F(h, h->rho_cur), ones.data(), n * 4, hipMemcpyHostToDevice));
    } else {
      std::vector<double> ones(n, 1.0);
      HIPCHK(hipMemcpy(F(h, h->rho_cur), ones.data(), n * 8, hipMemcpyHostToDevice));
    }
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  h->step = 0;
  h->growth = 0;
  h->divv_valid = false;
  h->p_valid = false;
  h->r_valid = true;
  h->l_yspec = h->r_yspec = h->p_yspec = false;
  h->ready = true;
  return QSM_OK;
}

namespace {

// Everything a step's kernel arguments depend on: the buffers behind the
// ping-pong fields and the phase flags of step n = h->step + 1.  The stb step
// reads no divergence or pressure (quasi-stb.cpp:1273-1282 are commented
// out), so its key leaves out the divv_switch flags: one signature serves
// every step after the first.
std::vector<long long> step_key(const qsm_handle_s* h) {
  const long long n = h->step + 1;
  const qsm_params& p = h->p;
  const bool quasi = p.variant == QSM_QUASI;
  return {(long long)h->psi_cur, (long long)h->rho_cur, (long long)(uintptr_t)h->psihat,
          (long long)h->rhoy, n == 1, n >= 2, quasi && n > p.divv_switch,
          quasi && n > p.divv_switch + 1};
}

int one_step(qsm_handle_s* h) {
  return h->p.variant == QSM_STB ? stb_step(h) : quasi_step(h);
}

// One step: replay its graph if this signature was captured before; else
// capture it (the capture also performs the host-side bookkeeping), then
// launch.  Timing (per-launch events) needs eager launches.
int graph_step(qsm_handle_s* h) {
  if (!h->use_graphs || h->timer.on) return one_step(h);
  const auto key = step_key(h);
  auto it = h->graphs.find(key);
  if (it != h->graphs.end()) {
    h->dry = true;
    const int rc = one_step(h);  // bookkeeping only
    h->dry = false;
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(it->second, h->stream));
    return QSM_OK;
  }
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = one_step(h);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(QSM_ERR_HIP, "step capture: %s", hipGetErrorString(e));
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return fail(QSM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  h->graphs[key] = ex;
  HIPCHK(hipGraphLaunch(ex, h->stream));
  return QSM_OK;
}

}  // namespace

extern "C" int qsm_step(qsm_handle h, int nsteps) {
  if (!h) return fail(QSM_ERR_ARG, "qsm_step: null handle");
  if (!h->ready) return fail(QSM_ERR_STATE, "qsm_step: call qsm_set_psi first");
  if (nsteps < 0) return fail(QSM_ERR_ARG, "qsm_step: negative step count");
  for (int s = 0; s < nsteps; ++s) {
    int rc = graph_step(h);
    if (rc) return rc;
  }
  if (h->p.variant == QSM_STB && nsteps > 0) {
    int rc = stb_growth(h);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return QSM_OK;
}

extern "C" int qsm_set_graphs(qsm_handle h, int enable) {
  if (!h) return fail(QSM_ERR_ARG, "qsm_set_graphs: null handle");
  h->use_graphs = enable != 0;
  return QSM_OK;
}

extern "C" int qsm_get_diag(qsm_handle h, qsm_diag* d) {
  if (!h || !d) return fail(QSM_ERR_ARG, "qsm_get_diag: null argument");
  if (!h->ready) return fail(QSM_ERR_STATE, "qsm_get_diag: no state");
  const qsm_params& p = h->p;
  int nb = (int)std::min<long>(1024, (long)(h->n + 255) / 256);
  if (nb > h->part_n) nb = (int)h->part_n;
  hipError_t e = launch_diag(h->fp32, F(h, h->psi_cur), F(h, h->psi_old), F(h, h->rho_cur), F(h, F_F),
                             (long)h->n, h->phys, h->part, nb, h->dsc + D_DIAG, h->stream);
  if (e != hipSuccess) return fail(QSM_ERR_HIP, "diag launch: %s", hipGetErrorString(e));
  double s[4];
  HIPCHK(hipMemcpyAsync(s, h->dsc + D_DIAG, 32, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const double dx = 2.0 * M_PI / (p.q0 * p.nw);
  d->L1 = s[0] / (p.dt * s[1]);
  d->mass = s[2] * dx * dx * dx;
  d->energy = s[3] * dx * dx * dx;
  const double vol = (double)p.nx * p.ny * p.nz * dx * dx * dx;
  const double rs = p.kp * p.amp + p.rho0;
  const double Lz = p.nz * dx;
  d->interface_z = (rs != p.rho0) ? Lz * (d->mass / vol - p.rho0) / (rs - p.rho0) : 0.0;
  d->growth = h->growth;
  d->step = h->step;
  if (!std::isfinite(d->L1) && h->step > 0) return fail(QSM_ERR_NONFINITE, "L1 is not finite at step %lld", h->step);
  if (!std::isfinite(d->mass) || !std::isfinite(d->energy))
    return fail(QSM_ERR_NONFINITE, "diagnostics not finite at step %lld", h->step);
  return QSM_OK;
}

extern "C" int qsm_run_l1(qsm_handle h, int max_steps, int step_l1, double lim_l1, int* steps_done) {
  if (!h || step_l1 <= 0 || max_steps < 0) return fail(QSM_ERR_ARG, "qsm_run_l1: bad argument");
  int done = 0;
  double L1 = 1.0;  // quasi.cpp:963
  while (L1 > lim_l1 && done < max_steps) {
    const int k = std::min(step_l1, max_steps - done);
    int rc = qsm_step(h, k);
    if (rc) { if (steps_done) *steps_done = done; return rc; }
    done += k;
    if (k == step_l1) {
      qsm_diag d;
      rc = qsm_get_diag(h, &d);
      if (rc) { if (steps_done) *steps_done = done; return rc; }
      L1 = d.L1;
    }
  }
  if (steps_done) *steps_done = done;
  return QSM_OK;
}

extern "C" int qsm_get_field(qsm_handle h, int which, double* host) {
  if (!h || !host) return fail(QSM_ERR_ARG, "qsm_get_field: null argument");
  const void* src = nullptr;
  bool is_double = !h->fp32;
  switch (which) {
    case QSM_PSI: src = F(h, h->psi_cur); break;
    case QSM_PSI_OLD: src = F(h, h->psi_old); break;
    case QSM_RHO: src = F(h, h->rho_cur); break;
    case QSM_VX: src = F(h, F_VX); break;
    case QSM_VY: src = F(h, F_VY); break;
    case QSM_VZ: src = F(h, F_VZ); break;
    case QSM_P: {
      if (h->p.variant == QSM_QUASI) {
        const int rc = pressure_output(h);
        if (rc) return rc;
      }
      src = F(h, F_P);
      break;
    }
    case QSM_DIVV: {
      const int rc = divv_output(h);
      if (rc) return rc;
      src = F(h, F_DIVV);
      break;
    }
    case QSM_PSIHAT: src = h->psihat; is_double = true; break;
    case QSM_PSIHAT_OLD: src = h->psihat_old; is_double = true; break;
    case QSM_F: src = F(h, F_F); break;
    case QSM_G: src = F(h, F_G); break;
    case QSM_GX: case QSM_GY: case QSM_GZ: src = F(h, F_GX + (which - QSM_GX)); break;
    case QSM_RX: case QSM_RY: case QSM_RZ:
      if (h->p.variant == QSM_QUASI && !h->r_valid)
        return fail(QSM_ERR_STATE,
                    "qsm_get_field: grad rho is formed at step 1 and after divv_switch only (step %lld)",
                    h->step);
      {
        const int rc = gradrho_output(h);
        if (rc) return rc;
      }
      src = F(h, F_RX + (which - QSM_RX));
      break;
    case QSM_LX: case QSM_LY: case QSM_LZ: {
      const int rc = lapgrad_output(h);
      if (rc) return rc;
      src = F(h, F_LX + (which - QSM_LX));
      break;
    }
    case QSM_RHOHAT: src = F(h, F_RHOHAT); break;
    case QSM_NHAT: src = F(h, F_NHAT_OLD); break;
    case QSM_SX: case QSM_SY: case QSM_SZ: src = F(h, F_SX + (which - QSM_SX)); break;
    case QSM_DHAT: src = F(h, F_DHAT); break;
    case QSM_DTPSI: {
      if (h->p.variant != QSM_STB)
        return fail(QSM_ERR_ARG, "qsm_get_field: dtpsi is an output of the stb variant only");
      if (h->step < 1) return fail(QSM_ERR_STATE, "qsm_get_field: dtpsi exists after a step");
      const int rc = dtpsi_output(h);
      if (rc) return rc;
      src = F(h, F_NTMP);
      break;
    }
    default: return fail(QSM_ERR_ARG, "qsm_get_field: unknown field %d", which);
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  if (is_double) {
    HIPCHK(hipMemcpy(host, src, h->n * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(h->n);
    HIPCHK(hipMemcpy(tmp.data(), src, h->n * 4, hipMemcpyDeviceToHost));
    for (size_t t = 0; t < h->n; ++t) host[t] = tmp[t];
  }
  // mixed precision keeps the (0,0,0) modes in FP64 beside the FP32 arrays
  if (h->fp32 && (which == QSM_RHOHAT || which == QSM_NHAT)) {
    double v;
    HIPCHK(hipMemcpy(&v, h->dsc + (which == QSM_RHOHAT ? D_RHOHAT0 : D_NHAT0), 8, hipMemcpyDeviceToHost));
    if (h->step > 0) host[0] = v;
  }
  return QSM_OK;
}

extern "C" int qsm_destroy(qsm_handle h) {
  if (h) hipStreamSynchronize(h->stream);
  free_handle(h);
  return QSM_OK;
}

extern "C" int qsm_timing_enable(qsm_handle h, int enable) {
  if (!h) return fail(QSM_ERR_ARG, "null handle");
  h->timer.reset();
  h->timer.on = enable != 0;
  h->timer.only = -1;
  return QSM_OK;
}

extern "C" int qsm_timing_select(qsm_handle h, int axis, int inv, int ld, int st, int variant) {
  if (!h) return fail(QSM_ERR_ARG, "null handle");
  if (axis < 0 || axis > 2 || ld < 0 || ld >= 64 || st < 0 || st >= 16 || variant < 0 ||
      variant >= kVariants)
    return fail(QSM_ERR_ARG, "qsm_timing_select: bad kernel (axis %d ld %d st %d variant %d)",
                axis, ld, st, variant);
  h->timer.reset();
  h->timer.on = true;
  h->timer.only = timer_key(axis, inv != 0, ld, st, variant);
  return QSM_OK;
}

extern "C" int qsm_timing_get(qsm_handle h, int which, double* total_ms, long long* launches,
                              double* alg_bytes) {
  if (!h || which < 0 || which > 6) return fail(QSM_ERR_ARG, "qsm_timing_get: bad argument");
  h->timer.drain();
  double ms = 0, by = 0;
  long long c = 0;
  for (const auto& kv : h->timer.acc) {
    const int axis = kv.first / (64 * 16 * 2) / 2, inv = (kv.first / (64 * 16 * 2)) % 2;
    const int q = 1 + axis + (inv ? 0 : 3);
    if (which == 0 || which == q) { ms += kv.second.ms; by += kv.second.bytes; c += kv.second.count; }
  }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = c;
  if (alg_bytes) *alg_bytes = by;
  return QSM_OK;
}

extern "C" int qsm_timing_kernel(qsm_handle h, int rank, int* axis, int* inv, int* ld, int* st,
                                 int* variant, double* total_ms, long long* launches,
                                 double* alg_bytes) {
  if (!h || rank < 0) return fail(QSM_ERR_ARG, "qsm_timing_kernel: bad argument");
  h->timer.drain();
  std::vector<std::pair<double, int>> order;
  for (const auto& kv : h->timer.acc) order.push_back({-kv.second.ms, kv.first});
  std::sort(order.begin(), order.end());
  if (rank >= (int)order.size()) return fail(QSM_ERR_ARG, "qsm_timing_kernel: only %zu kernels timed", order.size());
  const int key = order[rank].second;
  const auto& a = h->timer.acc[key];
  if (axis) *axis = key / (64 * 16 * kVariants) / 2;
  if (inv) *inv = (key / (64 * 16 * kVariants)) % 2;
  if (ld) *ld = (key / (16 * kVariants)) % 64;
  if (st) *st = (key / kVariants) % 16;
  if (variant) *variant = key % kVariants;
  if (total_ms) *total_ms = a.ms;
  if (launches) *launches = a.count;
  if (alg_bytes) *alg_bytes = a.bytes;
  return QSM_OK;
}

// ============================================================================
// ensembles: many independent instances (e.g. quasi-stb.cpp stability runs of
// different grid sizes) advanced together.  Every instance records its step's
// launches; launch q of all instances becomes one batched kernel per distinct
// pencil length, and the whole batched step is captured as one HIP graph per
// step signature (the same signature rules as single instances).

namespace {

struct EnsLaunch {
  int kind = 0, axis = 0, n = 0, inv = 0, ld = 0, st = 0, count = 0, variant = 0;
  int q = 0;                  // launch of the instances' step sequence this serves
  long total = 0;
  double bytes = 0;           // algorithmic HBM bytes over the launch's instances
  PassArgs* descs = nullptr;  // device
  int* prefix = nullptr;      // device
  const double** parts = nullptr;  // device (finish)
  long* nparts = nullptr;
  double** outs = nullptr;
};

struct EnsPlan {
  std::vector<EnsLaunch> launches;
  std::vector<void*> allocs;
  std::vector<hipEvent_t> events;  // fork and join events, one per launch
  hipGraphExec_t exec = nullptr;
};

// Largest number of steps one ensemble graph launch may chain.
constexpr int kMaxGraphSteps = 64;

// Side streams of an ensemble: the launches of one pass for different pencil
// lengths touch disjoint instances, so up to kEnsSide + 1 of them run
// side by side.  Default priority, like every stream here.
constexpr int kEnsSide = 4;

}  // namespace

struct qsm_ens_s {
  std::vector<qsm_handle_s*> inst;
  int prec = 0, fp32 = 0;
  hipStream_t stream = nullptr;
  hipStream_t side[kEnsSide] = {};
  bool use_graphs = true;
  std::map<std::vector<long long>, EnsPlan*> plans;
  // multi-step graphs: graph_steps steps per launch once the step sequence
  // repeats (keyed by the first step's signature and the step count); the
  // signatures of the last two steps run one at a time
  int graph_steps = 1;
  std::vector<long long> last_key[2];
  std::map<std::vector<long long>, hipGraphExec_t> blocks;
  // per-launch HIP-event timing (eager launches while on), keyed by
  // ens_timer_key: the pass kernel and its pencil length
  Timer timer;
  std::map<int, int> variant[3][2];  // tile-width variant of plain passes [axis][inv][length]
};

namespace {

template <typename X>
int upload(EnsPlan* pl, const std::vector<X>& v, X** out) {
  void* d = nullptr;
  HIPCHK(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(X)));
  pl->allocs.push_back(d);
  HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
  *out = (X*)d;
  return QSM_OK;
}

// timer key of an ensemble launch: the pass kernel (timer_key, with load
// code 63 for the batched FP64 finish) and its pencil length
constexpr int kEnsLenMod = 2048;
int ens_timer_key(const EnsLaunch& L) {
  const int base = L.kind == 0 ? timer_key(L.axis, L.inv != 0, L.ld, L.st, L.variant)
                               : timer_key(0, false, 63, 0, 0);
  return base * kEnsLenMod + (L.kind == 0 ? L.n : 0);
}

hipError_t ens_launch_raw(const qsm_ens_s* e, const EnsLaunch& L, hipStream_t s);

// one batched launch of a plan on stream s, with events around it while the
// ensemble's timer is on
hipError_t ens_launch(qsm_ens_s* e, const EnsLaunch& L, hipStream_t s) {
  Timer& T = e->timer;
  if (!T.on) return ens_launch_raw(e, L, s);
  hipEvent_t ea = T.get(), eb = T.get();
  (void)hipEventRecord(ea, s);
  const hipError_t err = ens_launch_raw(e, L, s);
  (void)hipEventRecord(eb, s);
  T.recs.push_back({ea, eb, ens_timer_key(L), L.bytes});
  return err;
}

hipError_t ens_launch_raw(const qsm_ens_s* e, const EnsLaunch& L, hipStream_t s) {
  if (L.kind == 0) {
    BatchLaunch B{L.descs, L.prefix, L.count, L.total};
    PassArgs dummy{};
    return launch_pass(e->fp32, L.n, L.axis, L.inv != 0, L.ld, L.st, dummy, nullptr, s, &B,
                       L.variant);
  }
  return launch_finish_batch(L.parts, L.nparts, L.count, L.outs, s);
}

// The launches of one step, in order.  Those serving one launch q of the
// instances' step sequence (one per pencil length) touch disjoint instances:
// the first runs on the ensemble stream, up to kEnsSide more on side streams
// forked from it, and all are joined back before the next launch starts.
int ens_issue(qsm_ens_s* e, const EnsPlan* pl) {
  const size_t nl = pl->launches.size();
  size_t a = 0;
  while (a < nl) {
    size_t b = a + 1;
    while (b < nl && pl->launches[b].q == pl->launches[a].q && b - a <= (size_t)kEnsSide) ++b;
    hipError_t err = hipSuccess;
    if (b - a > 1) err = hipEventRecord(pl->events[a], e->stream);  // fork
    for (size_t i = a + 1; i < b && err == hipSuccess; ++i) {
      const hipStream_t side = e->side[i - a - 1];
      err = hipStreamWaitEvent(side, pl->events[a], 0);
      if (err == hipSuccess) err = ens_launch(e, pl->launches[i], side);
      if (err == hipSuccess) err = hipEventRecord(pl->events[i], side);  // join
    }
    if (err == hipSuccess) err = ens_launch(e, pl->launches[a], e->stream);
    for (size_t i = a + 1; i < b && err == hipSuccess; ++i)
      err = hipStreamWaitEvent(e->stream, pl->events[i], 0);
    if (err != hipSuccess) return fail(QSM_ERR_HIP, "ensemble launch: %s", hipGetErrorString(err));
    a = b;
  }
  return QSM_OK;
}

int ens_build(qsm_ens_s* e, EnsPlan* pl) {
  const int count = (int)e->inst.size();
  std::vector<std::vector<PassRec>> recs(count);
  for (int i = 0; i < count; ++i) {
    e->inst[i]->rec = &recs[i];
    int rc = one_step(e->inst[i]);
    e->inst[i]->rec = nullptr;
    if (rc) return rc;
  }
  const size_t Q = recs[0].size();
  for (int i = 1; i < count; ++i) {
    if (recs[i].size() != Q) return fail(QSM_ERR_STATE, "ensemble instances disagree on the step sequence");
    for (size_t q = 0; q < Q; ++q) {
      const PassRec &a = recs[0][q], &b = recs[i][q];
      if (a.kind != b.kind || a.axis != b.axis || a.inv != b.inv || a.ld != b.ld || a.st != b.st)
        return fail(QSM_ERR_STATE, "ensemble instances disagree on launch %zu", q);
    }
  }
  for (size_t q = 0; q < Q; ++q) {
    const PassRec& r0 = recs[0][q];
    if (r0.kind == 1) {
      std::vector<const double*> parts(count);
      std::vector<long> np(count);
      std::vector<double*> outs(count);
      for (int i = 0; i < count; ++i) { parts[i] = recs[i][q].part; np[i] = recs[i][q].nparts; outs[i] = recs[i][q].out; }
      EnsLaunch L;
      L.kind = 1; L.count = count; L.q = (int)q;
      int rc = upload(pl, parts, &L.parts);
      if (!rc) rc = upload(pl, np, &L.nparts);
      if (!rc) rc = upload(pl, outs, &L.outs);
      if (rc) return rc;
      pl->launches.push_back(L);
      continue;
    }
    std::map<int, std::vector<int>> groups;  // pencil length -> instances
    for (int i = 0; i < count; ++i) groups[recs[i][q].n].push_back(i);
    for (auto& g : groups) {
      const bool plain = r0.ld == LD_PLAIN && r0.st == ST_PLAIN;
      const int var = plain ? e->variant[r0.axis][r0.inv ? 1 : 0][g.first] : 0;
      std::vector<PassArgs> descs;
      std::vector<int> prefix(1, 0);
      for (int i : g.second) {
        const PassRec& r = recs[i][q];
        descs.push_back(r.A);
        prefix.push_back(prefix.back() + (int)pass_tiles(e->fp32, r.n, r.axis, r.A.geo, var));
      }
      EnsLaunch L;
      L.kind = 0; L.axis = r0.axis; L.n = g.first; L.inv = r0.inv; L.ld = r0.ld; L.st = r0.st;
      L.q = (int)q;
      L.variant = var;
      L.count = (int)g.second.size(); L.total = prefix.back();
      for (int i : g.second) L.bytes += recs[i][q].bytes;
      int rc = upload(pl, descs, &L.descs);
      if (!rc) rc = upload(pl, prefix, &L.prefix);
      if (rc) return rc;
      pl->launches.push_back(L);
    }
  }
  for (size_t i = 0; i < pl->launches.size(); ++i) {
    hipEvent_t ev = nullptr;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    pl->events.push_back(ev);
  }
  return QSM_OK;
}

void ens_free_plan(EnsPlan* pl) {
  if (pl->exec) hipGraphExecDestroy(pl->exec);
  for (hipEvent_t ev : pl->events) hipEventDestroy(ev);
  for (void* d : pl->allocs) hipFree(d);
  delete pl;
}

int ens_one_step(qsm_ens_s* e) {
  const auto key = step_key(e->inst[0]);
  e->last_key[0] = e->last_key[1];
  e->last_key[1] = key;
  auto it = e->plans.find(key);
  if (it != e->plans.end()) {
    for (auto* h : e->inst) {
      h->dry = true;
      const int rc = one_step(h);
      h->dry = false;
      if (rc) return rc;
    }
    EnsPlan* pl = it->second;
    if (pl->exec && !e->timer.on) { HIPCHK(hipGraphLaunch(pl->exec, e->stream)); return QSM_OK; }
    return ens_issue(e, pl);
  }
  auto* pl = new EnsPlan;
  int rc = ens_build(e, pl);  // also advances every instance's bookkeeping
  if (rc) { ens_free_plan(pl); return rc; }
  e->plans[key] = pl;
  if (!e->use_graphs || e->timer.on) return ens_issue(e, pl);
  HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  rc = ens_issue(e, pl);
  hipGraph_t g = nullptr;
  hipError_t err = hipStreamEndCapture(e->stream, &g);
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  if (err != hipSuccess) return fail(QSM_ERR_HIP, "ensemble capture: %s", hipGetErrorString(err));
  err = hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (err != hipSuccess) return fail(QSM_ERR_HIP, "ensemble graph instantiate: %s", hipGetErrorString(err));
  HIPCHK(hipGraphLaunch(pl->exec, e->stream));
  return QSM_OK;
}

// Signature bits of step n that depend on n alone (the rest of step_key is
// buffer pointers, which alternate with period 2 from step 2 on).
int phase_bits(const qsm_params& p, long long n) {
  const bool quasi = p.variant == QSM_QUASI;
  return (n == 1 ? 1 : 0) | (n >= 2 ? 2 : 0) | (quasi && n > p.divv_switch ? 4 : 0) |
         (quasi && n > p.divv_switch + 1 ? 8 : 0);
}

void ens_free_blocks(qsm_ens_s* e) {
  for (auto& kv : e->blocks) (void)hipGraphExecDestroy(kv.second);
  e->blocks.clear();
}

// K steps (K even) as one graph launch, when the next K steps repeat the
// last two: the signature of the coming step equals that of two steps ago
// and no phase bit changes from two steps ago to the end of the block.  Then
// the block's steps alternate between the two captured single-step plans,
// and the block graph chains their launches K/2 times.  *done: the steps
// advanced (0: not applicable, the caller steps one at a time).
int ens_block(qsm_ens_s* e, int K, int* done) {
  *done = 0;
  qsm_handle_s* h0 = e->inst[0];
  const long long n0 = h0->step + 1;
  if (K < 2 || (K & 1) || n0 < 3 || e->last_key[0].empty()) return QSM_OK;
  const auto key = step_key(h0);
  if (key != e->last_key[0]) return QSM_OK;
  const int ph = phase_bits(h0->p, n0 - 2);
  for (long long n = n0 - 1; n < n0 + K; ++n)
    if (phase_bits(h0->p, n) != ph) return QSM_OK;
  auto pa = e->plans.find(e->last_key[0]), pb = e->plans.find(e->last_key[1]);
  if (pa == e->plans.end() || pb == e->plans.end()) return QSM_OK;
  std::vector<long long> bkey = key;
  bkey.push_back(K);
  auto dry_step = [&]() -> int {
    for (auto* h : e->inst) {
      h->dry = true;
      const int rc = one_step(h);
      h->dry = false;
      if (rc) return rc;
    }
    return QSM_OK;
  };
  auto it = e->blocks.find(bkey);
  if (it == e->blocks.end()) {
    // capture: the bookkeeping of every step (no HIP calls) and the launches
    // of its plan
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = QSM_OK;
    for (int j = 0; j < K && rc == QSM_OK; ++j) {
      rc = dry_step();
      if (rc == QSM_OK) rc = ens_issue(e, (j & 1) ? pb->second : pa->second);
    }
    hipGraph_t g = nullptr;
    hipError_t err = hipStreamEndCapture(e->stream, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (err != hipSuccess) return fail(QSM_ERR_HIP, "ensemble block capture: %s", hipGetErrorString(err));
    hipGraphExec_t ex = nullptr;
    err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (err != hipSuccess) return fail(QSM_ERR_HIP, "ensemble block instantiate: %s", hipGetErrorString(err));
    it = e->blocks.emplace(bkey, ex).first;
  } else {
    for (int j = 0; j < K; ++j) {
      const int rc = dry_step();
      if (rc) return rc;
    }
  }
  HIPCHK(hipGraphLaunch(it->second, e->stream));
  // the signatures of the block's last two steps, for the next block
  e->last_key[0] = (K & 1) ? pb->first : pa->first;
  e->last_key[1] = (K & 1) ? pa->first : pb->first;
  *done = K;
  return QSM_OK;
}

}  // namespace

extern "C" int qsm_ens_create(const qsm_params* params, int count, int prec, qsm_ens* out) {
  if (!params || !out || count <= 0) return fail(QSM_ERR_ARG, "qsm_ens_create: bad argument");
  *out = nullptr;
  for (int i = 1; i < count; ++i)
    if (params[i].variant != params[0].variant || params[i].divv_switch != params[0].divv_switch ||
        params[i].hydro != params[0].hydro)
      return fail(QSM_ERR_ARG, "qsm_ens_create: instances must share variant, divv_switch and hydro");
  auto* e = new qsm_ens_s;
  e->prec = prec;
  e->fp32 = prec == QSM_MIXED;
  for (int i = 0; i < count; ++i) {
    qsm_handle h = nullptr;
    int rc = qsm_create(&params[i], prec, &h);
    if (rc) { qsm_ens_destroy(e); return rc; }
    e->inst.push_back(h);
  }
  if (hipStreamCreate(&e->stream) != hipSuccess) { qsm_ens_destroy(e); return fail(QSM_ERR_HIP, "stream"); }
  for (auto& side : e->side)
    if (hipStreamCreate(&side) != hipSuccess) { qsm_ens_destroy(e); return fail(QSM_ERR_HIP, "stream"); }
  *out = e;
  return QSM_OK;
}

extern "C" int qsm_ens_set_psi(qsm_ens e, int index, const double* psi) {
  if (!e || index < 0 || index >= (int)e->inst.size()) return fail(QSM_ERR_ARG, "qsm_ens_set_psi: bad index");
  for (auto* h : e->inst)
    if (h->step != 0) return fail(QSM_ERR_STATE, "qsm_ens_set_psi: the ensemble has already stepped");
  return qsm_set_psi(e->inst[index], psi);
}

extern "C" int qsm_ens_step(qsm_ens e, int nsteps) {
  if (!e || nsteps < 0) return fail(QSM_ERR_ARG, "qsm_ens_step: bad argument");
  for (auto* h : e->inst)
    if (!h->ready) return fail(QSM_ERR_STATE, "qsm_ens_step: set psi of every instance first");
  for (int s = 0; s < nsteps;) {
    if (e->use_graphs && !e->timer.on && e->graph_steps > 1 && nsteps - s >= e->graph_steps) {
      int done = 0;
      const int rc = ens_block(e, e->graph_steps, &done);
      if (rc) return rc;
      if (done) { s += done; continue; }
    }
    const int rc = ens_one_step(e);
    if (rc) return rc;
    ++s;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  if (nsteps > 0 && e->inst[0]->p.variant == QSM_STB)
    for (auto* h : e->inst) {
      int rc = stb_growth(h);
      if (rc) return rc;
    }
  return QSM_OK;
}

extern "C" int qsm_ens_growth(qsm_ens e, double* out) {
  if (!e || !out) return fail(QSM_ERR_ARG, "qsm_ens_growth: null argument");
  for (size_t i = 0; i < e->inst.size(); ++i) out[i] = e->inst[i]->growth;
  return QSM_OK;
}

extern "C" int qsm_ens_get_field(qsm_ens e, int index, int which, double* host) {
  if (!e || index < 0 || index >= (int)e->inst.size()) return fail(QSM_ERR_ARG, "qsm_ens_get_field: bad index");
  HIPCHK(hipStreamSynchronize(e->stream));
  return qsm_get_field(e->inst[index], which, host);
}

extern "C" int qsm_ens_get_diag(qsm_ens e, int index, qsm_diag* d) {
  if (!e || index < 0 || index >= (int)e->inst.size()) return fail(QSM_ERR_ARG, "qsm_ens_get_diag: bad index");
  HIPCHK(hipStreamSynchronize(e->stream));
  return qsm_get_diag(e->inst[index], d);
}

extern "C" int qsm_ens_set_graphs(qsm_ens e, int enable) {
  if (!e) return fail(QSM_ERR_ARG, "qsm_ens_set_graphs: null ensemble");
  e->use_graphs = enable != 0;
  return QSM_OK;
}

extern "C" int qsm_ens_timing_enable(qsm_ens e, int enable) {
  if (!e) return fail(QSM_ERR_ARG, "qsm_ens_timing_enable: null ensemble");
  e->timer.reset();
  e->timer.on = enable != 0;
  return QSM_OK;
}

extern "C" int qsm_ens_timing_kernel(qsm_ens e, int rank, int* axis, int* inv, int* ld, int* st,
                                     int* variant, int* n, double* total_ms, long long* launches,
                                     double* alg_bytes) {
  if (!e || rank < 0) return fail(QSM_ERR_ARG, "qsm_ens_timing_kernel: bad argument");
  e->timer.drain();
  std::vector<std::pair<double, int>> order;
  for (const auto& kv : e->timer.acc) order.push_back({-kv.second.ms, kv.first});
  std::sort(order.begin(), order.end());
  if (rank >= (int)order.size())
    return fail(QSM_ERR_ARG, "qsm_ens_timing_kernel: only %zu kernels timed", order.size());
  const int full = order[rank].second, key = full / kEnsLenMod;
  const auto& a = e->timer.acc[full];
  if (axis) *axis = key / (64 * 16 * kVariants) / 2;
  if (inv) *inv = (key / (64 * 16 * kVariants)) % 2;
  if (ld) *ld = (key / (16 * kVariants)) % 64;
  if (st) *st = (key / kVariants) % 16;
  if (variant) *variant = key % kVariants;
  if (n) *n = full % kEnsLenMod;
  if (total_ms) *total_ms = a.ms;
  if (launches) *launches = a.count;
  if (alg_bytes) *alg_bytes = a.bytes;
  return QSM_OK;
}

extern "C" int qsm_ens_set_graph_steps(qsm_ens e, int steps) {
  if (!e || steps < 1 || (steps > 1 && (steps & 1)) || steps > kMaxGraphSteps)
    return fail(QSM_ERR_ARG, "qsm_ens_set_graph_steps: steps must be 1 or even, at most %d",
                kMaxGraphSteps);
  e->graph_steps = steps;
  return QSM_OK;
}

extern "C" int qsm_ens_destroy(qsm_ens e) {
  if (!e) return QSM_OK;
  if (e->stream) hipStreamSynchronize(e->stream);
  ens_free_blocks(e);
  for (auto& kv : e->plans) ens_free_plan(kv.second);
  for (auto* h : e->inst) qsm_destroy(h);
  for (auto side : e->side)
    if (side) { hipStreamSynchronize(side); hipStreamDestroy(side); }
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return QSM_OK;
}

// ============================================================================
// full-state checkpoint.  The reference restarts from psi<rank>.dat text files
// (6 significant digits, psi only: quasi.cpp:692-717, 2051-2062), which loses
// psihat_old, N_old, the lagged rhohat, rho_old, v and the step count.  This
// writes every buffer the step reads, in logical order, so a resumed run
// continues bit for bit.

namespace {

const char kCkptMagic[8] = {'Q', 'S', 'M', 'C', 'K', 'P', 'T', '1'};

struct CkptHeader {
  char magic[8];
  qsm_params p;
  int prec;
  int pad_;
  long long step;
  long long npts;
};

}  // namespace

extern "C" int qsm_checkpoint_write(qsm_handle h, const char* path) {
  if (!h || !path) return fail(QSM_ERR_ARG, "qsm_checkpoint_write: null argument");
  if (!h->ready) return fail(QSM_ERR_STATE, "qsm_checkpoint_write: no state");
  // every field in real space (the reader takes them so)
  int lrc = lapgrad_output(h);
  if (!lrc) lrc = gradrho_output(h);
  if (!lrc && h->p_yspec) lrc = pressure_output(h);
  if (lrc) return lrc;
  HIPCHK(hipStreamSynchronize(h->stream));
  FILE* fp = fopen(path, "wb");
  if (!fp) return fail(QSM_ERR_ARG, "qsm_checkpoint_write: cannot open %s", path);
  CkptHeader hd{};
  memcpy(hd.magic, kCkptMagic, 8);
  hd.p = h->p; hd.prec = h->prec; hd.step = h->step; hd.npts = (long long)h->n;
  bool ok = fwrite(&hd, sizeof hd, 1, fp) == 1;
  std::vector<char> buf(h->n * 8);
  auto put = [&](const void* d, size_t bytes) {
    if (!ok) return;
    ok = hipMemcpy(buf.data(), d, bytes, hipMemcpyDeviceToHost) == hipSuccess &&
         fwrite(buf.data(), 1, bytes, fp) == bytes;
  };
  put(h->psihat, h->n * 8);
  put(h->psihat_old, h->n * 8);
  put(h->dsc, D_COUNT * 8);
  const int order[] = {h->psi_cur, h->psi_old, h->rho_cur, h->rho_old};
  for (int f : order) put(h->f[f], h->n * h->esz);
  for (int f = 0; f < F_CKPT; ++f)
    if (f != F_PSI_A && f != F_PSI_B && f != F_RHO_A && f != F_RHO_B) put(h->f[f], h->n * h->esz);
  ok = (fclose(fp) == 0) && ok;
  return ok ? QSM_OK : fail(QSM_ERR_HIP, "qsm_checkpoint_write: write failed (%s)", path);
}

extern "C" int qsm_checkpoint_read(qsm_handle h, const char* path) {
  if (!h || !path) return fail(QSM_ERR_ARG, "qsm_checkpoint_read: null argument");
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(QSM_ERR_ARG, "qsm_checkpoint_read: cannot open %s", path);
  CkptHeader hd{};
  bool ok = fread(&hd, sizeof hd, 1, fp) == 1 && memcmp(hd.magic, kCkptMagic, 8) == 0;
  if (!ok) { fclose(fp); return fail(QSM_ERR_ARG, "qsm_checkpoint_read: %s is not a checkpoint", path); }
  if (memcmp(&hd.p, &h->p, sizeof(qsm_params)) != 0 || hd.prec != h->prec || hd.npts != (long long)h->n) {
    fclose(fp);
    return fail(QSM_ERR_ARG, "qsm_checkpoint_read: parameters or precision differ from the handle's");
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<char> buf(h->n * 8);
  auto get = [&](void* d, size_t bytes) {
    if (!ok) return;
    ok = fread(buf.data(), 1, bytes, fp) == bytes &&
         hipMemcpy(d, buf.data(), bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  get(h->psihat, h->n * 8);
  get(h->psihat_old, h->n * 8);
  get(h->dsc, D_COUNT * 8);
  const int order[] = {h->psi_cur, h->psi_old, h->rho_cur, h->rho_old};
  for (int f : order) get(h->f[f], h->n * h->esz);
  for (int f = 0; f < F_CKPT; ++f)
    if (f != F_PSI_A && f != F_PSI_B && f != F_RHO_A && f != F_RHO_B) get(h->f[f], h->n * h->esz);
  fclose(fp);
  if (!ok) return fail(QSM_ERR_ARG, "qsm_checkpoint_read: %s is truncated", path);
  if (h->fp32) HIPCHK(launch_round32(h->psihat, h->psihat32, (long)h->n, h->stream));
  {
    // psi's (and in the quasi variant rho's) passed fields, by the passes
    // the step runs
    int rc = psi_xy(h);
    h->rhoy = 0;
    if (!rc && h->p.variant == QSM_QUASI) rc = rho_inverse(h, nullptr);
    if (rc) return rc;
  }
  h->step = hd.step;
  h->divv_valid = false;
  h->p_valid = false;
  h->r_valid = grad_rho_formed(h->p, h->step);
  h->l_yspec = h->r_yspec = h->p_yspec = false;  // written in real space (qsm_checkpoint_write)
  h->ready = true;
  if (h->p.variant == QSM_STB) return stb_growth(h);
  return QSM_OK;
}

// ============================================================================
// autotuner: for every plain pass kind (axis x direction) of a handle, or
// every plain batched pass group (axis x direction x length) of an ensemble,
// time both tile-width variants on the scratch field (never the state) with
// HIP events and keep the faster.  Choices are cached in a flat JSON object
// {"<gpu>|<precision>|<grid>|<axis>|<dir>": variant} when a path is given.

namespace {

std::map<std::string, int> cache_read(const char* path) {
  std::map<std::string, int> m;
  if (!path) return m;
  FILE* fp = fopen(path, "r");
  if (!fp) return m;
  std::string txt;
  char buf[4096];
  size_t r;
  while ((r = fread(buf, 1, sizeof buf, fp)) > 0) txt.append(buf, r);
  fclose(fp);
  size_t pos = 0;
  while ((pos = txt.find('"', pos)) != std::string::npos) {
    const size_t end = txt.find('"', pos + 1);
    if (end == std::string::npos) break;
    const std::string key = txt.substr(pos + 1, end - pos - 1);
    const size_t colon = txt.find(':', end);
    if (colon == std::string::npos) break;
    m[key] = atoi(txt.c_str() + colon + 1);
    pos = txt.find_first_of(",}", colon);
    if (pos == std::string::npos) break;
  }
  return m;
}

void cache_write(const char* path, const std::map<std::string, int>& m) {
  if (!path) return;
  FILE* fp = fopen(path, "w");
  if (!fp) return;
  fprintf(fp, "{\n");
  size_t i = 0;
  for (const auto& kv : m) fprintf(fp, "  \"%s\": %d%s\n", kv.first.c_str(), kv.second, ++i < m.size() ? "," : "");
  fprintf(fp, "}\n");
  fclose(fp);
}

std::string gpu_name() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "gpu";
  std::string s = prop.gcnArchName;
  for (char& c : s) if (c == '|' || c == '"' || c == ':') c = '_';
  return s;
}

// Each variant is timed in kTuneRounds interleaved rounds and keeps its
// fastest; the default (variant 0) is replaced only by a variant at least 3%
// faster, the fastest of those, so run-to-run noise does not flip the
// choice.  Variants a length does not have are timed as +inf.
constexpr int kTuneRounds = 3;
int tune_pick(const double* ms, int nv) {
  int best = 0;
  for (int v = 1; v < nv; ++v)
    if (ms[v] < 0.97 * ms[0] && ms[v] < ms[best]) best = v;
  return best;
}

// average ms of `reps` launches of fn on stream s (after one warm-up launch)
template <class F>
double time_launches(hipStream_t s, int reps, F fn) {
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  fn();
  (void)hipEventRecord(a, s);
  for (int r = 0; r < reps; ++r) fn();
  (void)hipEventRecord(b, s);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

}  // namespace

namespace {

// QSM_PLAN_MEASURE: each axis's pass is timed in place on a scratch field
// in every variant its length has, in interleaved rounds, and the plan keeps
// the pick of tune_pick (the counterpart of FFTW_MEASURE at quasi.cpp:449-477,
// which also times candidate algorithms on scratch data when planning).
int plan_measure(qsm_plan_s* p) {
  const size_t bytes = (size_t)p->n[0] * p->n[1] * p->n[2] * (p->fp32 ? 4 : 8);
  void* d = nullptr;
  HIPCHK(hipMalloc(&d, bytes));
  hipStream_t s = nullptr;
  int rc = QSM_OK;
  if (hipMemsetAsync(d, 0, bytes, s) != hipSuccess) rc = fail(QSM_ERR_HIP, "plan_measure: clear");
  for (int ax = 0; ax < 3 && rc == QSM_OK; ++ax) {
    PassArgs A{};
    A.geo = Geo{p->n[0], p->n[1], p->n[2], 0};
    A.ld.code = LD_PLAIN; A.ld.a[0] = d; A.st.code = ST_PLAIN; A.st.d[0] = d;
    A.tw = p->tw[ax];
    A.dst = kind_dst(p->kind[ax]);
    double ms[kVariants];
    for (double& m : ms) m = 1e30;
    for (int round = 0; round < kTuneRounds && rc == QSM_OK; ++round)
      for (int v = 0; v < kVariants; ++v) {
        if (!pass_variant_ok(p->fp32, p->n[ax], v, ax, kind_inv(p->kind[ax]))) continue;
        hipError_t err = hipSuccess;
        const double t = time_launches(s, 3, [&] {
          hipError_t e2 = launch_pass(p->fp32, p->n[ax], ax, kind_inv(p->kind[ax]), LD_PLAIN,
                                      ST_PLAIN, A, nullptr, s, nullptr, v);
          if (e2 != hipSuccess) err = e2;
        });
        if (err != hipSuccess) { rc = fail(QSM_ERR_HIP, "plan_measure: %s", hipGetErrorString(err)); break; }
        ms[v] = std::min(ms[v], t);
      }
    if (rc == QSM_OK) p->variant[ax] = tune_pick(ms, kVariants);
  }
  if (hipStreamSynchronize(s) != hipSuccess && rc == QSM_OK) rc = fail(QSM_ERR_HIP, "plan_measure: sync");
  (void)hipFree(d);
  return rc;
}

}  // namespace

extern "C" int qsm_autotune(qsm_handle h, const char* cache_path) {
  if (!h) return fail(QSM_ERR_ARG, "qsm_autotune: null handle");
  auto cache = cache_read(cache_path);
  const std::string base = gpu_name() + "|" + (h->fp32 ? "f32" : "f64") + "|" + std::to_string(h->p.nx) +
                           "x" + std::to_string(h->p.ny) + "x" + std::to_string(h->p.nz);
  const int nn[3] = {h->p.nx, h->p.ny, h->p.nz};
  void* scratch = h->f[F_NTMP];  // a transform intermediate: rewritten before every use
  for (int axis = 0; axis < 3; ++axis)
    for (int inv = 0; inv < 2; ++inv) {
      const std::string key = base + "|" + std::to_string(axis) + "|" + (inv ? "inv" : "fwd");
      auto it = cache.find(key);
      if (it != cache.end()) { h->variant[axis][inv] = it->second; continue; }
      PassArgs A{};
      A.geo = h->geo; A.waves = h->waves; A.phys = h->phys; A.tw = h->tw[axis];
      A.ld.code = LD_PLAIN; A.ld.a[0] = scratch; A.st.code = ST_PLAIN; A.st.d[0] = scratch;
      double ms[kVariants];
      for (double& m : ms) m = 1e30;
      for (int round = 0; round < kTuneRounds; ++round)
        for (int v = 0; v < kVariants; ++v) {
          if (!pass_variant_ok(h->fp32, nn[axis], v, axis, inv != 0)) continue;
          hipError_t err = hipSuccess;
          const double t = time_launches(h->stream, 3, [&] {
            hipError_t e2 = launch_pass(h->fp32, nn[axis], axis, inv != 0, LD_PLAIN, ST_PLAIN, A,
                                        nullptr, h->stream, nullptr, v);
            if (e2 != hipSuccess) err = e2;
          });
          if (err != hipSuccess) return fail(QSM_ERR_HIP, "qsm_autotune: %s", hipGetErrorString(err));
          ms[v] = std::min(ms[v], t);
        }
      const int bv = tune_pick(ms, kVariants);
      h->variant[axis][inv] = bv;
      cache[key] = bv;
    }
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.second);
  h->graphs.clear();
  cache_write(cache_path, cache);
  return QSM_OK;
}

extern "C" int qsm_ens_autotune(qsm_ens e, const char* cache_path) {
  if (!e) return fail(QSM_ERR_ARG, "qsm_ens_autotune: null ensemble");
  auto cache = cache_read(cache_path);
  // lengths present on each axis, with the instances that have them
  std::map<int, std::vector<int>> by_n[3];
  for (int i = 0; i < (int)e->inst.size(); ++i) {
    const qsm_params& p = e->inst[i]->p;
    by_n[0][p.nx].push_back(i); by_n[1][p.ny].push_back(i); by_n[2][p.nz].push_back(i);
  }
  const std::string base = gpu_name() + "|" + (e->fp32 ? "f32" : "f64") + "|ens" + std::to_string(e->inst.size());
  for (int axis = 0; axis < 3; ++axis)
    for (auto& g : by_n[axis])
      for (int inv = 0; inv < 2; ++inv) {
        const std::string key = base + "|" + std::to_string(g.first) + "|" + std::to_string(axis) + "|" + (inv ? "inv" : "fwd");
        auto it = cache.find(key);
        if (it != cache.end()) { e->variant[axis][inv][g.first] = it->second; continue; }
        double mv[kBatchVariants];
        for (double& m : mv) m = 1e30;
        for (int rv = 0; rv < kBatchVariants * kTuneRounds; ++rv) {
          const int v = rv % kBatchVariants;  // rounds interleave the variants
          if (!pass_variant_ok(e->fp32, g.first, v, axis, inv != 0)) continue;
          std::vector<PassArgs> descs;
          std::vector<int> prefix(1, 0);
          for (int i : g.second) {
            qsm_handle_s* h = e->inst[i];
            PassArgs A{};
            A.geo = h->geo; A.waves = h->waves; A.phys = h->phys; A.tw = h->tw[axis];
            A.ld.code = LD_PLAIN; A.ld.a[0] = h->f[F_NTMP]; A.st.code = ST_PLAIN; A.st.d[0] = h->f[F_NTMP];
            descs.push_back(A);
            prefix.push_back(prefix.back() + (int)pass_tiles(e->fp32, g.first, axis, h->geo, v));
          }
          EnsPlan tmp;
          PassArgs* dd = nullptr;
          int* dp = nullptr;
          int rc = upload(&tmp, descs, &dd);
          if (!rc) rc = upload(&tmp, prefix, &dp);
          if (rc) { for (void* d : tmp.allocs) (void)hipFree(d); return rc; }
          BatchLaunch B{dd, dp, (int)descs.size(), prefix.back()};
          hipError_t err = hipSuccess;
          const double ms = time_launches(e->stream, 3, [&] {
            PassArgs dummy{};
            hipError_t e2 = launch_pass(e->fp32, g.first, axis, inv != 0, LD_PLAIN, ST_PLAIN, dummy, nullptr,
                                        e->stream, &B, v);
            if (e2 != hipSuccess) err = e2;
          });
          (void)hipStreamSynchronize(e->stream);
          for (void* d : tmp.allocs) (void)hipFree(d);
          tmp.allocs.clear();
          if (err != hipSuccess) return fail(QSM_ERR_HIP, "qsm_ens_autotune: %s", hipGetErrorString(err));
          mv[v] = std::min(mv[v], ms);
        }
        const int bv = tune_pick(mv, kBatchVariants);
        e->variant[axis][inv][g.first] = bv;
        cache[key] = bv;
      }
  ens_free_blocks(e);
  for (auto& kv : e->plans) ens_free_plan(kv.second);
  e->plans.clear();
  cache_write(cache_path, cache);
  return QSM_OK;
}

extern "C" int qsm_set_variant(qsm_handle h, int axis, int inv, int variant) {
  if (!h || axis < 0 || axis > 2 || (inv != 0 && inv != 1) || variant < 0 || variant >= kVariants)
    return fail(QSM_ERR_ARG, "qsm_set_variant: bad argument");
  const int n = axis == 0 ? h->p.nx : (axis == 1 ? h->p.ny : h->p.nz);
  if (!pass_variant_ok(h->fp32, n, variant, axis, inv != 0))
    return fail(QSM_ERR_UNSUPPORTED, "qsm_set_variant: length %d has no variant %d", n, variant);
  h->variant[axis][inv] = variant;
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.second);
  h->graphs.clear();
  return QSM_OK;
}
