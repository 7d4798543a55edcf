/*
 * qsm.h -- C ABI of the MI355X-native smectic-A / isotropic quasi-incompressible
 * time step (library libqsm.so).
 *
 * The reference (evitral/smaiso-quasi) is one C++ main() per program that calls
 * FFTW-MPI directly; its hot path is the while-loop body of quasi/quasi.cpp
 * (and the same kernels in quasi/quasi-stb.cpp).  This header replaces two
 * boundaries of that program:
 *
 *   1. the FFTW r2r boundary (transform level):
 *        fftw_mpi_plan_r2r_3d  quasi/quasi.cpp:449-495   -> qsm_r2r3d_plan
 *        fftw_execute          quasi/quasi.cpp:908...1964 -> qsm_r2r3d_execute
 *        fftw_destroy_plan     quasi/quasi.cpp:2310-2319 -> qsm_r2r3d_destroy
 *      with FFTW's unnormalised REDFT10/REDFT01/RODFT10/RODFT01 definitions
 *      and FFTW's numeric kind values, so existing call sites keep their
 *      constants;
 *   2. the step loop of main() (step level): parameters (quasi.cpp:136-219),
 *      setup (quasi.cpp:875-951), the step (quasi.cpp:1014-1965 /
 *      quasi-stb.cpp:983-1968), diagnostics (quasi.cpp:1969-2031) and the
 *      fields main() writes out.
 *
 * Conventions
 *   - Fields are row-major [nx][ny][nz] (z fastest), the reference layout with
 *     the rank slabs concatenated in rank order.  Host arrays are caller-owned
 *     FP64 and copied in/out; device memory is owned by the library.
 *   - Every function returns 0 (QSM_OK) or a negative QSM_ERR_* code; the
 *     message of the last failure on the calling thread is qsm_last_error().
 *   - A handle or plan is bound to one HIP stream; calls on one handle must
 *     not overlap; distinct handles may be used from distinct threads.
 *   - One GPU per process: the current HIP device at create time is used.
 */
#ifndef QSM_H
#define QSM_H

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define QSM_OK 0
#define QSM_ERR_ARG (-1)          /* invalid argument */
#define QSM_ERR_UNSUPPORTED (-2)  /* size or option without a compiled kernel */
#define QSM_ERR_HIP (-3)          /* HIP runtime failure (message names it) */
#define QSM_ERR_NONFINITE (-4)    /* a diagnostic became NaN/Inf (blow-up) */
#define QSM_ERR_STATE (-5)        /* call out of order (e.g. step before set_psi) */
#define QSM_ERR_NOMEM (-6)        /* device allocation failed */

/* r2r kinds: the values of FFTW's fftw_r2r_kind (fftw3.h) */
#define QSM_REDFT01 4 /* DCT-III: Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k+1/2)/N)     */
#define QSM_REDFT10 5 /* DCT-II:  Y_k = 2 sum_j X_j cos(pi (j+1/2) k/N)                */
#define QSM_RODFT01 8 /* DST-III: Y_k = (-1)^k X_{N-1} + 2 sum_{j<N-1} X_j sin(pi (j+1)(k+1/2)/N) */
#define QSM_RODFT10 9 /* DST-II:  Y_k = 2 sum_j X_j sin(pi (j+1/2)(k+1)/N)             */

/* arithmetic precision */
#define QSM_FP64 0  /* every field and transform in FP64 */
#define QSM_MIXED 1 /* FP32 fields, transforms and nonlinear terms; FP64 psihat
                       state, coefficients, (0,0,0) modes and reductions */

/* model variant */
#define QSM_QUASI 0 /* quasi/quasi.cpp      */
#define QSM_STB 1   /* quasi/quasi-stb.cpp  */

/* Parameters: every constant the reference step reads.  Names follow
 * quasi/quasi.cpp:159-219 (quasi-stb.cpp:149-220). */
typedef struct {
  int nx, ny, nz;      /* grid (quasi.cpp:164: 256^3)                              */
  int variant;         /* QSM_QUASI or QSM_STB                                      */
  int compat_ranks;    /* 0: exact stencils.  P > 0: reproduce the halo slab that an
                          `mpirun -np P` reference run reads for S_x in the v_y/v_z
                          projection (quasi.cpp:1639-1650); needs nx % P == 0        */
  int nw;              /* points per wavelength: L = N 2 pi/(q0 nw) (quasi.cpp:202) */
  int divv_switch;     /* quasi.cpp:161 (200): D enters after this step              */
  int hydro;           /* stb advection switch (quasi-stb.cpp:155, 1)               */
  double q0, alpha, beta, gamma, eps, nu, kp, rho0, amp, rho_m, dt;
  double filter;       /* density filter exponent 1.57*1.57*8 (quasi.cpp:1261)      */
  double growth_q;     /* stb perturbation wavenumber Qi (quasi-stb.cpp:157)        */
} qsm_params;

typedef struct {
  double L1;        /* sum|psi-psi_old| / (dt sum|psi|)          quasi.cpp:2010 */
  double mass;      /* sum rho dx^3                               quasi.cpp:2016 */
  double energy;    /* sum rho(-eps psi^2/2 + alpha F^2/2 - beta psi^4/4 + gamma psi^6/6) dx^3 */
  double interface_z; /* Lz (<rho> - rho0)/(rho_s - rho0): smectic fraction x Lz */
  double growth;    /* stb: growth rate of mode (i*,0,k*) of the last step (stb:2147-2158) */
  long long step;   /* steps taken since qsm_set_psi */
} qsm_diag;

/* fields of qsm_get_field (FP64 host copies; spectral ones in FFTW layout) */
#define QSM_PSI 0
#define QSM_RHO 1
#define QSM_VX 2
#define QSM_VY 3
#define QSM_VZ 4
#define QSM_P 5
#define QSM_DIVV 6
#define QSM_PSIHAT 7
#define QSM_F 8
#define QSM_G 9
#define QSM_GX 10
#define QSM_GY 11
#define QSM_GZ 12
#define QSM_RX 13
#define QSM_RY 14
#define QSM_RZ 15
#define QSM_LX 16
#define QSM_LY 17
#define QSM_LZ 18
#define QSM_RHOHAT 19
#define QSM_NHAT 20
#define QSM_PSIHAT_OLD 21
#define QSM_PSI_OLD 22
#define QSM_SX 23
#define QSM_SY 24
#define QSM_SZ 25
#define QSM_DHAT 26
#define QSM_DTPSI 27 /* stb only: dtpsi = -rho mu - v.g of the last step, the field
                        quasi-stb.cpp:1904-1913 computes for dtpsi.dat; formed on request */

typedef struct qsm_plan_s* qsm_plan;
typedef struct qsm_handle_s* qsm_handle;

const char* qsm_version(void);
const char* qsm_last_error(void);

/* Reference defaults of one variant (quasi.cpp:159-219 with argv = {nu=1,
 * eps*1000=0...}); the caller overrides what the run needs. */
int qsm_default_params(int variant, qsm_params* p);

/* ---- transform level: replaces fftw_mpi_plan_r2r_3d / fftw_execute /
 * fftw_destroy_plan (quasi.cpp:449-495, 908-1964, 2310-2319).  One kind per
 * axis, unnormalised.  prec: QSM_FP64 (double arrays) or QSM_MIXED (float
 * arrays).  flags: 0, or QSM_PLAN_MEASURE (the reference plans with
 * FFTW_MEASURE, quasi.cpp:449-477): time every tile and radix variant of
 * each axis's pass on a scratch field at plan time and keep the fastest.
 * Without it every axis takes variant 0.  in == out is allowed. */
#define QSM_PLAN_MEASURE 1u
int qsm_r2r3d_plan(int nx, int ny, int nz, int kx, int ky, int kz, int prec, unsigned flags,
                   qsm_plan* out);
/* the variant of one axis's pass by hand (see qsm_set_variant);
 * QSM_ERR_UNSUPPORTED where that axis's length has no such variant.
 * Variants 2 and 3 round differently from 0 and 1 (other radix order). */
int qsm_r2r3d_set_variant(qsm_plan p, int axis, int variant);
int qsm_r2r3d_execute(qsm_plan p, const void* d_in, void* d_out, void* hip_stream);
/* same with host arrays (copies in and out; the plan's precision decides
 * the element type) */
int qsm_r2r3d_execute_host(qsm_plan p, const void* h_in, void* h_out);
/* The same as qsm_r2r3d_execute_host without any device allocation, for
 * callers that cannot allocate (allocation-free harnesses, tight loops that
 * call a transform per fftw_execute of the reference).  No counterpart in the
 * reference.  The field is staged in a 128 MiB buffer the library holds per
 * GPU (at most 2^24 FP64 or 2^25 FP32 points: 256^3 FP64); larger fields
 * return QSM_ERR_UNSUPPORTED.  Calls on one GPU share that buffer: calls that
 * overlap on one GPU (from several threads) race and must be serialised by
 * the caller. */
int qsm_r2r3d_execute_host_noalloc(qsm_plan p, const void* h_in, void* h_out);
int qsm_r2r3d_destroy(qsm_plan p);

/* ---- step level: replaces the body of main() around the time loop */
int qsm_create(const qsm_params* p, int prec, qsm_handle* out);
/* setup of quasi.cpp:905-951 from psi(x) (FP64, [nx][ny][nz]) */
int qsm_set_psi(qsm_handle h, const double* psi);
/* advance nsteps steps (quasi.cpp:1017-1965 per step) */
int qsm_step(qsm_handle h, int nsteps);
/* steps are replayed as captured HIP graphs (one per buffer/phase signature,
 * captured on first use); enable = 0 launches every kernel eagerly */
int qsm_set_graphs(qsm_handle h, int enable);
/* run until L1 <= lim_l1 (checked every step_l1 steps, quasi.cpp:1014, 1972,
 * 2010) or max_steps; *steps_done receives the steps taken */
int qsm_run_l1(qsm_handle h, int max_steps, int step_l1, double lim_l1, int* steps_done);
/* copy field `which` of the current step to host (caller-owned, nx*ny*nz
 * doubles).  Fields no later stage of the step reads are formed here on
 * request: D (QSM_DIVV) always, p (QSM_P) up to step divv_switch + 1.  The
 * lagged grad rho (QSM_RX..RZ) exists at step 1 and after divv_switch only;
 * other steps return QSM_ERR_STATE. */
int qsm_get_field(qsm_handle h, int which, double* host);
/* diagnostics of the current state (compensated FP64 sums over the fields
 * of the last step); QSM_ERR_NONFINITE if any is NaN/Inf */
int qsm_get_diag(qsm_handle h, qsm_diag* d);
int qsm_destroy(qsm_handle h);

/* Full-state binary checkpoint (every buffer the step reads, the step count
 * and the parameters) for a bit-exact resume.  Replaces the reference's
 * psi<rank>.dat restart (quasi.cpp:692-717, 2051-2062), which keeps psi only
 * at 6 digits.  Reading requires a handle created with the same parameters
 * and precision. */
int qsm_checkpoint_write(qsm_handle h, const char* path);
int qsm_checkpoint_read(qsm_handle h, const char* path);

/* ---- ensembles: independent instances (e.g. the stability runs of
 * quasi-stb.cpp at different grid sizes) advanced together.  Each launch of
 * the step runs as one batched kernel per distinct pencil length over all
 * instances, and a whole step is replayed as one captured HIP graph.
 * params: count parameter sets (same variant, divv_switch and hydro). */
typedef struct qsm_ens_s* qsm_ens;
int qsm_ens_create(const qsm_params* params, int count, int prec, qsm_ens* out);
int qsm_ens_set_psi(qsm_ens e, int index, const double* psi);
int qsm_ens_step(qsm_ens e, int nsteps);
/* growth rate of every instance after the last step (quasi-stb.cpp:2147-2158) */
int qsm_ens_growth(qsm_ens e, double* out);
int qsm_ens_get_field(qsm_ens e, int index, int which, double* host);
int qsm_ens_get_diag(qsm_ens e, int index, qsm_diag* d);
int qsm_ens_set_graphs(qsm_ens e, int enable);
/* steps chained in one graph launch (1, or even up to 64; default 1).  Once
 * the step sequence repeats (from step 4 on, between phase changes such as
 * divv_switch), `steps` steps are captured once as one graph and replayed
 * per launch; results are bit for bit those of one graph per step. */
int qsm_ens_set_graph_steps(qsm_ens e, int steps);
/* device time of an ensemble's batched launches (no counterpart in the
 * reference): while enabled, steps launch eagerly with HIP events around
 * every batched kernel on the stream it runs on (launches of one pass for
 * different pencil lengths overlap on side streams, so their times add up
 * to more than the wall time).  Enabling resets the counters.  The rank-th
 * kernel (0 = most device time): the pass kernel as qsm_timing_kernel
 * reports it (load code 63: the batched FP64 finish of the mean-mode sums),
 * its pencil length n, total time, launches and algorithmic bytes. */
int qsm_ens_timing_enable(qsm_ens e, int enable);
int qsm_ens_timing_kernel(qsm_ens e, int rank, int* axis, int* inv, int* ld, int* st, int* variant,
                          int* n, double* total_ms, long long* launches, double* alg_bytes);
int qsm_ens_destroy(qsm_ens e);

/* ---- autotuner: times the variants of the plain transform passes on the
 * handle's (or ensemble's) own grid(s), using a scratch field, in
 * interleaved rounds, and keeps per axis and direction (and pencil length for
 * ensembles) the fastest variant that is at least 3% faster than the
 * default, else the default.  Variants: 0 default tile and radix plan, 1
 * half-width tiles, 2 and 3 alternative radix plans (pencil lengths 32-512;
 * they round differently, at the level of the transform's round-off), 4 and
 * 5 the software-pipelined forms of 0 and 2 (inverse passes whose tile fills
 * a CU's LDS alone: FP64 x at 512, FP32 x at 1024; single handles and
 * plans, not ensembles).
 * cache_path (may be NULL): JSON file of earlier choices, read and updated. */
int qsm_autotune(qsm_handle h, const char* cache_path);
/* set the variant (0-5, as above) of the plain passes along axis (0 x, 1 y,
 * 2 z) in direction inv (1: DCT-III/DST-III) by hand; QSM_ERR_UNSUPPORTED
 * where that axis's length and direction have no such variant */
int qsm_set_variant(qsm_handle h, int axis, int inv, int variant);
int qsm_ens_autotune(qsm_ens e, const char* cache_path);

/* ---- measurement support (no counterpart in the reference, which only
 * prints total wall time at quasi.cpp:2292-2305): device time of the pass
 * kernels, measured with HIP events recorded around every launch on the
 * handle's stream while enabled.  which: 0 = every pass; 1/2/3 = the
 * inverse (DCT-III/DST-III) x/y/z passes; 4/5/6 = the forward x/y/z passes.
 * alg_bytes: algorithmic HBM bytes of those launches (fields read + fields
 * written, at their element size).  qsm_timing_enable resets the counters. */
int qsm_timing_enable(qsm_handle h, int enable);
int qsm_timing_get(qsm_handle h, int which, double* total_ms, long long* launches,
                   double* alg_bytes);
/* the rank-th kernel (0 = most device time) among those timed: the pass
 * kernel k_pass<T, N, axis, inv, ld, st, variant> (k_pass_sx for the store
 * operators that also run the x inverse of what they store) with its load
 * and store operator codes and tile variant, its total time, launches and
 * algorithmic bytes */
int qsm_timing_kernel(qsm_handle h, int rank, int* axis, int* inv, int* ld, int* st, int* variant,
                      double* total_ms, long long* launches, double* alg_bytes);
/* reset the counters and time only the launches of one kernel, named by the
 * fields qsm_timing_kernel reports (axis, inv, ld, st, variant), so that a
 * measured run records two events per launch of that kernel instead of two
 * per launch of every kernel (about 3 us of device time per event pair);
 * qsm_timing_enable clears the selection */
int qsm_timing_select(qsm_handle h, int axis, int inv, int ld, int st, int variant);

#ifdef __cplusplus
}
#endif
#endif /* QSM_H */
