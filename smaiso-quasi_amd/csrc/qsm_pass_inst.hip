// qsm_pass_inst.hip -- instantiates the pass kernels of one field precision
// (QSM_T) and one axis (QSM_AX) for every compiled pencil length; the build
// compiles this file once per (precision, axis) pair.
#include "qsm_pass_impl.hpp"

#ifndef QSM_T
#error "QSM_T (float or double) must be defined"
#endif
#ifndef QSM_AX
#error "QSM_AX (0, 1 or 2) must be defined"
#endif

#define QSM_CAT2(a, b, c) a##_##b##_##c
#define QSM_CAT(a, b, c) QSM_CAT2(a, b, c)

namespace qsm {

hipError_t QSM_CAT(launch_pass, QSM_T, QSM_AX)(int n, bool inv, int ld, int st, int variant,
                                               const PassArgs& A, long* nb, hipStream_t s,
                                               const BatchLaunch* B) {
  switch (n) {
#define QSM_CASE(NN) \
  case NN:           \
    return dispatch_ops<QSM_T, QSM_AX, NN, QSM_IS_STEP_N(NN)>(inv, ld, st, variant, A, nb, s, B);
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return hipErrorInvalidValue;
  }
}

#if QSM_AX == 0
template <int N, bool STEP>
static hipError_t launch_group_if(int group, const GroupArgs& G, long* nb, hipStream_t s) {
  if constexpr (STEP) return launch_group_n<QSM_T, N, QSM_AX>(group, G, nb, s);
  return hipErrorInvalidValue;
}

hipError_t QSM_CAT(launch_group, QSM_T, QSM_AX)(int n, int group, const GroupArgs& G, long* nb,
                                                hipStream_t s) {
  switch (n) {
#define QSM_CASE(NN) \
  case NN:           \
    return launch_group_if<NN, QSM_IS_STEP_N(NN)>(group, G, nb, s);
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return hipErrorInvalidValue;
  }
}
#endif

#if QSM_AX == 2
template <int N, bool STEP>
static hipError_t launch_multi_if(const MultiArgs& A, long* nb, hipStream_t s) {
  if constexpr (STEP) return launch_multi_n<QSM_T, N>(A, nb, s);
  return hipErrorInvalidValue;
}

hipError_t QSM_CAT(launch_multi, QSM_T, n)(int n, const MultiArgs& A, long* nb, hipStream_t s) {
  switch (n) {
#define QSM_CASE(NN) \
  case NN:           \
    return launch_multi_if<NN, QSM_IS_STEP_N(NN)>(A, nb, s);
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return hipErrorInvalidValue;
  }
}

bool QSM_CAT(multi_ok, QSM_T, n)(int n) {
  switch (n) {
#define QSM_CASE(NN) \
  case NN:           \
    return QSM_IS_STEP_N(NN) && multi_ok<QSM_T, NN>();
    QSM_FOR_EACH_N(QSM_CASE)
#undef QSM_CASE
    default:
      return false;
  }
}
#endif

}  // namespace qsm
