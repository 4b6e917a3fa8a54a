// Solver state shared by the real and complex solver sources.
#pragma once

#include <rocblas/rocblas.h>

#include <vector>

#include "common.hpp"
#include "operator.hpp"

#define CHFSI_RB(...)                                                                    \
  do {                                                                                   \
    rocblas_status rs_ = (__VA_ARGS__);                                                  \
    if (rs_ != rocblas_status_success) {                                                 \
      ::chfsi::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #__VA_ARGS__,              \
                         rocblas_status_to_string(rs_));                                 \
      return CHFSI_ERR_BLAS;                                                             \
    }                                                                                    \
  } while (0)

struct chfsi_solver {
  chfsi_params params{};  // zero unless set (entry points that build a temporary solver)
  double a0 = 0, bLow = 0, bUp = 0;
  rocblas_handle blas = nullptr;
  chfsi::Scratch Y64, X32, Y32, HXb, Hp, rot, evals, info, part, lam, v0, split;
  // mixed-precision dense steps: FP32 copies of X, of an HX block and of the
  // N x N factor, FP32 split-K partials and rotation panel
  chfsi::Scratch Xf, HXf, Qf, split32, rot32;
  // defect form of the FP32 filter: FP32 defect source, coefficient tables
  chfsi::Scratch R32, dcoef;
  // Lanczos: dot products and norms kept on the device until the last step
  chfsi::Scratch lz;
  // generalised Rayleigh-Ritz: L^{-1} and a product buffer (2 N x N)
  chfsi::Scratch gep;
  // tile table of the FP64 gram kernel
  chfsi::Scratch gtiles;
  // rotation panels: copies of finished panels back into X run on a second
  // stream, overlapped with the next panel's rotation (two panel buffers)
  hipStream_t copy_stream = nullptr;
  hipEvent_t rot_done[2] = {nullptr, nullptr}, copy_done[2] = {nullptr, nullptr};
  // mixed filter: the FP64 top columns run on a second stream beside the
  // low-precision columns, their HX projections in their own scratch
  hipStream_t filter_stream = nullptr;
  hipEvent_t filter_fork = nullptr, filter_join = nullptr;
  chfsi::Scratch top_hx;
  bool have_v0 = false;
  // largest residual norm over the lowest 90% of states of the last pass that
  // computed residuals, and the filter precision the last pass used
  double last_res_lowest = 1e300;
  int last_precision = CHFSI_FP64;
  ~chfsi_solver() {
    if (blas) rocblas_destroy_handle(blas);
    for (hipEvent_t e : {rot_done[0], rot_done[1], copy_done[0], copy_done[1], filter_fork,
                         filter_join})
      if (e) (void)hipEventDestroy(e);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (filter_stream) (void)hipStreamDestroy(filter_stream);
    for (chfsi::Scratch *s : {&Y64, &X32, &Y32, &HXb, &Hp, &rot, &evals, &info, &part, &lam, &v0,
                              &split, &Xf, &HXf, &Qf, &split32, &rot32, &R32, &dcoef, &gep,
                              &gtiles, &top_hx, &lz})
      s->release();
  }
};

namespace chfsi {

// Degree lookup on the upper bound (…SolverDevice.cc:29-46, :96-105).
unsigned cheby_order_lookup(double upper);
// Degree of a solve pass: configured, or looked up (halved for all-electron
// runs), times the first-SCF factor for pseudopotential runs
// (…SolverDevice.cc:307-326).
unsigned solve_degree(const chfsi_params &P, double bUp, bool is_first_scf);
chfsi_status ensure_blas(chfsi_solver *s, hipStream_t st);
// x . y over n doubles (synchronises the stream)
chfsi_status dot(chfsi_solver *s, const double *x, const double *y, int64_t n, double *out,
                 hipStream_t st);
// eigenvalues of a small symmetric matrix (cyclic Jacobi), ascending
std::vector<double> small_sym_eigvals(std::vector<double> A, int n);
// seeded uniform [0,1) Lanczos start of length M (host)
std::vector<double> default_lanczos_start(int64_t M);

}  // namespace chfsi
