// Solver side of the C ABI: Chebyshev filter, Lanczos bounds, CholGS,
// Rayleigh-Ritz, residual norms and the one-pass driver.
//
// Reference: chebyshevOrthogonalizedSubspaceIterationSolverDevice
// (src/solvers/eigenSolvers/chebyshevOrthogonalizedSubspaceIterationSolverDevice.cc:155-736),
// linearAlgebraOperationsDevice (src/linAlg/linearAlgebraOperationsDevice.cc:340-727,
// :1832-2242, :3078-3240, :4610-4766), pseudoGSDevice.cc:81-463,
// rayleighRitzDevice.cc:80-353.  The dense N x N Cholesky and eigensolve run on
// the GPU (rocSOLVER) instead of ELPA/ScaLAPACK on the host.

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "dense.hpp"
#include "kernels.hpp"
#include "operator.hpp"
#include "solver.hpp"

using namespace chfsi;

namespace chfsi {

unsigned cheby_order_lookup(double upper) {
  static const unsigned table[][2] = {{500, 24},    {750, 30},    {1000, 39},   {1500, 50},
                                      {2000, 53},   {3000, 57},   {4000, 62},   {5000, 69},
                                      {9000, 77},   {14000, 104}, {20000, 119}, {30000, 162},
                                      {50000, 300}, {80000, 450}, {100000, 550}, {200000, 700},
                                      {500000, 1000}};
  const unsigned u = (unsigned)upper;
  for (auto &row : table)
    if (u <= row[0]) return row[1];
  return 1250;
}

// Degree of a solve pass (…SolverDevice.cc:307-326): the configured degree,
// or the table entry for bUp halved for all-electron runs (orthogType CGS, the
// device's only orthogonalisation); then the first-SCF factor, applied only to
// pseudopotential runs.  Both products truncate, as the reference's unsigned
// arithmetic does.
unsigned solve_degree(const chfsi_params &P, double bUp, bool is_first_scf) {
  unsigned m = (unsigned)std::max(0, P.cheby_degree);
  if (m == 0) {
    m = cheby_order_lookup(bUp);
    if (!P.is_pseudopotential) m = (unsigned)(m * 0.5);
  }
  if (is_first_scf && P.is_pseudopotential) m = (unsigned)(m * P.first_scf_degree_factor);
  return m;
}

chfsi_status ensure_blas(chfsi_solver *s, hipStream_t st) {
  if (!s->blas) CHFSI_RB(rocblas_create_handle(&s->blas));
  CHFSI_RB(rocblas_set_stream(s->blas, st));
  return CHFSI_OK;
}

// Sum of `n` doubles of a device array (synchronises the stream).
static chfsi_status host_sum(const double *d, int n, double *out, hipStream_t st) {
  std::vector<double> h(n);
  CHFSI_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  double s = 0;
  for (double x : h) s += x;
  *out = s;
  return CHFSI_OK;
}

chfsi_status dot(chfsi_solver *s, const double *x, const double *y, int64_t n, double *out,
                 hipStream_t st) {
  const int nch = 256;
  CHFSI_TRY(s->part.ensure(nch * sizeof(double)));
  CHFSI_HIP(launch_dot_partial(x, y, n, static_cast<double *>(s->part.ptr), nch, st));
  return host_sum(static_cast<double *>(s->part.ptr), nch, out, st);
}

std::vector<double> small_sym_eigvals(std::vector<double> A, int n) {
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - sn * akq;
          A[k * n + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - sn * aqk;
          A[q * n + k] = sn * apk + c * aqk;
        }
      }
  }
  std::vector<double> ev(n);
  for (int i = 0; i < n; ++i) ev[i] = A[i * n + i];
  std::sort(ev.begin(), ev.end());
  return ev;
}

// seeded uniform [0,1) start (the reference seeds rand() with the rank)
std::vector<double> default_lanczos_start(int64_t M) {
  std::vector<double> hv(M);
  uint64_t state = 0x9E3779B97F4A7C15ull;
  for (int64_t i = 0; i < M; ++i) {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    hv[i] = double(state >> 11) * (1.0 / 9007199254740992.0);
  }
  return hv;
}

}  // namespace chfsi

namespace {

template <typename T>
T *sptr(Scratch &s) {
  return static_cast<T *>(s.ptr);
}

// Scaled Chebyshev recurrence on [a, b] with a0 scaling (Zhou-Saad), fused into
// the HX launches.  Degree 1:  Y = (sigma1/e)(H~X - cX); degree k >= 2:
//   X_new = alpha1 (H~ Y - c Y) + alpha2 X
// which is one hx_fused call (alpha = alpha1, gamma = -c alpha1, beta = alpha2).
// Coefficients: linearAlgebraOperationsDevice.cc:547-726 / Opt.cc:286-351.
struct ChebState {
  double e, c, sigma, gamma;
};

ChebState cheb_init(double a, double b, double a0) {
  ChebState s;
  s.e = (b - a) / 2.0;
  s.c = (b + a) / 2.0;
  s.sigma = s.e / (a0 - s.c);
  s.gamma = 2.0 / s.sigma;
  return s;
}

// degree 1 from X into Y
template <typename T>
chfsi_status cheb_first(chfsi_operator *op, const T *X, int64_t ldX, T *Y, int64_t ldY, int B,
                        const ChebState &cs, hipStream_t st, bool bf16, void *hx_ws = nullptr) {
  return hx_fused<T>(op, X, ldX, Y, ldY, B, T(cs.sigma / cs.e), T(0), T(-cs.c * cs.sigma / cs.e),
                     1, true, st, bf16, nullptr, 0, nullptr, hx_ws);
}

// degrees d0..d1 (d0 >= 2); *cur holds the latest iterate, *prev the one before;
// hx_ws (nullable): projection scratch of the HX calls in place of the operator's
template <typename T>
chfsi_status cheb_steps(chfsi_operator *op, T **cur, int64_t *ldc, T **prev, int64_t *ldp, int B,
                        int d0, int d1, ChebState &cs, hipStream_t st, bool bf16,
                        void *hx_ws = nullptr) {
  for (int degree = d0; degree <= d1; ++degree) {
    const double sigma2 = 1.0 / (cs.gamma - cs.sigma);
    const double alpha1 = 2.0 * sigma2 / cs.e, alpha2 = -(cs.sigma * sigma2);
    CHFSI_TRY(hx_fused<T>(op, *cur, *ldc, *prev, *ldp, B, T(alpha1), T(alpha2),
                          T(-cs.c * alpha1), 0, true, st, bf16, nullptr, 0, nullptr, hx_ws));
    std::swap(*cur, *prev);
    std::swap(*ldc, *ldp);
    cs.sigma = sigma2;
  }
  return CHFSI_OK;
}

// Columns one filter recurrence runs on: whole blocks of B, all N unless the
// filter scratch (8 bytes per DOF and column in every precision: the FP64
// iterate, or the two FP32 iterates) would exceed 16 GiB.  A wider recurrence
// fills more of each colour launch (the per-launch ramp and drain are shared by
// more items) and gives every column the same arithmetic as block-by-block.
int filter_group_columns(const chfsi_params &P, int64_t M, int N, int B) {
  if (P.filter_columns > 0) return std::min(N, std::max(B, P.filter_columns / B * B));
  const int64_t budget = int64_t(16) << 30;
  const int64_t fit = std::max<int64_t>(1, budget / (8 * std::max<int64_t>(M, 1)) / B);
  return (int)std::min<int64_t>(N, fit * B);
}

// Defect form of the FP32 filter on the B-column block X (FP64, ld):
//   theta_v = <x_v, H~x_v> / <x_v, x_v>,  R = H~X - X diag(theta)  (one FP64 apply),
//   Y_k = X diag(D_k) + E_k  with  D_1 = (sigma/e)(theta - c),  E_0 = 0,  E_1 = (sigma/e) R,
//   D_k+1 = alpha1 (theta - c) D_k + alpha2 D_k-1,
//   E_k+1 = alpha1 (H~ - c) E_k + alpha2 E_k-1 + alpha1 D_k R,
// the scalars of the plain recurrence (cheb_steps).  E runs in FP32 through the
// fused HX (the R term enters its first-touch combine); the rounding then
// scales with |R| instead of |X|, so near convergence the filtered block keeps
// the FP64 eigenvectors' accuracy.  Equal to the plain filter in exact
// arithmetic.
chfsi_status filter_block_defect(chfsi_solver *s, chfsi_operator *op, double *X, int64_t ld,
                                 int B, int m, double a, double b, double a0, hipStream_t st) {
  const int64_t M = op->M;
  const ChebState cs = cheb_init(a, b, a0);
  CHFSI_TRY(s->HXb.ensure(sizeof(double) * M * B));
  double *HX = sptr<double>(s->HXb);
  CHFSI_TRY(hx_fused<double>(op, X, ld, HX, B, B, 1.0, 0.0, 0.0, 1, true, st));
  if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, HX, B, B, st));
  const int nch = 256;
  CHFSI_TRY(s->part.ensure(sizeof(double) * nch * 2 * B));
  double *part = sptr<double>(s->part);
  CHFSI_HIP(launch_colstats_partial(HX, B, X, ld, M, B, part, nch, st));
  std::vector<double> h(size_t(nch) * 2 * B);
  CHFSI_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  // host tables: theta (B), D_m (B), then the FP32 coefficient rows
  // alpha1 * D_(d-1) of degrees d = 2..m
  std::vector<double> th(2 * size_t(B));
  for (int v = 0; v < B; ++v) {
    double xh = 0, xx = 0;
    for (int k = 0; k < nch; ++k) {
      xh += h[(size_t(k) * 2) * B + v];
      xx += h[(size_t(k) * 2 + 1) * B + v];
    }
    th[v] = xx > 0 ? xh / xx : 0.0;
  }
  std::vector<float> rc(size_t(m + 1) * B, 0.0f);
  std::vector<double> Dp(B, 1.0), D(B), Dn(B);
  for (int v = 0; v < B; ++v) D[v] = cs.sigma / cs.e * (th[v] - cs.c);
  double sigma = cs.sigma;
  for (int degree = 2; degree <= m; ++degree) {
    const double sigma2 = 1.0 / (cs.gamma - sigma);
    const double alpha1 = 2.0 * sigma2 / cs.e, alpha2 = -(sigma * sigma2);
    for (int v = 0; v < B; ++v) {
      rc[size_t(degree) * B + v] = (float)(alpha1 * D[v]);
      Dn[v] = alpha1 * (th[v] - cs.c) * D[v] + alpha2 * Dp[v];
    }
    std::swap(Dp, D);
    std::swap(D, Dn);
    sigma = sigma2;
  }
  for (int v = 0; v < B; ++v) th[B + v] = D[v];
  CHFSI_TRY(s->dcoef.ensure(sizeof(double) * 2 * B + sizeof(float) * rc.size()));
  double *dth = sptr<double>(s->dcoef), *dD = dth + B;
  float *drc = reinterpret_cast<float *>(dth + 2 * B);
  CHFSI_HIP(hipMemcpyAsync(dth, th.data(), sizeof(double) * th.size(), hipMemcpyHostToDevice, st));
  CHFSI_HIP(hipMemcpyAsync(drc, rc.data(), sizeof(float) * rc.size(), hipMemcpyHostToDevice, st));

  CHFSI_TRY(s->X32.ensure(sizeof(float) * M * B));
  CHFSI_TRY(s->Y32.ensure(sizeof(float) * M * B));
  CHFSI_TRY(s->R32.ensure(sizeof(float) * M * B));
  float *Ep = sptr<float>(s->X32), *Ec = sptr<float>(s->Y32), *R = sptr<float>(s->R32);
  CHFSI_HIP(hipMemsetAsync(Ep, 0, sizeof(float) * M * B, st));
  CHFSI_HIP(launch_defect_source(HX, B, X, ld, dth, cs.sigma / cs.e, M, B, R, Ec, st));
  sigma = cs.sigma;
  for (int degree = 2; degree <= m; ++degree) {
    const double sigma2 = 1.0 / (cs.gamma - sigma);
    const double alpha1 = 2.0 * sigma2 / cs.e, alpha2 = -(sigma * sigma2);
    CHFSI_TRY(hx_fused<float>(op, Ec, B, Ep, B, B, float(alpha1), float(alpha2),
                              float(-cs.c * alpha1), 0, true, st, false, R, B,
                              drc + size_t(degree) * B));
    std::swap(Ec, Ep);
    sigma = sigma2;
  }
  CHFSI_HIP(launch_defect_finish(X, ld, dD, Ec, B, M, B, st));
  // the host tables are read by the copies above: wait before they go
  CHFSI_HIP(hipStreamSynchronize(st));
  return CHFSI_OK;
}

// In-place filter of the B-column block X (leading dimension ld).  FP32 and
// BF16 forms cast the block to FP32 once; with tail > 0 the last `tail`
// degrees run in FP64 from the cast-back pair of iterates, so that the
// rounding of the low-precision steps is itself filtered.  hx_ws (nullable,
// FP64 form only): projection scratch of its HX calls in place of the
// operator's own.
chfsi_status filter_block(chfsi_solver *s, chfsi_operator *op, double *X, int64_t ld, int B, int m,
                          double a, double b, double a0, int precision, int tail,
                          hipStream_t st, void *hx_ws = nullptr) {
  const int64_t M = op->M;
  ChebState cs = cheb_init(a, b, a0);
  CHFSI_TRY(s->Y64.ensure(sizeof(double) * M * B));
  double *Y = sptr<double>(s->Y64);
  if (precision == CHFSI_FP64) {
    CHFSI_HIP(hipMemsetAsync(Y, 0, sizeof(double) * M * B, st));
    CHFSI_TRY(cheb_first<double>(op, X, ld, Y, B, B, cs, st, false, hx_ws));
    double *cur = Y, *prev = X;
    int64_t ldc = B, ldp = ld;
    CHFSI_TRY(cheb_steps<double>(op, &cur, &ldc, &prev, &ldp, B, 2, m, cs, st, false, hx_ws));
    if (cur != X) CHFSI_HIP(launch_copy_block<double, double>(cur, ldc, X, ld, M, B, st));
    return CHFSI_OK;
  }
  if (precision != CHFSI_FP32 && precision != CHFSI_BF16) {
    set_error("unsupported filter precision %d", precision);
    return CHFSI_ERR_ARG;
  }
  const bool bf16 = precision == CHFSI_BF16;
  if (!bf16 && s->params.mixed_prec_defect_filter)
    return filter_block_defect(s, op, X, ld, B, m, a, b, a0, st);
  tail = std::max(0, std::min(tail, m - 1));
  const int mlow = m - tail;
  CHFSI_TRY(s->X32.ensure(sizeof(float) * M * B));
  CHFSI_TRY(s->Y32.ensure(sizeof(float) * M * B));
  float *X32 = sptr<float>(s->X32), *Y32 = sptr<float>(s->Y32);
  CHFSI_HIP(launch_copy_block<double, float>(X, ld, X32, B, M, B, st));
  CHFSI_HIP(hipMemsetAsync(Y32, 0, sizeof(float) * M * B, st));
  CHFSI_TRY(cheb_first<float>(op, X32, B, Y32, B, B, cs, st, bf16));
  float *cur = Y32, *prev = X32;
  int64_t ldc = B, ldp = B;
  CHFSI_TRY(cheb_steps<float>(op, &cur, &ldc, &prev, &ldp, B, 2, mlow, cs, st, bf16));
  if (tail == 0) {
    CHFSI_HIP(launch_copy_block<float, double>(cur, ldc, X, ld, M, B, st));
    return CHFSI_OK;
  }
  // FP64 tail: latest iterate -> X, previous -> Y
  CHFSI_HIP(launch_copy_block<float, double>(cur, ldc, X, ld, M, B, st));
  CHFSI_HIP(launch_copy_block<float, double>(prev, ldp, Y, B, M, B, st));
  double *c64 = X, *p64 = Y;
  int64_t ldc64 = ld, ldp64 = B;
  CHFSI_TRY(cheb_steps<double>(op, &c64, &ldc64, &p64, &ldp64, B, mlow + 1, m, cs, st, false));
  if (c64 != X) CHFSI_HIP(launch_copy_block<double, double>(c64, ldc64, X, ld, M, B, st));
  return CHFSI_OK;
}

chfsi_status lanczos_impl(chfsi_solver *s, chfsi_operator *op, const double *v0, int steps,
                          int reproducible, double *lower, double *upper, hipStream_t st) {
  const int64_t M = op->M;
  CHFSI_ARG(steps >= 1, "lanczos steps must be >= 1");
  CHFSI_TRY(s->HXb.ensure(sizeof(double) * M * 3));
  double *v = sptr<double>(s->HXb), *f = v + M, *vold = v + 2 * M;
  CHFSI_HIP(hipMemsetAsync(v, 0, sizeof(double) * M * 3, st));
  CHFSI_HIP(hipMemcpyAsync(v, v0, sizeof(double) * M, hipMemcpyDeviceToDevice, st));
  if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, v, 1, 1, st));
  double nrm2;
  CHFSI_TRY(dot(s, v, v, M, &nrm2, st));
  CHFSI_ARG(nrm2 > 0, "lanczos start vector is zero");
  CHFSI_HIP(launch_axpby<double>(M, 0.0, v, 1.0 / std::sqrt(nrm2), v, st));
  auto apply = [&](const double *x, double *y) -> chfsi_status {
    CHFSI_TRY(hx_fused<double>(op, x, 1, y, 1, 1, 1.0, 0.0, 0.0, 1, true, st));
    if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, y, 1, 1, st));
    return CHFSI_OK;
  };
  // The recurrence's scalars stay on the device: each dot product's partials
  // are summed into a slot (alpha_0 at 0; for step j, |f|^2 at 2j - 1 and
  // alpha_j at 2j; the last |f|^2 at 2 steps - 1) that the vector updates read,
  // and the host reads the slots once at the end.  The sums run in the order
  // of a host loop and the updates form -alpha, 1 / beta and -beta as the
  // host would, so the bounds equal those of a host-driven recurrence.
  const int nch = 256;
  CHFSI_TRY(s->part.ensure(nch * sizeof(double)));
  CHFSI_TRY(s->lz.ensure(sizeof(double) * 2 * steps));
  double *part = sptr<double>(s->part), *slot = sptr<double>(s->lz);
  auto dev_dot = [&](const double *x, const double *y, int k) -> chfsi_status {
    CHFSI_HIP(launch_dot_partial(x, y, M, part, nch, st));
    CHFSI_HIP(launch_sum_slot(part, nch, slot + k, st));
    return CHFSI_OK;
  };
  CHFSI_TRY(apply(v, f));
  CHFSI_TRY(dev_dot(f, v, 0));
  CHFSI_HIP(launch_axpby_slot(M, slot, 0, v, 1.0, f, st));  // f <- f - alpha_0 v
  for (int j = 1; j < steps; ++j) {
    CHFSI_TRY(dev_dot(f, f, 2 * j - 1));
    std::swap(v, vold);                                                    // vold <- v
    CHFSI_HIP(launch_axpby_slot(M, slot + 2 * j - 1, 1, f, 0.0, v, st));    // v <- f / beta
    CHFSI_TRY(apply(v, f));
    CHFSI_HIP(launch_axpby_slot(M, slot + 2 * j - 1, 2, vold, 1.0, f, st));  // f -= beta vold
    CHFSI_TRY(dev_dot(f, v, 2 * j));
    CHFSI_HIP(launch_axpby_slot(M, slot + 2 * j, 0, v, 1.0, f, st));        // f -= alpha v
  }
  CHFSI_TRY(dev_dot(f, f, 2 * steps - 1));
  std::vector<double> hs(2 * size_t(steps));
  CHFSI_HIP(hipMemcpyAsync(hs.data(), slot, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  std::vector<double> T(size_t(steps) * steps, 0.0);
  T[0] = hs[0];
  for (int j = 1; j < steps; ++j) {
    const double beta = std::sqrt(hs[2 * j - 1]);
    T[size_t(j) * steps + j - 1] = beta;
    T[size_t(j - 1) * steps + j] = beta;
    T[size_t(j) * steps + j] = hs[2 * j];
  }
  std::vector<double> theta = small_sym_eigvals(T, steps);
  const double fnorm = std::sqrt(hs[2 * steps - 1]);
  *lower = std::floor(theta.front());
  *upper = std::ceil(theta.back() + (reproducible ? fnorm : fnorm / 10.0));
  return CHFSI_OK;
}

// C (m x n, column-major, ldc) = A B^T with A m x K (lda) and B n x K (ldb).
// The contraction length K (the DOF count) is split into chunks computed as
// one strided-batched GEMM, so that the small m x n output still spreads over
// the whole GPU, then the partial products are summed.
chfsi_status gemm_nt_splitk(chfsi_solver *s, const double *A, int64_t lda, const double *B,
                            int64_t ldb, int m, int n, int64_t K, double *C, int64_t ldc,
                            hipStream_t st) {
  const int nk = (int)std::max<int64_t>(1, std::min<int64_t>(64, K / 8192));
  const int64_t Kc = K / nk;
  const int64_t rem = K - Kc * nk;
  const int slots = nk + (rem ? 1 : 0);
  CHFSI_TRY(s->split.ensure(sizeof(double) * slots * m * n));
  double *P = sptr<double>(s->split);
  const double one = 1.0, zero = 0.0;
  CHFSI_RB(rocblas_dgemm_strided_batched(
      s->blas, rocblas_operation_none, rocblas_operation_transpose, m, n, (rocblas_int)Kc, &one,
      A, (rocblas_int)lda, Kc * lda, B, (rocblas_int)ldb, Kc * ldb, &zero, P, m, (int64_t)m * n,
      nk));
  if (rem)
    CHFSI_RB(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_transpose, m, n,
                           (rocblas_int)rem, &one, A + nk * Kc * lda, (rocblas_int)lda,
                           B + nk * Kc * ldb, (rocblas_int)ldb, &zero, P + (int64_t)nk * m * n,
                           m));
  CHFSI_HIP(launch_sum_partials(P, slots, m, n, C, ldc, st));
  return CHFSI_OK;
}

// The same product from FP32 operands (FP32 split-K partials, summed in FP64).
chfsi_status sgemm_nt_splitk(chfsi_solver *s, const float *A, int64_t lda, const float *B,
                             int64_t ldb, int m, int n, int64_t K, double *C, int64_t ldc,
                             hipStream_t st) {
  const int nk = (int)std::max<int64_t>(1, std::min<int64_t>(64, K / 8192));
  const int64_t Kc = K / nk;
  const int64_t rem = K - Kc * nk;
  const int slots = nk + (rem ? 1 : 0);
  CHFSI_TRY(s->split32.ensure(sizeof(float) * slots * m * n));
  float *P = sptr<float>(s->split32);
  const float one = 1.0f, zero = 0.0f;
  CHFSI_RB(rocblas_sgemm_strided_batched(
      s->blas, rocblas_operation_none, rocblas_operation_transpose, m, n, (rocblas_int)Kc, &one,
      A, (rocblas_int)lda, Kc * lda, B, (rocblas_int)ldb, Kc * ldb, &zero, P, m, (int64_t)m * n,
      nk));
  if (rem)
    CHFSI_RB(rocblas_sgemm(s->blas, rocblas_operation_none, rocblas_operation_transpose, m, n,
                           (rocblas_int)rem, &one, A + nk * Kc * lda, (rocblas_int)lda,
                           B + nk * Kc * ldb, (rocblas_int)ldb, &zero, P + (int64_t)nk * m * n,
                           m));
  CHFSI_HIP(launch_sum_partials_f32(P, slots, m, n, C, ldc, st));
  return CHFSI_OK;
}

// FP32 copy of the node-major M x N block (for the mixed-precision products).
chfsi_status cast_x(chfsi_solver *s, const double *X, int64_t M, int N, float **out,
                    hipStream_t st) {
  CHFSI_TRY(s->Xf.ensure(sizeof(float) * M * N));
  *out = sptr<float>(s->Xf);
  CHFSI_HIP(launch_copy_block<double, float>(X, N, *out, N, M, N, st));
  return CHFSI_OK;
}

// Rotation panels: rows of DOFs per panel so that one panel holds 2^27
// entries (1 GiB in FP64).
int64_t panel_rows(int64_t M, int N) {
  return std::max<int64_t>(1, std::min<int64_t>(M, (int64_t(1) << 27) / N));
}

// X <- X Q, i.e. Xc <- Q^T Xc with Xc = X^T (N x M, ld N), by DOF panels
// (subspaceRotationScalapack, linearAlgebraOperationsDevice.cc:1832-2242).
// Mixed (subspaceRotationRRMixedPrecScalapack, :2660-3077): the bs x bs diagonal
// blocks of Q in FP64, the rest from FP32 copies of Q and X.
// C = A^T B over M rows with the library's FP64 MFMA kernel (lower-triangle
// tiles i >= j + diag_off, or all with diag_off = -1)
chfsi_status gram_fp64(chfsi_solver *s, const double *A, int64_t lda, const double *B,
                       int64_t ldb, int m, int n, int64_t M, int diag_off, double *C,
                       int64_t ldc, hipStream_t st) {
  const size_t pb = gram_partial_bytes(M, m, n, diag_off);
  if (pb == 0) return CHFSI_OK;
  CHFSI_TRY(s->split.ensure(pb));
  CHFSI_TRY(s->gtiles.ensure(sizeof(int2) * kGramMaxTiles));
  CHFSI_HIP(launch_gram(A, lda, B, ldb, m, n, M, diag_off, C, ldc, sptr<double>(s->split),
                        static_cast<int2 *>(s->gtiles.ptr), st));
  return CHFSI_OK;
}

// X <- X R in row panels with the library's FP64 MFMA kernel; R(i, j) =
// Rm[i + j ldr] (trans 0) or Rm[j + i ldr] (trans 1), upper triangular when upper
// Panel p is rotated into buffer p % 2 on st; its copy back into X runs on the
// solver's copy stream while panel p + 1 is rotated (the rotation reads only
// its own panel of X).  Before panel p + 2 reuses the buffer, st waits for
// panel p's copy; at the end st waits for every copy.  upper: R is upper
// triangular with its lower part stored as zeros; the panel is rotated by
// column blocks of 256, block [c0, c0 + 256) over the inner extent
// c0 + 256 only, so every launch has work items of one length.
chfsi_status rotate_fp64(chfsi_solver *s, double *X, int64_t M, int N, const double *Rm,
                         int64_t ldr, int trans, int upper, hipStream_t st) {
  const int64_t chunk = panel_rows(M, N);
  const int64_t npanel = (M + chunk - 1) / chunk;
  const int nbuf = npanel > 1 ? 2 : 1;
  CHFSI_TRY(s->rot.ensure(sizeof(double) * chunk * N * nbuf));
  if (!s->copy_stream) CHFSI_HIP(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
  for (int b = 0; b < 2; ++b) {
    if (!s->rot_done[b])
      CHFSI_HIP(hipEventCreateWithFlags(&s->rot_done[b], hipEventDisableTiming));
    if (!s->copy_done[b])
      CHFSI_HIP(hipEventCreateWithFlags(&s->copy_done[b], hipEventDisableTiming));
  }
  // one panel; on a failure the copies already queued are waited for before
  // the error returns, so that none of them writes X after the call
  auto panel = [&](int64_t p) -> chfsi_status {
    const int b = (int)(p & 1);
    const int64_t n0 = p * chunk, nc = std::min(chunk, M - n0);
    double *tmp = sptr<double>(s->rot) + (int64_t)b * chunk * N;
    if (p >= 2) CHFSI_HIP(hipStreamWaitEvent(st, s->copy_done[b], 0));
    if (upper) {
      constexpr int CB = 256;
      for (int c0 = 0; c0 < N; c0 += CB) {
        const int bw = std::min(CB, N - c0);
        // R(i, c0 + j): Rm[c0 + j + i ldr] (trans) or Rm[i + (c0 + j) ldr]
        const double *Rc = trans ? Rm + c0 : Rm + (int64_t)c0 * ldr;
        CHFSI_HIP(launch_rotate(X + n0 * N, N, Rc, ldr, trans, c0 + bw, bw, nc, tmp + c0, N, st));
      }
    } else {
      CHFSI_HIP(launch_rotate(X + n0 * N, N, Rm, ldr, trans, N, N, nc, tmp, N, st));
    }
    CHFSI_HIP(hipEventRecord(s->rot_done[b], st));
    CHFSI_HIP(hipStreamWaitEvent(s->copy_stream, s->rot_done[b], 0));
    CHFSI_HIP(hipMemcpyAsync(X + n0 * N, tmp, sizeof(double) * nc * N, hipMemcpyDeviceToDevice,
                             s->copy_stream));
    CHFSI_HIP(hipEventRecord(s->copy_done[b], s->copy_stream));
    return CHFSI_OK;
  };
  for (int64_t p = 0; p < npanel; ++p) {
    const chfsi_status stl = panel(p);
    if (stl != CHFSI_OK) {
      (void)hipStreamSynchronize(s->copy_stream);
      return stl;
    }
  }
  for (int b = 0; b < nbuf; ++b) CHFSI_HIP(hipStreamWaitEvent(st, s->copy_done[b], 0));
  return CHFSI_OK;
}

chfsi_status rotate_dense(chfsi_solver *s, double *X, const float *Xf, int64_t M, int N,
                          const double *Q, int bs, hipStream_t st) {
  if (!Xf && s->params.dense_rocblas == 0) return rotate_fp64(s, X, M, N, Q, N, 0, 0, st);
  const int64_t chunk = panel_rows(M, N);
  CHFSI_TRY(s->rot.ensure(sizeof(double) * chunk * N));
  double *tmp = sptr<double>(s->rot);
  const double one = 1.0, zero = 0.0;
  float *Qf = nullptr, *tmpf = nullptr;
  if (Xf) {
    CHFSI_TRY(s->Qf.ensure(sizeof(float) * N * N));
    CHFSI_TRY(s->rot32.ensure(sizeof(float) * chunk * N));
    Qf = sptr<float>(s->Qf);
    tmpf = sptr<float>(s->rot32);
    CHFSI_HIP(launch_cast_blocks(Q, N, N, bs, 1, Qf, st));
  }
  const float onef = 1.0f, zerof = 0.0f;
  for (int64_t n0 = 0; n0 < M; n0 += chunk) {
    const int64_t nc = std::min(chunk, M - n0);
    if (!Xf) {
      CHFSI_RB(rocblas_dgemm(s->blas, rocblas_operation_transpose, rocblas_operation_none, N,
                             (rocblas_int)nc, N, &one, Q, N, X + n0 * N, N, &zero, tmp, N));
    } else {
      CHFSI_RB(rocblas_sgemm(s->blas, rocblas_operation_transpose, rocblas_operation_none, N,
                             (rocblas_int)nc, N, &onef, Qf, N, Xf + n0 * N, N, &zerof, tmpf, N));
      for (int i0 = 0; i0 < N; i0 += bs) {
        const int bw = std::min(bs, N - i0);
        CHFSI_RB(rocblas_dgemm(s->blas, rocblas_operation_transpose, rocblas_operation_none, bw,
                               (rocblas_int)nc, bw, &one, Q + i0 + (int64_t)i0 * N, N,
                               X + n0 * N + i0, N, &zero, tmp + i0, N));
      }
      CHFSI_HIP(launch_add_f32_to_f64(tmpf, N, tmp, N, N, nc, st));
    }
    CHFSI_HIP(hipMemcpyAsync(X + n0 * N, tmp, sizeof(double) * nc * N, hipMemcpyDeviceToDevice,
                             st));
  }
  return CHFSI_OK;
}

// Xc <- Linv Xc with Linv lower triangular, by DOF panels and row blocks of bs
// (only the blocks on and below the diagonal are multiplied).  Mixed
// (subspaceRotationCGSMixedPrecScalapack, :2243-2659): the diagonal blocks in
// FP64, the blocks left of them from FP32 copies.
chfsi_status rotate_lower(chfsi_solver *s, double *X, const float *Xf, int64_t M, int N,
                          const double *Linv, int bs, hipStream_t st) {
  // X L^{-T}: R(i, j) = Linv[j + i N], upper triangular
  if (!Xf && s->params.dense_rocblas == 0) return rotate_fp64(s, X, M, N, Linv, N, 1, 1, st);
  const int64_t chunk = panel_rows(M, N);
  CHFSI_TRY(s->rot.ensure(sizeof(double) * chunk * N));
  double *tmp = sptr<double>(s->rot);
  const double one = 1.0, zero = 0.0;
  const float onef = 1.0f, zerof = 0.0f;
  float *Lf = nullptr, *tmpf = nullptr;
  if (Xf) {
    CHFSI_TRY(s->Qf.ensure(sizeof(float) * N * N));
    CHFSI_TRY(s->rot32.ensure(sizeof(float) * chunk * bs));
    Lf = sptr<float>(s->Qf);
    tmpf = sptr<float>(s->rot32);
    CHFSI_HIP(launch_cast_blocks(Linv, N, N, bs, 0, Lf, st));
  }
  for (int64_t n0 = 0; n0 < M; n0 += chunk) {
    const int64_t nc = std::min(chunk, M - n0);
    for (int i0 = 0; i0 < N; i0 += bs) {
      const int bw = std::min(bs, N - i0);
      if (!Xf) {
        CHFSI_RB(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_none, bw,
                               (rocblas_int)nc, i0 + bw, &one, Linv + i0, N, X + n0 * N, N, &zero,
                               tmp + i0, N));
      } else {
        CHFSI_RB(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_none, bw,
                               (rocblas_int)nc, bw, &one, Linv + i0 + (int64_t)i0 * N, N,
                               X + n0 * N + i0, N, &zero, tmp + i0, N));
        if (i0 > 0) {
          CHFSI_RB(rocblas_sgemm(s->blas, rocblas_operation_none, rocblas_operation_none, bw,
                                 (rocblas_int)nc, i0, &onef, Lf + i0, N, Xf + n0 * N, N, &zerof,
                                 tmpf, bw));
          CHFSI_HIP(launch_add_f32_to_f64(tmpf, bw, tmp + i0, N, bw, nc, st));
        }
      }
    }
    CHFSI_HIP(hipMemcpyAsync(X + n0 * N, tmp, sizeof(double) * nc * N, hipMemcpyDeviceToDevice,
                             st));
  }
  return CHFSI_OK;
}

// Lower block triangle of S = Xc Xc^T by column blocks of bs; mixed
// (fillParallelOverlapMatMixedPrecScalapack, :3831-4231): the diagonal blocks
// in FP64, the blocks below them from the FP32 copy.
chfsi_status overlap_lower(chfsi_solver *s, const double *X, const float *Xf, int64_t M, int N,
                           int bs, double *S, hipStream_t st) {
  if (!Xf && s->params.dense_rocblas != 1) return gram_fp64(s, X, N, X, N, N, N, M, 0, S, N, st);
  for (int j = 0; j < N; j += bs) {
    const int bw = std::min(bs, N - j);
    double *Sj = S + j + (int64_t)j * N;
    if (!Xf) {
      CHFSI_TRY(gemm_nt_splitk(s, X + j, N, X + j, N, N - j, bw, M, Sj, N, st));
    } else {
      CHFSI_TRY(gemm_nt_splitk(s, X + j, N, X + j, N, bw, bw, M, Sj, N, st));
      if (N - j - bw > 0)
        CHFSI_TRY(sgemm_nt_splitk(s, Xf + j + bw, N, Xf + j, N, N - j - bw, bw, M, Sj + bw, N,
                                  st));
    }
  }
  return CHFSI_OK;
}

// S = X^T X (lower), S = L L^T, X <- X L^{-T}.  Column-major view: Xc = X^T is
// N x M with leading dimension N, so S = Xc Xc^T and Xc <- L^{-1} Xc, applied
// as block-triangular GEMMs with the explicit triangular inverse.
chfsi_status cholgs_impl(chfsi_solver *s, double *X, int64_t M, int N, int bs, bool mixed_o,
                         bool mixed_sr, hipStream_t st) {
  CHFSI_TRY(ensure_blas(s, st));
  CHFSI_TRY(s->Hp.ensure(sizeof(double) * N * N));
  CHFSI_TRY(s->info.ensure(sizeof(rocblas_int)));
  double *S = sptr<double>(s->Hp);
  float *Xf = nullptr;
  if (mixed_o || mixed_sr) CHFSI_TRY(cast_x(s, X, M, N, &Xf, st));
  CHFSI_TRY(overlap_lower(s, X, mixed_o ? Xf : nullptr, M, N, bs, S, st));
  rocblas_int *dinfo = sptr<rocblas_int>(s->info);
  CHFSI_RB(rocsolver_dpotrf(s->blas, rocblas_fill_lower, N, S, N, dinfo));
  rocblas_int hinfo = 0;
  CHFSI_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) {
    set_error("overlap matrix is not positive definite (potrf info %d)", (int)hinfo);
    return CHFSI_ERR_NOT_SPD;
  }
  CHFSI_RB(rocsolver_dtrtri(s->blas, rocblas_fill_lower, rocblas_diagonal_non_unit, N, S, N,
                            dinfo));
  CHFSI_HIP(launch_zero_upper(S, N, N, st));
  return rotate_lower(s, X, mixed_sr ? Xf : nullptr, M, N, S, bs, st);
}

// Lower triangle of Hp = X^T H~ X by column blocks of B (XtHX,
// kohnShamDFTOperatorDevice.cc:4000-4157).  Mixed (XtHXMixedPrec, :4161-):
// the diagonal blocks in FP64, the blocks below them from the FP32 copy Xf of
// X and an FP32 copy of the HX block.
chfsi_status xthx_lower(chfsi_solver *s, chfsi_operator *op, const double *X, int64_t M, int N,
                        int B, double *Hp, const float *Xf, hipStream_t st) {
  const int G = filter_group_columns(s->params, M, N, B);
  CHFSI_TRY(s->HXb.ensure(sizeof(double) * M * G));
  double *HXg = sptr<double>(s->HXb);
  float *HXf = nullptr;
  if (Xf) {
    CHFSI_TRY(s->HXf.ensure(sizeof(float) * M * B));
    HXf = sptr<float>(s->HXf);
  }
  int g0 = 0, gw = 0;  // columns [g0, g0 + gw) of H~X are in HXg (ld gw)
  if (!Xf && s->params.dense_rocblas != 1) {
    // per group of columns: Hp[:, g0:g0+gw] = X^T (H~X)[:, g0:g0+gw], the
    // tiles that reach the lower triangle
    for (g0 = 0; g0 < N; g0 += G) {
      gw = std::min(G, N - g0);
      CHFSI_TRY(hx_fused<double>(op, X + g0, N, HXg, gw, gw, 1.0, 0.0, 0.0, 1, true, st));
      if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, HXg, gw, gw, st));
      CHFSI_TRY(gram_fp64(s, X, N, HXg, gw, N, gw, M, g0, Hp + (int64_t)g0 * N, N, st));
    }
    return CHFSI_OK;
  }
  for (int j = 0; j < N; j += B) {
    const int bw = std::min(B, N - j);
    if (j >= g0 + gw) {
      g0 = j;
      gw = std::min(G, N - j);
      CHFSI_TRY(hx_fused<double>(op, X + j, N, HXg, gw, gw, 1.0, 0.0, 0.0, 1, true, st));
      if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, HXg, gw, gw, st));
    }
    const double *HX = HXg + (j - g0);
    double *Hj = Hp + j + (int64_t)j * N;
    if (!Xf) {
      // Hp[j:N, j:j+bw] = Xc[j:N, :] * HXc^T
      CHFSI_TRY(gemm_nt_splitk(s, X + j, N, HX, gw, N - j, bw, M, Hj, N, st));
    } else {
      CHFSI_TRY(gemm_nt_splitk(s, X + j, N, HX, gw, bw, bw, M, Hj, N, st));
      if (N - j - bw > 0) {
        CHFSI_HIP(launch_copy_block<double, float>(HX, gw, HXf, bw, M, bw, st));
        CHFSI_TRY(sgemm_nt_splitk(s, Xf + j + bw, N, HXf, bw, N - j - bw, bw, M, Hj + bw, N, st));
      }
    }
  }
  return CHFSI_OK;
}

// Eigendecomposition of the symmetric N x N matrix A (lower triangle used) in
// place: A <- Q, ascending eigenvalues to the host.
chfsi_status sym_eig(chfsi_solver *s, double *A, int N, double *eig_host, hipStream_t st) {
  CHFSI_TRY(s->evals.ensure(sizeof(double) * 2 * N));
  CHFSI_TRY(s->info.ensure(sizeof(rocblas_int)));
  double *D = sptr<double>(s->evals), *E = D + N;
  rocblas_int *dinfo = sptr<rocblas_int>(s->info);
  CHFSI_RB(rocsolver_dsyevd(s->blas, rocblas_evect_original, rocblas_fill_lower, N, A, N, D, E,
                            dinfo));
  rocblas_int hinfo = 0;
  CHFSI_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipMemcpyAsync(eig_host, D, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) {
    set_error("projected eigensolve failed (syevd info %d)", (int)hinfo);
    return CHFSI_ERR_EIG;
  }
  return CHFSI_OK;
}

// Hp = X^T H~ X (lower triangle by column blocks), eig, X <- X Q
// (rayleighRitz, rayleighRitzDevice.cc:80-353).  mixed_xthx / mixed_rot: the
// reference's USE MIXED PREC XTHX / RR_SR forms.
chfsi_status rr_impl(chfsi_solver *s, chfsi_operator *op, double *X, int64_t M, int N, int B,
                     double *eig_host, bool mixed_xthx, bool mixed_rot, hipStream_t st) {
  CHFSI_TRY(ensure_blas(s, st));
  CHFSI_TRY(s->Hp.ensure(sizeof(double) * N * N));
  double *Hp = sptr<double>(s->Hp);
  float *Xf = nullptr;
  if (mixed_xthx || mixed_rot) CHFSI_TRY(cast_x(s, X, M, N, &Xf, st));
  CHFSI_TRY(xthx_lower(s, op, X, M, N, B, Hp, mixed_xthx ? Xf : nullptr, st));
  CHFSI_TRY(sym_eig(s, Hp, N, eig_host, st));
  return rotate_dense(s, X, mixed_rot ? Xf : nullptr, M, N, Hp, B, st);
}

// Rayleigh-Ritz in the generalised form on a filtered, not orthonormalised X
// (rayleighRitzGEP, rayleighRitzDevice.cc:355-819): S = X^T X = L L^T,
// Hp = X^T H~ X, Hhat = L^{-1} Hp L^{-T} = Q diag(lambda) Q^T, and
// X <- X (L^{-T} Q).  Mixed forms as rr_impl, with USE MIXED PREC CGS O for S.
chfsi_status gep_impl(chfsi_solver *s, chfsi_operator *op, double *X, int64_t M, int N, int B,
                      double *eig_host, bool mixed_o, bool mixed_xthx, bool mixed_rot,
                      hipStream_t st) {
  CHFSI_TRY(ensure_blas(s, st));
  CHFSI_TRY(s->Hp.ensure(sizeof(double) * N * N));
  CHFSI_TRY(s->gep.ensure(sizeof(double) * 2 * N * N));
  CHFSI_TRY(s->info.ensure(sizeof(rocblas_int)));
  double *Hp = sptr<double>(s->Hp), *Linv = sptr<double>(s->gep), *W = Linv + (int64_t)N * N;
  float *Xf = nullptr;
  if (mixed_o || mixed_xthx || mixed_rot) CHFSI_TRY(cast_x(s, X, M, N, &Xf, st));
  // overlap and its Cholesky factor (pseudoGSDevice.cc:121-301 steps)
  CHFSI_TRY(overlap_lower(s, X, mixed_o ? Xf : nullptr, M, N, B, Linv, st));
  rocblas_int *dinfo = sptr<rocblas_int>(s->info);
  CHFSI_RB(rocsolver_dpotrf(s->blas, rocblas_fill_lower, N, Linv, N, dinfo));
  rocblas_int hinfo = 0;
  CHFSI_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) {
    set_error("overlap matrix is not positive definite (potrf info %d)", (int)hinfo);
    return CHFSI_ERR_NOT_SPD;
  }
  CHFSI_RB(rocsolver_dtrtri(s->blas, rocblas_fill_lower, rocblas_diagonal_non_unit, N, Linv, N,
                            dinfo));
  CHFSI_HIP(launch_zero_upper(Linv, N, N, st));
  // projected Hamiltonian of the unrotated X, then Hhat = Linv Hp Linv^T
  CHFSI_TRY(xthx_lower(s, op, X, M, N, B, Hp, mixed_xthx ? Xf : nullptr, st));
  const double one = 1.0, zero = 0.0;
  CHFSI_RB(rocblas_dsymm(s->blas, rocblas_side_right, rocblas_fill_lower, N, N, &one, Hp, N,
                         Linv, N, &zero, W, N));
  CHFSI_RB(rocblas_dtrmm(s->blas, rocblas_side_right, rocblas_fill_lower,
                         rocblas_operation_transpose, rocblas_diagonal_non_unit, N, N, &one, Linv,
                         N, W, N, Hp, N));
  CHFSI_TRY(sym_eig(s, Hp, N, eig_host, st));
  // rotation matrix L^{-T} Q
  CHFSI_RB(rocblas_dtrmm(s->blas, rocblas_side_left, rocblas_fill_lower,
                         rocblas_operation_transpose, rocblas_diagonal_non_unit, N, N, &one, Linv,
                         N, Hp, N, W, N));
  return rotate_dense(s, X, mixed_rot ? Xf : nullptr, M, N, W, B, st);
}

// With rescale (nullable) each column block of X is multiplied by rescale[g]
// row-wise in place once its residuals are formed (X must then be writable;
// the blocks not yet applied stay untouched).
chfsi_status residual_impl(chfsi_solver *s, chfsi_operator *op, const double *X, int64_t ldX,
                           int64_t M, int N, int B, const double *eig_host, double *res_host,
                           hipStream_t st, const double *rescale = nullptr) {
  const int G = filter_group_columns(s->params, M, N, B);
  CHFSI_TRY(s->HXb.ensure(sizeof(double) * M * G));
  CHFSI_TRY(s->lam.ensure(sizeof(double) * N));
  const int nch = 512;
  CHFSI_TRY(s->part.ensure(sizeof(double) * nch * B));
  double *HX = sptr<double>(s->HXb), *lam = sptr<double>(s->lam), *part = sptr<double>(s->part);
  CHFSI_HIP(hipMemcpyAsync(lam, eig_host, sizeof(double) * N, hipMemcpyHostToDevice, st));
  std::vector<double> h(size_t(nch) * B);
  int g0 = 0, gw = 0;  // columns [g0, g0 + gw) of H~X are in HX (ld gw)
  for (int j = 0; j < N; j += B) {
    const int bw = std::min(B, N - j);
    if (j >= g0 + gw) {
      g0 = j;
      gw = std::min(G, N - j);
      CHFSI_TRY(hx_fused<double>(op, X + j, ldX, HX, gw, gw, 1.0, 0.0, 0.0, 1, true, st));
      if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, HX, gw, gw, st));
    }
    CHFSI_HIP(launch_residual_partial(HX + (j - g0), gw, X + j, ldX, lam + j, M, bw, part, nch,
                                      st, rescale ? const_cast<double *>(X) + j : nullptr,
                                      rescale));
    CHFSI_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nch * bw, hipMemcpyDeviceToHost, st));
    CHFSI_HIP(hipStreamSynchronize(st));
    for (int v = 0; v < bw; ++v) {
      double acc = 0;
      for (int k = 0; k < nch; ++k) acc += h[size_t(k) * bw + v];
      res_host[j + v] = std::sqrt(acc);
    }
  }
  return CHFSI_OK;
}

// Spectrum-split Rayleigh-Ritz on an orthonormal X (rayleighRitzGEPSpectrumSplitDirect,
// rayleighRitzDevice.cc:822-1278, after its Cholesky step): Hp = X^T H~ X (lower
// triangle by column blocks), full eigendecomposition, and only the top
// Nfr = N - Noc pairs kept: Xfrac = X Q[:, Noc:N] (M x Nfr, leading dimension
// ldf), eigenvalues ascending into eig_host[0..Nfr).  X itself is not
// rotated.  Mixed (useMixedPrecXTHXSpectrumSplit, XtHXMixedPrec,
// kohnShamDFTOperatorDevice.cc:4552-5081): column blocks that lie entirely
// within the first Noc states are formed from FP32 copies of X and of the HX
// block; the blocks reaching the top Nfr states stay FP64.
chfsi_status rr_split_impl(chfsi_solver *s, chfsi_operator *op, const double *X, int64_t M, int N,
                           int Noc, int B, double *Xfrac, int64_t ldf, double *eig_host,
                           bool mixed_xthx, hipStream_t st) {
  const int Nfr = N - Noc;
  CHFSI_TRY(ensure_blas(s, st));
  CHFSI_TRY(s->Hp.ensure(sizeof(double) * N * N));
  const int G = filter_group_columns(s->params, M, N, B);
  CHFSI_TRY(s->HXb.ensure(sizeof(double) * M * G));
  CHFSI_TRY(s->evals.ensure(sizeof(double) * 2 * N));
  CHFSI_TRY(s->info.ensure(sizeof(rocblas_int)));
  double *Hp = sptr<double>(s->Hp), *HXg = sptr<double>(s->HXb);
  float *Xf = nullptr, *HXf = nullptr;
  if (mixed_xthx && Noc >= B) {
    CHFSI_TRY(cast_x(s, X, M, N, &Xf, st));
    CHFSI_TRY(s->HXf.ensure(sizeof(float) * M * B));
    HXf = sptr<float>(s->HXf);
  }
  int g0 = 0, gw = 0;  // columns [g0, g0 + gw) of H~X are in HXg (ld gw)
  for (int j = 0; j < N; j += B) {
    const int bw = std::min(B, N - j);
    if (j >= g0 + gw) {
      g0 = j;
      gw = std::min(G, N - j);
      CHFSI_TRY(hx_fused<double>(op, X + j, N, HXg, gw, gw, 1.0, 0.0, 0.0, 1, true, st));
      if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, HXg, gw, gw, st));
    }
    const double *HX = HXg + (j - g0);
    double *Hj = Hp + j + (int64_t)j * N;
    if (Xf && j + bw <= Noc) {
      CHFSI_HIP(launch_copy_block<double, float>(HX, gw, HXf, bw, M, bw, st));
      CHFSI_TRY(sgemm_nt_splitk(s, Xf + j, N, HXf, bw, N - j, bw, M, Hj, N, st));
    } else {
      CHFSI_TRY(gemm_nt_splitk(s, X + j, N, HX, gw, N - j, bw, M, Hj, N, st));
    }
  }
  double *D = sptr<double>(s->evals), *E = D + N;
  rocblas_int *dinfo = sptr<rocblas_int>(s->info);
  CHFSI_RB(rocsolver_dsyevd(s->blas, rocblas_evect_original, rocblas_fill_lower, N, Hp, N, D, E,
                            dinfo));
  rocblas_int hinfo = 0;
  CHFSI_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipMemcpyAsync(eig_host, D + Noc, sizeof(double) * Nfr, hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) {
    set_error("projected eigensolve failed (syevd info %d)", (int)hinfo);
    return CHFSI_ERR_EIG;
  }
  // Xfrac^T (Nfr x M, ld ldf) = Q[:, Noc:N]^T (Nfr x N) * Xc (N x M, ld N)
  const double one = 1.0, zero = 0.0;
  CHFSI_RB(rocblas_dgemm(s->blas, rocblas_operation_transpose, rocblas_operation_none, Nfr,
                         (rocblas_int)M, N, &one, Hp + (int64_t)Noc * N, N, X, N, &zero, Xfrac,
                         (rocblas_int)ldf));
  return CHFSI_OK;
}

}  // namespace

extern "C" void chfsi_default_params(chfsi_params *p) {
  if (!p) return;
  p->cheby_block_size = 400;            // utils/dftParameters.cc:1232
  p->cheby_degree = 0;
  p->lanczos_steps = 20;
  p->reproducible = 0;
  p->reuse_lanczos_upper = 0;
  p->first_scf_degree_factor = 1.34;    // utils/dftParameters.cc:872-876
  p->filter_precision = CHFSI_FP64;    // USE MIXED PREC CHEBY false, utils/dftParameters.cc:1262
  p->use_shep = 1;                      // utils/dftParameters.cc:934-938
  p->fp64_tail_degrees = 0;
  p->mixed_prec_overlap = 0;            // utils/dftParameters.cc:1246-1249
  p->mixed_prec_cgs_rotation = 0;
  p->mixed_prec_xthx = 0;
  p->mixed_prec_rr_rotation = 0;
  p->mixed_prec_xthx_spectrum_split = 0;
  p->filter_columns = 0;
  p->mixed_prec_switch_residual = 0.0;
  p->mixed_prec_defect_filter = 0;
  p->is_pseudopotential = 0;            // utils/dftParameters.cc:1173
  p->lanczos_full_residual_margin = 0;
  p->mixed_prec_fp64_top_columns = 0;
  p->dense_rocblas = 0;
}

extern "C" chfsi_status chfsi_chebyshev_degree(const chfsi_params *params, double bUp,
                                               int32_t is_first_scf, int32_t *degree) {
  CHFSI_ARG(params && degree, "null argument");
  *degree = (int32_t)solve_degree(*params, bUp, is_first_scf != 0);
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_solver_set_params(chfsi_solver *s, const chfsi_params *params) {
  CHFSI_ARG(s && params, "null argument");
  s->params = *params;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_solver_create(const chfsi_params *params, double a0, double bLow,
                                            double bUp, chfsi_solver **out) {
  CHFSI_ARG(out, "null argument");
  auto s = new chfsi_solver();
  if (params)
    s->params = *params;
  else
    chfsi_default_params(&s->params);
  s->a0 = a0;
  s->bLow = bLow;
  s->bUp = bUp;
  *out = s;
  return CHFSI_OK;
}

extern "C" void chfsi_solver_destroy(chfsi_solver *s) { delete s; }

extern "C" chfsi_status chfsi_solver_reinit_bounds(chfsi_solver *s, double a0, double bLow,
                                                   double bUp) {
  CHFSI_ARG(s, "null solver");
  s->a0 = a0;
  s->bLow = bLow;
  s->bUp = bUp;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_solver_get_bounds(const chfsi_solver *s, double *a0, double *bLow,
                                                double *bUp) {
  CHFSI_ARG(s && a0 && bLow && bUp, "null argument");
  *a0 = s->a0;
  *bLow = s->bLow;
  *bUp = s->bUp;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_solver_last_filter_precision(const chfsi_solver *s,
                                                           int32_t *precision) {
  CHFSI_ARG(s && precision, "null argument");
  *precision = s->last_precision;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_solver_set_lanczos_start(chfsi_solver *s, const double *v0_dev,
                                                       int64_t M) {
  CHFSI_ARG(s, "null solver");
  s->have_v0 = false;
  s->v0.release();
  if (!v0_dev) return CHFSI_OK;
  CHFSI_ARG(M > 0, "bad length");
  CHFSI_TRY(s->v0.ensure(sizeof(double) * M));
  CHFSI_HIP(hipMemcpy(s->v0.ptr, v0_dev, sizeof(double) * M, hipMemcpyDeviceToDevice));
  s->have_v0 = true;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_chebyshev_filter(chfsi_operator *op, double *X, int64_t ld,
                                               int32_t B, int32_t m, double a, double b,
                                               double a0, int32_t precision, void *stream) {
  CHFSI_ARG(op && X && B > 0 && ld >= B && m >= 1, "bad argument");
  chfsi_solver tmp;
  return filter_block(&tmp, op, X, ld, B, m, a, b, a0, precision, 0, as_stream(stream));
}

// Reverse pass of the scaled Chebyshev recurrence
//   X_1 = k1 (H~ - c) X_0,   X_k = a1_k (H~ - c) X_{k-1} + a2_k X_{k-2}  (k = 2..m),  Y = X_m.
// With L_k = dL/dX_k, going from k = m down to 1:
//   g     += a1_k * adjV(L_k, X_{k-1})           (dL/dV through the HX of step k)
//   L_{k-1} += a1_k (H~ - c) L_k                  (one fused HX, beta = 1)
//   L_{k-2} += a2_k L_k
// and X0bar = L_0.  Forward iterates come from checkpoints (pairs X_{k-1}, X_k
// every `every` degrees) re-run one segment at a time.  T is the arithmetic of
// the iterates and adjoints (FP64, or FP32 on the FP32 apply); g is always
// accumulated in FP64 from FP64 interpolations of the FP32 vectors.
namespace {
template <typename T>
chfsi_status filter_backward_impl(chfsi_operator *op, const double *X0, int64_t ldx,
                                  const double *Ybar, int64_t ldy, int B, int m, double a,
                                  double b, double a0, int every, double *Xbar, int64_t ldxb,
                                  double *g, hipStream_t st) {
  const int64_t M = op->M;
  const size_t vec = size_t(M) * B;
  every = std::max(1, std::min<int>(every, m));
  ChebState cs = cheb_init(a, b, a0);
  std::vector<double> a1(m + 1, 0.0), a2(m + 1, 0.0);
  a1[1] = cs.sigma / cs.e;
  {
    ChebState t = cs;
    for (int k = 2; k <= m; ++k) {
      const double sigma2 = 1.0 / (t.gamma - t.sigma);
      a1[k] = 2.0 * sigma2 / t.e;
      a2[k] = -(t.sigma * sigma2);
      t.sigma = sigma2;
    }
  }
  const T c = T(cs.c);
  // storage: X_0 copy, checkpoint pairs, one segment of iterates, 3 adjoint buffers
  const int nseg = (m + every - 1) / every;
  Scratch ck, seg, adj;
  CHFSI_TRY(ck.ensure(sizeof(T) * vec * (1 + 2 * nseg)));
  CHFSI_TRY(seg.ensure(sizeof(T) * vec * (every + 2)));
  CHFSI_TRY(adj.ensure(sizeof(T) * vec * 3));
  T *ckp = sptr<T>(ck), *segp = sptr<T>(seg), *adjp = sptr<T>(adj);
  auto ckv = [&](int i) { return ckp + vec * i; };
  CHFSI_HIP(launch_copy_block<double, T>(X0, ldx, ckv(0), B, M, B, st));
  CHFSI_HIP(hipMemsetAsync(segp, 0, sizeof(T) * vec * (every + 2), st));
  // forward sweep storing (X_{k0-1}, X_{k0}) for every segment start k0 = s*every >= 1
  {
    T *prev = segp, *cur = segp + vec;
    CHFSI_HIP(launch_copy_block<T, T>(ckv(0), B, prev, B, M, B, st));
    int64_t ldc = B, ldp = B;
    CHFSI_TRY(hx_fused<T>(op, prev, B, cur, B, B, T(a1[1]), T(0), T(-c * a1[1]), 1, true, st));
    for (int k = 1; k <= m; ++k) {
      if (k % every == 0 && k < m) {
        const int sg = k / every;
        CHFSI_HIP(launch_copy_block<T, T>(prev, ldp, ckv(2 * sg - 1), B, M, B, st));
        CHFSI_HIP(launch_copy_block<T, T>(cur, ldc, ckv(2 * sg), B, M, B, st));
      }
      if (k == m) break;
      CHFSI_TRY(hx_fused<T>(op, cur, ldc, prev, ldp, B, T(a1[k + 1]), T(a2[k + 1]),
                            T(-c * a1[k + 1]), 0, true, st));
      std::swap(cur, prev);
      std::swap(ldc, ldp);
    }
  }
  // adjoint sweep: Lk = L_k, Lm1 = L_{k-1}, Lm2 = L_{k-2}
  T *Lk = adjp, *Lm1 = adjp + vec, *Lm2 = adjp + 2 * vec;
  CHFSI_HIP(launch_copy_block<double, T>(Ybar, ldy, Lk, B, M, B, st));
  CHFSI_HIP(hipMemsetAsync(Lm1, 0, sizeof(T) * vec * 2, st));
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int k0 = sg * every, k1 = std::min(m, k0 + every);
    // iterates X_{k0-1 .. k1-1} into seg[0 .. ]: seg[i] = X_{k0 - 1 + i}
    auto X = [&](int k) { return segp + vec * (k - k0 + 1); };
    if (k0 == 0) {
      CHFSI_HIP(launch_copy_block<T, T>(ckv(0), B, X(0), B, M, B, st));
    } else {
      CHFSI_HIP(launch_copy_block<T, T>(ckv(2 * sg - 1), B, X(k0 - 1), B, M, B, st));
      CHFSI_HIP(launch_copy_block<T, T>(ckv(2 * sg), B, X(k0), B, M, B, st));
    }
    for (int k = k0 + 1; k <= k1 - 1; ++k) {
      if (k == 1)
        CHFSI_TRY(hx_fused<T>(op, X(0), B, X(1), B, B, T(a1[1]), T(0), T(-c * a1[1]), 1, true,
                              st));
      else {
        CHFSI_HIP(launch_copy_block<T, T>(X(k - 2), B, X(k), B, M, B, st));
        CHFSI_TRY(hx_fused<T>(op, X(k - 1), B, X(k), B, B, T(a1[k]), T(a2[k]), T(-c * a1[k]), 0,
                              true, st));
      }
    }
    for (int k = k1; k >= k0 + 1; --k) {
      CHFSI_TRY(quad_product<T>(op, Lk, B, X(k - 1), B, B, true, op->d_jxw, nullptr, a1[k], g, 1,
                                st, 0));
      CHFSI_TRY(hx_fused<T>(op, Lk, B, Lm1, B, B, T(a1[k]), T(1), T(-c * a1[k]), 0, true, st));
      if (k >= 2) CHFSI_HIP(launch_axpby<T>((int64_t)vec, T(a2[k]), Lk, T(1), Lm2, st));
      // shift: (L_k, L_{k-1}, L_{k-2}) <- (L_{k-1}, L_{k-2}, 0)
      T *t = Lk;
      Lk = Lm1;
      Lm1 = Lm2;
      Lm2 = t;
      CHFSI_HIP(hipMemsetAsync(Lm2, 0, sizeof(T) * vec, st));
    }
  }
  CHFSI_HIP(launch_copy_block<T, double>(Lk, B, Xbar, ldxb, M, B, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  return CHFSI_OK;
}

}  // namespace

extern "C" chfsi_status chfsi_chebyshev_filter_backward(
    chfsi_operator *op, const double *X0, int64_t ldx, const double *Ybar, int64_t ldy, int32_t B,
    int32_t m, double a, double b, double a0, int32_t every, double *Xbar, int64_t ldxb,
    double *g, int32_t precision, void *stream) {
  CHFSI_ARG(op && X0 && Ybar && Xbar && g && B > 0 && m >= 1, "bad argument");
  CHFSI_ARG(ldx >= B && ldy >= B && ldxb >= B, "bad leading dimension");
  CHFSI_ARG(op->d_Nq && op->d_jxw, "operator was created without shape data / JxW");
  hipStream_t st = as_stream(stream);
  if (precision == CHFSI_FP64)
    return filter_backward_impl<double>(op, X0, ldx, Ybar, ldy, B, m, a, b, a0, every, Xbar, ldxb,
                                        g, st);
  CHFSI_ARG(precision == CHFSI_FP32, "filter backward precision must be CHFSI_FP64 or CHFSI_FP32");
  return filter_backward_impl<float>(op, X0, ldx, Ybar, ldy, B, m, a, b, a0, every, Xbar, ldxb, g,
                                     st);
}

extern "C" chfsi_status chfsi_lanczos_bounds(chfsi_operator *op, const double *v0, int32_t steps,
                                             int32_t reproducible, double *lower, double *upper,
                                             void *stream) {
  CHFSI_ARG(op && v0 && lower && upper, "null argument");
  chfsi_solver tmp;
  return lanczos_impl(&tmp, op, v0, steps, reproducible, lower, upper, as_stream(stream));
}

extern "C" chfsi_status chfsi_cholgs(chfsi_operator *op, double *X, int64_t M, int32_t N,
                                     void *stream) {
  CHFSI_ARG(op && X && M == op->M && N > 0, "bad argument");
  chfsi_solver tmp;
  return cholgs_impl(&tmp, X, M, N, std::min<int>(N, 256), false, false, as_stream(stream));
}

extern "C" chfsi_status chfsi_rayleigh_ritz(chfsi_operator *op, double *X, int64_t M, int32_t N,
                                            int32_t block, double *eig_host, void *stream) {
  CHFSI_ARG(op && X && eig_host && M == op->M && N > 0 && block > 0, "bad argument");
  chfsi_solver tmp;
  return rr_impl(&tmp, op, X, M, N, std::min<int>(block, N), eig_host, false, false,
                 as_stream(stream));
}

extern "C" chfsi_status chfsi_residual_norms(chfsi_operator *op, const double *X, int64_t M,
                                             int32_t N, int32_t block, const double *eig_host,
                                             double *res_host, void *stream) {
  CHFSI_ARG(op && X && eig_host && res_host && M == op->M && N > 0 && block > 0, "bad argument");
  chfsi_solver tmp;
  return residual_impl(&tmp, op, X, N, M, N, std::min<int>(block, N), eig_host, res_host,
                       as_stream(stream));
}

extern "C" chfsi_status chfsi_rayleigh_ritz_spectrum_split(chfsi_operator *op, double *X, int64_t M,
                                                           int32_t N, int32_t n_core,
                                                           double *Xfrac, int64_t ld_frac,
                                                           int32_t block, double *eig_host,
                                                           int32_t mixed_xthx, void *stream) {
  CHFSI_ARG(op && X && Xfrac && eig_host && M == op->M && N > 0 && block > 0, "bad argument");
  CHFSI_ARG(n_core >= 0 && n_core < N && ld_frac >= N - n_core, "need 0 <= n_core < N, ld_frac >= N - n_core");
  hipStream_t st = as_stream(stream);
  chfsi_solver tmp;
  const int B = std::min<int>(block, N);
  CHFSI_TRY(cholgs_impl(&tmp, X, M, N, B, false, false, st));
  return rr_split_impl(&tmp, op, X, M, N, n_core, B, Xfrac, ld_frac, eig_host, mixed_xthx != 0, st);
}

namespace {
// Start of a pass shared by both solve forms (…SolverDevice.cc:241-450): Lanczos
// bounds when needed, degree, M^{1/2} scaling into the Loewdin basis and the
// blocked filter.  Records ev[1] after Lanczos and ev[2] after the filter.
// Lanczos bounds from the supplied start vector, or the seeded default one.
chfsi_status run_lanczos(chfsi_solver *s, chfsi_operator *op, int64_t M, double *lo, double *up,
                         hipStream_t st) {
  const chfsi_params &P = s->params;
  const double *v0 = s->have_v0 ? sptr<double>(s->v0) : nullptr;
  std::vector<double> hv;
  Scratch gen;
  if (!v0) {
    hv = default_lanczos_start(M);
    CHFSI_TRY(gen.ensure(sizeof(double) * M));
    CHFSI_HIP(hipMemcpy(gen.ptr, hv.data(), sizeof(double) * M, hipMemcpyHostToDevice));
    v0 = sptr<double>(gen);
  }
  chfsi_status stl = lanczos_impl(s, op, v0, P.lanczos_steps,
                                  P.reproducible || P.lanczos_full_residual_margin, lo, up, st);
  gen.release();
  return stl;
}

// Filter all N columns of the Loewdin-scaled X in groups with degree m; the
// precision is the filter precision of a mixed pass unless the residual-driven
// switch has handed the passes to FP64.
chfsi_status filter_all(chfsi_solver *s, chfsi_operator *op, double *X, int64_t M, int N, int m,
                        int32_t use_mixed_prec_overall, hipStream_t st) {
  const chfsi_params &P = s->params;
  const int B = std::min<int>(P.cheby_block_size > 0 ? P.cheby_block_size : N, N);
  const bool switched =
      P.mixed_prec_switch_residual > 0 && s->last_res_lowest < P.mixed_prec_switch_residual;
  const int prec = use_mixed_prec_overall && !switched ? P.filter_precision : CHFSI_FP64;
  s->last_precision = prec;
  const int G = filter_group_columns(P, M, N, B);
  // the top columns in FP64 when the filter runs in a lower precision
  const int ktop = prec == CHFSI_FP64 ? 0 : std::max(0, std::min(P.mixed_prec_fp64_top_columns, N));
  const int nlow = N - ktop;
  const int tail = use_mixed_prec_overall ? P.fp64_tail_degrees : 0;
  // The FP64 top columns and the low-precision columns are independent
  // recurrences: the top groups run on the solver's filter stream beside the
  // low groups on st, so that the colour launches of each fill the other's
  // ramp and drain.  The top groups' HX projections use their own scratch
  // (top_hx), so the defect form's FP64 apply on st keeps the operator's
  // buffers; the low groups never touch Y64 unless they run an FP64 tail, in
  // which case both sets stay on st.
  const bool side = ktop > 0 && nlow > 0 &&
                    ((prec == CHFSI_FP32 && P.mixed_prec_defect_filter) || tail == 0);
  hipStream_t top_st = st;
  void *top_ws = nullptr;
  if (side) {
    if (!s->filter_stream)
      CHFSI_HIP(hipStreamCreateWithFlags(&s->filter_stream, hipStreamNonBlocking));
    for (hipEvent_t *e : {&s->filter_fork, &s->filter_join})
      if (!*e) CHFSI_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    if (op->n_slots > 0) {
      CHFSI_TRY(s->top_hx.ensure(
          noalloc_scratch_bytes(sizeof(double), op->n_slots, op->n_proj, std::min(G, ktop))));
      top_ws = s->top_hx.ptr;
    }
    top_st = s->filter_stream;
    CHFSI_HIP(hipEventRecord(s->filter_fork, st));
    CHFSI_HIP(hipStreamWaitEvent(top_st, s->filter_fork, 0));
    // st waits for whatever the filter stream was given, also when a group
    // fails to enqueue, so that no work on X outlives the call unordered
    auto join = [&](chfsi_status stl) -> chfsi_status {
      const hipError_t e1 = hipEventRecord(s->filter_join, top_st);
      const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(st, s->filter_join, 0) : e1;
      if (stl == CHFSI_OK && e2 != hipSuccess) {
        set_error("HIP error %s while joining the filter stream", hipGetErrorString(e2));
        return CHFSI_ERR_HIP;
      }
      if (e2 != hipSuccess) (void)hipStreamSynchronize(top_st);
      return stl;
    };
    chfsi_status stl = CHFSI_OK;
    for (int j = nlow; j < N && stl == CHFSI_OK; j += G) {
      const int bw = std::min(G, N - j);
      stl = filter_block(s, op, X + j, N, bw, m, s->bLow, s->bUp, s->a0, CHFSI_FP64, 0, top_st,
                         top_ws);
    }
    for (int j = 0; j < nlow && stl == CHFSI_OK; j += G) {
      const int bw = std::min(G, nlow - j);
      stl = filter_block(s, op, X + j, N, bw, m, s->bLow, s->bUp, s->a0, prec, tail, st);
    }
    return join(stl);
  }
  for (int j = 0; j < nlow; j += G) {
    const int bw = std::min(G, nlow - j);
    CHFSI_TRY(filter_block(s, op, X + j, N, bw, m, s->bLow, s->bUp, s->a0, prec, tail, st));
  }
  for (int j = nlow; j < N; j += G) {
    const int bw = std::min(G, N - j);
    CHFSI_TRY(filter_block(s, op, X + j, N, bw, m, s->bLow, s->bUp, s->a0, CHFSI_FP64, 0, st));
  }
  return CHFSI_OK;
}

chfsi_status lanczos_scale_filter(chfsi_solver *s, chfsi_operator *op, double *X, int64_t M, int N,
                                  int32_t is_first_filtering_call, int32_t use_mixed_prec_overall,
                                  int32_t is_first_scf, int *B_out, hipEvent_t *ev,
                                  hipStream_t st) {
  const chfsi_params &P = s->params;
  if (is_first_filtering_call || !P.reuse_lanczos_upper) {
    double lo, up;
    CHFSI_TRY(run_lanczos(s, op, M, &lo, &up, st));
    if (is_first_filtering_call) {
      s->a0 = lo;
      s->bUp = up;
      s->bLow = lo + (up - lo) * N / double(M) * (P.reproducible ? 10.0 : 200.0);
    } else {
      s->bUp = up;
    }
  }
  CHFSI_HIP(hipEventRecord(ev[1], st));

  const unsigned m = solve_degree(P, s->bUp, is_first_scf != 0);

  CHFSI_HIP(launch_scale_rows<double>(X, N, M, N, 1.0, op->d_sqrt_mass, st));
  const int B = std::min<int>(P.cheby_block_size > 0 ? P.cheby_block_size : N, N);
  if (is_first_filtering_call) s->last_res_lowest = 1e300;
  CHFSI_TRY(filter_all(s, op, X, M, N, (int)m, use_mixed_prec_overall, st));
  CHFSI_HIP(hipEventRecord(ev[2], st));
  *B_out = B;
  return CHFSI_OK;
}

struct PassEvents {
  hipEvent_t ev[6] = {};
  ~PassEvents() {
    for (hipEvent_t e : ev)
      if (e) hipEventDestroy(e);
  }
  chfsi_status create() {
    for (auto &e : ev) CHFSI_HIP(hipEventCreate(&e));
    return CHFSI_OK;
  }
  // per-phase times {lanczos, filter, cholgs, rayleigh_ritz, residual, total}
  chfsi_status report(double *timings_ms) {
    CHFSI_HIP(hipEventSynchronize(ev[5]));
    if (!timings_ms) return CHFSI_OK;
    float t;
    for (int k = 0; k < 4; ++k) {
      CHFSI_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      timings_ms[k] = t;
    }
    CHFSI_HIP(hipEventElapsedTime(&t, ev[4], ev[5]));
    timings_ms[4] = t;
    CHFSI_HIP(hipEventElapsedTime(&t, ev[0], ev[5]));
    timings_ms[5] = t;
    return CHFSI_OK;
  }
};
}  // namespace

extern "C" chfsi_status chfsi_solve(chfsi_solver *s, chfsi_operator *op, double *X, int64_t M,
                                    int32_t N, double *eig_host, double *res_host,
                                    int32_t is_first_filtering_call, int32_t compute_residual,
                                    int32_t use_mixed_prec_overall, int32_t is_first_scf,
                                    double *bUp_out, double *timings_ms, void *stream) {
  CHFSI_ARG(s && op && X && eig_host, "null argument");
  CHFSI_ARG(M == op->M && N > 0, "X must be n_dofs x N");
  CHFSI_ARG(!compute_residual || res_host, "res_host required when computing residuals");
  hipStream_t st = as_stream(stream);
  const chfsi_params &P = s->params;
  PassEvents pe;
  CHFSI_TRY(pe.create());
  CHFSI_HIP(hipEventRecord(pe.ev[0], st));
  int B = 0;
  CHFSI_TRY(lanczos_scale_filter(s, op, X, M, N, is_first_filtering_call, use_mixed_prec_overall,
                                 is_first_scf, &B, pe.ev, st));

  const bool mx = use_mixed_prec_overall != 0;
  if (P.use_shep) {
    // CholGS + Rayleigh-Ritz (SHEP route, …SolverDevice.cc:609-643)
    CHFSI_TRY(cholgs_impl(s, X, M, N, B, mx && P.mixed_prec_overlap,
                          mx && P.mixed_prec_cgs_rotation, st));
    CHFSI_HIP(hipEventRecord(pe.ev[3], st));
    CHFSI_TRY(rr_impl(s, op, X, M, N, B, eig_host, mx && P.mixed_prec_xthx,
                      mx && P.mixed_prec_rr_rotation, st));
  } else {
    // generalised Rayleigh-Ritz without CholGS (…SolverDevice.cc:645-664)
    CHFSI_HIP(hipEventRecord(pe.ev[3], st));
    CHFSI_TRY(gep_impl(s, op, X, M, N, B, eig_host, mx && P.mixed_prec_overlap,
                       mx && P.mixed_prec_xthx, mx && P.mixed_prec_rr_rotation, st));
  }
  CHFSI_HIP(hipEventRecord(pe.ev[4], st));

  // back to the FE basis: folded into the residuals' read of X when they are
  // computed
  if (compute_residual) {
    CHFSI_TRY(residual_impl(s, op, X, N, M, N, B, eig_host, res_host, st, op->d_inv_sqrt_mass));
    const int nlow = std::max(1, (9 * N) / 10);
    s->last_res_lowest = *std::max_element(res_host, res_host + nlow);
  } else {
    CHFSI_HIP(launch_scale_rows<double>(X, N, M, N, 1.0, op->d_inv_sqrt_mass, st));
  }
  CHFSI_HIP(hipEventRecord(pe.ev[5], st));
  CHFSI_TRY(pe.report(timings_ms));
  if (bUp_out) *bUp_out = s->bUp;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_solve_no_rr(chfsi_solver *s, chfsi_operator *op, double *X,
                                          int64_t M, int32_t N, int32_t n_passes,
                                          int32_t scale_degree, int32_t use_mixed_prec_overall,
                                          void *stream) {
  CHFSI_ARG(s && op && X, "null argument");
  CHFSI_ARG(M == op->M && N > 0, "X must be n_dofs x N");
  CHFSI_ARG(n_passes >= 1, "n_passes must be at least 1");
  hipStream_t st = as_stream(stream);
  const chfsi_params &P = s->params;
  // upper bound only (…SolverDevice.cc:812-824); a0 and bLow as last set
  if (!P.reuse_lanczos_upper) {
    double lo, up;
    CHFSI_TRY(run_lanczos(s, op, M, &lo, &up, st));
    s->bUp = up;
  }
  // degree (…SolverDevice.cc:826-839): the scaling factor applies when scale_degree
  unsigned m = P.cheby_degree > 0 ? (unsigned)P.cheby_degree : cheby_order_lookup(s->bUp);
  if (scale_degree) m = (unsigned)(m * P.first_scf_degree_factor);
  const int B = std::min<int>(P.cheby_block_size > 0 ? P.cheby_block_size : N, N);
  const bool mx = use_mixed_prec_overall != 0;
  CHFSI_HIP(launch_scale_rows<double>(X, N, M, N, 1.0, op->d_sqrt_mass, st));
  for (int pass = 0; pass < n_passes; ++pass) {
    CHFSI_TRY(filter_all(s, op, X, M, N, (int)m, use_mixed_prec_overall, st));
    CHFSI_TRY(cholgs_impl(s, X, M, N, B, mx && P.mixed_prec_overlap,
                          mx && P.mixed_prec_cgs_rotation, st));
  }
  CHFSI_HIP(launch_scale_rows<double>(X, N, M, N, 1.0, op->d_inv_sqrt_mass, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_solve_spectrum_split(
    chfsi_solver *s, chfsi_operator *op, double *X, int64_t M, int32_t N, int32_t n_core,
    double *Xfrac, int64_t ld_frac, double *eig_host, double *res_host,
    int32_t is_first_filtering_call, int32_t compute_residual, int32_t use_mixed_prec_overall,
    int32_t is_first_scf, double *bUp_out, double *timings_ms, void *stream) {
  CHFSI_ARG(s && op && X && Xfrac && eig_host, "null argument");
  CHFSI_ARG(M == op->M && N > 0, "X must be n_dofs x N");
  CHFSI_ARG(n_core >= 0 && n_core < N && ld_frac >= N - n_core,
            "need 0 <= n_core < N, ld_frac >= N - n_core");
  CHFSI_ARG(!compute_residual || res_host, "res_host required when computing residuals");
  hipStream_t st = as_stream(stream);
  const chfsi_params &P = s->params;
  const int Nfr = N - n_core;
  PassEvents pe;
  CHFSI_TRY(pe.create());
  CHFSI_HIP(hipEventRecord(pe.ev[0], st));
  int B = 0;
  CHFSI_TRY(lanczos_scale_filter(s, op, X, M, N, is_first_filtering_call, use_mixed_prec_overall,
                                 is_first_scf, &B, pe.ev, st));

  // Rayleigh-Ritz in the generalised form with the spectrum split
  // (…SolverDevice.cc:584-607): Cholesky orthonormalisation, projected
  // Hamiltonian, top Nfr pairs into Xfrac
  const bool mx = use_mixed_prec_overall != 0;
  CHFSI_TRY(cholgs_impl(s, X, M, N, B, mx && P.mixed_prec_overlap, mx && P.mixed_prec_cgs_rotation,
                        st));
  CHFSI_HIP(hipEventRecord(pe.ev[3], st));
  CHFSI_TRY(rr_split_impl(s, op, X, M, N, n_core, B, Xfrac, ld_frac, eig_host,
                          mx && P.mixed_prec_xthx_spectrum_split, st));
  CHFSI_HIP(hipEventRecord(pe.ev[4], st));

  // residuals of the Nfr Ritz pairs (…SolverDevice.cc:676-690)
  if (compute_residual)
    CHFSI_TRY(residual_impl(s, op, Xfrac, ld_frac, M, Nfr, std::min(B, Nfr), eig_host, res_host,
                            st));

  // both back to the FE basis (…SolverDevice.cc:718-732)
  CHFSI_HIP(launch_scale_rows<double>(X, N, M, N, 1.0, op->d_inv_sqrt_mass, st));
  CHFSI_HIP(launch_scale_rows<double>(Xfrac, ld_frac, M, Nfr, 1.0, op->d_inv_sqrt_mass, st));
  CHFSI_HIP(hipEventRecord(pe.ev[5], st));
  CHFSI_TRY(pe.report(timings_ms));
  if (bUp_out) *bUp_out = s->bUp;
  return CHFSI_OK;
}
