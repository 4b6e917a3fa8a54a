// Device kernels of libchfsi and their host-side launchers (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chfsi {

// Rows of the cell GEMM are padded to a multiple of 16 (one MFMA block) and
// the K dimension is streamed in tiles of 16.
constexpr int kTile = 16;

// Arguments of one fused HX launch over one colour class of cells.
//
//   dst[g,v] <- base(g,v) + alpha * s_out[g] * sum_k Htile(c)[i,k] * xin(k,v)
//
// where for the cell-local node i mapping to DOF g:
//   base = (first touch of g in colour order)
//            ? beta' * dst[g,v] + gamma * src[g,v]    (beta' = beta or beta*s_out[g])
//            : dst[g,v]
// except on a constrained row (s_out[g] == 0, a hanging node), which collects
// the plain sums alpha * sum_k Htile[i,k] xin(k,v) (base 0 at first touch)
// for the slave -> master pass after the colour launches
// and xin(k,v) = s_in[g_k] * src[g_k, v] for the n_e nodes, followed by the
// cell's nonlocal slots whose rows come from W[proj(slot), v] (already scaled
// by the coupling constants).
template <typename T>
struct HxLaunch {
  const int *cells;        // cell ids of this colour
  int ncells;
  int ntiles;              // column tiles of TN
  int per_xcd;             // work items per XCD slice
  const uint32_t *map;     // C*n_e: DOF | (first-touch flag << 31)
  int ne, ne_pad;
  const T *Ht;             // C * ne_pad * ne_pad, tiled [kt][row][16]
  const T *src;
  int64_t ld_src;
  T *dst;
  int64_t ld_dst;
  const T *s_in;           // nullable (== 1)
  const T *s_out;          // nullable (== 1)
  T alpha, beta, gamma;
  int beta_mode;           // 0: beta*dst, 1: dst not read (beta = 0), 2: beta*s_out*dst
  int B;
  const int *nl_off;       // C+1 slot offsets (multiples of 16); nullable
  const T *nl_C;           // per cell [kt][row][16] blocks, slot-major offsets nl_off[c]*ne_pad
  const int *slot_proj;    // slot -> projector (-1 pad)
  const T *W;              // n_proj x ldW
  int64_t ldW;
  int variant;             // tile variant (see hx_variants)
  // defect source of the filter's defect form (nullable): first-touch rows
  // add rcoef[v] * R[g, v] after the Chebyshev combine
  const T *R;
  int64_t ldR;
  const T *rcoef;
  // 1: the cell matrices are exactly symmetric (built on the device); the
  // one-column kernel then reads the slices on and below the diagonal only
  int h_sym;
};

// Arguments of the projection C_e^T (s_in * src)_e for all cells with slots.
template <typename T>
struct ProjLaunch {
  const int *cells;        // cells that own slots
  int ncells;
  int ntiles;
  int per_xcd;             // work items per XCD slice
  const uint32_t *map;
  int ne, ne_pad;
  const int *nl_off;
  const T *nl_Ct;          // per cell [kt][slot][16] over nodes, offset nl_off[c]*ne_pad
  const T *src;
  int64_t ld_src;
  const T *s_in;
  T *partial;              // n_slots x ldP
  int64_t ldP;
  int B;
  int max_srb;             // largest slot count of any cell / 16 (1..8)
  int variant;             // projection variant (see proj_variants)
  // owned-row form (variant 4): rows own_map[c * ne_pad + r], r < own_n[c],
  // replace the cell's nodes, and nl_Ct holds the assembled values on them
  const uint32_t *own_map;
  const int *own_n;
};

template <typename T>
hipError_t launch_hx(const HxLaunch<T> &a, hipStream_t s);
int hx_variants(bool fp64);
int hx_bf16_variants();
// BF16 cell matrices / inputs, FP32 accumulation and multivectors
hipError_t launch_hx_bf16(const HxLaunch<float> &a, const uint16_t *Ht16, const uint16_t *nlC16,
                          hipStream_t s);
hipError_t launch_f64_to_bf16(const double *x, int64_t n, uint16_t *y, hipStream_t s);
template <typename T>
hipError_t launch_proj(const ProjLaunch<T> &a, hipStream_t s);
int proj_variants();
template <typename T>
hipError_t launch_proj_reduce(const int *proj_ptr, const int *proj_slots, const double *proj_V,
                              int n_proj, const T *partial, int64_t ldP, T *W, int64_t ldW,
                              int B, hipStream_t s);

// H_e = 1/2 K_e + N^T diag(vJxW_c) N, written tiled/padded in FP64, FP32 and BF16
// (Ht32, Ht16 nullable); K is one n_e x n_e stiffness for all cells, or with
// K_per_cell one per cell (C x n_e x n_e, the cells' gradient integrals).
// With gJxW (C*n_q*3, nullable) and the shape gradients G (n_q*n_e*3) the GGA
// term 2 sum_q gJxW(q).(grad N_i(q) N_j(q) + N_i(q) grad N_j(q)) is added; G
// holds the physical gradients (jac = 0), or the reference cell's with
// per-cell inverse Jacobians invJ: C x 3 diagonals (jac = 3, Cartesian cells)
// or C x 9 row-major matrices (jac = 9), grad N = J grad_ref N.
hipError_t launch_build_h(const double *K, const double *Nq, const double *vJxW, const double *G,
                          const double *gJxW, int C, int ne, int ne_pad, int nq, double *Ht64,
                          float *Ht32, uint16_t *Ht16, hipStream_t s, bool K_per_cell = false,
                          const double *invJ = nullptr, int jac = 0);
// tiled padded -> dense C x ne x ne (row-major) and back
hipError_t launch_untile_h(const double *Ht, int C, int ne, int ne_pad, double *H, hipStream_t s);
hipError_t launch_tile_h(const double *H, int C, int ne, int ne_pad, double *Ht64, float *Ht32,
                         hipStream_t s);

// x[g, v] *= a * s[g]  (stridedBlockScale, DeviceKernelsGeneric.cc:255-278)
template <typename T>
hipError_t launch_scale_rows(T *x, int64_t ld, int64_t M, int B, double a, const double *s,
                             hipStream_t st);
// dst[g, v] = (TD) src[g, v]  with independent leading dimensions and types
template <typename TS, typename TD>
hipError_t launch_copy_block(const TS *src, int64_t ld_src, TD *dst, int64_t ld_dst, int64_t M,
                             int B, hipStream_t st);
// y = a*x + b*y on contiguous arrays
template <typename T>
hipError_t launch_axpby(int64_t n, T a, const T *x, T b, T *y, hipStream_t st);
// column sums of squares of (Y - X*diag(lam)) into partial[chunk][B]; with
// X_rescaled (== X) and rescale, X[g, v] *= rescale[g] once it has been read
hipError_t launch_residual_partial(const double *HX, int64_t ldH, const double *X, int64_t ldX,
                                   const double *lam, int64_t M, int B, double *partial,
                                   int nchunks, hipStream_t st, double *X_rescaled = nullptr,
                                   const double *rescale = nullptr);
// column sums of X*HX and X*X into partial[chunk][2][B] (Rayleigh quotients)
hipError_t launch_colstats_partial(const double *HX, int64_t ldH, const double *X, int64_t ldX,
                                   int64_t M, int B, double *partial, int nchunks, hipStream_t st);
// defect source of the filter: r = HX - X*diag(theta) in FP64, R = (float) r,
// E = (float)(s * r)
hipError_t launch_defect_source(const double *HX, int64_t ldH, const double *X, int64_t ldX,
                                const double *theta, double s, int64_t M, int B, float *R,
                                float *E, hipStream_t st);
// X[g, v] = X[g, v] * D[v] + (double) E[g, v]
hipError_t launch_defect_finish(double *X, int64_t ldX, const double *D, const float *E,
                                int64_t ldE, int64_t M, int B, hipStream_t st);
// dot products x.y into partial[chunk]; then the host sums
hipError_t launch_dot_partial(const double *x, const double *y, int64_t n, double *partial,
                              int nchunks, hipStream_t st);
// *slot = sum of partial[0..n) in index order, on the device
hipError_t launch_sum_slot(const double *partial, int n, double *slot, hipStream_t st);
// y = a*x + b*y with a = -c (form 0), 1/sqrt(c) (1) or -sqrt(c) (2), c = *slot
// read on the device
hipError_t launch_axpby_slot(int64_t n, const double *slot, int form, const double *x, double b,
                             double *y, hipStream_t st);

// g[c*nqc + q] (+)= wscale JxW[c*nqc + q] sum_v w_{v >> wshift} (N(q).(s a_v)|_c)(N(q).(s b_v)|_c)
// s, w and JxW nullable (== 1); wshift = 1 shares each weight between the two
// real columns of an interleaved complex column
hipError_t launch_quad_product(const uint32_t *map, int C, int ne, int nqc, const double *Nq,
                               const double *JxW, const double *a, int64_t lda, const double *b,
                               int64_t ldb, int B, const double *s_in, const double *w,
                               double wscale, double *g, int accumulate, hipStream_t s,
                               int wshift = 0);
// the same with FP32 inputs (products and sums in FP64)
hipError_t launch_quad_product(const uint32_t *map, int C, int ne, int nqc, const double *Nq,
                               const double *JxW, const float *a, int64_t lda, const float *b,
                               int64_t ldb, int B, const double *s_in, const double *w,
                               double wscale, double *g, int accumulate, hipStream_t s,
                               int wshift = 0);

// C[i + j*ldc] = sum_k P[k*m*n + i + j*m]  (split-K reduction, column-major m x n)
hipError_t launch_sum_partials(const double *P, int nk, int m, int n, double *C, int64_t ldc,
                               hipStream_t st);
// the same with FP32 partials (mixed-precision off-diagonal blocks)
hipError_t launch_sum_partials_f32(const float *P, int nk, int m, int n, double *C, int64_t ldc,
                                   hipStream_t st);
// C[i + j*ldc] += A[i + j*lda], column-major m x n, FP32 into FP64
hipError_t launch_add_f32_to_f64(const float *A, int64_t lda, double *C, int64_t ldc, int m,
                                 int64_t n, hipStream_t st);
// B (n x n, ld n) = (float) A, diagonal bs x bs blocks zeroed when zero_diag
hipError_t launch_cast_blocks(const double *A, int n, int64_t lda, int bs, int zero_diag,
                              float *B, hipStream_t st);
// A[i + j*lda] = 0 for i < j (clear the strict upper triangle of an n x n matrix)
hipError_t launch_zero_upper(double *A, int n, int64_t lda, hipStream_t st);

// constraints: CSR over slave rows
hipError_t launch_distribute(const int64_t *rows, const int64_t *ptr, const int64_t *cols,
                             const double *vals, int64_t nrows, double *X, int64_t ld, int B,
                             hipStream_t st);
hipError_t launch_slave_to_master(const int64_t *m_rows, const int64_t *m_ptr,
                                  const int64_t *m_slaves, const double *m_vals, int64_t nm,
                                  double *X, int64_t ld, int B, hipStream_t st);
hipError_t launch_zero_rows(const int64_t *rows, int64_t nrows, double *X, int64_t ld, int B,
                            hipStream_t st);
// x[rows[r], v] = sum_j vals[j] g[cols[j]] x[cols[j], v]  (g nullable: 1)
template <typename T>
hipError_t launch_distribute_scaled(const int64_t *rows, const int64_t *ptr, const int64_t *cols,
                                    const double *vals, int64_t nrows, const T *g, T *X,
                                    int64_t ld, int B, hipStream_t st);
// x[m, v] += so[m] sum_j vals[j] x[slave_j, v]  (master-major CSR; so nullable: 1)
template <typename T>
hipError_t launch_slave_to_master_scaled(const int64_t *m_rows, const int64_t *m_ptr,
                                         const int64_t *m_slaves, const double *m_vals,
                                         int64_t nm, const T *so, T *X, int64_t ld, int B,
                                         hipStream_t st);
template <typename T>
hipError_t launch_zero_rows_t(const int64_t *rows, int64_t nrows, T *X, int64_t ld, int B,
                              hipStream_t st);

}  // namespace chfsi
