// Device-resident operator state shared by the operator and solver sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <vector>

#include "common.hpp"

namespace chfsi {

// Grow-only device buffer.
struct Scratch {
  void *ptr = nullptr;
  size_t size = 0;
  chfsi_status ensure(size_t bytes);
  void release();
};

// One fused apply  dst <- base + alpha * S H S src  over all colour classes
// (see HxLaunch for base / beta_mode).  bf16 selects the BF16-input kernel
// (T = float only).  R (nullable, ld ldR) adds rcoef[v] * R[:, v] to every row
// once (the defect form of the filter; FP32, not with bf16).
// Layout of a caller-provided HX scratch: projection partials (n_slots x B),
// then the V-scaled projections (n_proj x B) at the next 256-byte boundary.
inline size_t noalloc_w_offset(size_t elem, int n_slots, int B) {
  return (elem * size_t(n_slots) * size_t(B) + 255) / 256 * 256;
}
inline size_t noalloc_scratch_bytes(size_t elem, int n_slots, int n_proj, int B) {
  return noalloc_w_offset(elem, n_slots, B) + elem * size_t(n_proj) * size_t(B);
}

template <typename T>
chfsi_status hx_fused(chfsi_operator *op, const T *src, int64_t ld_src, T *dst, int64_t ld_dst,
                      int B, T alpha, T beta, T gamma, int beta_mode, bool scaled, hipStream_t s,
                      bool bf16 = false, const T *R = nullptr, int64_t ldR = 0,
                      const T *rcoef = nullptr, void *ext_scratch = nullptr);

// g (+)= wscale JxW sum_v w_v (N.(s a_v))(N.(s b_v)) per (cell, quadrature
// point), s = M^{-1/2} when scaled; with general constraints the rows of a and
// b are distributed for the interpolation and cleared after (a and b are
// written).  See launch_quad_product.
template <typename T>
chfsi_status quad_product(chfsi_operator *op, const T *a, int64_t lda, const T *b, int64_t ldb,
                          int B, bool scaled, const double *jxw, const double *w, double wscale,
                          double *g, int accumulate, hipStream_t st, int wshift);

}  // namespace chfsi

struct chfsi_operator {
  int device = 0;
  int64_t M = 0;
  int C = 0, ne = 0, ne_pad = 0, nq = 0;
  // cell -> DOF map with first-touch flag (bit 31), cells sorted by colour
  uint32_t *d_map = nullptr;
  int *d_cells_by_color = nullptr;
  static constexpr int kMaxColors = 128;
  int ncolors = 0;
  int color_off[kMaxColors + 1] = {0};
  // Loewdin scalings
  double *d_sqrt_mass = nullptr, *d_inv_sqrt_mass = nullptr;
  float *d_sqrt_mass32 = nullptr, *d_inv_sqrt_mass32 = nullptr;
  // shape data and cell Hamiltonians (padded K-tiled layout)
  double *d_K = nullptr, *d_Nq = nullptr, *d_Gq = nullptr, *d_jxw = nullptr;
  // per-cell geometry: stiffness per cell (d_K then holds C*n_e*n_e) and the
  // cells' inverse Jacobians (jac = 3 diagonal, 9 full; 0: none)
  bool K_per_cell = false;
  double *d_invJ = nullptr;
  int jac = 0;
  double *d_Ht64 = nullptr;
  float *d_Ht32 = nullptr;
  uint16_t *d_Ht16 = nullptr, *d_nlC16 = nullptr;  // BF16 copies
  bool has_H = false;
  bool h_sym = false;  // cell matrices exactly symmetric (built on the device)
  // nonlocal projector slots
  int n_proj = 0, n_slots = 0, n_nl_cells = 0;
  int *d_nl_off = nullptr, *d_slot_proj = nullptr, *d_proj_ptr = nullptr,
      *d_proj_slots = nullptr, *d_nl_cells = nullptr;
  double *d_nlC64 = nullptr, *d_nlCt64 = nullptr, *d_proj_V = nullptr;
  float *d_nlC32 = nullptr, *d_nlCt32 = nullptr;
  // owned-row form of the projection (projection variant 4): per nonlocal
  // cell the DOFs whose (DOF, projector) terms that cell forms (own_n[c] of
  // them, own_map[c * ne_pad + r]) and the assembled projector values on
  // those rows, in the slot layout of nlCt
  uint32_t *d_own_map = nullptr;
  int *d_own_n = nullptr;
  double *d_nlCto64 = nullptr;
  float *d_nlCto32 = nullptr;
  chfsi::Scratch partial64, partial32, W64, W32, tmp;
  // HX tile and projection variants per precision (chosen by chfsi_operator_autotune)
  int hx_variant64 = 0, hx_variant32 = 0;
  int hx_variant16 = 0;  // BF16 apply: 0 K = 32 ring, 1 register-staged K = 16
  int proj_variant64 = 0, proj_variant32 = 0;
  int max_srb = 1;  // largest slot count of any cell / 16
  // extra cell ids C .. C + n_extra_cells - 1: further projector groups of
  // cells with more than 128 slots (zero cell matrix, own colours, the nodes
  // of their real cell)
  int n_extra_cells = 0;
  // constraints (slave-major and master-major CSR)
  int64_t nc = 0, nm = 0;
  int64_t *d_c_rows = nullptr, *d_c_ptr = nullptr, *d_c_cols = nullptr;
  double *d_c_vals = nullptr;
  int64_t *d_m_rows = nullptr, *d_m_ptr = nullptr, *d_m_slaves = nullptr;
  double *d_m_vals = nullptr;
  // general (hanging-node) rows: not folded into the map; every apply
  // distributes into the constrained rows of its source (gathered with scale
  // 1: s_gather = M^{-1/2} with 1 on constrained rows) and sends the rows' sums
  // to their masters (output scale 0 marks them; ones_cons = 1 with 0 on
  // constrained rows for the unscaled apply)
  bool general_constraints = false;
  double *d_s_gather64 = nullptr, *d_ones_cons64 = nullptr;
  float *d_s_gather32 = nullptr, *d_ones_cons32 = nullptr;
  // Bloch k-points: complex cell matrices and projector tiles per k-point in
  // the paired (Re, Im) tile layout of the complex kernels
  int nk = 0, active_k = 0;
  std::vector<double> kpts;  // 3 per k-point
  double *d_Mc = nullptr, *d_Gd = nullptr;
  double *d_Hk64 = nullptr;
  float *d_Hk32 = nullptr;
  double *d_nlCk64 = nullptr, *d_nlCtk64 = nullptr;
  float *d_nlCk32 = nullptr, *d_nlCtk32 = nullptr;
  // owned-row form of the complex projector tiles (the rows of d_own_map),
  // per k-point; zproj_variant 0: cell form, 1: owned rows (the default)
  double *d_nlCtko64 = nullptr;
  float *d_nlCtko32 = nullptr;
  int zproj_variant = 1;
  int zhx_variant64 = 0, zhx_variant32 = 0;
  chfsi::Scratch ztab_dev, ztab_W, ztab_P;  // tables and scratch of single complex applies

  ~chfsi_operator();
};

namespace chfsi {
// Rebuilds the complex cell matrices of every k-point from the real tiles.
chfsi_status rebuild_kpoint_matrices(chfsi_operator *op, hipStream_t s);
}  // namespace chfsi
