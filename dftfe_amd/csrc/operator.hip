// Operator side of the C ABI: device-resident FE operator data (cell maps,
// colour classes, mass vectors, cell Hamiltonians, nonlocal projector slots,
// constraints) and the HX / HXCheby / constraint entry points.
//
// Reference: kohnShamDFTOperatorDeviceClass (include/kohnShamDFTOperatorDevice.h:43-782,
// src/dftOperator/kohnShamDFTOperatorDevice.cc:492-1057, :3609-3997) and
// constraintMatrixInfoDevice (utils/constraintMatrixInfoDevice.cc).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "bloch.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "operator.hpp"

namespace chfsi {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

template <typename T>
static chfsi_status dev_upload(T **d, const T *h, size_t n) {
  *d = nullptr;
  if (n == 0) return CHFSI_OK;
  if (hipMalloc(reinterpret_cast<void **>(d), n * sizeof(T)) != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed", n * sizeof(T));
    return CHFSI_ERR_NOMEM;
  }
  if (h) CHFSI_HIP(hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return CHFSI_OK;
}

template <typename T>
static chfsi_status dev_alloc(T **d, size_t n) {
  return dev_upload<T>(d, nullptr, n);
}

chfsi_status Scratch::ensure(size_t bytes) {
  if (bytes <= size) return CHFSI_OK;
  if (ptr) hipFree(ptr);
  ptr = nullptr;
  size = 0;
  if (hipMalloc(&ptr, bytes) != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed", bytes);
    return CHFSI_ERR_NOMEM;
  }
  size = bytes;
  return CHFSI_OK;
}

void Scratch::release() {
  if (ptr) hipFree(ptr);
  ptr = nullptr;
  size = 0;
}

}  // namespace chfsi

using namespace chfsi;

chfsi_operator::~chfsi_operator() {
  void *ptrs[] = {d_map, d_cells_by_color, d_sqrt_mass, d_inv_sqrt_mass, d_sqrt_mass32,
                  d_inv_sqrt_mass32, d_K, d_Nq, d_Gq, d_jxw, d_Ht64, d_Ht32, d_Ht16, d_nlC16, d_nl_off,
                  d_nlC64, d_nlC32,
                  d_nlCt64, d_nlCt32, d_slot_proj, d_proj_ptr, d_proj_slots, d_proj_V,
                  d_nl_cells, d_c_rows, d_c_ptr, d_c_cols, d_c_vals, d_m_rows, d_m_ptr,
                  d_m_slaves, d_m_vals, d_Mc, d_Gd, d_Hk64, d_Hk32, d_nlCk64, d_nlCtk64,
                  d_nlCk32, d_nlCtk32, d_invJ, d_s_gather64, d_ones_cons64, d_s_gather32,
                  d_ones_cons32, d_own_map, d_own_n, d_nlCto64, d_nlCto32, d_nlCtko64,
                  d_nlCtko32};
  for (void *p : ptrs)
    if (p) hipFree(p);
  ztab_dev.release();
  ztab_W.release();
  ztab_P.release();
  partial64.release();
  partial32.release();
  W64.release();
  W32.release();
  tmp.release();
}

extern "C" const char *chfsi_last_error(void) { return g_last_error.c_str(); }

extern "C" chfsi_status chfsi_operator_create(const chfsi_operator_desc *d, chfsi_operator **out) {
  CHFSI_ARG(d && out, "null argument");
  *out = nullptr;
  CHFSI_ARG(d->n_dofs > 0 && d->n_dofs < (int64_t(1) << 31), "n_dofs out of range");
  CHFSI_ARG(d->n_cells > 0 && d->nodes_per_cell > 0 && d->quad_per_cell > 0, "bad sizes");
  CHFSI_ARG(d->cell_map && d->cell_color && d->sqrt_mass && d->inv_sqrt_mass, "missing arrays");
  const int C = d->n_cells, ne = d->nodes_per_cell;
  const int ne_pad = (ne + 15) / 16 * 16;
  {
    const int nrb = ne_pad / 16;
    CHFSI_ARG(nrb == 1 || nrb == 2 || nrb == 4 || nrb == 8 || nrb == 14 || nrb == 22,
              "nodes_per_cell %d is not (p+1)^3 for an FE degree p in 1..6", ne);
  }
  const int64_t M = d->n_dofs;

  auto op = new chfsi_operator();
  std::unique_ptr<chfsi_operator> guard(op);
  CHFSI_HIP(hipGetDevice(&op->device));
  op->M = M;
  op->C = C;
  op->ne = ne;
  op->ne_pad = ne_pad;
  op->nq = d->quad_per_cell;

  // ---- constraints: periodic identities (single master, coefficient 1) are
  // folded into the cell map; general rows are kept for the explicit kernels.
  std::vector<int64_t> fold(M);
  for (int64_t g = 0; g < M; ++g) fold[g] = g;
  op->nc = d->n_constrained;
  bool all_identity = true;
  if (op->nc > 0) {
    CHFSI_ARG(d->constraint_rows && d->constraint_ptr && d->constraint_cols && d->constraint_vals,
              "missing constraint arrays");
    for (int64_t r = 0; r < op->nc; ++r) {
      const int64_t s = d->constraint_rows[r];
      CHFSI_ARG(s >= 0 && s < M, "constraint row out of range");
      const int64_t b = d->constraint_ptr[r], e = d->constraint_ptr[r + 1];
      for (int64_t j = b; j < e; ++j)
        CHFSI_ARG(d->constraint_cols[j] >= 0 && d->constraint_cols[j] < M, "master out of range");
      if (e - b == 1 && d->constraint_vals[b] == 1.0)
        fold[s] = d->constraint_cols[b];
      else
        all_identity = false;
    }
    if (!all_identity) {
      // general rows: nothing is folded; masters must be unconstrained
      op->general_constraints = true;
      for (int64_t g = 0; g < M; ++g) fold[g] = g;
      std::vector<char> is_c(M, 0);
      for (int64_t r = 0; r < op->nc; ++r) is_c[d->constraint_rows[r]] = 1;
      for (int64_t j = 0; j < d->constraint_ptr[op->nc]; ++j)
        CHFSI_ARG(!is_c[d->constraint_cols[j]], "master %lld of a constraint is itself constrained",
                  (long long)d->constraint_cols[j]);
      std::vector<double> sg(d->inv_sqrt_mass, d->inv_sqrt_mass + M), oc(M, 1.0);
      for (int64_t r = 0; r < op->nc; ++r) {
        sg[d->constraint_rows[r]] = 1.0;
        oc[d->constraint_rows[r]] = 0.0;
      }
      std::vector<float> sg32(sg.begin(), sg.end()), oc32(oc.begin(), oc.end());
      CHFSI_TRY(dev_upload(&op->d_s_gather64, sg.data(), M));
      CHFSI_TRY(dev_upload(&op->d_ones_cons64, oc.data(), M));
      CHFSI_TRY(dev_upload(&op->d_s_gather32, sg32.data(), M));
      CHFSI_TRY(dev_upload(&op->d_ones_cons32, oc32.data(), M));
    }
    const int64_t nnz = d->constraint_ptr[op->nc];
    std::vector<int64_t> ptr(d->constraint_ptr, d->constraint_ptr + op->nc + 1);
    CHFSI_TRY(dev_upload(&op->d_c_rows, d->constraint_rows, op->nc));
    CHFSI_TRY(dev_upload(&op->d_c_ptr, ptr.data(), op->nc + 1));
    CHFSI_TRY(dev_upload(&op->d_c_cols, d->constraint_cols, nnz));
    CHFSI_TRY(dev_upload(&op->d_c_vals, d->constraint_vals, nnz));
    // transpose to master-major CSR for the atomic-free slave->master kernel
    std::map<int64_t, std::vector<std::pair<int64_t, double>>> bym;
    for (int64_t r = 0; r < op->nc; ++r)
      for (int64_t j = d->constraint_ptr[r]; j < d->constraint_ptr[r + 1]; ++j)
        bym[d->constraint_cols[j]].push_back({d->constraint_rows[r], d->constraint_vals[j]});
    std::vector<int64_t> mrows, mptr{0}, mslaves;
    std::vector<double> mvals;
    for (auto &kv : bym) {
      mrows.push_back(kv.first);
      for (auto &sv : kv.second) {
        mslaves.push_back(sv.first);
        mvals.push_back(sv.second);
      }
      mptr.push_back((int64_t)mslaves.size());
    }
    op->nm = (int64_t)mrows.size();
    CHFSI_TRY(dev_upload(&op->d_m_rows, mrows.data(), mrows.size()));
    CHFSI_TRY(dev_upload(&op->d_m_ptr, mptr.data(), mptr.size()));
    CHFSI_TRY(dev_upload(&op->d_m_slaves, mslaves.data(), mslaves.size()));
    CHFSI_TRY(dev_upload(&op->d_m_vals, mvals.data(), mvals.size()));
  }

  // ---- projector pairs per cell.  The fused kernels carry at most kCellSlots
  // projector slots per cell; a cell with more keeps its first pairs and hands
  // each further group to an extra cell id (C, C + 1, ...) that shares its
  // nodes and has a zero cell matrix, in a colour after all real colours, so
  // the apply adds the group's C_e W term on a later, non-first touch.
  constexpr int kCellSlots = 128;
  const int P = d->n_pairs > 0 ? d->proj_per_pair : 0;
  std::vector<std::vector<int>> pairs_of(C);
  std::vector<int> src_cell(C);  // the real cell of every cell id
  for (int c = 0; c < C; ++c) src_cell[c] = c;
  if (d->n_pairs > 0) {
    CHFSI_ARG(P > 0 && P <= kCellSlots && d->pair_cell && d->pair_proj0 && d->pair_C && d->proj_V,
              "missing nonlocal arrays (or more than %d projectors per pair)", kCellSlots);
    for (int p = 0; p < d->n_pairs; ++p) {
      CHFSI_ARG(d->pair_cell[p] >= 0 && d->pair_cell[p] < C, "pair_cell out of range");
      CHFSI_ARG(d->pair_proj0[p] >= 0 && d->pair_proj0[p] + P <= d->n_proj,
                "pair projector out of range");
      pairs_of[d->pair_cell[p]].push_back(p);
    }
    const size_t per = kCellSlots / P;  // pairs one cell id carries
    for (int c = 0; c < C; ++c) {
      if (pairs_of[c].size() <= per) continue;
      const std::vector<int> all = pairs_of[c];
      pairs_of[c].resize(per);
      for (size_t b = per; b < all.size(); b += per) {
        pairs_of.emplace_back(all.begin() + b, all.begin() + std::min(all.size(), b + per));
        src_cell.push_back(c);
      }
    }
  }
  const int Ctot = (int)pairs_of.size();  // real and extra cell ids
  op->n_extra_cells = Ctot - C;
  CHFSI_ARG(op->n_extra_cells == 0 || d->n_kpoints == 0,
            "with k-points a cell overlaps at most %d projector slots", kCellSlots);

  // ---- colour classes and first-touch flags
  constexpr int KC = chfsi_operator::kMaxColors;
  std::vector<std::vector<int>> by_color(KC);
  for (int c = 0; c < C; ++c) {
    const int col = d->cell_color[c];
    CHFSI_ARG(col >= 0 && col < KC, "cell colour must be in [0, %d)", KC);
    by_color[col].push_back(c);
    op->ncolors = std::max(op->ncolors, col + 1);
  }
  {
    // group j of a cell of colour k goes to colour j * ncolors_real + k: cells
    // of one real colour share no DOF, so neither do their extra cells
    const int nreal = op->ncolors;
    std::vector<int> group(C, 0);
    for (int c = C; c < Ctot; ++c) {
      const int rc = src_cell[c];
      const int col = ++group[rc] * nreal + d->cell_color[rc];
      CHFSI_ARG(col < KC, "cell %d overlaps too many projectors (%d colours needed, %d supported)",
                rc, col + 1, KC);
      by_color[col].push_back(c);
      op->ncolors = std::max(op->ncolors, col + 1);
    }
  }
  std::vector<int> cells_sorted;
  op->color_off[0] = 0;
  for (int k = 0; k < op->ncolors; ++k) {
    cells_sorted.insert(cells_sorted.end(), by_color[k].begin(), by_color[k].end());
    op->color_off[k + 1] = (int)cells_sorted.size();
  }
  std::vector<uint32_t> map(size_t(Ctot) * ne);
  std::vector<int> touched_color(M, -1);
  std::vector<int> touched_cell(M, -1);
  for (int k = 0; k < op->ncolors; ++k) {
    for (int c : by_color[k]) {
      for (int i = 0; i < ne; ++i) {
        const int64_t g0 = d->cell_map[size_t(src_cell[c]) * ne + i];
        CHFSI_ARG(g0 >= 0 && g0 < M, "cell_map entry out of range");
        const int64_t g = fold[g0];
        uint32_t w = (uint32_t)g;
        if (touched_color[g] == k) {
          set_error("cells %d and %d of colour %d share DOF %lld: colouring invalid",
                    touched_cell[g], c, k, (long long)g);
          return CHFSI_ERR_ARG;
        }
        if (touched_color[g] < 0) w |= 0x80000000u;
        touched_color[g] = k;
        touched_cell[g] = c;
        map[size_t(c) * ne + i] = w;
      }
    }
  }
  CHFSI_TRY(dev_upload(&op->d_map, map.data(), map.size()));
  CHFSI_TRY(dev_upload(&op->d_cells_by_color, cells_sorted.data(), cells_sorted.size()));

  // ---- mass vectors
  std::vector<float> f32(M);
  CHFSI_TRY(dev_upload(&op->d_sqrt_mass, d->sqrt_mass, M));
  CHFSI_TRY(dev_upload(&op->d_inv_sqrt_mass, d->inv_sqrt_mass, M));
  for (int64_t g = 0; g < M; ++g) f32[g] = (float)d->sqrt_mass[g];
  CHFSI_TRY(dev_upload(&op->d_sqrt_mass32, f32.data(), M));
  for (int64_t g = 0; g < M; ++g) f32[g] = (float)d->inv_sqrt_mass[g];
  CHFSI_TRY(dev_upload(&op->d_inv_sqrt_mass32, f32.data(), M));

  // ---- shape data for the cell Hamiltonian build
  if (d->shape_values && (d->stiffness || d->cell_stiffness)) {
    CHFSI_TRY(dev_upload(&op->d_Nq, d->shape_values, size_t(op->nq) * ne));
    if (d->cell_stiffness) {
      op->K_per_cell = true;
      CHFSI_TRY(dev_upload(&op->d_K, d->cell_stiffness, size_t(C) * ne * ne));
    } else {
      CHFSI_TRY(dev_upload(&op->d_K, d->stiffness, size_t(ne) * ne));
    }
    if (d->shape_gradients)
      CHFSI_TRY(dev_upload(&op->d_Gq, d->shape_gradients, size_t(op->nq) * ne * 3));
  }
  if (d->inverse_jacobians) {
    CHFSI_ARG(d->jacobian_kind == 3 || d->jacobian_kind == 9,
              "jacobian_kind must be 3 (Cartesian) or 9 (affine) with inverse_jacobians");
    op->jac = d->jacobian_kind;
    CHFSI_TRY(dev_upload(&op->d_invJ, d->inverse_jacobians, size_t(C) * op->jac));
  } else {
    CHFSI_ARG(d->jacobian_kind == 0, "jacobian_kind set without inverse_jacobians");
  }
  if (d->jxw) CHFSI_TRY(dev_upload(&op->d_jxw, d->jxw, size_t(C) * op->nq));
  if (d->n_kpoints > 0) {
    CHFSI_ARG(d->kpoints && d->cell_mass && d->cell_grad, "k-points need kpoints, cell_mass, cell_grad");
    CHFSI_ARG(d->n_pairs == 0 || d->pair_C_kpoints, "k-points need pair_C_kpoints");
    CHFSI_ARG(d->shape_values && d->stiffness, "k-points need shape data for the cell matrices");
    CHFSI_ARG(!d->cell_stiffness && !d->inverse_jacobians,
              "the k-point cell terms are built for one cell geometry (uniform meshes)");
    CHFSI_ARG(!op->general_constraints, "k-points support periodic constraints only");
    op->nk = d->n_kpoints;
    op->kpts.assign(d->kpoints, d->kpoints + 3 * size_t(op->nk));
    CHFSI_TRY(dev_upload(&op->d_Mc, d->cell_mass, size_t(ne) * ne));
    CHFSI_TRY(dev_upload(&op->d_Gd, d->cell_grad, size_t(3) * ne * ne));
    const size_t per_k = size_t(C) * 2 * ne_pad * ne_pad;
    CHFSI_TRY(dev_alloc(&op->d_Hk64, per_k * op->nk));
    CHFSI_TRY(dev_alloc(&op->d_Hk32, per_k * op->nk));
  }
  CHFSI_TRY(dev_alloc(&op->d_Ht64, size_t(Ctot) * ne_pad * ne_pad));
  CHFSI_TRY(dev_alloc(&op->d_Ht32, size_t(Ctot) * ne_pad * ne_pad));
  CHFSI_TRY(dev_alloc(&op->d_Ht16, size_t(Ctot) * ne_pad * ne_pad));
  if (Ctot > C) {
    // the extra cells' matrices stay zero (the builds write cells 0..C-1)
    const size_t off = size_t(C) * ne_pad * ne_pad, n = size_t(Ctot - C) * ne_pad * ne_pad;
    CHFSI_HIP(hipMemset(op->d_Ht64 + off, 0, n * sizeof(double)));
    CHFSI_HIP(hipMemset(op->d_Ht32 + off, 0, n * sizeof(float)));
    CHFSI_HIP(hipMemset(op->d_Ht16 + off, 0, n * sizeof(uint16_t)));
  }

  // ---- nonlocal slots: per cell the (pair, alpha) columns, padded to 16
  op->n_proj = d->n_proj;
  if (d->n_pairs > 0) {
    std::vector<int> nl_off(Ctot + 1, 0);
    for (int c = 0; c < Ctot; ++c) {
      const int cnt = (int)pairs_of[c].size() * P;
      nl_off[c + 1] = nl_off[c] + (cnt + 15) / 16 * 16;  // at most kCellSlots
    }
    const int S = nl_off[Ctot];
    op->n_slots = S;
    for (int c = 0; c < Ctot; ++c)
      op->max_srb = std::max(op->max_srb, (nl_off[c + 1] - nl_off[c]) / 16);
    std::vector<int> slot_proj(S, -1);
    std::vector<double> nlC(size_t(S) * ne_pad, 0.0), nlCt(size_t(S) * ne_pad, 0.0);
    std::vector<std::vector<int>> slots_of(d->n_proj);
    std::vector<int> nl_cells;
    for (int c = 0; c < Ctot; ++c) {
      const int s0 = nl_off[c], ns = nl_off[c + 1] - s0;
      if (ns == 0) continue;
      nl_cells.push_back(c);
      int local = 0;
      for (int p : pairs_of[c]) {
        for (int al = 0; al < P; ++al, ++local) {
          const int s = s0 + local;
          slot_proj[s] = d->pair_proj0[p] + al;
          slots_of[slot_proj[s]].push_back(s);
          for (int i = 0; i < ne; ++i) {
            const double v = d->pair_C[(size_t(p) * ne + i) * P + al];
            // main-kernel layout: block s0*ne_pad, [kts][row i][16 slots]
            nlC[size_t(s0) * ne_pad + size_t(local / 16) * ne_pad * 16 + size_t(i) * 16 +
                (local % 16)] = v;
            // projection layout: [kt over nodes][slot][16 nodes]
            nlCt[size_t(s0) * ne_pad + size_t(i / 16) * ns * 16 + size_t(local) * 16 + (i % 16)] =
                v;
          }
        }
      }
    }
    std::vector<int> proj_ptr(d->n_proj + 1, 0), proj_slots;
    for (int a = 0; a < d->n_proj; ++a) {
      proj_slots.insert(proj_slots.end(), slots_of[a].begin(), slots_of[a].end());
      proj_ptr[a + 1] = (int)proj_slots.size();
    }
    op->n_nl_cells = (int)nl_cells.size();
    // ---- owned-row form.  C^T x = sum_g c[g] x[g] with c[g] the projector
    // value assembled over every cell node mapped to DOF g; the cell form
    // above reads x[g] once per cell node (343 nodes per cell at p = 6 for
    // 216 DOFs).  Here each DOF g of the nonlocal cells is read by as few of
    // its cells as cover its projectors: the cell whose projectors cover most
    // of the ones left takes row g and, in its existing slots, their assembled
    // values (zero in slots whose projector another cell took).  Slots, their
    // reduction and the apply are unchanged.
    {
      std::vector<int64_t> gptr(M + 1, 0);
      for (int c : nl_cells)
        for (int i = 0; i < ne; ++i) ++gptr[(map[size_t(c) * ne + i] & 0x7fffffffu) + 1];
      for (int64_t g = 0; g < M; ++g) gptr[g + 1] += gptr[g];
      std::vector<int> inc_cell(gptr[M]), inc_node(gptr[M]);
      {
        std::vector<int64_t> fill(gptr.begin(), gptr.end() - 1);
        for (int c : nl_cells)
          for (int i = 0; i < ne; ++i) {
            const int64_t k = fill[map[size_t(c) * ne + i] & 0x7fffffffu]++;
            inc_cell[k] = c;
            inc_node[k] = i;
          }
      }
      std::vector<uint32_t> own_map(size_t(Ctot) * ne_pad, 0u);
      std::vector<int> own_n(Ctot, 0);
      std::vector<double> nlCto(nlCt.size(), 0.0);
      // with k-points the same rows carry each k-point's assembled complex
      // values in the paired (Re, Im) tile layout of the complex projection
      const int nk = op->nk;
      const size_t per_k = size_t(2) * S * ne_pad;
      std::vector<double> zcto(per_k * nk, 0.0);
      std::vector<int> keys, cands;
      std::vector<double> vals, valsk;  // valsk: [key][k-point][alpha][Re, Im]
      std::vector<char> done;
      for (int64_t g = 0; g < M; ++g) {
        if (gptr[g] == gptr[g + 1]) continue;
        keys.clear();
        vals.clear();
        valsk.clear();
        cands.clear();
        for (int64_t k = gptr[g]; k < gptr[g + 1]; ++k) {
          const int c = inc_cell[k], i = inc_node[k];
          if (std::find(cands.begin(), cands.end(), c) == cands.end()) cands.push_back(c);
          for (int p : pairs_of[c]) {
            const int key = d->pair_proj0[p];
            size_t q = std::find(keys.begin(), keys.end(), key) - keys.begin();
            if (q == keys.size()) {
              keys.push_back(key);
              vals.resize(vals.size() + P, 0.0);
              valsk.resize(valsk.size() + size_t(nk) * P * 2, 0.0);
            }
            for (int al = 0; al < P; ++al) vals[q * P + al] += d->pair_C[(size_t(p) * ne + i) * P + al];
            for (int kp = 0; kp < nk; ++kp)
              for (int al = 0; al < P; ++al)
                for (int h = 0; h < 2; ++h)
                  valsk[((q * nk + kp) * P + al) * 2 + h] +=
                      d->pair_C_kpoints[2 * (((size_t(kp) * d->n_pairs + p) * ne + i) * P + al) + h];
          }
        }
        std::sort(cands.begin(), cands.end());
        done.assign(keys.size(), 1);
        int left = 0;
        for (size_t q = 0; q < keys.size(); ++q) {
          bool nz = false;
          for (int al = 0; al < P; ++al) nz = nz || vals[q * P + al] != 0.0;
          for (size_t e = 0; e < size_t(nk) * P * 2; ++e) nz = nz || valsk[q * nk * P * 2 + e] != 0.0;
          if (nz) {
            done[q] = 0;
            ++left;
          }
        }
        while (left > 0) {
          int best = -1, best_cnt = 0;
          for (int c : cands) {
            int cnt = 0;
            for (int p : pairs_of[c]) {
              const size_t q = std::find(keys.begin(), keys.end(), d->pair_proj0[p]) - keys.begin();
              cnt += !done[q];
            }
            if (cnt > best_cnt) {
              best_cnt = cnt;
              best = c;
            }
          }
          const int r = own_n[best]++;
          own_map[size_t(best) * ne_pad + r] = (uint32_t)g;
          const int s0 = nl_off[best], ns = nl_off[best + 1] - s0;
          int local = 0;
          for (int p : pairs_of[best]) {
            const size_t q = std::find(keys.begin(), keys.end(), d->pair_proj0[p]) - keys.begin();
            if (!done[q]) {
              for (int al = 0; al < P; ++al)
                nlCto[size_t(s0) * ne_pad + size_t(r / 16) * ns * 16 + size_t(local + al) * 16 +
                      (r % 16)] = vals[q * P + al];
              for (int kp = 0; kp < nk; ++kp)
                for (int al = 0; al < P; ++al)
                  for (int h = 0; h < 2; ++h)
                    zcto[per_k * kp + size_t(s0) * 2 * ne_pad +
                         size_t(2 * (r / 16) + h) * ns * 16 + size_t(local + al) * 16 + (r % 16)] =
                        valsk[((q * nk + kp) * P + al) * 2 + h];
              done[q] = 1;
              --left;
            }
            local += P;
          }
        }
      }
      CHFSI_TRY(dev_upload(&op->d_own_map, own_map.data(), own_map.size()));
      CHFSI_TRY(dev_upload(&op->d_own_n, own_n.data(), own_n.size()));
      CHFSI_TRY(dev_upload(&op->d_nlCto64, nlCto.data(), nlCto.size()));
      std::vector<float> o32(nlCto.begin(), nlCto.end());
      CHFSI_TRY(dev_upload(&op->d_nlCto32, o32.data(), o32.size()));
      if (nk > 0) {
        CHFSI_TRY(dev_upload(&op->d_nlCtko64, zcto.data(), zcto.size()));
        std::vector<float> z32(zcto.begin(), zcto.end());
        CHFSI_TRY(dev_upload(&op->d_nlCtko32, z32.data(), z32.size()));
      }
    }
    std::vector<float> tmp32(nlC.size());
    CHFSI_TRY(dev_upload(&op->d_nl_off, nl_off.data(), nl_off.size()));
    CHFSI_TRY(dev_upload(&op->d_slot_proj, slot_proj.data(), slot_proj.size()));
    CHFSI_TRY(dev_upload(&op->d_nlC64, nlC.data(), nlC.size()));
    for (size_t i = 0; i < nlC.size(); ++i) tmp32[i] = (float)nlC[i];
    CHFSI_TRY(dev_upload(&op->d_nlC32, tmp32.data(), tmp32.size()));
    CHFSI_TRY(dev_upload(&op->d_nlCt64, nlCt.data(), nlCt.size()));
    for (size_t i = 0; i < nlCt.size(); ++i) tmp32[i] = (float)nlCt[i];
    CHFSI_TRY(dev_upload(&op->d_nlCt32, tmp32.data(), tmp32.size()));
    if (op->nk > 0) {
      // complex projector tiles per k-point, paired (Re, Im) layouts
      const size_t per_k = size_t(2) * S * ne_pad;
      std::vector<double> zC(per_k * op->nk, 0.0), zCt(per_k * op->nk, 0.0);
      for (int k = 0; k < op->nk; ++k) {
        double *zc = zC.data() + per_k * k, *zct = zCt.data() + per_k * k;
        for (int c = 0; c < C; ++c) {
          const int s0 = nl_off[c], ns = nl_off[c + 1] - s0;
          int local = 0;
          for (int p : pairs_of[c]) {
            for (int al = 0; al < P; ++al, ++local) {
              for (int i = 0; i < ne; ++i) {
                const size_t src = ((size_t(k) * d->n_pairs + p) * ne + i) * P + al;
                for (int h = 0; h < 2; ++h) {
                  const double v = d->pair_C_kpoints[2 * src + h];
                  zc[size_t(s0) * 2 * ne_pad + size_t(2 * (local / 16) + h) * ne_pad * 16 +
                     size_t(i) * 16 + (local % 16)] = v;
                  zct[size_t(s0) * 2 * ne_pad + size_t(2 * (i / 16) + h) * ns * 16 +
                      size_t(local) * 16 + (i % 16)] = v;
                }
              }
            }
          }
        }
      }
      std::vector<float> z32(zC.size());
      CHFSI_TRY(dev_upload(&op->d_nlCk64, zC.data(), zC.size()));
      for (size_t i = 0; i < zC.size(); ++i) z32[i] = (float)zC[i];
      CHFSI_TRY(dev_upload(&op->d_nlCk32, z32.data(), z32.size()));
      CHFSI_TRY(dev_upload(&op->d_nlCtk64, zCt.data(), zCt.size()));
      for (size_t i = 0; i < zCt.size(); ++i) z32[i] = (float)zCt[i];
      CHFSI_TRY(dev_upload(&op->d_nlCtk32, z32.data(), z32.size()));
    }
    CHFSI_TRY(dev_alloc(&op->d_nlC16, nlC.size()));
    CHFSI_HIP(launch_f64_to_bf16(op->d_nlC64, (int64_t)nlC.size(), op->d_nlC16, 0));
    CHFSI_TRY(dev_upload(&op->d_proj_ptr, proj_ptr.data(), proj_ptr.size()));
    CHFSI_TRY(dev_upload(&op->d_proj_slots, proj_slots.data(), proj_slots.size()));
    CHFSI_TRY(dev_upload(&op->d_proj_V, d->proj_V, d->n_proj));
    CHFSI_TRY(dev_upload(&op->d_nl_cells, nl_cells.data(), nl_cells.size()));
  }
  CHFSI_HIP(hipDeviceSynchronize());
  *out = guard.release();
  return CHFSI_OK;
}

extern "C" void chfsi_operator_destroy(chfsi_operator *op) { delete op; }

extern "C" chfsi_status chfsi_operator_set_potential(chfsi_operator *op, const double *vJxW,
                                                     void *stream) {
  CHFSI_ARG(op && vJxW, "null argument");
  CHFSI_ARG(op->d_Nq && op->d_K, "operator was created without shape data");
  CHFSI_HIP(launch_build_h(op->d_K, op->d_Nq, vJxW, nullptr, nullptr, op->C, op->ne, op->ne_pad,
                           op->nq, op->d_Ht64, op->d_Ht32, op->d_Ht16, as_stream(stream),
                           op->K_per_cell));
  op->has_H = true;
  op->h_sym = true;  // the build writes each off-diagonal pair block from one sum
  return rebuild_kpoint_matrices(op, as_stream(stream));
}

extern "C" chfsi_status chfsi_operator_set_potential_gga(chfsi_operator *op, const double *vJxW,
                                                         const double *gJxW, void *stream) {
  CHFSI_ARG(op && vJxW && gJxW, "null argument");
  CHFSI_ARG(op->d_Nq && op->d_K, "operator was created without shape data");
  CHFSI_ARG(op->d_Gq, "operator was created without shape gradients");
  CHFSI_HIP(launch_build_h(op->d_K, op->d_Nq, vJxW, op->d_Gq, gJxW, op->C, op->ne, op->ne_pad,
                           op->nq, op->d_Ht64, op->d_Ht32, op->d_Ht16, as_stream(stream),
                           op->K_per_cell, op->d_invJ, op->jac));
  op->has_H = true;
  op->h_sym = true;
  return rebuild_kpoint_matrices(op, as_stream(stream));
}

extern "C" chfsi_status chfsi_operator_set_cell_matrices(chfsi_operator *op, const double *H,
                                                         void *stream) {
  CHFSI_ARG(op && H, "null argument");
  CHFSI_HIP(launch_tile_h(H, op->C, op->ne, op->ne_pad, op->d_Ht64, op->d_Ht32,
                          as_stream(stream)));
  CHFSI_HIP(launch_f64_to_bf16(op->d_Ht64, int64_t(op->C) * op->ne_pad * op->ne_pad, op->d_Ht16,
                               as_stream(stream)));
  op->has_H = true;
  op->h_sym = false;  // caller's matrices: not assumed symmetric
  return rebuild_kpoint_matrices(op, as_stream(stream));
}

extern "C" chfsi_status chfsi_operator_get_cell_matrices(chfsi_operator *op, double *H,
                                                         void *stream) {
  CHFSI_ARG(op && H, "null argument");
  if (!op->has_H) {
    set_error("cell matrices not set");
    return CHFSI_ERR_STATE;
  }
  CHFSI_HIP(launch_untile_h(op->d_Ht64, op->C, op->ne, op->ne_pad, H, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" const double *chfsi_operator_sqrt_mass(const chfsi_operator *op) {
  return op ? op->d_sqrt_mass : nullptr;
}
extern "C" const double *chfsi_operator_inv_sqrt_mass(const chfsi_operator *op) {
  return op ? op->d_inv_sqrt_mass : nullptr;
}
extern "C" int64_t chfsi_operator_n_dofs(const chfsi_operator *op) { return op ? op->M : 0; }

namespace chfsi {

// ---------------------------------------------------------------------------
// Fixed device scratch of the no-allocation HX entry point
// ---------------------------------------------------------------------------
// One array per GPU, resident for the life of the process.  In host code the
// symbol is only a stand-in: its device address is looked up once per device
// (with that device current) and every use goes through that address.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_addr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device (looked up on first use).
static chfsi_status noalloc_scratch(void **out) {
  int dev = 0;
  CHFSI_HIP(hipGetDevice(&dev));
  CHFSI_ARG(dev >= 0 && dev < kMaxDevices, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                hipGetErrorString(e));
      return CHFSI_ERR_HIP;
    }
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return CHFSI_OK;
}

template <typename T>
struct OpData;
template <>
struct OpData<double> {
  static const double *Ht(chfsi_operator *op) { return op->d_Ht64; }
  static const double *nlC(chfsi_operator *op) { return op->d_nlC64; }
  static const double *nlCt(chfsi_operator *op) { return op->d_nlCt64; }
  static const double *nlCto(chfsi_operator *op) { return op->d_nlCto64; }
  static const double *s(chfsi_operator *op) { return op->d_inv_sqrt_mass; }
  static Scratch &partial(chfsi_operator *op) { return op->partial64; }
  static Scratch &W(chfsi_operator *op) { return op->W64; }
  static const double *s_gather(chfsi_operator *op) { return op->d_s_gather64; }
  static const double *ones_cons(chfsi_operator *op) { return op->d_ones_cons64; }
};
template <>
struct OpData<float> {
  static const float *Ht(chfsi_operator *op) { return op->d_Ht32; }
  static const float *nlC(chfsi_operator *op) { return op->d_nlC32; }
  static const float *nlCt(chfsi_operator *op) { return op->d_nlCt32; }
  static const float *nlCto(chfsi_operator *op) { return op->d_nlCto32; }
  static const float *s(chfsi_operator *op) { return op->d_inv_sqrt_mass32; }
  static Scratch &partial(chfsi_operator *op) { return op->partial32; }
  static Scratch &W(chfsi_operator *op) { return op->W32; }
  static const float *s_gather(chfsi_operator *op) { return op->d_s_gather32; }
  static const float *ones_cons(chfsi_operator *op) { return op->d_ones_cons32; }
};

template <typename T>
chfsi_status hx_fused(chfsi_operator *op, const T *src, int64_t ld_src, T *dst, int64_t ld_dst,
                      int B, T alpha, T beta, T gamma, int beta_mode, bool scaled,
                      hipStream_t s, bool bf16, const T *R, int64_t ldR, const T *rcoef,
                      void *ext_scratch) {
  CHFSI_ARG(!bf16 || sizeof(T) == 4, "the BF16 apply works on FP32 multivectors");
  CHFSI_ARG(!R || (!bf16 && sizeof(T) == 4 && rcoef && ldR >= B), "bad defect source");
  if (!op->has_H) {
    set_error("cell matrices not set (call chfsi_operator_set_potential)");
    return CHFSI_ERR_STATE;
  }
  CHFSI_ARG(B > 0 && ld_src >= B && ld_dst >= B, "bad block width / leading dimension");
  CHFSI_ARG(src != dst, "HX needs distinct source and destination");
  const T *sc = scaled ? OpData<T>::s(op) : nullptr;
  // input (gather) and output (scatter) scales; the same M^{-1/2} unless the
  // operator has general constraints
  const T *s_in = sc, *s_out = sc;
  T *src_w = const_cast<T *>(src);
  if (op->general_constraints) {
    // constraintMatrixInfoDevice::distribute before the cell products
    // (kohnShamDFTOperatorDevice.cc:3806-3807): the constrained rows of src
    // take sum_j c_j s_j src_j (scaled masters) and are gathered with scale 1;
    // in the scatter a zero output scale marks them, so they collect the plain
    // cell sums for the slave -> master pass below
    CHFSI_HIP(launch_distribute_scaled<T>(op->d_c_rows, op->d_c_ptr, op->d_c_cols, op->d_c_vals,
                                          op->nc, sc, src_w, ld_src, B, s));
    s_in = scaled ? OpData<T>::s_gather(op) : nullptr;
    s_out = scaled ? sc : OpData<T>::ones_cons(op);
  }
  T *W = nullptr;
  if (op->n_slots > 0) {
    // projection partials (n_slots x B) and V-scaled projections (n_proj x B):
    // the operator's own buffers, or a caller-provided scratch laid out as
    // [partials | W] with W at the next 256-byte boundary
    T *Pp = nullptr;
    if (ext_scratch) {
      Pp = static_cast<T *>(ext_scratch);
      W = reinterpret_cast<T *>(static_cast<char *>(ext_scratch) +
                                noalloc_w_offset(sizeof(T), op->n_slots, B));
    } else {
      Scratch &P = OpData<T>::partial(op);
      Scratch &Wb = OpData<T>::W(op);
      CHFSI_TRY(P.ensure(sizeof(T) * size_t(op->n_slots) * B));
      CHFSI_TRY(Wb.ensure(sizeof(T) * size_t(op->n_proj) * B));
      Pp = static_cast<T *>(P.ptr);
      W = static_cast<T *>(Wb.ptr);
    }
    ProjLaunch<T> pl{};
    pl.cells = op->d_nl_cells;
    pl.ncells = op->n_nl_cells;
    pl.map = op->d_map;
    pl.ne = op->ne;
    pl.ne_pad = op->ne_pad;
    pl.nl_off = op->d_nl_off;
    pl.nl_Ct = OpData<T>::nlCt(op);
    pl.src = src;
    pl.ld_src = ld_src;
    pl.s_in = s_in;
    pl.partial = Pp;
    pl.ldP = B;
    pl.B = B;
    pl.max_srb = op->max_srb;
    pl.variant = sizeof(T) == 8 ? op->proj_variant64 : op->proj_variant32;
    if (pl.variant >= 4) {
      pl.nl_Ct = OpData<T>::nlCto(op);
      pl.own_map = op->d_own_map;
      pl.own_n = op->d_own_n;
    }
    CHFSI_HIP(launch_proj<T>(pl, s));
    CHFSI_HIP(launch_proj_reduce<T>(op->d_proj_ptr, op->d_proj_slots, op->d_proj_V, op->n_proj,
                                    Pp, B, W, B, B, s));
  }
  HxLaunch<T> a{};
  a.map = op->d_map;
  a.ne = op->ne;
  a.ne_pad = op->ne_pad;
  a.Ht = OpData<T>::Ht(op);
  a.src = src;
  a.ld_src = ld_src;
  a.dst = dst;
  a.ld_dst = ld_dst;
  a.s_in = s_in;
  a.s_out = s_out;
  a.alpha = alpha;
  a.beta = beta;
  a.gamma = gamma;
  a.beta_mode = beta_mode;
  a.B = B;
  a.variant = sizeof(T) == 8 ? op->hx_variant64 : (bf16 ? op->hx_variant16 : op->hx_variant32);
  a.R = R;
  a.ldR = ldR;
  a.rcoef = rcoef;
  a.h_sym = op->h_sym ? 1 : 0;
  if (op->n_slots > 0) {
    a.nl_off = op->d_nl_off;
    a.nl_C = OpData<T>::nlC(op);
    a.slot_proj = op->d_slot_proj;
    a.W = W;
    a.ldW = B;
  }
  for (int k = 0; k < op->ncolors; ++k) {
    a.cells = op->d_cells_by_color + op->color_off[k];
    a.ncells = op->color_off[k + 1] - op->color_off[k];
    if (a.ncells == 0) continue;
    if constexpr (sizeof(T) == 4) {
      if (bf16) {
        CHFSI_HIP(launch_hx_bf16(a, op->d_Ht16, op->d_nlC16, s));
        continue;
      }
    }
    CHFSI_HIP(launch_hx<T>(a, s));
  }
  if (op->general_constraints) {
    // distribute_slave_to_master (kohnShamDFTOperatorDevice.cc:3832-3838):
    // dst[m] += s_out[m] sum_s c_sm dst[s], then the constrained rows of dst
    // and src are cleared
    CHFSI_HIP(launch_slave_to_master_scaled<T>(op->d_m_rows, op->d_m_ptr, op->d_m_slaves,
                                               op->d_m_vals, op->nm, s_out, dst, ld_dst, B, s));
    CHFSI_HIP(launch_zero_rows_t<T>(op->d_c_rows, op->nc, dst, ld_dst, B, s));
    CHFSI_HIP(launch_zero_rows_t<T>(op->d_c_rows, op->nc, src_w, ld_src, B, s));
  }
  return CHFSI_OK;
}

template chfsi_status hx_fused<double>(chfsi_operator *, const double *, int64_t, double *,
                                       int64_t, int, double, double, double, int, bool,
                                       hipStream_t, bool, const double *, int64_t,
                                       const double *, void *);
template chfsi_status hx_fused<float>(chfsi_operator *, const float *, int64_t, float *, int64_t,
                                      int, float, float, float, int, bool, hipStream_t, bool,
                                      const float *, int64_t, const float *, void *);

}  // namespace chfsi

extern "C" chfsi_status chfsi_hx(chfsi_operator *op, const double *src, int64_t ld_src,
                                 double *dst, int64_t ld_dst, int32_t B, int32_t scale_flag,
                                 double scalar, void *stream) {
  CHFSI_ARG(op && src && dst, "null argument");
  // dst <- (scale_flag ? dst : S dst) + scalar * S H S src
  CHFSI_TRY(hx_fused<double>(op, src, ld_src, dst, ld_dst, B, scalar, 1.0, 0.0,
                             scale_flag ? 0 : 2, true, as_stream(stream)));
  if (op->nc > 0) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, dst, ld_dst, B, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_hx_noalloc(chfsi_operator *op, const double *src, int64_t ld_src,
                                         double *dst, int64_t ld_dst, int32_t B,
                                         int32_t scale_flag, double scalar, void *stream) {
  CHFSI_ARG(op && src && dst, "null argument");
  CHFSI_ARG(B > 0, "bad block width");
  void *ws = nullptr;
  if (op->n_slots > 0) {
    const size_t need = noalloc_scratch_bytes(sizeof(double), op->n_slots, op->n_proj, B);
    CHFSI_ARG(need <= kNoAllocScratchBytes,
              "chfsi_hx_noalloc: %d columns need %zu bytes of projection scratch "
              "(%d slots, %d projectors); the fixed scratch holds %zu",
              B, need, op->n_slots, op->n_proj, kNoAllocScratchBytes);
    CHFSI_TRY(noalloc_scratch(&ws));
  }
  CHFSI_TRY(hx_fused<double>(op, src, ld_src, dst, ld_dst, B, scalar, 1.0, 0.0,
                             scale_flag ? 0 : 2, true, as_stream(stream), false, nullptr, 0,
                             nullptr, ws));
  if (op->nc > 0) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, dst, ld_dst, B, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_hx_fp32(chfsi_operator *op, const float *src, int64_t ld_src,
                                      float *dst, int64_t ld_dst, int32_t B, int32_t scale_flag,
                                      float scalar, void *stream) {
  CHFSI_ARG(op && src && dst, "null argument");
  CHFSI_TRY(hx_fused<float>(op, src, ld_src, dst, ld_dst, B, scalar, 1.0f, 0.0f,
                            scale_flag ? 0 : 2, true, as_stream(stream)));
  if (op->nc > 0)
    CHFSI_HIP(launch_zero_rows_t<float>(op->d_c_rows, op->nc, dst, ld_dst, B, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_hx_bf16(chfsi_operator *op, const float *src, int64_t ld_src,
                                      float *dst, int64_t ld_dst, int32_t B, int32_t scale_flag,
                                      float scalar, void *stream) {
  CHFSI_ARG(op && src && dst, "null argument");
  CHFSI_TRY(hx_fused<float>(op, src, ld_src, dst, ld_dst, B, scalar, 1.0f, 0.0f,
                            scale_flag ? 0 : 2, true, as_stream(stream), true));
  if (op->nc > 0)
    CHFSI_HIP(launch_zero_rows_t<float>(op->d_c_rows, op->nc, dst, ld_dst, B, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_hx_cheby(chfsi_operator *op, const double *src, int64_t ld_src,
                                       double *dst, int64_t ld_dst, int32_t B, void *stream) {
  CHFSI_ARG(op && src && dst, "null argument");
  return hx_fused<double>(op, src, ld_src, dst, ld_dst, B, 1.0, 1.0, 0.0, 0, false,
                          as_stream(stream));
}

extern "C" chfsi_status chfsi_constraints_distribute(chfsi_operator *op, double *X, int64_t ld,
                                                     int32_t B, void *stream) {
  CHFSI_ARG(op && X && ld >= B, "bad argument");
  CHFSI_HIP(launch_distribute(op->d_c_rows, op->d_c_ptr, op->d_c_cols, op->d_c_vals, op->nc, X,
                              ld, B, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_constraints_slave_to_master(chfsi_operator *op, double *X,
                                                          int64_t ld, int32_t B, void *stream) {
  CHFSI_ARG(op && X && ld >= B, "bad argument");
  CHFSI_HIP(launch_slave_to_master(op->d_m_rows, op->d_m_ptr, op->d_m_slaves, op->d_m_vals,
                                   op->nm, X, ld, B, as_stream(stream)));
  CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, X, ld, B, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_constraints_set_zero(chfsi_operator *op, double *X, int64_t ld,
                                                   int32_t B, void *stream) {
  CHFSI_ARG(op && X && ld >= B, "bad argument");
  CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, X, ld, B, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_operator_set_hx_variant(chfsi_operator *op, int32_t precision,
                                                     int32_t variant) {
  CHFSI_ARG(op, "null operator");
  if (precision == CHFSI_BF16) {
    CHFSI_ARG(variant >= 0 && variant < hx_bf16_variants(),
              "bad BF16 variant (0, 2-5: K = 32 ring, 1: K = 16)");
    op->hx_variant16 = variant;
    return CHFSI_OK;
  }
  CHFSI_ARG(precision == CHFSI_FP64 || precision == CHFSI_FP32, "bad precision");
  CHFSI_ARG(variant >= 0 && variant < hx_variants(precision == CHFSI_FP64), "bad variant");
  if (precision == CHFSI_FP64)
    op->hx_variant64 = variant;
  else
    op->hx_variant32 = variant;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_operator_set_proj_variant(chfsi_operator *op, int32_t precision,
                                                       int32_t variant) {
  CHFSI_ARG(op, "null operator");
  CHFSI_ARG(precision == CHFSI_FP64 || precision == CHFSI_FP32, "bad precision");
  CHFSI_ARG(variant >= 0 && variant < proj_variants(), "bad variant");
  (precision == CHFSI_FP64 ? op->proj_variant64 : op->proj_variant32) = variant;
  return CHFSI_OK;
}

namespace {
template <typename T>
chfsi_status time_variants(chfsi_operator *op, int B, int *best, float *times, hipStream_t st) {
  // times[0..15]: tile variants, times[16..23]: projection variants
  const int nv = hx_variants(sizeof(T) == 8) < 16 ? hx_variants(sizeof(T) == 8) : 16;
  Scratch x, y;
  CHFSI_TRY(x.ensure(sizeof(T) * op->M * B));
  CHFSI_TRY(y.ensure(sizeof(T) * op->M * B));
  CHFSI_HIP(hipMemsetAsync(x.ptr, 0, sizeof(T) * op->M * B, st));
  CHFSI_HIP(hipMemsetAsync(y.ptr, 0, sizeof(T) * op->M * B, st));
  hipEvent_t e0, e1;
  CHFSI_HIP(hipEventCreate(&e0));
  CHFSI_HIP(hipEventCreate(&e1));
  int &var = sizeof(T) == 8 ? op->hx_variant64 : op->hx_variant32;
  float best_t = 1e30f;
  for (int v = 0; v < nv; ++v) {
    var = v;
    const T *xs = static_cast<const T *>(x.ptr);
    T *ys = static_cast<T *>(y.ptr);
    chfsi_status stt = hx_fused<T>(op, xs, B, ys, B, B, T(1), T(1), T(0), 0, true, st);
    if (stt == CHFSI_OK) {
      CHFSI_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < 3 && stt == CHFSI_OK; ++r)
        stt = hx_fused<T>(op, xs, B, ys, B, B, T(1), T(1), T(0), 0, true, st);
      CHFSI_HIP(hipEventRecord(e1, st));
      CHFSI_HIP(hipEventSynchronize(e1));
    }
    float ms = 1e30f;
    if (stt == CHFSI_OK) CHFSI_HIP(hipEventElapsedTime(&ms, e0, e1));
    times[v] = stt == CHFSI_OK ? ms / 3 : -1.0f;
    if (stt == CHFSI_OK && ms < best_t) {
      best_t = ms;
      *best = v;
    }
  }
  var = *best;
  // projection variants with the chosen tile variant, timed as whole applies
  if (op->n_slots > 0) {
    int &pv = sizeof(T) == 8 ? op->proj_variant64 : op->proj_variant32;
    float best_p = 1e30f;
    int bp = 0;
    for (int v = 0; v < proj_variants() && v < 8; ++v) {
      pv = v;
      const T *xs = static_cast<const T *>(x.ptr);
      T *ys = static_cast<T *>(y.ptr);
      chfsi_status stt = hx_fused<T>(op, xs, B, ys, B, B, T(1), T(1), T(0), 0, true, st);
      float ms = 1e30f;
      if (stt == CHFSI_OK) {
        CHFSI_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 3 && stt == CHFSI_OK; ++r)
          stt = hx_fused<T>(op, xs, B, ys, B, B, T(1), T(1), T(0), 0, true, st);
        CHFSI_HIP(hipEventRecord(e1, st));
        CHFSI_HIP(hipEventSynchronize(e1));
        if (stt == CHFSI_OK) CHFSI_HIP(hipEventElapsedTime(&ms, e0, e1));
      }
      times[16 + v] = stt == CHFSI_OK ? ms / 3 : -1.0f;
      if (stt == CHFSI_OK && ms < best_p) {
        best_p = ms;
        bp = v;
      }
    }
    pv = bp;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  x.release();
  y.release();
  return CHFSI_OK;
}
}  // namespace

extern "C" chfsi_status chfsi_operator_autotune(chfsi_operator *op, int32_t B, int32_t *best64,
                                                int32_t *best32, float *times_ms, void *stream) {
  CHFSI_ARG(op && B > 0, "bad argument");
  hipStream_t st = as_stream(stream);
  float t[48];
  for (float &v : t) v = -1.0f;
  int b64 = 0, b32 = 0;
  CHFSI_TRY(time_variants<double>(op, B, &b64, t, st));
  CHFSI_TRY(time_variants<float>(op, B, &b32, t + 24, st));
  if (best64) *best64 = b64;
  if (best32) *best32 = b32;
  if (times_ms)
    for (int i = 0; i < 48; ++i) times_ms[i] = t[i];
  return CHFSI_OK;
}

namespace chfsi {

template <typename T>
chfsi_status quad_product(chfsi_operator *op, const T *a, int64_t lda, const T *b, int64_t ldb,
                          int B, bool scaled, const double *jxw, const double *w, double wscale,
                          double *g, int accumulate, hipStream_t st, int wshift) {
  const double *s_in = scaled ? op->d_inv_sqrt_mass : nullptr;
  T *aw = const_cast<T *>(a), *bw = const_cast<T *>(b);
  if (op->general_constraints) {
    // interpolation needs the constrained rows' values: distributed from the
    // (scaled) masters into the rows, gathered with scale 1, cleared after
    const T *sc = scaled ? OpData<T>::s(op) : nullptr;
    CHFSI_HIP(launch_distribute_scaled<T>(op->d_c_rows, op->d_c_ptr, op->d_c_cols, op->d_c_vals,
                                          op->nc, sc, aw, lda, B, st));
    if (bw != aw)
      CHFSI_HIP(launch_distribute_scaled<T>(op->d_c_rows, op->d_c_ptr, op->d_c_cols,
                                            op->d_c_vals, op->nc, sc, bw, ldb, B, st));
    if (scaled) s_in = op->d_s_gather64;
  }
  CHFSI_HIP(launch_quad_product(op->d_map, op->C, op->ne, op->nq, op->d_Nq, jxw, a, lda, b, ldb,
                                B, s_in, w, wscale, g, accumulate, st, wshift));
  if (op->general_constraints) {
    CHFSI_HIP(launch_zero_rows_t<T>(op->d_c_rows, op->nc, aw, lda, B, st));
    if (bw != aw) CHFSI_HIP(launch_zero_rows_t<T>(op->d_c_rows, op->nc, bw, ldb, B, st));
  }
  return CHFSI_OK;
}
template chfsi_status quad_product<double>(chfsi_operator *, const double *, int64_t,
                                           const double *, int64_t, int, bool, const double *,
                                           const double *, double, double *, int, hipStream_t,
                                           int);
template chfsi_status quad_product<float>(chfsi_operator *, const float *, int64_t, const float *,
                                          int64_t, int, bool, const double *, const double *,
                                          double, double *, int, hipStream_t, int);

}  // namespace chfsi

extern "C" chfsi_status chfsi_density(chfsi_operator *op, const double *X, int64_t ld, int32_t N,
                                      const double *w_dev, double *rho_dev, int32_t accumulate,
                                      void *stream) {
  CHFSI_ARG(op && X && rho_dev && N > 0 && ld >= N, "bad argument");
  CHFSI_ARG(op->d_Nq, "operator was created without shape data");
  return quad_product<double>(op, X, ld, X, ld, N, false, nullptr, w_dev, 1.0, rho_dev,
                              accumulate, as_stream(stream), 0);
}

extern "C" chfsi_status chfsi_density_complex(chfsi_operator *op, const void *X, int64_t ld,
                                              int32_t N, const double *w_dev, double *rho_dev,
                                              int32_t accumulate, void *stream) {
  CHFSI_ARG(op && X && rho_dev && N > 0 && ld >= N, "bad argument");
  CHFSI_ARG(op->d_Nq, "operator was created without shape data");
  // |psi|^2 = Re^2 + Im^2: the real contraction over the 2N interleaved real
  // columns, each weight shared by the pair
  const double *Xr = static_cast<const double *>(X);
  return quad_product<double>(op, Xr, 2 * ld, Xr, 2 * ld, 2 * N, false, nullptr, w_dev, 1.0,
                              rho_dev, accumulate, as_stream(stream), 1);
}

extern "C" chfsi_status chfsi_hx_adjoint_v(chfsi_operator *op, const double *Ybar, int64_t ldy,
                                           const double *X, int64_t ldx, int32_t B,
                                           double scalar, double *g_dev, int32_t accumulate,
                                           void *stream) {
  CHFSI_ARG(op && Ybar && X && g_dev && B > 0 && ldy >= B && ldx >= B, "bad argument");
  CHFSI_ARG(op->d_Nq && op->d_jxw, "operator was created without shape data / JxW");
  return quad_product<double>(op, Ybar, ldy, X, ldx, B, true, op->d_jxw, nullptr, scalar, g_dev,
                              accumulate, as_stream(stream), 0);
}
