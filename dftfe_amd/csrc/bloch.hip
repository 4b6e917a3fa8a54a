// Complex (Bloch k-point) side of the C ABI: active k-point, complex HX,
// complex Chebyshev filter, the batched multi-k-point filter (plain launches
// or one captured HIP graph), the complex dense steps and the complex pass.
//
// Reference: kohnShamDFTOperatorDeviceClass with dataTypes::number complex
// (kohnShamDFTOperatorDevice.cc:1037-1057 reinitkPointSpinIndex, :3765-3860 HX),
// linearAlgebraOperationsDevice.cc:340-727 (Lanczos, filter), pseudoGSDevice.cc:81-463,
// rayleighRitzDevice.cc:80-353, chebyshevOrthogonalizedSubspaceIterationSolverDevice.cc:155-736,
// and the per-k-point loop of src/dft/dft.cc:3077-3130.
//
// A complex block of Bc columns is handled as 2*Bc real columns everywhere a
// real kernel applies unchanged (scaling, casts, copies, axpy, dot products,
// residual sums); only the cell GEMM and the projections need the complex
// kernels, and the dense N x N steps use the complex BLAS / LAPACK routines.

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "bloch.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "operator.hpp"
#include "solver.hpp"

namespace chfsi {

chfsi_status rebuild_kpoint_matrices(chfsi_operator *op, hipStream_t s) {
  if (op->nk == 0) return CHFSI_OK;
  const size_t per_k = size_t(op->C) * 2 * op->ne_pad * op->ne_pad;
  for (int k = 0; k < op->nk; ++k)
    CHFSI_HIP(launch_build_hk(op->d_Ht64, op->d_Mc, op->d_Gd, op->C, op->ne, op->ne_pad,
                              op->kpts[3 * k], op->kpts[3 * k + 1], op->kpts[3 * k + 2],
                              op->d_Hk64 + per_k * k, op->d_Hk32 + per_k * k, s));
  return CHFSI_OK;
}

namespace {

template <typename T>
struct ZOp;
template <>
struct ZOp<double> {
  static const double *Hk(chfsi_operator *op, int k) {
    return op->d_Hk64 + size_t(op->C) * 2 * op->ne_pad * op->ne_pad * k;
  }
  static const double *nlC(chfsi_operator *op, int k) {
    return op->d_nlCk64 ? op->d_nlCk64 + size_t(2) * op->n_slots * op->ne_pad * k : nullptr;
  }
  static const double *nlCt(chfsi_operator *op, int k) {
    const double *t = op->zproj_variant == 1 ? op->d_nlCtko64 : op->d_nlCtk64;
    return t ? t + size_t(2) * op->n_slots * op->ne_pad * k : nullptr;
  }
  static const double *s(chfsi_operator *op) { return op->d_inv_sqrt_mass; }
  static int &variant(chfsi_operator *op) { return op->zhx_variant64; }
};
template <>
struct ZOp<float> {
  static const float *Hk(chfsi_operator *op, int k) {
    return op->d_Hk32 + size_t(op->C) * 2 * op->ne_pad * op->ne_pad * k;
  }
  static const float *nlC(chfsi_operator *op, int k) {
    return op->d_nlCk32 ? op->d_nlCk32 + size_t(2) * op->n_slots * op->ne_pad * k : nullptr;
  }
  static const float *nlCt(chfsi_operator *op, int k) {
    const float *t = op->zproj_variant == 1 ? op->d_nlCtko32 : op->d_nlCtk32;
    return t ? t + size_t(2) * op->n_slots * op->ne_pad * k : nullptr;
  }
  static const float *s(chfsi_operator *op) { return op->d_inv_sqrt_mass32; }
  static int &variant(chfsi_operator *op) { return op->zhx_variant32; }
};

// One complex apply of a batch entry (widths and leading dimensions in
// complex elements):  dst <- base + alpha * S H_k S src  (base as HxLaunch).
template <typename T>
struct ZApply {
  int k;
  const T *src;
  int64_t ld_src;
  T *dst;
  int64_t ld_dst;
  int Bc;
  double alpha, beta, gamma;
  int beta_mode;
};

// Item tables of `ndeg` rounds of `n` entries; the projection scratch of an
// entry is shared by its rounds (its width must not change between rounds).
template <typename T>
chfsi_status ztable_build(chfsi_operator *op, const std::vector<ZApply<T>> &req, int n,
                          Scratch &W, Scratch &P, std::vector<ZHxItem<T>> &host) {
  const int ndeg = (int)req.size() / n;
  std::vector<size_t> woff(n), poff(n);
  size_t wtot = 0, ptot = 0;
  for (int e = 0; e < n; ++e) {
    const size_t B = 2 * size_t(req[e].Bc);
    woff[e] = wtot;
    poff[e] = ptot;
    wtot += size_t(op->n_proj) * B;
    ptot += size_t(op->n_slots) * B;
  }
  if (op->n_slots > 0) {
    CHFSI_TRY(W.ensure(sizeof(T) * std::max<size_t>(wtot, 1)));
    CHFSI_TRY(P.ensure(sizeof(T) * std::max<size_t>(ptot, 1)));
  }
  host.assign(req.size(), ZHxItem<T>{});
  for (int d = 0; d < ndeg; ++d)
    for (int e = 0; e < n; ++e) {
      const ZApply<T> &r = req[size_t(d) * n + e];
      ZHxItem<T> &it = host[size_t(d) * n + e];
      it.Ht = ZOp<T>::Hk(op, r.k);
      it.nlC = ZOp<T>::nlC(op, r.k);
      it.nlCt = ZOp<T>::nlCt(op, r.k);
      it.src = r.src;
      it.dst = r.dst;
      it.B = 2 * r.Bc;
      it.ld_src = 2 * r.ld_src;
      it.ld_dst = 2 * r.ld_dst;
      if (op->n_slots > 0) {
        it.W = static_cast<T *>(W.ptr) + woff[e];
        it.partial = static_cast<T *>(P.ptr) + poff[e];
      }
      it.ldW = it.B;
      it.ldP = it.B;
      it.alpha = T(r.alpha);
      it.beta = T(r.beta);
      it.gamma = T(r.gamma);
      it.beta_mode = r.beta_mode;
    }
  return CHFSI_OK;
}

// Upload after the stream has drained (a previous round may still read the table).
template <typename T>
chfsi_status ztable_upload(const std::vector<ZHxItem<T>> &host, Scratch &dev, hipStream_t s) {
  CHFSI_TRY(dev.ensure(sizeof(ZHxItem<T>) * host.size()));
  CHFSI_HIP(hipStreamSynchronize(s));
  CHFSI_HIP(hipMemcpy(dev.ptr, host.data(), sizeof(ZHxItem<T>) * host.size(),
                      hipMemcpyHostToDevice));
  return CHFSI_OK;
}

// One round: projection + reduction + eight colour launches over all entries.
template <typename T>
chfsi_status zapply_round(chfsi_operator *op, const ZHxItem<T> *dev_items,
                          const ZHxItem<T> *host_items, int n, int variant, hipStream_t s) {
  const T *sc = ZOp<T>::s(op);
  if (op->n_slots > 0) {
    int maxB = 0;
    for (int e = 0; e < n; ++e) maxB = std::max(maxB, host_items[e].B);
    ZProjLaunch<T> pl{};
    pl.items = dev_items;
    pl.nk = n;
    pl.cells = op->d_nl_cells;
    pl.ncells = op->n_nl_cells;
    pl.max_srb = op->max_srb;
    pl.map = op->d_map;
    pl.ne = op->ne;
    pl.ne_pad = op->ne_pad;
    pl.s_in = sc;
    pl.nl_off = op->d_nl_off;
    if (op->zproj_variant == 1) {
      pl.own_map = op->d_own_map;
      pl.own_n = op->d_own_n;
    }
    CHFSI_HIP(launch_zproj<T>(pl, host_items, s));
    CHFSI_HIP(launch_zproj_reduce<T>(dev_items, n, maxB, op->d_proj_ptr, op->d_proj_slots,
                                     op->d_proj_V, op->n_proj, s));
  }
  ZHxLaunch<T> a{};
  a.items = dev_items;
  a.nk = n;
  a.map = op->d_map;
  a.ne = op->ne;
  a.ne_pad = op->ne_pad;
  a.s_in = sc;
  a.s_out = sc;
  if (op->n_slots > 0) {
    a.nl_off = op->d_nl_off;
    a.slot_proj = op->d_slot_proj;
  }
  for (int c = 0; c < op->ncolors; ++c) {
    a.cells = op->d_cells_by_color + op->color_off[c];
    a.ncells = op->color_off[c + 1] - op->color_off[c];
    if (a.ncells == 0) continue;
    CHFSI_HIP(launch_zhx<T>(a, host_items, variant, s));
  }
  return CHFSI_OK;
}

chfsi_status check_complex(chfsi_operator *op) {
  if (!op) {
    set_error("null operator");
    return CHFSI_ERR_ARG;
  }
  if (op->nk == 0) {
    set_error("operator was created without k-points");
    return CHFSI_ERR_STATE;
  }
  if (!op->has_H) {
    set_error("cell matrices not set (call chfsi_operator_set_potential)");
    return CHFSI_ERR_STATE;
  }
  return CHFSI_OK;
}

// Single complex apply at k-point k (tables in the operator's scratch).
template <typename T>
chfsi_status zhx(chfsi_operator *op, int k, const T *src, int64_t ld_src, T *dst, int64_t ld_dst,
                 int Bc, double alpha, double beta, double gamma, int beta_mode, hipStream_t s) {
  CHFSI_TRY(check_complex(op));
  CHFSI_ARG(Bc > 0 && ld_src >= Bc && ld_dst >= Bc, "bad block width / leading dimension");
  CHFSI_ARG(src != dst, "HX needs distinct source and destination");
  std::vector<ZApply<T>> req{{k, src, ld_src, dst, ld_dst, Bc, alpha, beta, gamma, beta_mode}};
  std::vector<ZHxItem<T>> host;
  CHFSI_TRY(ztable_build<T>(op, req, 1, op->ztab_W, op->ztab_P, host));
  CHFSI_TRY(ztable_upload<T>(host, op->ztab_dev, s));
  return zapply_round<T>(op, static_cast<const ZHxItem<T> *>(op->ztab_dev.ptr), host.data(), 1,
                         ZOp<T>::variant(op), s);
}

struct Cheb {
  double e, c, sigma, gamma;
};

}  // namespace
}  // namespace chfsi

using namespace chfsi;

// ---------------------------------------------------------------------------
// Batched filter object
// ---------------------------------------------------------------------------
struct chfsi_batch {
  chfsi_operator *op = nullptr;
  int n = 0, m = 0, precision = CHFSI_FP64;
  std::vector<int> kidx, Nc;
  std::vector<double *> X;
  std::vector<int64_t> ld;
  std::vector<double> a, b, a0;
  std::vector<size_t> off;  // per entry offset (reals) into the work buffers
  Scratch work0, work1;     // FP64: Y; FP32: X32 and Y32
  Scratch tab, W, P;
  std::vector<ZHxItem<double>> h64;
  std::vector<ZHxItem<float>> h32;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  ~chfsi_batch() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    if (cap) hipStreamDestroy(cap);
    for (Scratch *s : {&work0, &work1, &tab, &W, &P}) s->release();
  }
};

namespace {

// Work buffer pointers of entry e: P0 holds the input (and even iterates),
// P1 the odd iterates.
template <typename T>
void batch_buffers(chfsi_batch *bt, int e, T **P0, int64_t *ld0, T **P1) {
  if constexpr (sizeof(T) == 8) {
    *P0 = reinterpret_cast<T *>(bt->X[e]);
    *ld0 = bt->ld[e];
    *P1 = static_cast<T *>(bt->work0.ptr) + bt->off[e];
  } else {
    *P0 = static_cast<T *>(bt->work0.ptr) + bt->off[e];
    *ld0 = bt->Nc[e];
    *P1 = static_cast<T *>(bt->work1.ptr) + bt->off[e];
  }
}

template <typename T>
std::vector<ZHxItem<T>> &batch_host(chfsi_batch *bt) {
  if constexpr (sizeof(T) == 8)
    return bt->h64;
  else
    return bt->h32;
}

// Item table of all m degrees: degree d, entry e at row (d-1)*n + e.
// Degree 1: P1 <- k1 (H~ - c) P0; degree d >= 2: dst <- a1 (H~ - c) src + a2 dst
// with (src, dst) = (P1, P0) for even d and (P0, P1) for odd d
// (linearAlgebraOperationsDevice.cc:547-726).
template <typename T>
chfsi_status batch_table(chfsi_batch *bt) {
  const int n = bt->n, m = bt->m;
  std::vector<ZApply<T>> req(size_t(m) * n);
  for (int e = 0; e < n; ++e) {
    T *P0, *P1;
    int64_t ld0;
    batch_buffers<T>(bt, e, &P0, &ld0, &P1);
    const int64_t ld1 = bt->Nc[e];
    Cheb cs;
    cs.e = (bt->b[e] - bt->a[e]) / 2.0;
    cs.c = (bt->b[e] + bt->a[e]) / 2.0;
    cs.sigma = cs.e / (bt->a0[e] - cs.c);
    cs.gamma = 2.0 / cs.sigma;
    const double k1 = cs.sigma / cs.e;
    req[e] = ZApply<T>{bt->kidx[e], P0, ld0, P1, ld1, bt->Nc[e], k1, 0.0, -cs.c * k1, 1};
    for (int d = 2; d <= m; ++d) {
      const double sigma2 = 1.0 / (cs.gamma - cs.sigma);
      const double a1 = 2.0 * sigma2 / cs.e, a2 = -(cs.sigma * sigma2);
      ZApply<T> r{bt->kidx[e], nullptr, 0, nullptr, 0, bt->Nc[e], a1, a2, -cs.c * a1, 0};
      if (d % 2 == 0) {
        r.src = P1;
        r.ld_src = ld1;
        r.dst = P0;
        r.ld_dst = ld0;
      } else {
        r.src = P0;
        r.ld_src = ld0;
        r.dst = P1;
        r.ld_dst = ld1;
      }
      req[size_t(d - 1) * n + e] = r;
      cs.sigma = sigma2;
    }
  }
  CHFSI_TRY(ztable_build<T>(bt->op, req, n, bt->W, bt->P, batch_host<T>(bt)));
  return CHFSI_OK;
}

// The whole filter as a stream of launches (recorded into a graph or run directly).
template <typename T>
chfsi_status batch_record(chfsi_batch *bt, hipStream_t s) {
  chfsi_operator *op = bt->op;
  const int64_t M = op->M;
  const int n = bt->n;
  for (int e = 0; e < n; ++e) {
    T *P0, *P1;
    int64_t ld0;
    batch_buffers<T>(bt, e, &P0, &ld0, &P1);
    if constexpr (sizeof(T) == 4)
      CHFSI_HIP(launch_copy_block<double, float>(bt->X[e], 2 * bt->ld[e], P0, 2 * ld0, M,
                                                 2 * bt->Nc[e], s));
    CHFSI_HIP(launch_fill_zero<T>(P1, M * 2 * int64_t(bt->Nc[e]), s));
  }
  const ZHxItem<T> *dev = static_cast<const ZHxItem<T> *>(bt->tab.ptr);
  const std::vector<ZHxItem<T>> &host = batch_host<T>(bt);
  for (int d = 0; d < bt->m; ++d)
    CHFSI_TRY(zapply_round<T>(op, dev + size_t(d) * n, host.data() + size_t(d) * n, n,
                              ZOp<T>::variant(op), s));
  for (int e = 0; e < n; ++e) {
    T *P0, *P1;
    int64_t ld0;
    batch_buffers<T>(bt, e, &P0, &ld0, &P1);
    const T *res = (bt->m % 2 == 1) ? P1 : P0;
    const int64_t ldr = (bt->m % 2 == 1) ? bt->Nc[e] : ld0;
    if constexpr (sizeof(T) == 4)
      CHFSI_HIP(launch_copy_block<float, double>(res, 2 * ldr, bt->X[e], 2 * bt->ld[e], M,
                                                 2 * bt->Nc[e], s));
    else if (res != P0)
      CHFSI_HIP(launch_copy_block<double, double>(res, 2 * ldr, bt->X[e], 2 * bt->ld[e], M,
                                                  2 * bt->Nc[e], s));
  }
  return CHFSI_OK;
}

chfsi_status batch_init(chfsi_operator *op, const chfsi_kbatch *kb, chfsi_batch *bt) {
  CHFSI_TRY(check_complex(op));
  CHFSI_ARG(kb && kb->n >= 1 && kb->n <= kMaxBatch, "batch size must be in [1, %d]", kMaxBatch);
  CHFSI_ARG(kb->kpoint && kb->X && kb->ld && kb->n_vec && kb->a && kb->b && kb->a0,
            "missing batch arrays");
  CHFSI_ARG(kb->degree >= 1, "degree must be >= 1");
  CHFSI_ARG(kb->precision == CHFSI_FP64 || kb->precision == CHFSI_FP32,
            "batched filter precision must be CHFSI_FP64 or CHFSI_FP32");
  CHFSI_ARG(kb->precision == CHFSI_FP64 || op->nc == 0,
            "FP32 complex filter supports unconstrained (periodic-folded) layouts only");
  bt->op = op;
  bt->n = kb->n;
  bt->m = kb->degree;
  bt->precision = kb->precision;
  size_t tot = 0;
  for (int e = 0; e < kb->n; ++e) {
    CHFSI_ARG(kb->kpoint[e] >= 0 && kb->kpoint[e] < op->nk, "k-point index out of range");
    CHFSI_ARG(kb->X[e] && kb->n_vec[e] > 0 && kb->ld[e] >= kb->n_vec[e], "bad batch entry %d", e);
    bt->kidx.push_back(kb->kpoint[e]);
    bt->X.push_back(static_cast<double *>(kb->X[e]));
    bt->ld.push_back(kb->ld[e]);
    bt->Nc.push_back(kb->n_vec[e]);
    bt->a.push_back(kb->a[e]);
    bt->b.push_back(kb->b[e]);
    bt->a0.push_back(kb->a0[e]);
    bt->off.push_back(tot);
    tot += size_t(op->M) * 2 * kb->n_vec[e];
  }
  const size_t es = kb->precision == CHFSI_FP64 ? sizeof(double) : sizeof(float);
  CHFSI_TRY(bt->work0.ensure(es * tot));
  if (kb->precision == CHFSI_FP32) CHFSI_TRY(bt->work1.ensure(es * tot));
  return CHFSI_OK;
}

chfsi_status batch_tables(chfsi_batch *bt, hipStream_t s) {
  if (bt->precision == CHFSI_FP64) {
    CHFSI_TRY(batch_table<double>(bt));
    return ztable_upload<double>(bt->h64, bt->tab, s);
  }
  CHFSI_TRY(batch_table<float>(bt));
  return ztable_upload<float>(bt->h32, bt->tab, s);
}

chfsi_status batch_launch(chfsi_batch *bt, hipStream_t s) {
  return bt->precision == CHFSI_FP64 ? batch_record<double>(bt, s) : batch_record<float>(bt, s);
}

}  // namespace

extern "C" chfsi_status chfsi_operator_set_kpoint(chfsi_operator *op, int32_t k) {
  CHFSI_ARG(op, "null operator");
  CHFSI_ARG(k >= 0 && k < op->nk, "k-point %d out of range [0, %d)", k, op->nk);
  op->active_k = k;
  return CHFSI_OK;
}

extern "C" int32_t chfsi_operator_n_kpoints(const chfsi_operator *op) { return op ? op->nk : 0; }

extern "C" chfsi_status chfsi_operator_set_hx_variant_complex(chfsi_operator *op,
                                                             int32_t precision, int32_t variant) {
  CHFSI_ARG(op, "null operator");
  CHFSI_ARG(precision == CHFSI_FP64 || precision == CHFSI_FP32, "bad precision");
  CHFSI_ARG(variant >= 0 && variant < zhx_variants(precision == CHFSI_FP64), "bad variant");
  (precision == CHFSI_FP64 ? op->zhx_variant64 : op->zhx_variant32) = variant;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_operator_set_proj_variant_complex(chfsi_operator *op,
                                                               int32_t variant) {
  CHFSI_ARG(op, "null operator");
  CHFSI_ARG(variant == 0 || variant == 1, "bad variant (0: cell form, 1: owned rows)");
  CHFSI_ARG(variant == 0 || op->n_slots == 0 || op->d_nlCtko64,
            "operator was created without k-points");
  CHFSI_ARG(op->nk > 0, "operator was created without k-points");
  op->zproj_variant = variant;
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_hx_complex(chfsi_operator *op, const void *src, int64_t ld_src,
                                         void *dst, int64_t ld_dst, int32_t B, int32_t scale_flag,
                                         double scalar, int32_t precision, void *stream) {
  CHFSI_ARG(op && src && dst, "null argument");
  hipStream_t st = as_stream(stream);
  const int mode = scale_flag ? 0 : 2;
  if (precision == CHFSI_FP64) {
    CHFSI_TRY(zhx<double>(op, op->active_k, static_cast<const double *>(src), ld_src,
                          static_cast<double *>(dst), ld_dst, B, scalar, 1.0, 0.0, mode, st));
    if (op->nc > 0)
      CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, static_cast<double *>(dst), 2 * ld_dst,
                                 2 * B, st));
    return CHFSI_OK;
  }
  CHFSI_ARG(precision == CHFSI_FP32, "complex HX precision must be CHFSI_FP64 or CHFSI_FP32");
  CHFSI_ARG(op->nc == 0, "FP32 complex HX supports unconstrained (periodic-folded) layouts only");
  return zhx<float>(op, op->active_k, static_cast<const float *>(src), ld_src,
                    static_cast<float *>(dst), ld_dst, B, scalar, 1.0, 0.0, mode, st);
}

extern "C" chfsi_status chfsi_chebyshev_filter_complex(chfsi_operator *op, void *X, int64_t ld,
                                                       int32_t B, int32_t m, double a, double b,
                                                       double a0, int32_t precision,
                                                       void *stream) {
  CHFSI_ARG(op && X, "null argument");
  const int32_t k = op->active_k;
  chfsi_kbatch kb{1, &k, &X, &ld, &B, &a, &b, &a0, m, precision};
  return chfsi_batch_filter(op, &kb, stream);
}

extern "C" chfsi_status chfsi_batch_filter(chfsi_operator *op, const chfsi_kbatch *batch,
                                           void *stream) {
  hipStream_t st = as_stream(stream);
  std::unique_ptr<chfsi_batch> bt(new chfsi_batch());
  CHFSI_TRY(batch_init(op, batch, bt.get()));
  CHFSI_TRY(batch_tables(bt.get(), st));
  CHFSI_TRY(batch_launch(bt.get(), st));
  CHFSI_HIP(hipStreamSynchronize(st));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_batch_filter_create(chfsi_operator *op, const chfsi_kbatch *batch,
                                                  void *stream, chfsi_batch **out) {
  CHFSI_ARG(out, "null argument");
  *out = nullptr;
  hipStream_t st = as_stream(stream);
  std::unique_ptr<chfsi_batch> bt(new chfsi_batch());
  CHFSI_TRY(batch_init(op, batch, bt.get()));
  CHFSI_TRY(batch_tables(bt.get(), st));
  CHFSI_HIP(hipStreamCreateWithFlags(&bt->cap, hipStreamNonBlocking));
  CHFSI_HIP(hipStreamBeginCapture(bt->cap, hipStreamCaptureModeThreadLocal));
  chfsi_status rec = batch_launch(bt.get(), bt->cap);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(bt->cap, &g);
  if ((rec != CHFSI_OK || e != hipSuccess) && g) hipGraphDestroy(g);
  CHFSI_TRY(rec);
  CHFSI_HIP(e);
  bt->graph = g;
  CHFSI_HIP(hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0));
  *out = bt.release();
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_batch_filter_set_bounds(chfsi_batch *bt, const double *a,
                                                      const double *b, const double *a0) {
  CHFSI_ARG(bt && a && b && a0, "null argument");
  for (int e = 0; e < bt->n; ++e) {
    bt->a[e] = a[e];
    bt->b[e] = b[e];
    bt->a0[e] = a0[e];
  }
  // the table is read by the graph: wait for every replay before rewriting it
  CHFSI_HIP(hipDeviceSynchronize());
  return batch_tables(bt, nullptr);
}

extern "C" chfsi_status chfsi_batch_filter_run(chfsi_batch *bt, void *stream) {
  CHFSI_ARG(bt && bt->exec, "batch has no captured graph");
  CHFSI_HIP(hipGraphLaunch(bt->exec, as_stream(stream)));
  return CHFSI_OK;
}

extern "C" chfsi_status chfsi_batch_filter_launch(chfsi_batch *bt, void *stream) {
  CHFSI_ARG(bt, "null batch");
  return batch_launch(bt, as_stream(stream));
}

extern "C" void chfsi_batch_filter_destroy(chfsi_batch *bt) {
  if (bt) hipDeviceSynchronize();
  delete bt;
}

extern "C" chfsi_status chfsi_operator_autotune_complex(chfsi_operator *op,
                                                        const chfsi_kbatch *batch, int32_t *best64,
                                                        int32_t *best32, float *times_ms,
                                                        void *stream) {
  CHFSI_TRY(check_complex(op));
  CHFSI_ARG(batch, "null batch");
  hipStream_t st = as_stream(stream);
  if (times_ms)
    for (int i = 0; i < 16; ++i) times_ms[i] = -1.0f;
  CHFSI_ARG(batch->n >= 1 && batch->n <= kMaxBatch && batch->X && batch->ld && batch->n_vec,
            "bad batch");
  // time on copies of the blocks: the caller's data stays untouched
  Scratch copies;
  std::vector<void *> xs(batch->n);
  std::vector<int64_t> lds(batch->n);
  {
    size_t tot = 0;
    for (int e = 0; e < batch->n; ++e) tot += size_t(op->M) * batch->n_vec[e];
    CHFSI_TRY(copies.ensure(sizeof(double) * 2 * tot));
    size_t off = 0;
    for (int e = 0; e < batch->n; ++e) {
      double *dst = static_cast<double *>(copies.ptr) + 2 * off;
      CHFSI_ARG(batch->X[e] && batch->ld[e] >= batch->n_vec[e], "bad batch entry %d", e);
      CHFSI_HIP(launch_copy_block<double, double>(static_cast<const double *>(batch->X[e]),
                                                  2 * batch->ld[e], dst, 2 * batch->n_vec[e],
                                                  op->M, 2 * batch->n_vec[e], st));
      xs[e] = dst;
      lds[e] = batch->n_vec[e];
      off += size_t(op->M) * batch->n_vec[e];
    }
  }
  hipEvent_t e0, e1;
  CHFSI_HIP(hipEventCreate(&e0));
  CHFSI_HIP(hipEventCreate(&e1));
  chfsi_status status = CHFSI_OK;
  for (int prec : {CHFSI_FP64, CHFSI_FP32}) {
    chfsi_kbatch kb = *batch;
    kb.X = xs.data();
    kb.ld = lds.data();
    kb.precision = prec;
    kb.degree = 2;
    std::unique_ptr<chfsi_batch> bt(new chfsi_batch());
    status = batch_init(op, &kb, bt.get());
    if (status == CHFSI_OK) status = batch_tables(bt.get(), st);
    if (status != CHFSI_OK) break;
    const bool fp64 = prec == CHFSI_FP64;
    int &var = fp64 ? op->zhx_variant64 : op->zhx_variant32;
    const int saved = var;
    float best = 1e30f;
    int bestv = saved;
    for (int v = 0; v < zhx_variants(fp64) && status == CHFSI_OK; ++v) {
      var = v;
      status = batch_launch(bt.get(), st);  // warm-up
      if (status != CHFSI_OK) break;
      if (hipEventRecord(e0, st) != hipSuccess) status = CHFSI_ERR_HIP;
      for (int r = 0; r < 3 && status == CHFSI_OK; ++r) status = batch_launch(bt.get(), st);
      if (status != CHFSI_OK) break;
      float ms = 0;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        status = CHFSI_ERR_HIP;
        break;
      }
      ms /= 3.0f;
      if (times_ms) times_ms[(fp64 ? 0 : 8) + v] = ms;
      if (ms < best) {
        best = ms;
        bestv = v;
      }
    }
    var = status == CHFSI_OK ? bestv : saved;
    if (fp64 && best64) *best64 = var;
    if (!fp64 && best32) *best32 = var;
    if (status != CHFSI_OK) break;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamSynchronize(st);
  return status;
}

// ---------------------------------------------------------------------------
// Complex dense steps
// ---------------------------------------------------------------------------
namespace {

using zc = rocblas_double_complex;

template <typename T>
T *sp(Scratch &s) {
  return static_cast<T *>(s.ptr);
}

// C (m x n, column-major complex, ldc) = A B^H with A m x K (lda), B n x K (ldb),
// K split into chunks of one strided-batched ZGEMM, partials summed as reals.
chfsi_status zgemm_nc_splitk(chfsi_solver *s, const zc *A, int64_t lda, const zc *B, int64_t ldb,
                             int m, int n, int64_t K, zc *C, int64_t ldc, hipStream_t st) {
  const int nk = (int)std::max<int64_t>(1, std::min<int64_t>(64, K / 8192));
  const int64_t Kc = K / nk;
  const int64_t rem = K - Kc * nk;
  const int slots = nk + (rem ? 1 : 0);
  CHFSI_TRY(s->split.ensure(sizeof(zc) * slots * m * n));
  zc *P = sp<zc>(s->split);
  const zc one{1.0, 0.0}, zero{0.0, 0.0};
  CHFSI_RB(rocblas_zgemm_strided_batched(
      s->blas, rocblas_operation_none, rocblas_operation_conjugate_transpose, m, n,
      (rocblas_int)Kc, &one, A, (rocblas_int)lda, Kc * lda, B, (rocblas_int)ldb, Kc * ldb, &zero,
      P, m, (int64_t)m * n, nk));
  if (rem)
    CHFSI_RB(rocblas_zgemm(s->blas, rocblas_operation_none, rocblas_operation_conjugate_transpose,
                           m, n, (rocblas_int)rem, &one, A + nk * Kc * lda, (rocblas_int)lda,
                           B + nk * Kc * ldb, (rocblas_int)ldb, &zero, P + (int64_t)nk * m * n,
                           m));
  CHFSI_HIP(launch_sum_partials(reinterpret_cast<const double *>(P), slots, 2 * m, n,
                                reinterpret_cast<double *>(C), 2 * ldc, st));
  return CHFSI_OK;
}

// Xc <- op(Q) Xc with Xc = X^T (N x M complex, ld N), through a scratch panel.
chfsi_status zrotate(chfsi_solver *s, zc *X, int64_t M, int N, const zc *Q,
                     rocblas_operation opQ, hipStream_t st) {
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(M, (int64_t(1) << 26) / N));
  CHFSI_TRY(s->rot.ensure(sizeof(zc) * chunk * N));
  zc *tmp = sp<zc>(s->rot);
  const zc one{1.0, 0.0}, zero{0.0, 0.0};
  for (int64_t n0 = 0; n0 < M; n0 += chunk) {
    const int64_t nc = std::min(chunk, M - n0);
    CHFSI_RB(rocblas_zgemm(s->blas, opQ, rocblas_operation_none, N, (rocblas_int)nc, N, &one, Q,
                           N, X + n0 * N, N, &zero, tmp, N));
    CHFSI_HIP(hipMemcpyAsync(X + n0 * N, tmp, sizeof(zc) * nc * N, hipMemcpyDeviceToDevice, st));
  }
  return CHFSI_OK;
}

// X^H X = L L^H.  Column-major, G = Xc Xc^H = conj(X^H X) = L' L'^H with
// L' = conj(L), and X L^{-H} in the Xc view is L'^{-1} Xc.
chfsi_status zcholgs_impl(chfsi_solver *s, zc *X, int64_t M, int N, hipStream_t st) {
  CHFSI_TRY(ensure_blas(s, st));
  CHFSI_TRY(s->Hp.ensure(sizeof(zc) * N * N));
  CHFSI_TRY(s->info.ensure(sizeof(rocblas_int)));
  zc *S = sp<zc>(s->Hp);
  CHFSI_TRY(zgemm_nc_splitk(s, X, N, X, N, N, N, M, S, N, st));
  rocblas_int *dinfo = sp<rocblas_int>(s->info);
  CHFSI_RB(rocsolver_zpotrf(s->blas, rocblas_fill_lower, N, S, N, dinfo));
  rocblas_int hinfo = 0;
  CHFSI_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) {
    set_error("overlap matrix is not positive definite (zpotrf info %d)", (int)hinfo);
    return CHFSI_ERR_NOT_SPD;
  }
  CHFSI_RB(rocsolver_ztrtri(s->blas, rocblas_fill_lower, rocblas_diagonal_non_unit, N, S, N,
                            dinfo));
  CHFSI_HIP(launch_zero_upper_complex(reinterpret_cast<double *>(S), N, N, st));
  return zrotate(s, X, M, N, S, rocblas_operation_none, st);
}

// HX of a complex block at the active k-point (S H S, constrained rows zeroed).
chfsi_status zapply(chfsi_operator *op, const double *x, int64_t ldx, double *y, int64_t ldy,
                    int Bc, hipStream_t st) {
  CHFSI_TRY(zhx<double>(op, op->active_k, x, ldx, y, ldy, Bc, 1.0, 0.0, 0.0, 1, st));
  if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, y, 2 * ldy, 2 * Bc, st));
  return CHFSI_OK;
}

// Lower triangle of G = Xc HXc^H = conj(X^H H~ X) by column blocks of B.
chfsi_status zxthx_lower(chfsi_solver *s, chfsi_operator *op, const zc *X, int64_t M, int N,
                         int B, zc *Hp, hipStream_t st) {
  CHFSI_TRY(s->HXb.ensure(sizeof(zc) * M * B));
  zc *HX = sp<zc>(s->HXb);
  for (int j = 0; j < N; j += B) {
    const int bw = std::min(B, N - j);
    CHFSI_TRY(zapply(op, reinterpret_cast<const double *>(X + j), N,
                     reinterpret_cast<double *>(HX), bw, bw, st));
    CHFSI_TRY(zgemm_nc_splitk(s, X + j, N, HX, bw, N - j, bw, M, Hp + j + (int64_t)j * N, N, st));
  }
  return CHFSI_OK;
}

// Hermitian eigendecomposition in place (lower triangle used): A <- Q,
// ascending eigenvalues to the host.
chfsi_status zeig(chfsi_solver *s, zc *A, int N, double *eig_host, hipStream_t st) {
  CHFSI_TRY(s->evals.ensure(sizeof(double) * 2 * N));
  CHFSI_TRY(s->info.ensure(sizeof(rocblas_int)));
  double *D = sp<double>(s->evals), *E = D + N;
  rocblas_int *dinfo = sp<rocblas_int>(s->info);
  CHFSI_RB(rocsolver_zheevd(s->blas, rocblas_evect_original, rocblas_fill_lower, N, A, N, D, E,
                            dinfo));
  rocblas_int hinfo = 0;
  CHFSI_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipMemcpyAsync(eig_host, D, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) {
    set_error("projected eigensolve failed (zheevd info %d)", (int)hinfo);
    return CHFSI_ERR_EIG;
  }
  return CHFSI_OK;
}

// Hp = X^H H~ X by column blocks; G = Xc HXc^H = conj(Hp) = Q' D Q'^H with
// Q' = conj(Q); X Q in the Xc view is Q'^H Xc.
chfsi_status zrr_impl(chfsi_solver *s, chfsi_operator *op, zc *X, int64_t M, int N, int B,
                      double *eig_host, hipStream_t st) {
  CHFSI_TRY(ensure_blas(s, st));
  CHFSI_TRY(s->Hp.ensure(sizeof(zc) * N * N));
  zc *Hp = sp<zc>(s->Hp);
  CHFSI_TRY(zxthx_lower(s, op, X, M, N, B, Hp, st));
  CHFSI_TRY(zeig(s, Hp, N, eig_host, st));
  return zrotate(s, X, M, N, Hp, rocblas_operation_conjugate_transpose, st);
}

// Generalised Rayleigh-Ritz without CholGS (rayleighRitzGEP,
// rayleighRitzDevice.cc:355-819, complex).  In the Xc view the overlap is
// Xc Xc^H = conj(X^H X) = L' L'^H and the projected Hamiltonian G = conj(Hp);
// Ghat = L'^{-1} G L'^{-H} = Q' D Q'^H, and X <- X L^{-H} Q is
// Xc <- (L'^{-H} Q')^H Xc.
chfsi_status zgep_impl(chfsi_solver *s, chfsi_operator *op, zc *X, int64_t M, int N, int B,
                       double *eig_host, hipStream_t st) {
  CHFSI_TRY(ensure_blas(s, st));
  CHFSI_TRY(s->Hp.ensure(sizeof(zc) * N * N));
  CHFSI_TRY(s->gep.ensure(sizeof(zc) * 2 * N * N));
  CHFSI_TRY(s->info.ensure(sizeof(rocblas_int)));
  zc *G = sp<zc>(s->Hp), *Li = sp<zc>(s->gep), *W = Li + (int64_t)N * N;
  CHFSI_TRY(zgemm_nc_splitk(s, X, N, X, N, N, N, M, Li, N, st));
  rocblas_int *dinfo = sp<rocblas_int>(s->info);
  CHFSI_RB(rocsolver_zpotrf(s->blas, rocblas_fill_lower, N, Li, N, dinfo));
  rocblas_int hinfo = 0;
  CHFSI_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  CHFSI_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) {
    set_error("overlap matrix is not positive definite (zpotrf info %d)", (int)hinfo);
    return CHFSI_ERR_NOT_SPD;
  }
  CHFSI_RB(rocsolver_ztrtri(s->blas, rocblas_fill_lower, rocblas_diagonal_non_unit, N, Li, N,
                            dinfo));
  CHFSI_HIP(launch_zero_upper_complex(reinterpret_cast<double *>(Li), N, N, st));
  CHFSI_TRY(zxthx_lower(s, op, X, M, N, B, G, st));
  const zc one{1.0, 0.0}, zero{0.0, 0.0};
  CHFSI_RB(rocblas_zhemm(s->blas, rocblas_side_right, rocblas_fill_lower, N, N, &one, G, N, Li,
                         N, &zero, W, N));
  CHFSI_RB(rocblas_ztrmm(s->blas, rocblas_side_right, rocblas_fill_lower,
                         rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit, N, N,
                         &one, Li, N, W, N, G, N));
  CHFSI_TRY(zeig(s, G, N, eig_host, st));
  CHFSI_RB(rocblas_ztrmm(s->blas, rocblas_side_left, rocblas_fill_lower,
                         rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit, N, N,
                         &one, Li, N, G, N, W, N));
  return zrotate(s, X, M, N, W, rocblas_operation_conjugate_transpose, st);
}

// |H~ x_j - lambda_j x_j| over 2N real columns with each eigenvalue repeated
chfsi_status zresidual_impl(chfsi_solver *s, chfsi_operator *op, const zc *X, int64_t M, int N,
                            int B, const double *eig_host, double *res_host, hipStream_t st) {
  CHFSI_TRY(s->HXb.ensure(sizeof(zc) * M * B));
  CHFSI_TRY(s->lam.ensure(sizeof(double) * 2 * N));
  const int nch = 512;
  CHFSI_TRY(s->part.ensure(sizeof(double) * nch * 2 * B));
  double *HX = sp<double>(s->HXb), *lam = sp<double>(s->lam), *part = sp<double>(s->part);
  std::vector<double> lam2(2 * size_t(N));
  for (int v = 0; v < N; ++v) lam2[2 * v] = lam2[2 * v + 1] = eig_host[v];
  CHFSI_HIP(hipMemcpyAsync(lam, lam2.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice, st));
  std::vector<double> h(size_t(nch) * 2 * B);
  const double *Xr = reinterpret_cast<const double *>(X);
  for (int j = 0; j < N; j += B) {
    const int bw = std::min(B, N - j);
    CHFSI_TRY(zapply(op, Xr + 2 * j, N, HX, bw, bw, st));
    CHFSI_HIP(launch_residual_partial(HX, 2 * bw, Xr + 2 * j, 2 * (int64_t)N, lam + 2 * j, M,
                                      2 * bw, part, nch, st));
    CHFSI_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nch * 2 * bw,
                             hipMemcpyDeviceToHost, st));
    CHFSI_HIP(hipStreamSynchronize(st));
    for (int v = 0; v < bw; ++v) {
      double acc = 0;
      for (int k = 0; k < nch; ++k) acc += h[size_t(k) * 2 * bw + 2 * v] + h[size_t(k) * 2 * bw + 2 * v + 1];
      res_host[j + v] = std::sqrt(acc);
    }
  }
  return CHFSI_OK;
}

// Lanczos on the hermitian H~ (complex vectors as 2M reals: Re v^H f and |f|^2
// are real dot products of the interleaved arrays).
chfsi_status zlanczos_impl(chfsi_solver *s, chfsi_operator *op, const double *v0c, int steps,
                           int reproducible, double *lower, double *upper, hipStream_t st) {
  const int64_t M = op->M, M2 = 2 * M;
  CHFSI_ARG(steps >= 1, "lanczos steps must be >= 1");
  CHFSI_TRY(s->HXb.ensure(sizeof(double) * M2 * 3));
  double *v = sp<double>(s->HXb), *f = v + M2, *vold = v + 2 * M2;
  CHFSI_HIP(hipMemsetAsync(v, 0, sizeof(double) * M2 * 3, st));
  CHFSI_HIP(hipMemcpyAsync(v, v0c, sizeof(double) * M2, hipMemcpyDeviceToDevice, st));
  if (op->nc) CHFSI_HIP(launch_zero_rows(op->d_c_rows, op->nc, v, 2, 2, st));
  double nrm2;
  CHFSI_TRY(dot(s, v, v, M2, &nrm2, st));
  CHFSI_ARG(nrm2 > 0, "lanczos start vector is zero");
  CHFSI_HIP(launch_axpby<double>(M2, 0.0, v, 1.0 / std::sqrt(nrm2), v, st));
  std::vector<double> T(size_t(steps) * steps, 0.0);
  CHFSI_TRY(zapply(op, v, 1, f, 1, 1, st));
  double alpha;
  CHFSI_TRY(dot(s, f, v, M2, &alpha, st));
  CHFSI_HIP(launch_axpby<double>(M2, -alpha, v, 1.0, f, st));
  T[0] = alpha;
  for (int j = 1; j < steps; ++j) {
    double fn2;
    CHFSI_TRY(dot(s, f, f, M2, &fn2, st));
    const double beta = std::sqrt(fn2);
    std::swap(v, vold);
    CHFSI_HIP(launch_axpby<double>(M2, 1.0 / beta, f, 0.0, v, st));
    CHFSI_TRY(zapply(op, v, 1, f, 1, 1, st));
    CHFSI_HIP(launch_axpby<double>(M2, -beta, vold, 1.0, f, st));
    CHFSI_TRY(dot(s, f, v, M2, &alpha, st));
    CHFSI_HIP(launch_axpby<double>(M2, -alpha, v, 1.0, f, st));
    T[size_t(j) * steps + j - 1] = beta;
    T[size_t(j - 1) * steps + j] = beta;
    T[size_t(j) * steps + j] = alpha;
  }
  std::vector<double> theta = small_sym_eigvals(T, steps);
  double fn2;
  CHFSI_TRY(dot(s, f, f, M2, &fn2, st));
  const double fnorm = std::sqrt(fn2);
  *lower = std::floor(theta.front());
  *upper = std::ceil(theta.back() + (reproducible ? fnorm : fnorm / 10.0));
  return CHFSI_OK;
}

}  // namespace

extern "C" chfsi_status chfsi_lanczos_bounds_complex(chfsi_operator *op, const void *v0,
                                                     int32_t steps, int32_t reproducible,
                                                     double *lower, double *upper, void *stream) {
  CHFSI_ARG(op && v0 && lower && upper, "null argument");
  CHFSI_TRY(check_complex(op));
  chfsi_solver tmp;
  return zlanczos_impl(&tmp, op, static_cast<const double *>(v0), steps, reproducible, lower,
                       upper, as_stream(stream));
}

extern "C" chfsi_status chfsi_cholgs_complex(chfsi_operator *op, void *X, int64_t M, int32_t N,
                                             void *stream) {
  CHFSI_ARG(op && X && M == op->M && N > 0, "bad argument");
  chfsi_solver tmp;
  return zcholgs_impl(&tmp, static_cast<zc *>(X), M, N, as_stream(stream));
}

extern "C" chfsi_status chfsi_rayleigh_ritz_complex(chfsi_operator *op, void *X, int64_t M,
                                                    int32_t N, int32_t block, double *eig_host,
                                                    void *stream) {
  CHFSI_ARG(op && X && eig_host && M == op->M && N > 0 && block > 0, "bad argument");
  CHFSI_TRY(check_complex(op));
  chfsi_solver tmp;
  return zrr_impl(&tmp, op, static_cast<zc *>(X), M, N, std::min<int>(block, N), eig_host,
                  as_stream(stream));
}

extern "C" chfsi_status chfsi_residual_norms_complex(chfsi_operator *op, const void *X, int64_t M,
                                                     int32_t N, int32_t block,
                                                     const double *eig_host, double *res_host,
                                                     void *stream) {
  CHFSI_ARG(op && X && eig_host && res_host && M == op->M && N > 0 && block > 0, "bad argument");
  CHFSI_TRY(check_complex(op));
  chfsi_solver tmp;
  return zresidual_impl(&tmp, op, static_cast<const zc *>(X), M, N, std::min<int>(block, N),
                        eig_host, res_host, as_stream(stream));
}

extern "C" chfsi_status chfsi_solve_complex(chfsi_solver *s, chfsi_operator *op, void *Xv,
                                            int64_t M, int32_t N, double *eig_host,
                                            double *res_host, int32_t is_first_filtering_call,
                                            int32_t compute_residual,
                                            int32_t use_mixed_prec_overall, int32_t is_first_scf,
                                            double *bUp_out, double *timings_ms, void *stream) {
  CHFSI_ARG(s && op && Xv && eig_host, "null argument");
  CHFSI_TRY(check_complex(op));
  CHFSI_ARG(M == op->M && N > 0, "X must be n_dofs x N");
  CHFSI_ARG(!compute_residual || res_host, "res_host required when computing residuals");
  hipStream_t st = as_stream(stream);
  const chfsi_params &P = s->params;
  zc *X = static_cast<zc *>(Xv);
  double *Xr = static_cast<double *>(Xv);
  hipEvent_t ev[6];
  for (auto &e : ev) CHFSI_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t *e;
    ~EvGuard() {
      for (int i = 0; i < 6; ++i) hipEventDestroy(e[i]);
    }
  } guard{ev};
  CHFSI_HIP(hipEventRecord(ev[0], st));

  if (is_first_filtering_call || !P.reuse_lanczos_upper) {
    // real start vector placed in the real parts (the reference draws rand() values)
    Scratch gen;
    CHFSI_TRY(gen.ensure(sizeof(double) * 2 * M));
    if (s->have_v0) {
      CHFSI_HIP(launch_fill_zero<double>(sp<double>(gen), 2 * M, st));
      CHFSI_HIP(launch_copy_block<double, double>(sp<double>(s->v0), 1, sp<double>(gen), 2, M, 1,
                                                  st));
    } else {
      const std::vector<double> hv = default_lanczos_start(M);
      std::vector<double> h2(2 * size_t(M), 0.0);
      for (int64_t i = 0; i < M; ++i) h2[2 * i] = hv[i];
      CHFSI_HIP(hipStreamSynchronize(st));
      CHFSI_HIP(hipMemcpy(gen.ptr, h2.data(), sizeof(double) * 2 * M, hipMemcpyHostToDevice));
    }
    double lo, up;
    chfsi_status stl = zlanczos_impl(s, op, sp<double>(gen), P.lanczos_steps,
                                     P.reproducible || P.lanczos_full_residual_margin, &lo, &up,
                                     st);
    CHFSI_HIP(hipStreamSynchronize(st));
    gen.release();
    CHFSI_TRY(stl);
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

  CHFSI_HIP(launch_scale_rows<double>(Xr, 2 * (int64_t)N, M, 2 * N, 1.0, op->d_sqrt_mass, st));
  const int B = std::min<int>(P.cheby_block_size > 0 ? P.cheby_block_size : N, N);
  int prec = use_mixed_prec_overall ? P.filter_precision : CHFSI_FP64;
  if (prec == CHFSI_BF16) prec = CHFSI_FP32;  // no BF16 complex form
  {
    // all column blocks of this k-point filtered as one batch
    const int nb = (N + B - 1) / B;
    CHFSI_ARG(nb <= kMaxBatch, "at most %d filter blocks per pass", kMaxBatch);
    std::vector<int32_t> kk(nb, op->active_k), nv(nb);
    std::vector<void *> xs(nb);
    std::vector<int64_t> lds(nb, N);
    std::vector<double> av(nb, s->bLow), bv(nb, s->bUp), a0v(nb, s->a0);
    for (int i = 0; i < nb; ++i) {
      xs[i] = X + (int64_t)i * B;
      nv[i] = std::min(B, N - i * B);
    }
    chfsi_kbatch kb{nb, kk.data(), xs.data(), lds.data(), nv.data(), av.data(), bv.data(),
                    a0v.data(), (int32_t)m, prec};
    CHFSI_TRY(chfsi_batch_filter(op, &kb, stream));
  }
  CHFSI_HIP(hipEventRecord(ev[2], st));
  if (P.use_shep) {
    CHFSI_TRY(zcholgs_impl(s, X, M, N, st));
    CHFSI_HIP(hipEventRecord(ev[3], st));
    CHFSI_TRY(zrr_impl(s, op, X, M, N, B, eig_host, st));
  } else {
    CHFSI_HIP(hipEventRecord(ev[3], st));
    CHFSI_TRY(zgep_impl(s, op, X, M, N, B, eig_host, st));
  }
  CHFSI_HIP(hipEventRecord(ev[4], st));
  if (compute_residual) CHFSI_TRY(zresidual_impl(s, op, X, M, N, B, eig_host, res_host, st));
  CHFSI_HIP(launch_scale_rows<double>(Xr, 2 * (int64_t)N, M, 2 * N, 1.0, op->d_inv_sqrt_mass,
                                      st));
  CHFSI_HIP(hipEventRecord(ev[5], st));
  CHFSI_HIP(hipEventSynchronize(ev[5]));
  if (timings_ms) {
    float t;
    for (int k = 0; k < 4; ++k) {
      CHFSI_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      timings_ms[k] = t;
    }
    CHFSI_HIP(hipEventElapsedTime(&t, ev[4], ev[5]));
    timings_ms[4] = t;
    CHFSI_HIP(hipEventElapsedTime(&t, ev[0], ev[5]));
    timings_ms[5] = t;
  }
  if (bUp_out) *bUp_out = s->bUp;
  return CHFSI_OK;
}
