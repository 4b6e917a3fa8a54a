// Complex (Bloch k-point) form of the fused HX: kernels and launchers.
//
// Complex multivectors are stored as std::complex pairs, node-major, so a
// block of B complex columns is a block of 2B real columns with (re, im)
// adjacent.  For a complex cell matrix H = A + iG the product is
//     Y = A X + G (J X),   (J X)[:, 2v] = -X[:, 2v+1],  (J X)[:, 2v+1] = X[:, 2v],
// i.e. one real GEMM over a K dimension of 2 n_e whose second half reads the
// same gathered block with the columns of each pair swapped and the real
// column negated.  The cell matrices are stored as K-tile pairs
// (Re slice t, Im slice t), so every gathered slice of X serves two tiles.
// The nonlocal term C V C^H uses the same pairing over the projector slots,
// and the projection C^H X the conjugate pairing (J' = -J).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace chfsi {

constexpr int kMaxBatch = 64;  // batch entries (k-points) per launch

// Per batch entry (one k-point block) of a complex launch.  Sizes in reals.
template <typename T>
struct ZHxItem {
  const T *Ht;        // [cell][2*nkt][ne_pad][16]: tile 2t = Re slice t, 2t+1 = Im slice t
  const T *nlC;       // per cell at nl_off[c]*2*ne_pad: [slot tile][Re|Im][ne_pad][16]
  const T *nlCt;      // per cell at nl_off[c]*2*ne_pad: [2*nkt][slot][16] (node tiles, Re|Im)
  const T *src;
  T *dst;
  T *W;               // n_proj x ldW: V-scaled projections
  T *partial;         // n_slots x ldP: per-slot projection partials
  int64_t ld_src, ld_dst, ldW, ldP;
  T alpha, beta, gamma;
  int beta_mode;      // as HxLaunch: 0 beta*dst, 1 dst not read, 2 beta*s_out*dst
  int B;              // real columns (2 x complex columns)
};

template <typename T>
struct ZHxLaunch {
  const ZHxItem<T> *items;   // device array of nk entries
  int nk;
  int kbegin[kMaxBatch + 1];  // first work item of each entry (filled by the launcher)
  const int *cells;
  int ncells;
  int per_xcd;
  const uint32_t *map;
  int ne, ne_pad;
  const T *s_in, *s_out;     // nullable (== 1)
  const int *nl_off;         // nullable: no nonlocal part
  const int *slot_proj;
};

template <typename T>
struct ZProjLaunch {
  const ZHxItem<T> *items;
  int nk;
  int kbegin[kMaxBatch + 1];
  const int *cells;          // cells that own slots
  int ncells;
  int per_xcd;
  int max_srb;               // largest slot count of any cell / 16 (1..8)
  const uint32_t *map;
  int ne, ne_pad;
  const T *s_in;
  const int *nl_off;
  // owned-row form (null: the cell form): rows own_map[c * ne_pad + r],
  // r < own_n[c], with the items' nlCt holding the assembled values
  const uint32_t *own_map;
  const int *own_n;
};

// host copies of the items give the per-entry widths for the work split
template <typename T>
hipError_t launch_zhx(ZHxLaunch<T> a, const ZHxItem<T> *host_items, int variant, hipStream_t s);
template <typename T>
hipError_t launch_zproj(ZProjLaunch<T> a, const ZHxItem<T> *host_items, hipStream_t s);
template <typename T>
hipError_t launch_zproj_reduce(const ZHxItem<T> *items, int nk, int maxB, const int *proj_ptr,
                               const int *proj_slots, const double *proj_V, int n_proj,
                               hipStream_t s);
int zhx_variants(bool fp64);
int zhx_tile_width(bool fp64, int variant);

// Complex cell matrices of one k-point from the real tiles H0 (V and kinetic):
//   Re = H0 + |k|^2/2 Mc,  Im = -(k . G),  G_d[i][j] = int N_i dN_j/dx_d
hipError_t launch_build_hk(const double *Ht64, const double *Mc, const double *Gd, int C, int ne,
                           int ne_pad, double k0, double k1, double k2, double *Hk64, float *Hk32,
                           hipStream_t s);
template <typename T>
hipError_t launch_fill_zero(T *x, int64_t n, hipStream_t s);
// A[i + j*lda] = 0 for i < j on a complex n x n matrix
hipError_t launch_zero_upper_complex(double *A, int n, int64_t lda, hipStream_t st);

}  // namespace chfsi
