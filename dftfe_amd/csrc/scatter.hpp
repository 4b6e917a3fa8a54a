// Epilogue shared by the fused cell-GEMM kernels (real and complex forms).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mfma.hpp"

#include <type_traits>
#include <utility>

namespace chfsi {

// Launch descriptions with a defect source (R, ldR, rcoef): the real HX.
template <typename P, typename = void>
struct HasDefect : std::false_type {};
template <typename P>
struct HasDefect<P, std::void_t<decltype(std::declval<P>().R)>> : std::true_type {};

// Loads and stores through the global address space: the multivector blocks
// are device memory, and pointers read from a batch table would otherwise be
// treated as generic (flat instructions, which also count against the LDS
// wait counter).
template <typename T>
__device__ __forceinline__ T gload(const T *p, int64_t i) {
  return ((const __attribute__((address_space(1))) T *)p)[i];
}
template <typename T>
__device__ __forceinline__ void gstore(T *p, int64_t i, T v) {
  ((__attribute__((address_space(1))) T *)p)[i] = v;
}

// Epilogue of the fused HX: scaled scatter of the wave's accumulators with the
// first-touch combine (rows rb = wr + WR*j, column blocks wc*CBW + cb).
// P is the launch (or batch-entry) description providing dst, src, ld_dst,
// ld_src, alpha, beta, gamma, beta_mode and B (real columns); ne is the
// number of local rows, smap the DOF and first-touch bit of each, sout its
// output scale (zero for a constrained row: see the store below).  The destination values of a batch of row blocks (ELEMS
// elements per thread) are loaded from clamped, always-valid addresses before
// any is used, so their latencies overlap; the gamma term reads the source
// rows of first-touch rows only.  Loads are never skipped by a branch; the
// combine is a select and the stores are predicated.
template <typename T, int NRB, int WR, int RBW, int CBW, int ELEMS = (sizeof(T) == 4 ? 16 : 8),
          typename P>
__device__ __forceinline__ void hx_scatter(const P &a, int ne,
                                           const typename V4<T>::t (&acc)[RBW][CBW],
                                           const uint32_t *smap, const T *sout, int wr, int wc,
                                           int lane, int v0) {
  // row blocks per batch: ELEMS elements per thread
  constexpr int JBMAX = ELEMS / (4 * CBW) > 0 ? ELEMS / (4 * CBW) : 1;
  constexpr int JB = RBW < JBMAX ? RBW : JBMAX;
  const bool any_src = a.gamma != T(0);
  bool any_r = false;
  const T *Rp = nullptr;
  int64_t ldR = 0;
  T rc[CBW];
  if constexpr (HasDefect<P>::value && sizeof(T) == 4) {
    any_r = a.R != nullptr;
    Rp = any_r ? a.R : a.dst;
    ldR = a.ldR;
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int v = v0 + (wc * CBW + cb) * 16 + (lane & 15);
      rc[cb] = any_r ? a.rcoef[v < a.B ? v : a.B - 1] : T(0);
    }
  }
#pragma unroll
  for (int j0 = 0; j0 < RBW; j0 += JB) {
    T dv[JB][CBW][4], sv[JB][CBW][4];
    uint32_t gg[JB][4];
    bool nsrc[JB][4], fst[JB][4];
#pragma unroll
    for (int jj = 0; jj < JB; ++jj) {
      const int rb = wr + WR * (j0 + jj);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + acc_row<T>(lane, r);
        const int rowc = (j0 + jj < RBW && rb < NRB && row < ne) ? row : 0;
        const uint32_t w = smap[rowc];
        gg[jj][r] = w & 0x7fffffffu;
        fst[jj][r] = (w >> 31) != 0;
        nsrc[jj][r] = any_src && fst[jj][r];
      }
    }
#pragma unroll
    for (int jj = 0; jj < JB; ++jj)
#pragma unroll
      for (int cb = 0; cb < CBW; ++cb) {
        const int v = v0 + (wc * CBW + cb) * 16 + (lane & 15);
        const int vc = v < a.B ? v : a.B - 1;
#pragma unroll
        for (int r = 0; r < 4; ++r) dv[jj][cb][r] = gload(a.dst, (int64_t)gg[jj][r] * a.ld_dst + vc);
      }
    if (any_src) {
#pragma unroll
      for (int jj = 0; jj < JB; ++jj)
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) {
          const int v = v0 + (wc * CBW + cb) * 16 + (lane & 15);
          const int vc = v < a.B ? v : a.B - 1;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // rows that need no source read element 0 of the block instead:
            // one cache line for all of them, and no branch between the loads
            const T x = gload(a.src, nsrc[jj][r] ? (int64_t)gg[jj][r] * a.ld_src + vc : 0);
            sv[jj][cb][r] = nsrc[jj][r] ? x : T(0);
          }
        }
    } else {
#pragma unroll
      for (int jj = 0; jj < JB; ++jj)
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) sv[jj][cb][r] = T(0);
    }
    if constexpr (HasDefect<P>::value && sizeof(T) == 4) {
      if (any_r) {
        // defect form: first-touch rows also take rcoef * R, folded into the
        // source term here (sv becomes gamma*src + rcoef*R); the other rows
        // read element 0 of the block, as for the source
#pragma unroll
        for (int jj = 0; jj < JB; ++jj)
#pragma unroll
          for (int cb = 0; cb < CBW; ++cb) {
            const int v = v0 + (wc * CBW + cb) * 16 + (lane & 15);
            const int vc = v < a.B ? v : a.B - 1;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const T x = gload(Rp, fst[jj][r] ? (int64_t)gg[jj][r] * ldR + vc : 0);
              sv[jj][cb][r] = fmadd(rc[cb], fst[jj][r] ? x : T(0), a.gamma * sv[jj][cb][r]);
            }
          }
      }
    }
#pragma unroll
    for (int jj = 0; jj < JB; ++jj) {
      const int j = j0 + jj;
      if (j >= RBW) continue;
      const int rb = wr + WR * j;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + acc_row<T>(lane, r);
        const bool rok = rb < NRB && row < ne;
        const int rowc = rok ? row : 0;
        const bool first = (smap[rowc] >> 31) != 0;
        const T so = sout[rowc];
        // a zero output scale marks a constrained row (hanging node): it
        // collects the plain cell sums, without the combine, for the
        // slave -> master pass that follows the colour launches
        const bool cons = so == T(0);
        const T sa = a.alpha * (cons ? T(1) : so);
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) {
          const int v = v0 + (wc * CBW + cb) * 16 + (lane & 15);
          const T d = dv[jj][cb][r];
          const T keep = a.beta_mode == 0 ? a.beta * d : (a.beta_mode == 2 ? (a.beta * so) * d : T(0));
          T base = first ? fmadd(a.gamma, sv[jj][cb][r], keep) : d;
          if constexpr (HasDefect<P>::value && sizeof(T) == 4) {
            if (any_r) base = first ? sv[jj][cb][r] + keep : d;
          }
          base = (first && cons) ? T(0) : base;
          if (rok && v < a.B)
            gstore(a.dst, (int64_t)gg[jj][r] * a.ld_dst + v, fmadd(sa, acc[j][cb][r], base));
        }
      }
    }
  }
}

}  // namespace chfsi
