// Drives the C++ drop-in classes (include/dftfe_chfsi.hpp) the way DFT-FE's
// kohnShamEigenSpaceCompute drives its device solver (src/dft/kohnShamEigenSolve.cc:505-655),
// on a problem written by the GPU tests as raw little-endian arrays in one
// directory:
//   meta.txt   "key value" lines: M C ne nq n_proj n_pairs N B m a0 bLow bUp n_split
//   map.i64 color.i8 sqrt_mass.f64 inv_sqrt_mass.f64 Nq.f64 K.f64 Gq.f64 jxw.f64
//   pair_cell.i32 pair_proj0.i32 pair_C.f64 proj_V.f64 vJxW.f64 gJxW.f64 X0.f64
// Writes into the same directory:
//   hx.f64          H~ X0 through kohnShamDFTOperatorDeviceClass::HX
//   eig1.f64 res1.f64 X1.f64   first solve (full Rayleigh-Ritz)
//   eig2.f64 res2.f64          second solve after the caller's bound update
//   Xnr.f64                    solveNoRR, one pass, from X0
//   eig_mixed_default.f64      solve from X0 with useMixedPrecOverall and default parameters
//   eig_gep.f64 res_gep.f64    solve from X0 on the generalised Rayleigh-Ritz route
//   eig_split.f64 Xfrac.f64    spectrum-split solve from X0 (top n_split states)
//   hx_gga.f64                 H~ X0 after a GGA rebuild
// Exit status 0 on success; any library error ends the program with a message.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "dftfe_chfsi.hpp"

namespace {

std::string g_dir;

template <typename T>
std::vector<T> load(const char *name, size_t n) {
  std::vector<T> v(n);
  const std::string path = g_dir + "/" + name;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  const size_t got = n ? std::fread(v.data(), sizeof(T), n, f) : 0;
  std::fclose(f);
  if (got != n) throw std::runtime_error("short read: " + path);
  return v;
}

template <typename T>
void save(const char *name, const T *p, size_t n) {
  const std::string path = g_dir + "/" + name;
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(p, sizeof(T), n, f);
  std::fclose(f);
}

std::vector<double> toHost(const double *d, size_t n) {
  std::vector<double> h(n);
  dftfe::internal::hipCheck(hipMemcpy(h.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost),
                            "hipMemcpy");
  return h;
}

void toDevice(double *d, const std::vector<double> &h) {
  dftfe::internal::hipCheck(
      hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice), "hipMemcpy");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <problem directory>\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  try {
    std::map<std::string, double> meta;
    {
      std::ifstream in(g_dir + "/meta.txt");
      std::string k;
      double v;
      while (in >> k >> v) meta[k] = v;
    }
    const int64_t M = (int64_t)meta.at("M");
    const int C = (int)meta.at("C"), ne = (int)meta.at("ne"), nq = (int)meta.at("nq");
    const int n_proj = (int)meta.at("n_proj"), n_pairs = (int)meta.at("n_pairs");
    const unsigned N = (unsigned)meta.at("N"), B = (unsigned)meta.at("B");
    const unsigned m = (unsigned)meta.at("m"), n_split = (unsigned)meta.at("n_split");

    auto map = load<int64_t>("map.i64", size_t(C) * ne);
    auto color = load<int8_t>("color.i8", C);
    auto sq = load<double>("sqrt_mass.f64", M);
    auto isq = load<double>("inv_sqrt_mass.f64", M);
    auto Nq = load<double>("Nq.f64", size_t(nq) * ne);
    auto K = load<double>("K.f64", size_t(ne) * ne);
    auto Gq = load<double>("Gq.f64", size_t(nq) * ne * 3);
    auto jxw = load<double>("jxw.f64", size_t(C) * nq);
    auto pc = load<int32_t>("pair_cell.i32", n_pairs);
    auto pp = load<int32_t>("pair_proj0.i32", n_pairs);
    auto pC = load<double>("pair_C.f64", size_t(n_pairs) * ne * 4);
    auto pV = load<double>("proj_V.f64", n_proj);

    chfsi_operator_desc d{};
    d.n_dofs = M;
    d.n_cells = C;
    d.nodes_per_cell = ne;
    d.quad_per_cell = nq;
    d.cell_map = map.data();
    d.cell_color = color.data();
    d.sqrt_mass = sq.data();
    d.inv_sqrt_mass = isq.data();
    d.shape_values = Nq.data();
    d.stiffness = K.data();
    d.shape_gradients = Gq.data();
    d.jxw = jxw.data();
    d.n_proj = n_proj;
    d.n_pairs = n_pairs;
    d.proj_per_pair = 4;
    d.pair_cell = pc.data();
    d.pair_proj0 = pp.data();
    d.pair_C = pC.data();
    d.proj_V = pV.data();

    // operator: creation (init / computeMassVector), V_eff, cell matrices
    dftfe::kohnShamDFTOperatorDeviceClass op(d);
    op.init();
    op.reinit(B, true);
    op.reinitkPointSpinIndex(0, 0);
    dftfe::distributedDeviceVec<double> vEff(C, nq), gEff(C, nq * 3);
    toDevice(vEff.begin(), load<double>("vJxW.f64", size_t(C) * nq));
    toDevice(gEff.begin(), load<double>("gJxW.f64", size_t(C) * nq * 3));
    op.setVEffJxW(vEff.begin());
    op.computeHamiltonianMatricesAllkpt(0);

    const auto X0 = load<double>("X0.f64", size_t(M) * N);
    dftfe::distributedDeviceVec<double> X(M, N), Y(M, N), projKet;
    toDevice(X.begin(), X0);
    Y.setValue(0.0);
    op.HX(X, projKet, (unsigned)M, N, false, 1.0, Y);
    save("hx.f64", toHost(Y.begin(), size_t(M) * N).data(), size_t(M) * N);

    // solver with the caller's bounds and parameters
    dftfe::dftParameters prm;
    prm.chebyWfcBlockSize = B;
    prm.wfcBlockSize = B;
    prm.chebyshevOrder = m;
    prm.reuseLanczosUpperBoundFromFirstCall = true;
    prm.isPseudopotential = false;
    MPI_Comm comm = 0;
    dftfe::elpaScalaManager elpa;
    dftfe::utils::DeviceCCLWrapper ccl;
    std::shared_ptr<dftfe::linearAlgebra::BLASWrapper<dftfe::utils::MemorySpace::DEVICE>> blas;
    dftfe::chebyshevOrthogonalizedSubspaceIterationSolverDevice solver(
        comm, comm, meta.at("a0"), meta.at("bLow"), meta.at("bUp"), prm);

    std::vector<double> eig, res;
    const double bUp = solver.solve(op, blas, elpa, X.begin(), nullptr, unsigned(M * N), N, eig,
                                    res, ccl, comm, false, true);
    save("eig1.f64", eig.data(), eig.size());
    save("res1.f64", res.data(), res.size());
    save("X1.f64", toHost(X.begin(), size_t(M) * N).data(), size_t(M) * N);

    // kohnShamEigenSolve.cc:483-498: a0 = lowest, bLow = highest Ritz value
    solver.reinitSpectrumBounds(eig.front(), eig.back(), bUp);
    solver.solve(op, blas, elpa, X.begin(), nullptr, unsigned(M * N), N, eig, res, ccl, comm,
                 false, true);
    save("eig2.f64", eig.data(), eig.size());
    save("res2.f64", res.data(), res.size());

    // solveNoRR from X0 with the first bounds
    solver.reinitSpectrumBounds(meta.at("a0"), meta.at("bLow"), meta.at("bUp"));
    toDevice(X.begin(), X0);
    std::vector<double> unused;
    solver.solveNoRR(op, blas, elpa, X.begin(), unsigned(M * N), N, unused, ccl, comm, 1, false);
    save("Xnr.f64", toHost(X.begin(), size_t(M) * N).data(), size_t(M) * N);

    // useMixedPrecOverall with the default parameters (USE MIXED PREC CHEBY
    // false): the filter stays FP64, as in the reference
    solver.reinitSpectrumBounds(meta.at("a0"), meta.at("bLow"), meta.at("bUp"));
    toDevice(X.begin(), X0);
    solver.solve(op, blas, elpa, X.begin(), nullptr, unsigned(M * N), N, eig, res, ccl, comm,
                 false, true, /*useMixedPrecOverall=*/true);
    save("eig_mixed_default.f64", eig.data(), eig.size());

    // SUBSPACE PROJ SHEP GPU false: generalised Rayleigh-Ritz without CholGS
    {
      dftfe::dftParameters prm_gep = prm;
      prm_gep.useSubspaceProjectedSHEPGPU = false;
      dftfe::chebyshevOrthogonalizedSubspaceIterationSolverDevice solver_gep(
          comm, comm, meta.at("a0"), meta.at("bLow"), meta.at("bUp"), prm_gep);
      toDevice(X.begin(), X0);
      solver_gep.solve(op, blas, elpa, X.begin(), nullptr, unsigned(M * N), N, eig, res, ccl,
                       comm, false, true);
      save("eig_gep.f64", eig.data(), eig.size());
      save("res_gep.f64", res.data(), res.size());
    }

    // spectrum split from X0: eigenValues sized to the kept states
    toDevice(X.begin(), X0);
    dftfe::distributedDeviceVec<double> Xfrac(M, n_split);
    std::vector<double> eig_s(n_split), res_s;
    solver.solve(op, blas, elpa, X.begin(), Xfrac.begin(), unsigned(M * N), N, eig_s, res_s, ccl,
                 comm, false, true);
    save("eig_split.f64", eig_s.data(), eig_s.size());
    save("Xfrac.f64", toHost(Xfrac.begin(), size_t(M) * n_split).data(), size_t(M) * n_split);

    // GGA rebuild, then H~ X0 again
    op.setVEffJxW(vEff.begin(), gEff.begin());
    op.computeHamiltonianMatricesAllkpt(0);
    toDevice(X.begin(), X0);
    Y.setValue(0.0);
    op.HX(X, projKet, (unsigned)M, N, false, 1.0, Y);
    save("hx_gga.f64", toHost(Y.begin(), size_t(M) * N).data(), size_t(M) * N);
    dftfe::internal::hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
  } catch (const std::exception &e) {
    std::fprintf(stderr, "dropin_solve: %s\n", e.what());
    return 1;
  }
  std::printf("dropin_solve: done\n");
  return 0;
}
