// Drives the C++ drop-in classes (include/dftfe_chfsi.hpp) built with
// USE_COMPLEX, so dataTypes::number is std::complex<double> as in the
// reference's k-point build (include/dftfeDataTypes.h:38-48), the way
// kohnShamEigenSpaceCompute drives its device solver for one k-point
// (src/dft/kohnShamEigenSolve.cc:505-655).  Problem directory (raw
// little-endian arrays written by the GPU test):
//   meta.txt   "key value" lines: M C ne nq n_proj n_pairs N B m nk kidx a0 bLow bUp
//   map.i64 color.i8 sqrt_mass.f64 inv_sqrt_mass.f64 Nq.f64 K.f64 jxw.f64
//   pair_cell.i32 pair_proj0.i32 pair_C.f64 proj_V.f64 vJxW.f64
//   kpts.f64 Mc.f64 Gc.f64 pair_C_k.f64 (complex, interleaved) Xc0.c128
// Writes:
//   hx_c.c128                 H~_k X0 through kohnShamDFTOperatorDeviceClass::HX
//   eigc1.f64 resc1.f64       first solve at k-point kidx
//   eigc2.f64 resc2.f64       second solve after the caller's bound update
// Exit status 0 on success.
#ifndef USE_COMPLEX
#define USE_COMPLEX
#endif

#include <complex>
#include <cstdio>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "dftfe_chfsi.hpp"

namespace {

std::string g_dir;
using cplx = std::complex<double>;

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

template <typename T>
std::vector<T> toHost(const T *d, size_t n) {
  std::vector<T> h(n);
  dftfe::internal::hipCheck(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost),
                            "hipMemcpy");
  return h;
}

template <typename T>
void toDevice(T *d, const std::vector<T> &h) {
  dftfe::internal::hipCheck(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice),
                            "hipMemcpy");
}

}  // namespace

int main(int argc, char **argv) {
  static_assert(sizeof(dftfe::dataTypes::number) == sizeof(cplx), "complex build expected");
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
    const unsigned m = (unsigned)meta.at("m");
    const int nk = (int)meta.at("nk"), kidx = (int)meta.at("kidx");

    auto map = load<int64_t>("map.i64", size_t(C) * ne);
    auto color = load<int8_t>("color.i8", C);
    auto sq = load<double>("sqrt_mass.f64", M);
    auto isq = load<double>("inv_sqrt_mass.f64", M);
    auto Nq = load<double>("Nq.f64", size_t(nq) * ne);
    auto K = load<double>("K.f64", size_t(ne) * ne);
    auto jxw = load<double>("jxw.f64", size_t(C) * nq);
    auto pc = load<int32_t>("pair_cell.i32", n_pairs);
    auto pp = load<int32_t>("pair_proj0.i32", n_pairs);
    auto pC = load<double>("pair_C.f64", size_t(n_pairs) * ne * 4);
    auto pV = load<double>("proj_V.f64", n_proj);
    auto kp = load<double>("kpts.f64", size_t(3) * nk);
    auto Mc = load<double>("Mc.f64", size_t(ne) * ne);
    auto Gc = load<double>("Gc.f64", size_t(3) * ne * ne);
    auto pCk = load<double>("pair_C_k.f64", size_t(nk) * n_pairs * ne * 4 * 2);

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
    d.jxw = jxw.data();
    d.n_proj = n_proj;
    d.n_pairs = n_pairs;
    d.proj_per_pair = 4;
    d.pair_cell = pc.data();
    d.pair_proj0 = pp.data();
    d.pair_C = pC.data();
    d.proj_V = pV.data();
    d.n_kpoints = nk;
    d.kpoints = kp.data();
    d.cell_mass = Mc.data();
    d.cell_grad = Gc.data();
    d.pair_C_kpoints = pCk.data();

    // operator with all k-points; V_eff; cell matrices of every k-point
    dftfe::kohnShamDFTOperatorDeviceClass op(d);
    op.init();
    op.reinit(B, true);
    dftfe::distributedDeviceVec<double> vEff(C, nq);
    toDevice(vEff.begin(), load<double>("vJxW.f64", size_t(C) * nq));
    op.setVEffJxW(vEff.begin());
    op.computeHamiltonianMatricesAllkpt(0);
    op.reinitkPointSpinIndex(kidx, 0);

    const auto X0 = load<cplx>("Xc0.c128", size_t(M) * N);
    dftfe::distributedDeviceVec<dftfe::dataTypes::number> X(M, N), Y(M, N), projKet;
    toDevice(X.begin(), X0);
    Y.setValue(cplx(0.0, 0.0));
    op.HX(X, projKet, (unsigned)M, N, false, 1.0, Y);
    save("hx_c.c128", toHost(Y.begin(), size_t(M) * N).data(), size_t(M) * N);

    dftfe::dftParameters prm;
    prm.chebyWfcBlockSize = B;
    prm.wfcBlockSize = B;
    prm.chebyshevOrder = m;
    prm.reuseLanczosUpperBoundFromFirstCall = true;
    MPI_Comm comm = 0;
    dftfe::elpaScalaManager elpa;
    dftfe::utils::DeviceCCLWrapper ccl;
    std::shared_ptr<dftfe::linearAlgebra::BLASWrapper<dftfe::utils::MemorySpace::DEVICE>> blas;
    dftfe::chebyshevOrthogonalizedSubspaceIterationSolverDevice solver(
        comm, comm, meta.at("a0"), meta.at("bLow"), meta.at("bUp"), prm);

    std::vector<double> eig, res;
    const double bUp = solver.solve(op, blas, elpa, X.begin(), nullptr, unsigned(M * N), N, eig,
                                    res, ccl, comm, false, true);
    save("eigc1.f64", eig.data(), eig.size());
    save("resc1.f64", res.data(), res.size());
    solver.reinitSpectrumBounds(eig.front(), eig.back(), bUp);
    solver.solve(op, blas, elpa, X.begin(), nullptr, unsigned(M * N), N, eig, res, ccl, comm,
                 false, true);
    save("eigc2.f64", eig.data(), eig.size());
    save("resc2.f64", res.data(), res.size());

    // the forms built for real wavefunctions refuse complex ones with a message
    bool refused = false;
    try {
      std::vector<double> unused;
      solver.solveNoRR(op, blas, elpa, X.begin(), unsigned(M * N), N, unused, ccl, comm, 1,
                       false);
    } catch (const std::runtime_error &) {
      refused = true;
    }
    if (!refused) throw std::runtime_error("solveNoRR accepted complex wavefunctions");
    dftfe::internal::hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
  } catch (const std::exception &e) {
    std::fprintf(stderr, "dropin_solve_complex: %s\n", e.what());
    return 1;
  }
  std::printf("dropin_solve_complex: done\n");
  return 0;
}
