/* C++ drop-in for DFT-FE's device ChFSI interface on one rank's whole domain,
 * built without deal.II, over the libchfsi C ABI (chfsi.h).
 *
 * Same class names, method names, argument meaning and defaults as
 *   dftfe::operatorDFTDeviceClass                 include/operatorDevice.h:43-420
 *   dftfe::chebyshevOrthogonalizedSubspaceIterationSolverDevice
 *       include/chebyshevOrthogonalizedSubspaceIterationSolverDevice.h:37-166
 * with dftfe::distributedDeviceVec a node-major device multivector
 * (X[node * numberComponents + vec], as linearAlgebra::MultiVector on device).
 *
 * One rank owns the whole domain: the MPI communicators, the ELPA/ScaLAPACK
 * manager, the device CCL wrapper and the BLAS wrapper are accepted and not
 * used (the library keeps its own rocBLAS/rocSOLVER handles and runs the
 * projected eigensolve on the GPU); ghost updates are no-ops.
 *
 * Errors: a status other than CHFSI_OK from the library, or a HIP error here,
 * throws std::runtime_error with the library's message (the reference raises
 * deal.II exceptions through AssertThrow).
 *
 * Compile with any C++14 compiler that sees the HIP runtime headers (hipcc,
 * or g++ with -I/opt/rocm/include -D__HIP_PLATFORM_AMD__) and link libchfsi.so
 * and libamdhip64.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "chfsi.h"

#ifndef MPI_VERSION
/* stand-in when mpi.h is not included: the communicators are not used */
typedef int MPI_Comm;
#endif

namespace dftfe {

namespace utils {
enum class MemorySpace { HOST, DEVICE };
/* collectives over the domain communicator (DeviceCCLWrapper): none on one rank */
class DeviceCCLWrapper {};
}  // namespace utils

namespace linearAlgebra {
/* the library owns its BLAS handles */
template <utils::MemorySpace memorySpace>
class BLASWrapper {};
}  // namespace linearAlgebra

/* dataTypes::number is real, or std::complex<double> when the including code
 * defines USE_COMPLEX, as in the reference (include/dftfeDataTypes.h:38-48).
 * The complex form runs on the operator's active k-point
 * (reinitkPointSpinIndex). */
namespace dataTypes {
#ifdef USE_COMPLEX
typedef std::complex<double> number;
typedef std::complex<float> numberFP32;
#else
typedef double number;
typedef float numberFP32;
#endif
}  // namespace dataTypes

/* the projected dense eigensolve runs on the GPU (rocSOLVER) instead of ELPA */
class elpaScalaManager {};

namespace internal {
inline void chfsiCheck(chfsi_status st, const char *what) {
  if (st != CHFSI_OK)
    throw std::runtime_error(std::string(what) + ": " + chfsi_last_error());
}
inline void hipCheck(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

/* the library call of each number type: real entry points, or the complex
 * ones at the operator's active k-point (interleaved std::complex layout) */
inline chfsi_status hxCall(chfsi_operator *op, const double *x, int64_t ldx, double *y,
                           int64_t ldy, int32_t B, int32_t scaleFlag, double scalar, void *st) {
  return chfsi_hx(op, x, ldx, y, ldy, B, scaleFlag, scalar, st);
}
inline chfsi_status hxCall(chfsi_operator *op, const std::complex<double> *x, int64_t ldx,
                           std::complex<double> *y, int64_t ldy, int32_t B, int32_t scaleFlag,
                           double scalar, void *st) {
  return chfsi_hx_complex(op, x, ldx, y, ldy, B, scaleFlag, scalar, CHFSI_FP64, st);
}
inline chfsi_status solveCall(chfsi_solver *s, chfsi_operator *op, double *X, int64_t M,
                              int32_t N, double *eig, double *res, int32_t first,
                              int32_t computeResidual, int32_t mixed, int32_t firstScf,
                              double *bUp, double *timings, void *st) {
  return chfsi_solve(s, op, X, M, N, eig, res, first, computeResidual, mixed, firstScf, bUp,
                     timings, st);
}
inline chfsi_status solveCall(chfsi_solver *s, chfsi_operator *op, std::complex<double> *X,
                              int64_t M, int32_t N, double *eig, double *res, int32_t first,
                              int32_t computeResidual, int32_t mixed, int32_t firstScf,
                              double *bUp, double *timings, void *st) {
  return chfsi_solve_complex(s, op, X, M, N, eig, res, first, computeResidual, mixed, firstScf,
                             bUp, timings, st);
}
constexpr bool isComplexNumber() {
  return sizeof(dataTypes::number) == 2 * sizeof(double);
}
}  // namespace internal

/* The dftParameters fields the device eigensolve reads (include/dftParameters.h),
 * with the reference's defaults (utils/dftParameters.cc). */
struct dftParameters {
  unsigned int chebyWfcBlockSize = 400;
  unsigned int wfcBlockSize = 400;
  unsigned int chebyshevOrder = 0;
  bool reproducible_output = false;
  bool reuseLanczosUpperBoundFromFirstCall = false;
  double chebyshevFilterPolyDegreeFirstScfScalingFactor = 1.34;
  bool isPseudopotential = false;
  bool useSubspaceProjectedSHEPGPU = true;
  bool useMixedPrecCheby = false;
  bool useMixedPrecCGS_O = false;
  bool useMixedPrecCGS_SR = false;
  bool useMixedPrecSubspaceRotRR = false;
  bool useMixedPrecXTHXSpectrumSplit = false;
  int verbosity = 0;
};

/* The library parameters a dftParameters set maps to (the remaining fields,
 * which have no reference counterpart, keep chfsi_default_params values). */
inline chfsi_params toChfsiParams(const dftParameters &p) {
  chfsi_params c;
  chfsi_default_params(&c);
  c.cheby_block_size = (int32_t)p.chebyWfcBlockSize;
  c.cheby_degree = (int32_t)p.chebyshevOrder;
  c.reproducible = p.reproducible_output ? 1 : 0;
  c.lanczos_steps = p.reproducible_output ? 40 : 20;  // linearAlgebraOperationsDevice.cc:353-354
  c.reuse_lanczos_upper = p.reuseLanczosUpperBoundFromFirstCall ? 1 : 0;
  c.first_scf_degree_factor = p.chebyshevFilterPolyDegreeFirstScfScalingFactor;
  c.use_shep = p.useSubspaceProjectedSHEPGPU ? 1 : 0;
  c.is_pseudopotential = p.isPseudopotential ? 1 : 0;
  /* the filter bulk leaves FP64 only under USE MIXED PREC CHEBY, and then only
   * in passes the caller runs with useMixedPrecOverall
   * (linearAlgebraOperationsDevice.cc:627).  The library's FP32 recurrence is
   * a stronger form than the reference's, which keeps the cell products in
   * FP64 and sends only the ghost exchange in FP32
   * (kohnShamDFTOperatorDevice.cc:3895-3926). */
  c.filter_precision = p.useMixedPrecCheby ? CHFSI_FP32 : CHFSI_FP64;
  c.mixed_prec_overlap = p.useMixedPrecCGS_O ? 1 : 0;
  c.mixed_prec_cgs_rotation = p.useMixedPrecCGS_SR ? 1 : 0;
  c.mixed_prec_rr_rotation = p.useMixedPrecSubspaceRotRR ? 1 : 0;
  c.mixed_prec_xthx_spectrum_split = p.useMixedPrecXTHXSpectrumSplit ? 1 : 0;
  return c;
}

/* Node-major device multivector: locallyOwnedSize() nodes x numVectors()
 * components, no ghosts (one rank owns the domain). */
template <typename T>
class distributedDeviceVec {
 public:
  distributedDeviceVec() = default;
  distributedDeviceVec(std::int64_t locallyOwnedSize, unsigned int numberComponents) {
    reinit(locallyOwnedSize, numberComponents);
  }
  distributedDeviceVec(const distributedDeviceVec &) = delete;
  distributedDeviceVec &operator=(const distributedDeviceVec &) = delete;
  distributedDeviceVec(distributedDeviceVec &&o) noexcept { swap(o); }
  distributedDeviceVec &operator=(distributedDeviceVec &&o) noexcept {
    swap(o);
    return *this;
  }
  ~distributedDeviceVec() {
    if (d_data) (void)hipFree(d_data);
  }

  void reinit(std::int64_t locallyOwnedSize, unsigned int numberComponents) {
    if (d_data) internal::hipCheck(hipFree(d_data), "hipFree");
    d_data = nullptr;
    d_nodes = locallyOwnedSize;
    d_components = numberComponents;
    const size_t bytes = sizeof(T) * size_t(locallyOwnedSize) * numberComponents;
    if (bytes) {
      internal::hipCheck(hipMalloc(&d_data, bytes), "hipMalloc");
      internal::hipCheck(hipMemset(d_data, 0, bytes), "hipMemset");
    }
  }
  T *begin() { return d_data; }
  const T *begin() const { return d_data; }
  T *data() { return d_data; }
  const T *data() const { return d_data; }
  std::int64_t locallyOwnedSize() const { return d_nodes; }
  std::int64_t ghostSize() const { return 0; }
  unsigned int numVectors() const { return d_components; }
  /* every entry set to val */
  void setValue(T val) {
    const size_t n = size_t(d_nodes) * d_components;
    if (!n) return;
    if (val == T(0)) {
      internal::hipCheck(hipMemset(d_data, 0, sizeof(T) * n), "hipMemset");
    } else {
      std::vector<T> h(n, val);
      internal::hipCheck(hipMemcpy(d_data, h.data(), sizeof(T) * n, hipMemcpyHostToDevice),
                         "hipMemcpy");
    }
  }
  void swap(distributedDeviceVec &o) noexcept {
    std::swap(d_data, o.d_data);
    std::swap(d_nodes, o.d_nodes);
    std::swap(d_components, o.d_components);
  }
  void updateGhostValues() {}
  void accumulateAddLocallyOwned() {}
  void zeroOutGhosts() {}

 private:
  T *d_data = nullptr;
  std::int64_t d_nodes = 0;
  unsigned int d_components = 0;
};

/* operatorDFTDeviceClass (operatorDevice.h:43-420): the members the device
 * eigensolve calls.  XtHX and its overlap variants are not separate entry
 * points: the solver's Rayleigh-Ritz forms X^T H~ X on the device itself. */
class operatorDFTDeviceClass {
 public:
  virtual ~operatorDFTDeviceClass() = default;
  virtual void init() = 0;
  virtual void createDeviceBlasHandle() = 0;
  virtual void destroyDeviceBlasHandle() = 0;
  virtual const double *getSqrtMassVec() = 0;
  virtual const double *getInvSqrtMassVec() = 0;
  virtual distributedDeviceVec<dataTypes::number> &getParallelChebyBlockVectorDevice() = 0;
  virtual distributedDeviceVec<dataTypes::number> &getParallelChebyBlockVector2Device() = 0;
  virtual distributedDeviceVec<dataTypes::number> &
  getParallelProjectorKetTimesBlockVectorDevice() = 0;
  virtual void reinit(const unsigned int wavefunBlockSize, bool flag) = 0;
  virtual void reinitkPointSpinIndex(const unsigned int kPointIndex,
                                     const unsigned int spinIndex) = 0;
  virtual void HX(distributedDeviceVec<dataTypes::number> &X,
                  distributedDeviceVec<dataTypes::number> &projectorKetTimesVector,
                  const unsigned int localVectorSize, const unsigned int numberComponents,
                  const bool scaleFlag, const double scalar,
                  distributedDeviceVec<dataTypes::number> &Y, const bool doUnscalingX = true,
                  const bool onlyHPrimePartForFirstOrderDensityMatResponse = false) = 0;
  virtual void HXCheby(distributedDeviceVec<dataTypes::number> &X,
                       distributedDeviceVec<dataTypes::numberFP32> &XTemp,
                       distributedDeviceVec<dataTypes::number> &projectorKetTimesVector,
                       const unsigned int localVectorSize, const unsigned int numberComponents,
                       distributedDeviceVec<dataTypes::number> &Y, bool mixPrecFlag = false,
                       bool returnBeforeCompressSkipUpdateSkipNonLocal = false,
                       bool returnBeforeCompressSkipUpdateSkipLocal = false) = 0;
  /* the library operator the solver runs on (no reference counterpart) */
  virtual chfsi_operator *chfsiOperator() = 0;
};

/* kohnShamDFTOperatorDeviceClass (kohnShamDFTOperatorDevice.cc) on libchfsi.
 * The constructor takes the description init/reinit/computeMassVector gather
 * from deal.II (kohnShamDFTOperatorDevice.cc:492-1031); its host arrays are
 * copied to the device and need not outlive the call. */
class kohnShamDFTOperatorDeviceClass : public operatorDFTDeviceClass {
 public:
  explicit kohnShamDFTOperatorDeviceClass(const chfsi_operator_desc &desc,
                                          hipStream_t stream = nullptr)
      : d_stream(stream) {
    internal::chfsiCheck(chfsi_operator_create(&desc, &d_op), "chfsi_operator_create");
    d_M = chfsi_operator_n_dofs(d_op);
  }
  ~kohnShamDFTOperatorDeviceClass() override { chfsi_operator_destroy(d_op); }
  kohnShamDFTOperatorDeviceClass(const kohnShamDFTOperatorDeviceClass &) = delete;
  kohnShamDFTOperatorDeviceClass &operator=(const kohnShamDFTOperatorDeviceClass &) = delete;

  /* the constructor has done the work of init and computeMassVector */
  void init() override {}
  void createDeviceBlasHandle() override {}
  void destroyDeviceBlasHandle() override {}
  const double *getSqrtMassVec() override { return chfsi_operator_sqrt_mass(d_op); }
  const double *getInvSqrtMassVec() override { return chfsi_operator_inv_sqrt_mass(d_op); }
  distributedDeviceVec<dataTypes::number> &getParallelChebyBlockVectorDevice() override {
    return d_cheby1;
  }
  distributedDeviceVec<dataTypes::number> &getParallelChebyBlockVector2Device() override {
    return d_cheby2;
  }
  distributedDeviceVec<dataTypes::number> &getParallelProjectorKetTimesBlockVectorDevice()
      override {
    return d_projKet;
  }

  /* scratch block vectors of the given width (kohnShamDFTOperatorDevice.cc:492-607) */
  void reinit(const unsigned int wavefunBlockSize, bool) override {
    d_cheby1.reinit(d_M, wavefunBlockSize);
    d_cheby2.reinit(d_M, wavefunBlockSize);
  }
  /* active Bloch k-point (kohnShamDFTOperatorDevice.cc:1037-1057); one spin channel */
  void reinitkPointSpinIndex(const unsigned int kPointIndex, const unsigned int) override {
    if (chfsi_operator_n_kpoints(d_op) > 0)
      internal::chfsiCheck(chfsi_operator_set_kpoint(d_op, (int32_t)kPointIndex),
                           "chfsi_operator_set_kpoint");
    else if (kPointIndex != 0)
      throw std::runtime_error("reinitkPointSpinIndex: the operator holds no k-points");
  }

  /* computeVEff output on the device: V_eff JxW (C x n_q) and, for GGA, the
   * derExcWithSigmaTimesGradRhoJxW vector (C x n_q x 3); read by
   * computeHamiltonianMatricesAllkpt */
  void setVEffJxW(const double *vEffJxW_dev, const double *derExcWithSigmaTimesGradRhoJxW_dev =
                                                 nullptr) {
    d_vEffJxW = vEffJxW_dev;
    d_gJxW = derExcWithSigmaTimesGradRhoJxW_dev;
  }
  /* rebuild every cell Hamiltonian on the device
   * (hamiltonianMatrixCalculatorFlattenedDevice.cc:1073-1375) */
  void computeHamiltonianMatricesAllkpt(const unsigned int = 0,
                                        const bool onlyHPrimePart = false) {
    if (onlyHPrimePart)
      throw std::runtime_error("computeHamiltonianMatricesAllkpt: H' only is not supported");
    if (!d_vEffJxW) throw std::runtime_error("computeHamiltonianMatricesAllkpt: no V_eff set");
    if (d_gJxW)
      internal::chfsiCheck(chfsi_operator_set_potential_gga(d_op, d_vEffJxW, d_gJxW, d_stream),
                           "chfsi_operator_set_potential_gga");
    else
      internal::chfsiCheck(chfsi_operator_set_potential(d_op, d_vEffJxW, d_stream),
                           "chfsi_operator_set_potential");
  }

  /* Y <- (scaleFlag ? Y : M^{-1/2} Y) + scalar M^{-1/2} H M^{-1/2} X
   * (kohnShamDFTOperatorDevice.cc:3765-3860); X is not modified */
  void HX(distributedDeviceVec<dataTypes::number> &X, distributedDeviceVec<dataTypes::number> &,
          const unsigned int, const unsigned int numberComponents, const bool scaleFlag,
          const double scalar, distributedDeviceVec<dataTypes::number> &Y, const bool = true,
          const bool onlyHPrimePartForFirstOrderDensityMatResponse = false) override {
    if (onlyHPrimePartForFirstOrderDensityMatResponse)
      throw std::runtime_error("HX: H' only is not supported");
    internal::chfsiCheck(internal::hxCall(d_op, X.begin(), numberComponents, Y.begin(),
                                          numberComponents, (int32_t)numberComponents,
                                          scaleFlag ? 1 : 0, scalar, d_stream),
                         "HX");
  }
  /* Y <- Y + H X with X already scaled (kohnShamDFTOperatorDevice.cc:3874-3997) */
  void HXCheby(distributedDeviceVec<dataTypes::number> &X,
               distributedDeviceVec<dataTypes::numberFP32> &,
               distributedDeviceVec<dataTypes::number> &, const unsigned int,
               const unsigned int numberComponents, distributedDeviceVec<dataTypes::number> &Y,
               bool = false, bool returnBeforeCompressSkipUpdateSkipNonLocal = false,
               bool returnBeforeCompressSkipUpdateSkipLocal = false) override {
    if (returnBeforeCompressSkipUpdateSkipNonLocal || returnBeforeCompressSkipUpdateSkipLocal)
      throw std::runtime_error("HXCheby: the split local / nonlocal forms are not supported");
#ifdef USE_COMPLEX
    (void)X;
    (void)numberComponents;
    (void)Y;
    throw std::runtime_error("HXCheby: the complex form is applied inside the solver's filter");
#else
    internal::chfsiCheck(chfsi_hx_cheby(d_op, X.begin(), numberComponents, Y.begin(),
                                        numberComponents, (int32_t)numberComponents, d_stream),
                         "chfsi_hx_cheby");
#endif
  }
  chfsi_operator *chfsiOperator() override { return d_op; }
  hipStream_t stream() const { return d_stream; }
  std::int64_t nDofs() const { return d_M; }

 private:
  chfsi_operator *d_op = nullptr;
  hipStream_t d_stream = nullptr;
  std::int64_t d_M = 0;
  const double *d_vEffJxW = nullptr, *d_gJxW = nullptr;
  distributedDeviceVec<dataTypes::number> d_cheby1, d_cheby2, d_projKet;
};

/* chebyshevOrthogonalizedSubspaceIterationSolverDevice
 * (chebyshevOrthogonalizedSubspaceIterationSolverDevice.h:37-166; .cc:155-1080).
 * As in the reference, the solver keeps a reference to dftParams, which must
 * outlive it; the library parameters are taken from it at construction. */
class chebyshevOrthogonalizedSubspaceIterationSolverDevice {
 public:
  chebyshevOrthogonalizedSubspaceIterationSolverDevice(const MPI_Comm &, const MPI_Comm &,
                                                       double lowerBoundWantedSpectrum,
                                                       double lowerBoundUnWantedSpectrum,
                                                       double upperBoundUnWantedSpectrum,
                                                       const dftParameters &dftParams)
      : d_dftParams(dftParams) {
    const chfsi_params p = toChfsiParams(dftParams);
    internal::chfsiCheck(chfsi_solver_create(&p, lowerBoundWantedSpectrum,
                                             lowerBoundUnWantedSpectrum,
                                             upperBoundUnWantedSpectrum, &d_solver),
                         "chfsi_solver_create");
  }
  ~chebyshevOrthogonalizedSubspaceIterationSolverDevice() { chfsi_solver_destroy(d_solver); }
  chebyshevOrthogonalizedSubspaceIterationSolverDevice(
      const chebyshevOrthogonalizedSubspaceIterationSolverDevice &) = delete;
  chebyshevOrthogonalizedSubspaceIterationSolverDevice &operator=(
      const chebyshevOrthogonalizedSubspaceIterationSolverDevice &) = delete;

  /* One ChFSI pass on eigenVectorsFlattenedDevice (node-major, FE basis,
   * flattenedSize = n_dofs * totalNumberWaveFunctions).  eigenValues empty or
   * of size N: full Rayleigh-Ritz, N eigenvalues and residuals.  Any other
   * size n < N: the spectrum-split route (…SolverDevice.cc:584-607), the top
   * n Ritz vectors into eigenVectorsRotFracDensityFlattenedDevice (n_dofs x n).
   * Returns the refreshed upper bound of the unwanted spectrum. */
  double solve(operatorDFTDeviceClass &operatorMatrix,
               const std::shared_ptr<linearAlgebra::BLASWrapper<utils::MemorySpace::DEVICE>> &,
               elpaScalaManager &, dataTypes::number *eigenVectorsFlattenedDevice,
               dataTypes::number *eigenVectorsRotFracDensityFlattenedDevice,
               const unsigned int flattenedSize, const unsigned int totalNumberWaveFunctions,
               std::vector<double> &eigenValues, std::vector<double> &residuals,
               utils::DeviceCCLWrapper &, const MPI_Comm &, const bool isFirstFilteringCall,
               const bool computeResidual, const bool useMixedPrecOverall = false,
               const bool isFirstScf = false) {
    const unsigned int N = totalNumberWaveFunctions;
    if (N == 0 || flattenedSize % N != 0)
      throw std::runtime_error("solve: flattenedSize is not a multiple of the wavefunction count");
    const std::int64_t M = flattenedSize / N;
    double bUp = 0.0;
    if (!eigenValues.empty() && eigenValues.size() != N) {
      if (internal::isComplexNumber())
        throw std::runtime_error("solve: the spectrum split is built for real wavefunctions");
      const unsigned int nfr = (unsigned int)eigenValues.size();
      if (nfr > N || !eigenVectorsRotFracDensityFlattenedDevice)
        throw std::runtime_error("solve: spectrum split needs n <= N and the XFrac array");
      residuals.assign(nfr, 0.0);
      internal::chfsiCheck(
          chfsi_solve_spectrum_split(d_solver, operatorMatrix.chfsiOperator(),
                                     reinterpret_cast<double *>(eigenVectorsFlattenedDevice), M,
                                     (int32_t)N, (int32_t)(N - nfr),
                                     reinterpret_cast<double *>(
                                         eigenVectorsRotFracDensityFlattenedDevice),
                                     nfr,
                                     eigenValues.data(), residuals.data(), isFirstFilteringCall,
                                     computeResidual, useMixedPrecOverall, isFirstScf, &bUp,
                                     d_timings, streamOf(operatorMatrix)),
          "chfsi_solve_spectrum_split");
    } else {
      eigenValues.assign(N, 0.0);
      residuals.assign(N, 0.0);
      internal::chfsiCheck(
          internal::solveCall(d_solver, operatorMatrix.chfsiOperator(),
                              eigenVectorsFlattenedDevice, M, (int32_t)N, eigenValues.data(),
                              residuals.data(), isFirstFilteringCall, computeResidual,
                              useMixedPrecOverall, isFirstScf, &bUp, d_timings,
                              streamOf(operatorMatrix)),
          "solve");
    }
    return bUp;
  }

  /* numberPasses of filter + CholGS without Rayleigh-Ritz (…SolverDevice.cc:742-1080);
   * eigenValues is not changed */
  void solveNoRR(operatorDFTDeviceClass &operatorMatrix,
                 const std::shared_ptr<linearAlgebra::BLASWrapper<utils::MemorySpace::DEVICE>> &,
                 elpaScalaManager &, dataTypes::number *eigenVectorsFlattenedDevice,
                 const unsigned int flattenedSize, const unsigned int totalNumberWaveFunctions,
                 std::vector<double> &, utils::DeviceCCLWrapper &, const MPI_Comm &,
                 const unsigned int numberPasses, const bool useMixedPrecOverall) {
    const unsigned int N = totalNumberWaveFunctions;
    if (N == 0 || flattenedSize % N != 0)
      throw std::runtime_error("solveNoRR: flattenedSize is not a multiple of the wavefunction count");
    if (internal::isComplexNumber())
      throw std::runtime_error("solveNoRR: built for real wavefunctions");
    internal::chfsiCheck(
        chfsi_solve_no_rr(d_solver, operatorMatrix.chfsiOperator(),
                          reinterpret_cast<double *>(eigenVectorsFlattenedDevice),
                          flattenedSize / N, (int32_t)N, (int32_t)numberPasses,
                          d_dftParams.isPseudopotential ? 1 : 0, useMixedPrecOverall,
                          streamOf(operatorMatrix)),
        "chfsi_solve_no_rr");
  }

  void reinitSpectrumBounds(double lowerBoundWantedSpectrum, double lowerBoundUnWantedSpectrum,
                            double upperBoundUnWantedSpectrum) {
    internal::chfsiCheck(chfsi_solver_reinit_bounds(d_solver, lowerBoundWantedSpectrum,
                                                    lowerBoundUnWantedSpectrum,
                                                    upperBoundUnWantedSpectrum),
                         "chfsi_solver_reinit_bounds");
  }

  /* no reference counterpart: library options without a dftParameters field,
   * the bounds, and the last pass's phase times {lanczos, filter, cholgs,
   * rayleigh_ritz, residual, total} in ms */
  void setLibraryParams(const chfsi_params &p) {
    internal::chfsiCheck(chfsi_solver_set_params(d_solver, &p), "chfsi_solver_set_params");
  }
  void getSpectrumBounds(double &a0, double &bLow, double &bUp) const {
    internal::chfsiCheck(chfsi_solver_get_bounds(d_solver, &a0, &bLow, &bUp),
                         "chfsi_solver_get_bounds");
  }
  const double *lastPassTimings() const { return d_timings; }

 private:
  static void *streamOf(operatorDFTDeviceClass &op) {
    auto *ks = dynamic_cast<kohnShamDFTOperatorDeviceClass *>(&op);
    return ks ? (void *)ks->stream() : nullptr;
  }
  const dftParameters &d_dftParams;
  chfsi_solver *d_solver = nullptr;
  double d_timings[6] = {0, 0, 0, 0, 0, 0};
};

}  // namespace dftfe
