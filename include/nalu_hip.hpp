// nalu_hip.hpp -- C++ host mirror of Nalu's LinearSystem / LinearSolver /
// element-algorithm interfaces over the C ABI in nalu_hip.h.
//
// Method names, argument meaning and error behaviour follow the reference:
//   LinearSystem   include/LinearSystem.h:39-167 (TpetraLinearSystem, src/TpetraLinearSystem.C)
//   solver config  src/LinearSolverConfig.C:34-108 (method, preconditioner, tolerance, ...)
// Logic errors throw std::runtime_error, as STK_ThrowRequire does in the
// reference; solve() returns 0 (converged) or 1 (iteration cap reached).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "nalu_hip.h"

namespace nalu_hip {

inline int check(int rc, const char* what) {
  if (rc < 0) throw std::runtime_error(std::string(what) + ": " + nalu_last_error());
  return rc;
}

struct SolverConfig {
  std::string method = "gmres";          // gmres | cg | fgmres
  std::string preconditioner = "sgs";    // none | jacobi | sgs | mt_sgs | learned | muelu
  double tolerance = 1e-8;
  double finalTolerance = 1e-8;
  int maxIterations = 1000;
  int kspace = 50;
  bool mixedPrecision = false;

  nalu_solver_config c_config() const {
    nalu_solver_config c{};
    if (method == "gmres") c.method = NALU_METHOD_GMRES;
    else if (method == "cg") c.method = NALU_METHOD_CG;
    else if (method == "fgmres") c.method = NALU_METHOD_FGMRES;
    else throw std::runtime_error("unsupported linear solver method: " + method);
    if (preconditioner == "none") c.precond = NALU_PREC_NONE;
    else if (preconditioner == "jacobi" || preconditioner == "default") c.precond = NALU_PREC_JACOBI;
    else if (preconditioner == "sgs" || preconditioner == "mt_sgs") c.precond = NALU_PREC_SGS;
    else if (preconditioner == "learned") c.precond = NALU_PREC_LEARNED;
    else if (preconditioner == "muelu") c.precond = NALU_PREC_MUELU;
    else throw std::runtime_error("unsupported preconditioner: " + preconditioner);
    c.tolerance = tolerance;
    c.final_tolerance = finalTolerance;
    c.max_iterations = maxIterations;
    c.kspace = kspace;
    c.precision = mixedPrecision ? NALU_PRECISION_MIXED : NALU_PRECISION_FP64;
    c.final_outer_iter = 0;
    return c;
  }
};

// A device-resident Hex8 block (coordinates, connectivity, optional global ids).
class Mesh {
 public:
  Mesh(int64_t numNodes, int64_t numElems, const int32_t* conn, const double* coords,
       const int64_t* globalIds = nullptr) {
    check(nalu_mesh_create(numNodes, numElems, conn, coords, globalIds, &m_), "nalu_mesh_create");
  }
  ~Mesh() { nalu_mesh_destroy(m_); }
  Mesh(const Mesh&) = delete;
  Mesh& operator=(const Mesh&) = delete;
  nalu_mesh* get() const { return m_; }

 private:
  nalu_mesh* m_ = nullptr;
};

// HIP-backed counterpart of TpetraLinearSystem.
class LinearSystem {
 public:
  LinearSystem(int numDof, const SolverConfig& cfg) : numDof_(numDof) {
    const nalu_solver_config c = cfg.c_config();
    check(nalu_linsys_create(numDof, &c, &ls_), "LinearSystem::create");
  }
  ~LinearSystem() { nalu_linsys_destroy(ls_); }
  LinearSystem(const LinearSystem&) = delete;
  LinearSystem& operator=(const LinearSystem&) = delete;

  nalu_linsys* get() const { return ls_; }
  int numDof() const { return numDof_; }

  // graph
  void buildElemToNodeGraph(const Mesh& mesh) {
    check(nalu_linsys_build_elem_to_node_graph(ls_, mesh.get()), "buildElemToNodeGraph");
  }
  void buildDirichletNodeGraph() {}  // no-op, as in the reference (include/LinearSystem.h:67-74)
  void finalizeLinearSystem() { check(nalu_linsys_finalize(ls_), "finalizeLinearSystem"); }

  // assembly
  void zeroSystem() { check(nalu_linsys_zero(ls_), "zeroSystem"); }
  // std::vector overload (src/TpetraLinearSystem.C:1524-1577): one element, host data
  void sumInto(const std::vector<int32_t>& entities, const std::vector<double>& rhs,
               const std::vector<double>& lhs) {
    const size_t R = entities.size() * (size_t)numDof_;
    if (rhs.size() != R || lhs.size() != R * R) throw std::runtime_error("sumInto: lhs/rhs size mismatch");
    check(nalu_linsys_sum_into(ls_, (int)entities.size(), entities.data(), rhs.data(), lhs.data()), "sumInto");
  }
  // batched device overload: element matrices of the graph mesh
  void sumInto(int64_t numElems, const int32_t* connDev, const double* lhsDev, const double* rhsDev) {
    check(nalu_linsys_sum_into_batch(ls_, numElems, connDev, lhsDev, rhsDev), "sumInto");
  }
  void applyDirichletBCs(int64_t numNodes, const int32_t* nodesDev, const double* solutionDev,
                         const double* bcValuesDev, unsigned beginPos, unsigned endPos) {
    check(nalu_linsys_apply_dirichlet(ls_, numNodes, nodesDev, solutionDev, bcValuesDev, (int)beginPos,
                                      (int)endPos), "applyDirichletBCs");
  }
  void resetRows(int64_t numNodes, const int32_t* nodesDev, unsigned beginPos, unsigned endPos,
                 double diagValue = 1.0, double rhsResidual = 0.0) {
    check(nalu_linsys_reset_rows(ls_, numNodes, nodesDev, (int)beginPos, (int)endPos, diagValue, rhsResidual),
          "resetRows");
  }
  void loadComplete() { check(nalu_linsys_load_complete(ls_), "loadComplete"); }
  // write_matrix_files (src/LinearSolverConfig.C:102); empty name turns it off
  void setWriteMatrixFiles(const std::string& baseName) {
    check(nalu_linsys_set_write_matrix_files(ls_, baseName.empty() ? nullptr : baseName.c_str()),
          "setWriteMatrixFiles");
  }

  // MueLu parameter list (muelu_xml_file_name, src/LinearSolverConfig.C:86-100) as a struct
  void setMueLuParams(const nalu_muelu_params& p) {
    check(nalu_linsys_set_muelu_params(ls_, &p), "setMueLuParams");
  }

  // solve; writes the delta solution (device array of numRows doubles)
  int solve(double* deltaDev) { return check(nalu_linsys_solve(ls_, deltaDev), "solve"); }
  int linearSolveIterations() const { return nalu_linsys_iterations(ls_); }
  double linearResidual() const { return nalu_linsys_linear_residual(ls_); }
  double nonLinearResidual() const { return nalu_linsys_nonlinear_residual(ls_); }
  double scaledNonLinearResidual() const { return nalu_linsys_scaled_nonlinear_residual(ls_); }

 private:
  nalu_linsys* ls_ = nullptr;
  int numDof_ = 1;
};

// Consolidated element algorithms (AssembleElemSolverAlgorithm with one kernel).
enum class Scatter { RowGather = 0, Atomic = 1, NodeTile = 2 };

inline void assembleScalarDiff(LinearSystem& ls, const Mesh& m, const double* qDev, const double* kDev,
                               bool shifted = false, Scatter mode = Scatter::NodeTile) {
  check(nalu_assemble_scalar_diff(ls.get(), m.get(), qDev, kDev, shifted, (int)mode), "ScalarDiffElemKernel");
}
inline void computeMdot(const Mesh& m, const double* rhoDev, const double* pDev, const double* uDev,
                        const double* gpdxDev, double projTimeScale, double* mdotDev, bool shiftMdot = false,
                        bool shiftPoisson = false, bool fp32Arithmetic = false) {
  auto fn = fp32Arithmetic ? nalu_compute_mdot_f32 : nalu_compute_mdot;
  check(fn(m.get(), rhoDev, pDev, uDev, gpdxDev, projTimeScale, shiftMdot, shiftPoisson, mdotDev, nullptr),
        "ComputeMdotElemAlgorithm");
}
inline void assembleContinuity(LinearSystem& ls, const Mesh& m, const double* rhoDev, const double* pDev,
                               const double* uDev, const double* gpdxDev, double projTimeScale,
                               bool shiftMdot = false, bool shiftPoisson = false, bool reducedSens = false,
                               Scatter mode = Scatter::NodeTile) {
  check(nalu_assemble_continuity(ls.get(), m.get(), rhoDev, pDev, uDev, gpdxDev, projTimeScale, shiftMdot,
                                 shiftPoisson, reducedSens, (int)mode), "ContinuityAdvElemKernel");
}
inline void assembleMomentum(LinearSystem& ls, const Mesh& m, const double* uDev, const double* muDev,
                             const double* mdotDev, double includeDivU = 0.0, bool skewSymmetric = false,
                             bool shiftedGradOp = false, Scatter mode = Scatter::NodeTile) {
  check(nalu_assemble_momentum(ls.get(), m.get(), uDev, muDev, mdotDev, includeDivU, skewSymmetric, shiftedGradOp,
                               (int)mode), "MomentumAdvDiffElemKernel");
}

}  // namespace nalu_hip
