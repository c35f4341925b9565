// nalu_hip.hpp -- C++ host mirror of Nalu's LinearSystem / LinearSolver /
// element-algorithm interfaces over the C ABI in nalu_hip.h.
//
// Method names, argument meaning and error behaviour follow the reference:
//   LinearSystem      include/LinearSystem.h:39-167 (TpetraLinearSystem, src/TpetraLinearSystem.C)
//   Kernel            include/kernel/Kernel.h:95-125
//   AssembleElemSolverAlgorithm  include/AssembleElemSolverAlgorithm.h:52-105
//   solver config     src/LinearSolverConfig.C:34-108 (method, preconditioner, tolerance, ...)
// The reference's STK types are replaced by this build's own: a Mesh (one
// device-resident Hex8 block) stands for the PartVector a graph builder
// receives, node indices stand for stk::mesh::Entity, and device arrays stand
// for STK fields.  Graph builders of assembly paths outside the Hex8 element
// path (node, face, edge, reduced-element, face-element, non-conformal and
// overset graphs) and the constraint preparation throw "unsupported".
// Logic errors throw std::runtime_error, as STK_ThrowRequire does in the
// reference; solve() returns 0 (converged) or 1 (iteration cap reached).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "nalu_hip.h"

namespace nalu_hip {

inline int check(int rc, const char* what) {
  if (rc < 0) throw std::runtime_error(std::string(what) + ": " + nalu_last_error());
  return rc;
}

[[noreturn]] inline void unsupported(const char* what) {
  throw std::runtime_error(std::string(what) + ": unsupported by the HIP linear system (Hex8 element path only)");
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
  const int32_t* connectivityDevice() const {
    const int32_t* c = nullptr;
    check(nalu_mesh_device_ptrs(m_, nullptr, &c), "nalu_mesh_device_ptrs");
    return c;
  }

 private:
  nalu_mesh* m_ = nullptr;
};

// Nodes of a Dirichlet-type boundary and the device fields it reads: the
// stand-in for (solutionField, bcValuesField, parts) of applyDirichletBCs.
struct DirichletBC {
  int64_t numNodes = 0;
  const int32_t* nodesDev = nullptr;       // node indices
  const double* solutionDev = nullptr;     // nodal [N][numDof]
  const double* bcValuesDev = nullptr;     // nodal [N][numDof]
};

// Device views of one batch of element contributions for the Kokkos sumInto
// overload: the stand-in for (entities, rhs, lhs, localIds, sortPermutation)
// of one team iteration, batched over numElems elements.
struct ElementBatch {
  int64_t numElems = 0;
  int numEntities = 8;
  const int32_t* entitiesDev = nullptr;    // [E][numEntities] node indices
  const double* rhsDev = nullptr;          // [E][R]
  const double* lhsDev = nullptr;          // [E][R][R]
  int32_t* localIdsDev = nullptr;          // [E][R] scratch, overwritten
  int32_t* sortPermutationDev = nullptr;   // [E][R] scratch, overwritten
};

// HIP-backed counterpart of TpetraLinearSystem: every pure virtual of
// sierra::nalu::LinearSystem (include/LinearSystem.h:54-127) is present.
class LinearSystem {
 public:
  LinearSystem(int numDof, const SolverConfig& cfg, std::string name = "")
      : numDof_(numDof), eqSysName_(std::move(name)) {
    const nalu_solver_config c = cfg.c_config();
    check(nalu_linsys_create(numDof, &c, &ls_), "LinearSystem::create");
  }
  ~LinearSystem() { nalu_linsys_destroy(ls_); }
  LinearSystem(const LinearSystem&) = delete;
  LinearSystem& operator=(const LinearSystem&) = delete;

  nalu_linsys* get() const { return ls_; }

  // ---- graph / matrix construction (include/LinearSystem.h:54-74)
  void buildNodeGraph(const Mesh&) { unsupported("buildNodeGraph"); }
  void buildFaceToNodeGraph(const Mesh&) { unsupported("buildFaceToNodeGraph"); }
  void buildEdgeToNodeGraph(const Mesh&) { unsupported("buildEdgeToNodeGraph"); }
  void buildElemToNodeGraph(const Mesh& mesh) {
    check(nalu_linsys_build_elem_to_node_graph(ls_, mesh.get()), "buildElemToNodeGraph");
  }
  void buildReducedElemToNodeGraph(const Mesh&) { unsupported("buildReducedElemToNodeGraph"); }
  void buildFaceElemToNodeGraph(const Mesh&) { unsupported("buildFaceElemToNodeGraph"); }
  void buildNonConformalNodeGraph(const Mesh&) { unsupported("buildNonConformalNodeGraph"); }
  void buildOversetNodeGraph(const Mesh&) { unsupported("buildOversetNodeGraph"); }
  void finalizeLinearSystem() { check(nalu_linsys_finalize(ls_), "finalizeLinearSystem"); }
  // no-ops, as in the reference (include/LinearSystem.h:67-74): Dirichlet rows
  // keep their element-graph structure
  void buildDirichletNodeGraph(const Mesh&) {}
  void buildDirichletNodeGraph(const std::vector<int32_t>&) {}

  // ---- assembly (include/LinearSystem.h:77-98)
  void zeroSystem() { check(nalu_linsys_zero(ls_), "zeroSystem"); }
  // Kokkos overload, a batch of elements on the device
  void sumInto(const ElementBatch& b, const char* /*trace_tag*/ = nullptr) {
    check(nalu_linsys_sum_into_elements(ls_, b.numElems, b.numEntities, b.entitiesDev, b.rhsDev, b.lhsDev,
                                        b.localIdsDev, b.sortPermutationDev),
          "sumInto");
  }
  // std::vector overload: one element, host data (scratch vectors are unused:
  // the call is queued and applied on the device)
  void sumInto(const std::vector<int32_t>& entities, std::vector<int>& /*scratchIds*/,
               std::vector<double>& /*scratchVals*/, const std::vector<double>& rhs, const std::vector<double>& lhs,
               const char* /*trace_tag*/ = nullptr) {
    const size_t R = entities.size() * (size_t)numDof_;
    if (rhs.size() != R || lhs.size() != R * R) throw std::runtime_error("sumInto: lhs/rhs size mismatch");
    check(nalu_linsys_sum_into(ls_, (int)entities.size(), entities.data(), rhs.data(), lhs.data()), "sumInto");
  }
  // element matrices of the graph mesh's own elements (element scatter table)
  void sumIntoMeshElements(int64_t numElems, const int32_t* connDev, const double* lhsDev, const double* rhsDev) {
    check(nalu_linsys_sum_into_batch(ls_, numElems, connDev, lhsDev, rhsDev), "sumInto");
  }

  // ---- boundary conditions / constraints (include/LinearSystem.h:100-120)
  void applyDirichletBCs(const DirichletBC& bc, unsigned beginPos, unsigned endPos) {
    check(nalu_linsys_apply_dirichlet(ls_, bc.numNodes, bc.nodesDev, bc.solutionDev, bc.bcValuesDev, (int)beginPos,
                                      (int)endPos),
          "applyDirichletBCs");
  }
  void prepareConstraints(unsigned /*beginPos*/, unsigned /*endPos*/) { unsupported("prepareConstraints"); }
  void resetRows(int64_t numNodes, const int32_t* nodesDev, unsigned beginPos, unsigned endPos,
                 double diagValue = 1.0, double rhsResidual = 0.0) {
    check(nalu_linsys_reset_rows(ls_, numNodes, nodesDev, (int)beginPos, (int)endPos, diagValue, rhsResidual),
          "resetRows");
  }

  // ---- solve (include/LinearSystem.h:123-127)
  // writes the delta solution (device array of numRows doubles), the
  // linearSolutionField of the reference
  int solve(double* deltaDev) {
    linearSolveIterations_ = 0;
    const int rc = check(nalu_linsys_solve(ls_, deltaDev), "solve");
    linearSolveIterations_ = nalu_linsys_iterations(ls_);
    linearResidual_ = nalu_linsys_linear_residual(ls_);
    nonLinearResidual_ = nalu_linsys_nonlinear_residual(ls_);
    scaledNonLinearResidual_ = nalu_linsys_scaled_nonlinear_residual(ls_);
    return rc;
  }
  void loadComplete() { check(nalu_linsys_load_complete(ls_), "loadComplete"); }
  void writeToFile(const char* filename, bool /*useOwned*/ = true) {
    check(nalu_linsys_write_to_file(ls_, filename), "writeToFile");
  }
  void writeSolutionToFile(const char* filename, bool /*useOwned*/ = true) {
    check(nalu_linsys_write_solution_to_file(ls_, filename), "writeSolutionToFile");
  }
  // write_matrix_files (src/LinearSolverConfig.C:102); empty name turns it off
  void setWriteMatrixFiles(const std::string& baseName) {
    check(nalu_linsys_set_write_matrix_files(ls_, baseName.empty() ? nullptr : baseName.c_str()),
          "setWriteMatrixFiles");
  }
  // MueLu parameter list (muelu_xml_file_name, src/LinearSolverConfig.C:86-100) as a struct
  void setMueLuParams(const nalu_muelu_params& p) { check(nalu_linsys_set_muelu_params(ls_, &p), "setMueLuParams"); }
  // base of the learned preconditioner: NALU_PREC_SGS or NALU_PREC_MUELU (no reference counterpart)
  void setLearnedBase(int precond) { check(nalu_linsys_set_learned_base(ls_, precond), "setLearnedBase"); }

  // ---- accessors (include/LinearSystem.h:128-141)
  unsigned numDof() const { return (unsigned)numDof_; }
  const int& linearSolveIterations() const { return linearSolveIterations_; }
  const double& linearResidual() const { return linearResidual_; }
  const double& nonLinearResidual() const { return nonLinearResidual_; }
  const double& scaledNonLinearResidual() const { return scaledNonLinearResidual_; }
  void setNonLinearResidual(double nlr) { nonLinearResidual_ = nlr; }
  const std::string name() const { return eqSysName_; }
  bool& recomputePreconditioner() { return recomputePreconditioner_; }
  bool& reusePreconditioner() { return reusePreconditioner_; }
  // preconditioner set-up time of the solves so far (the reference's precond timer)
  double get_timer_precond() const {
    double t[7] = {0};
    check(nalu_linsys_phase_times(ls_, t), "get_timer_precond");
    return t[5] - precondTimerZero_;
  }
  void zero_timer_precond() {
    double t[7] = {0};
    check(nalu_linsys_phase_times(ls_, t), "zero_timer_precond");
    precondTimerZero_ = t[5];
  }

 protected:
  // construction starts with buildElemToNodeGraph; errors are checked at every call
  void beginLinearSystemConstruction() {}
  void checkError(int errCode, const char* msg) { check(errCode, msg); }

 private:
  nalu_linsys* ls_ = nullptr;
  int numDof_ = 1;
  std::string eqSysName_;
  int linearSolveIterations_ = 0;
  double linearResidual_ = 0.0, nonLinearResidual_ = 0.0, scaledNonLinearResidual_ = 0.0;
  bool recomputePreconditioner_ = true, reusePreconditioner_ = false;
  double precondTimerZero_ = 0.0;
};

// ---------------------------------------------------------------- kernels
// TimeIntegrator: what Kernel::setup reads (dt and the BDF gammas).
struct TimeIntegrator {
  double dt = 1.0;
  double gamma1 = 1.0, gamma2 = -1.0, gamma3 = 0.0;
};

// Kernel (include/kernel/Kernel.h:95-125): a kind, its options and the device
// fields it reads; execution happens inside the algorithm's fused pass.
class Kernel {
 public:
  virtual ~Kernel() = default;
  virtual void setup(const TimeIntegrator&) {}
  const nalu_kernel_desc& desc() const { return d_; }

 protected:
  explicit Kernel(int kind) { d_.kind = kind; }
  nalu_kernel_desc d_{};
};

class ScalarDiffElemKernel : public Kernel {
 public:
  ScalarDiffElemKernel(const double* qDev, const double* diffFluxCoeffDev, bool shiftedGradOp = false)
      : Kernel(NALU_KERNEL_SCALAR_DIFF) {
    d_.field[0] = qDev;
    d_.field[1] = diffFluxCoeffDev;
    d_.shifted_grad_op = shiftedGradOp;
  }
};

class ContinuityAdvElemKernel : public Kernel {
 public:
  ContinuityAdvElemKernel(const double* rhoDev, const double* pDev, const double* uDev, const double* GpdxDev,
                          bool shiftedGradOp = false, bool shiftMdot = false, bool reducedSensitivities = false)
      : Kernel(NALU_KERNEL_CONTINUITY_ADV) {
    d_.field[0] = rhoDev;
    d_.field[1] = pDev;
    d_.field[2] = uDev;
    d_.field[3] = GpdxDev;
    d_.shifted_grad_op = shiftedGradOp;
    d_.shift_mdot = shiftMdot;
    d_.reduced_sens = reducedSensitivities;
  }
};

class MomentumAdvDiffElemKernel : public Kernel {
 public:
  MomentumAdvDiffElemKernel(const double* uDev, const double* muDev, const double* mdotDev,
                            double includeDivU = 0.0, bool skewSymmetric = false, bool shiftedGradOp = false)
      : Kernel(NALU_KERNEL_MOMENTUM_ADVDIFF) {
    d_.field[0] = uDev;
    d_.field[1] = muDev;
    d_.field[2] = mdotDev;
    d_.include_divu = includeDivU;
    d_.skew_symmetric = skewSymmetric;
    d_.shifted_grad_op = shiftedGradOp;
  }
};

class ContinuityMassElemKernel : public Kernel {
 public:
  ContinuityMassElemKernel(const double* rhoNm1, const double* rhoN, const double* rhoNp1, bool lumped = false)
      : Kernel(NALU_KERNEL_CONTINUITY_MASS) {
    d_.field[0] = rhoNm1;
    d_.field[1] = rhoN;
    d_.field[2] = rhoNp1;
    d_.lumped = lumped;
  }
};

class MomentumMassElemKernel : public Kernel {
 public:
  MomentumMassElemKernel(const double* rhoNm1, const double* rhoN, const double* rhoNp1, const double* uNm1,
                         const double* uN, const double* uNp1, const double* GjpDev, bool lumped = false,
                         bool balancedForce = false)
      : Kernel(NALU_KERNEL_MOMENTUM_MASS) {
    const double* f[7] = {rhoNm1, rhoN, rhoNp1, uNm1, uN, uNp1, GjpDev};
    for (int i = 0; i < 7; ++i) d_.field[i] = f[i];
    d_.lumped = lumped;
    d_.balanced_force = balancedForce;
  }
};

// AssembleElemSolverAlgorithm: owns its kernels (src/Algorithm.C:46-55);
// execute() sets every kernel up and runs the fused assembly of the list.
class AssembleElemSolverAlgorithm {
 public:
  AssembleElemSolverAlgorithm(LinearSystem& linsys, const Mesh& mesh) {
    check(nalu_elem_algorithm_create(linsys.get(), mesh.get(), &alg_), "AssembleElemSolverAlgorithm");
  }
  ~AssembleElemSolverAlgorithm() { nalu_elem_algorithm_destroy(alg_); }
  AssembleElemSolverAlgorithm(const AssembleElemSolverAlgorithm&) = delete;
  AssembleElemSolverAlgorithm& operator=(const AssembleElemSolverAlgorithm&) = delete;

  void addKernel(std::unique_ptr<Kernel> k) {
    check(nalu_elem_algorithm_add_kernel(alg_, &k->desc()), "activeKernels_.push_back");
    activeKernels_.push_back(std::move(k));
  }
  void execute(const TimeIntegrator& ti = TimeIntegrator()) {
    for (auto& k : activeKernels_) k->setup(ti);
    const double g[3] = {ti.gamma1, ti.gamma2, ti.gamma3};
    check(nalu_elem_algorithm_set_time_integrator(alg_, ti.dt, g), "TimeIntegrator");
    check(nalu_elem_algorithm_execute(alg_), "AssembleElemSolverAlgorithm::execute");
  }
  const std::vector<std::unique_ptr<Kernel>>& activeKernels() const { return activeKernels_; }

 private:
  nalu_elem_algorithm* alg_ = nullptr;
  std::vector<std::unique_ptr<Kernel>> activeKernels_;
};

// Single-kernel convenience entries (scatter mode: 0 row gather, 1 atomic,
// 2 node tile, 3 element-visit tile).
enum class Scatter { RowGather = 0, Atomic = 1, NodeTile = 2, ElementTile = 3 };

inline void computeMdot(const Mesh& m, const double* rhoDev, const double* pDev, const double* uDev,
                        const double* gpdxDev, double projTimeScale, double* mdotDev, bool shiftMdot = false,
                        bool shiftPoisson = false, bool fp32Arithmetic = false) {
  auto fn = fp32Arithmetic ? nalu_compute_mdot_f32 : nalu_compute_mdot;
  check(fn(m.get(), rhoDev, pDev, uDev, gpdxDev, projTimeScale, shiftMdot, shiftPoisson, mdotDev, nullptr),
        "ComputeMdotElemAlgorithm");
}

// ContinuityAdvElemKernel and ComputeMdotElemAlgorithm from one pass over the
// scs geometry (row gather): the continuity system of `linsys` and mdot [E][12].
inline void assembleContinuityAndMdot(LinearSystem& linsys, const Mesh& m, const double* rhoDev, const double* pDev,
                                      const double* uDev, const double* gpdxDev, double projTimeScale,
                                      double* mdotDev, bool shiftMdot = false, bool shiftPoisson = false,
                                      bool reducedSensitivities = false) {
  check(nalu_assemble_continuity_mdot(linsys.get(), m.get(), rhoDev, pDev, uDev, gpdxDev, projTimeScale, shiftMdot,
                                      shiftPoisson, reducedSensitivities, mdotDev),
        "ContinuityAdvElemKernel + ComputeMdotElemAlgorithm");
}

}  // namespace nalu_hip
