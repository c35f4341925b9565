// Consolidated element algorithm with a kernel list: the counterpart of
// AssembleElemSolverAlgorithm (include/AssembleElemSolverAlgorithm.h:52-105,
// src/AssembleElemSolverAlgorithm.C:75-100) and its Kernel interface
// (include/kernel/Kernel.h:95-125).
//
// add_kernel mirrors activeKernels_.push_back; execute() first calls every
// kernel's setup(TimeIntegrator) -- ContinuityAdvElemKernel takes
// projTimeScale = dt / gamma1 (src/kernel/ContinuityAdvElemKernel.C:79-84),
// the mass kernels dt and the BDF gammas (src/kernel/MomentumMassElemKernel.C:80-88,
// src/kernel/ContinuityMassElemKernel.C:67-75) -- and then assembles every
// kernel's element contributions into the system.  The reference sums the
// kernels' element matrices per element and applies them once; here each
// kernel is one fused gather + geometry + physics + scatter pass (the
// element-visit tile assembly for the advection / diffusion kernels, the
// owner-computes node pass for the mass kernels), all adding into the same
// rows, which gives the same sums up to rounding order -- except that a
// momentum advection-diffusion kernel and a momentum mass kernel of the same
// list share one node-tile pass (nalu_assemble_momentum_with_mass).
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

struct nalu_elem_algorithm {
  nalu_linsys* ls = nullptr;
  const nalu_mesh* mesh = nullptr;
  std::vector<nalu_kernel_desc> kernels;
  double dt = 1.0;
  double gamma[3] = {1.0, -1.0, 0.0};
};

namespace nalu {
namespace {

constexpr int kEtile = 3;  // scatter mode of the advection / diffusion kernels

int dof_of(int kind) {
  switch (kind) {
    case NALU_KERNEL_SCALAR_DIFF:
    case NALU_KERNEL_CONTINUITY_ADV:
    case NALU_KERNEL_CONTINUITY_MASS: return 1;
    case NALU_KERNEL_MOMENTUM_ADVDIFF:
    case NALU_KERNEL_MOMENTUM_MASS: return 3;
    default: return -1;
  }
}

int fields_of(int kind) {
  switch (kind) {
    case NALU_KERNEL_SCALAR_DIFF: return 2;       // q, diffusion coefficient
    case NALU_KERNEL_CONTINUITY_ADV: return 4;    // rho, p, u, Gpdx
    case NALU_KERNEL_MOMENTUM_ADVDIFF: return 3;  // u, mu, mdot
    case NALU_KERNEL_CONTINUITY_MASS: return 3;   // rho Nm1, N, Np1
    case NALU_KERNEL_MOMENTUM_MASS: return 7;     // rho Nm1, N, Np1, u Nm1, N, Np1, Gjp
    default: return 0;
  }
}

int run(nalu_elem_algorithm* a, const nalu_kernel_desc& k) {
  const double* const* f = k.field;
  nalu_linsys* ls = a->ls;
  const nalu_mesh* m = a->mesh;
  switch (k.kind) {
    case NALU_KERNEL_SCALAR_DIFF:
      return nalu_assemble_scalar_diff(ls, m, f[0], f[1], k.shifted_grad_op, kEtile);
    case NALU_KERNEL_CONTINUITY_ADV:
      return nalu_assemble_continuity(ls, m, f[0], f[1], f[2], f[3], a->dt / a->gamma[0], k.shift_mdot,
                                      k.shifted_grad_op, k.reduced_sens, kEtile);
    case NALU_KERNEL_MOMENTUM_ADVDIFF:
      return nalu_assemble_momentum(ls, m, f[0], f[1], f[2], k.include_divu, k.skew_symmetric, k.shifted_grad_op,
                                    kEtile);
    case NALU_KERNEL_CONTINUITY_MASS:
      return nalu_assemble_continuity_mass(ls, m, f[0], f[1], f[2], a->dt, a->gamma, k.lumped);
    case NALU_KERNEL_MOMENTUM_MASS:
      return nalu_assemble_momentum_mass(ls, m, f[0], f[1], f[2], f[3], f[4], f[5], f[6], a->dt, a->gamma, k.lumped,
                                         k.balanced_force);
    default: return NALU_ERR_ARG;
  }
}

}  // namespace
}  // namespace nalu

using namespace nalu;

extern "C" int nalu_elem_algorithm_create(nalu_linsys* ls, const nalu_mesh* m, nalu_elem_algorithm** out) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && m && out, NALU_ERR_ARG, "elem_algorithm_create: null argument");
    NALU_REQUIRE(ls->finalized && ls->mesh == m, NALU_ERR_STATE,
                 "elem_algorithm_create: the system's graph must be finalized on this mesh");
    auto* a = new nalu_elem_algorithm();
    a->ls = ls;
    a->mesh = m;
    *out = a;
    return NALU_OK;
  });
}

extern "C" int nalu_elem_algorithm_destroy(nalu_elem_algorithm* a) {
  delete a;  // the algorithm owns its kernels (src/Algorithm.C:46-55)
  return NALU_OK;
}

extern "C" int nalu_elem_algorithm_add_kernel(nalu_elem_algorithm* a, const nalu_kernel_desc* k) {
  return guarded([&]() -> int {
    NALU_REQUIRE(a && k, NALU_ERR_ARG, "add_kernel: null argument");
    const int nd = dof_of(k->kind);
    NALU_REQUIRE(nd > 0, NALU_ERR_ARG, "add_kernel: unknown kernel kind");
    NALU_REQUIRE(nd == a->ls->num_dof, NALU_ERR_ARG, "add_kernel: kernel and linear system numDof differ");
    for (int i = 0; i < fields_of(k->kind); ++i)
      NALU_REQUIRE(k->field[i] != nullptr, NALU_ERR_ARG, "add_kernel: missing field " + std::to_string(i));
    a->kernels.push_back(*k);
    return NALU_OK;
  });
}

extern "C" int nalu_elem_algorithm_set_time_integrator(nalu_elem_algorithm* a, double dt, const double* gamma) {
  return guarded([&]() -> int {
    NALU_REQUIRE(a && gamma, NALU_ERR_ARG, "set_time_integrator: null argument");
    NALU_REQUIRE(std::isfinite(dt) && dt > 0.0 && std::isfinite(gamma[0]) && gamma[0] != 0.0 &&
                     std::isfinite(gamma[1]) && std::isfinite(gamma[2]),
                 NALU_ERR_ARG, "set_time_integrator: dt must be positive, gamma1 nonzero");
    a->dt = dt;
    for (int i = 0; i < 3; ++i) a->gamma[i] = gamma[i];
    return NALU_OK;
  });
}

extern "C" int nalu_elem_algorithm_execute(nalu_elem_algorithm* a) {
  return guarded([&]() -> int {
    NALU_REQUIRE(a, NALU_ERR_ARG, "elem_algorithm_execute: null handle");
    // a momentum advection-diffusion kernel and a momentum mass kernel in the
    // same list evaluate in one pass over the elements (one LDS row segment per
    // node tile, one write of the matrix); every other kernel runs as its own pass
    int adv = -1, mass = -1;
    for (int i = 0; i < (int)a->kernels.size(); ++i) {
      if (a->kernels[i].kind == NALU_KERNEL_MOMENTUM_ADVDIFF && adv < 0) adv = i;
      if (a->kernels[i].kind == NALU_KERNEL_MOMENTUM_MASS && mass < 0) mass = i;
    }
    for (int i = 0; i < (int)a->kernels.size(); ++i) {
      int rc;
      if (adv >= 0 && mass >= 0 && (i == adv || i == mass)) {
        if (i != std::min(adv, mass)) continue;
        const nalu_kernel_desc& ka = a->kernels[adv];
        const nalu_kernel_desc& km = a->kernels[mass];
        const double* const* fa = ka.field;
        const double* const* fm = km.field;
        rc = nalu_assemble_momentum_with_mass(a->ls, a->mesh, fa[0], fa[1], fa[2], ka.include_divu,
                                              ka.skew_symmetric, ka.shifted_grad_op, fm[0], fm[1], fm[2], fm[3],
                                              fm[4], fm[5], fm[6], a->dt, a->gamma, km.lumped, km.balanced_force);
      } else {
        rc = run(a, a->kernels[i]);
      }
      if (rc < 0) return rc;  // the failing kernel's message is the last error
    }
    return NALU_OK;
  });
}
