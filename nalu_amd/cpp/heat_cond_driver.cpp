// Heat-conduction equation system driver in the reference's call order:
// HeatCondEquationSystem::initialize -> buildElemToNodeGraph + finalizeLinearSystem
// (src/HeatCondEquationSystem.C:1038-1042), then EquationSystem::assemble_and_solve
// (src/EquationSystem.C:264-307): zeroSystem, element assembly, Dirichlet,
// loadComplete, solve; then T += dT (src/HeatCondEquationSystem.C:1105-1110).
// Assembly goes through AssembleElemSolverAlgorithm with a ScalarDiffElemKernel,
// Dirichlet rows through applyDirichletBCs, as the reference's equation system does.
//
// Usage: heat_cond_driver [n=32] [jitter=0.1] [seed=1234]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "../../include/nalu_hip.hpp"

#define HIPOK(x)                                                                  \
  do {                                                                            \
    if ((x) != hipSuccess) {                                                      \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(hipGetLastError())); \
      return 2;                                                                   \
    }                                                                             \
  } while (0)

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 32;
  const double jitter = argc > 2 ? std::atof(argv[2]) : 0.1;
  const unsigned seed = argc > 3 ? (unsigned)std::atol(argv[3]) : 1234u;
  const int64_t nn = (int64_t)(n + 1) * (n + 1) * (n + 1), ne = (int64_t)n * n * n;
  std::vector<double> coords(3 * nn), k(nn), T(nn, 0.0), bcv(nn, 0.0);
  std::vector<int32_t> conn(8 * ne), bc;
  nalu_hip::check(nalu_lattice_generate(n, jitter, 1.0, 0.0, seed, coords.data(), conn.data()), "lattice");
  nalu_hip::check(nalu_lognormal_field(nn, 0.5, seed + 1, k.data()), "conductivity");
  for (int64_t a = 0; a < nn; ++a) {
    const int64_t i = a % (n + 1);
    if (i == 0 || i == n) {
      bc.push_back((int32_t)a);
      bcv[a] = (i == n) ? 1.0 : 0.0;  // T = 0 at x = 0, T = 1 at x = 1
    }
  }
  try {
    nalu_hip::Mesh mesh(nn, ne, conn.data(), coords.data());
    nalu_hip::SolverConfig cfg;  // gmres(50) + sgs, tolerance 1e-8
    nalu_hip::LinearSystem linsys(1, cfg);
    linsys.buildElemToNodeGraph(mesh);
    linsys.finalizeLinearSystem();

    double *dT, *dk, *dbcv, *dd;
    int32_t* dbc;
    HIPOK(hipMalloc(&dT, sizeof(double) * nn));
    HIPOK(hipMalloc(&dk, sizeof(double) * nn));
    HIPOK(hipMalloc(&dbcv, sizeof(double) * nn));
    HIPOK(hipMalloc(&dd, sizeof(double) * nn));
    HIPOK(hipMalloc(&dbc, sizeof(int32_t) * bc.size()));
    HIPOK(hipMemcpy(dT, T.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dk, k.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dbcv, bcv.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dbc, bc.data(), sizeof(int32_t) * bc.size(), hipMemcpyHostToDevice));

    // the equation system's consolidated element algorithm with its one kernel
    nalu_hip::AssembleElemSolverAlgorithm assemble(linsys, mesh);
    assemble.addKernel(std::make_unique<nalu_hip::ScalarDiffElemKernel>(dT, dk));
    const nalu_hip::DirichletBC bcT{(int64_t)bc.size(), dbc, dT, dbcv};

    const auto t0 = std::chrono::steady_clock::now();
    linsys.zeroSystem();
    assemble.execute();
    linsys.applyDirichletBCs(bcT, 0, 1);
    linsys.loadComplete();
    const int status = linsys.solve(dd);
    HIPOK(hipDeviceSynchronize());
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // T += dT
    std::vector<double> d(nn);
    HIPOK(hipMemcpy(d.data(), dd, sizeof(double) * nn, hipMemcpyDeviceToHost));
    double tmin = 1e300, tmax = -1e300;
    for (int64_t a = 0; a < nn; ++a) {
      T[a] += d[a];
      tmin = std::fmin(tmin, T[a]);
      tmax = std::fmax(tmax, T[a]);
    }
    std::printf("{\"n\": %d, \"status\": %d, \"iterations\": %d, \"linear_residual\": %.3e, "
                "\"nonlinear_residual\": %.3e, \"T_min\": %.6f, \"T_max\": %.6f, \"seconds\": %.4f}\n",
                n, status, linsys.linearSolveIterations(), linsys.linearResidual(), linsys.nonLinearResidual(),
                tmin, tmax, secs);
    (void)hipFree(dT);
    (void)hipFree(dk);
    (void)hipFree(dbcv);
    (void)hipFree(dd);
    (void)hipFree(dbc);
    return status == 0 ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
}
