// Host side of the C ABI (include/nalu_hip.h) under AddressSanitizer and
// UndefinedBehaviorSanitizer, on a machine without a GPU: the library is built
// host-only by tests/sanitize/Makefile, and this driver runs every path that
// needs no device -- the lattice and field generators, handle create/destroy
// over every solver configuration, parameter validation, and each entry point's
// argument and state checks (null handles, unfinalized systems, sizes out of
// range).  Every call must return its documented error code without a
// sanitizer report.  Exit status 0 means clean.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/nalu_hip.h"

static int g_fail = 0;

static void expect(bool ok, const char* what, int line) {
  if (!ok) {
    std::fprintf(stderr, "host_abi: %s (line %d): %s\n", what, line, nalu_last_error());
    g_fail = 1;
  }
}
#define EXPECT(c) expect((c), #c, __LINE__)

static nalu_solver_config cfg(int method, int precond, int precision = NALU_PRECISION_FP64) {
  nalu_solver_config c{};
  c.method = method;
  c.precond = precond;
  c.tolerance = 1e-8;
  c.final_tolerance = 1e-8;
  c.max_iterations = 100;
  c.kspace = 20;
  c.precision = precision;
  c.final_outer_iter = 0;
  return c;
}

int main() {
  EXPECT(nalu_device_available() == 0);

  // generators
  for (int n : {1, 2, 5}) {
    const size_t nn = (size_t)(n + 1) * (n + 1) * (n + 1), ne = (size_t)n * n * n;
    std::vector<double> X(3 * nn);
    std::vector<int32_t> conn(8 * ne);
    EXPECT(nalu_lattice_generate(n, 0.1, 1.0, 1.0, 7u, X.data(), conn.data()) == NALU_OK);
    EXPECT(nalu_lattice_generate(n, 0.0, 1.2, 100.0, 7u, X.data(), conn.data()) == NALU_OK);
    for (double v : X) EXPECT(std::isfinite(v) && v >= -1e-12 && v <= 1.0 + 1e-12);
    for (int32_t c : conn) EXPECT(c >= 0 && (size_t)c < nn);
  }
  {
    std::vector<double> X(24);
    std::vector<int32_t> conn(8);
    EXPECT(nalu_lattice_generate(0, 0.0, 1.0, 1.0, 1u, X.data(), conn.data()) == NALU_ERR_ARG);
    EXPECT(nalu_lattice_generate(1, 0.0, 1.0, 1.0, 1u, nullptr, conn.data()) == NALU_ERR_ARG);
    std::vector<double> f(1000);
    EXPECT(nalu_normal_field(1000, 1.0, 3u, f.data()) == NALU_OK);
    EXPECT(nalu_lognormal_field(1000, 0.5, 3u, f.data()) == NALU_OK);
    for (double v : f) EXPECT(v > 0.0 && std::isfinite(v));
    EXPECT(nalu_normal_field(10, 1.0, 3u, nullptr) == NALU_ERR_ARG);
  }

  // handle lifetime over every configuration; the learned preconditioner needs FGMRES
  for (int nd : {1, 3})
    for (int m : {NALU_METHOD_GMRES, NALU_METHOD_CG, NALU_METHOD_FGMRES})
      for (int p : {NALU_PREC_NONE, NALU_PREC_JACOBI, NALU_PREC_SGS, NALU_PREC_LEARNED, NALU_PREC_MUELU})
        for (int prec : {NALU_PRECISION_FP64, NALU_PRECISION_MIXED}) {
          const nalu_solver_config c = cfg(m, p, prec);
          nalu_linsys* ls = nullptr;
          const int rc = nalu_linsys_create(nd, &c, &ls);
          if (p == NALU_PREC_LEARNED && m != NALU_METHOD_FGMRES) {
            EXPECT(rc == NALU_ERR_ARG && ls == nullptr);
          } else {
            EXPECT(rc == NALU_OK && ls != nullptr);
            EXPECT(nalu_linsys_destroy(ls) == NALU_OK);
          }
        }
  {
    nalu_solver_config bad = cfg(7, NALU_PREC_SGS);
    nalu_linsys* ls = nullptr;
    EXPECT(nalu_linsys_create(1, &bad, &ls) == NALU_ERR_ARG);
    bad = cfg(NALU_METHOD_GMRES, 9);
    EXPECT(nalu_linsys_create(1, &bad, &ls) == NALU_ERR_ARG);
    bad = cfg(NALU_METHOD_GMRES, NALU_PREC_SGS);
    bad.kspace = 0;
    EXPECT(nalu_linsys_create(1, &bad, &ls) == NALU_ERR_ARG);
    bad = cfg(NALU_METHOD_CG, NALU_PREC_SGS);
    bad.max_iterations = 0;
    EXPECT(nalu_linsys_create(1, &bad, &ls) == NALU_ERR_ARG);
    bad = cfg(NALU_METHOD_CG, NALU_PREC_SGS);
    bad.tolerance = NAN;
    EXPECT(nalu_linsys_create(1, &bad, &ls) == NALU_ERR_ARG);
    const nalu_solver_config ok = cfg(NALU_METHOD_GMRES, NALU_PREC_SGS);
    EXPECT(nalu_linsys_create(0, &ok, &ls) == NALU_ERR_ARG);
    EXPECT(nalu_linsys_create(1, nullptr, &ls) == NALU_ERR_ARG);
    EXPECT(nalu_linsys_create(1, &ok, nullptr) == NALU_ERR_ARG);
  }

  // state and argument checks on a system with no graph
  const nalu_solver_config c = cfg(NALU_METHOD_GMRES, NALU_PREC_MUELU);
  nalu_linsys* ls = nullptr;
  EXPECT(nalu_linsys_create(3, &c, &ls) == NALU_OK);
  nalu_muelu_params mp{1000, 10, 4, 20.0, 1.1, 30, 0.0, 4.0 / 3.0};
  EXPECT(nalu_linsys_set_muelu_params(ls, &mp) == NALU_OK);
  for (int k = 0; k < 5; ++k) {
    nalu_muelu_params b = mp;
    if (k == 0) b.max_coarse_size = 0;
    if (k == 1) b.max_coarse_size = 1 << 20;
    if (k == 2) b.cheb_degree = 0;
    if (k == 3) b.cheb_ratio = 1.0;
    if (k == 4) b.drop_tol = -1.0;
    EXPECT(nalu_linsys_set_muelu_params(ls, &b) == NALU_ERR_ARG);
  }
  EXPECT(nalu_linsys_set_muelu_params(ls, nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_linsys_zero(ls) == NALU_ERR_STATE);
  EXPECT(nalu_linsys_finalize(ls) == NALU_ERR_STATE);
  EXPECT(nalu_linsys_build_elem_to_node_graph(ls, nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_linsys_load_complete(ls) < 0);
  EXPECT(nalu_linsys_remove_rhs_mean(ls) < 0);
  EXPECT(nalu_linsys_solve(ls, nullptr) < 0);
  EXPECT(nalu_linsys_precond_setup(ls) < 0);
  EXPECT(nalu_linsys_set_l2_scaling(ls, 0.5) == NALU_OK);
  EXPECT(nalu_linsys_set_l2_scaling(ls, -1.0) == NALU_ERR_ARG);
  EXPECT(nalu_linsys_iterations(ls) == 0);
  int64_t nr = -1, nz = -1;
  EXPECT(nalu_linsys_sizes(ls, &nr, &nz) < 0 || (nr == 0 && nz == 0));
  double t7[7];
  EXPECT(nalu_linsys_phase_times(ls, t7) == NALU_OK);
  EXPECT(nalu_linsys_phase_times(ls, nullptr) == NALU_ERR_ARG);
  {
    const int32_t ent[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    std::vector<double> rhs(24), lhs(24 * 24);
    EXPECT(nalu_linsys_sum_into(ls, 8, ent, rhs.data(), lhs.data()) < 0);
    EXPECT(nalu_linsys_sum_into(ls, 8, nullptr, rhs.data(), lhs.data()) < 0);
    EXPECT(nalu_linsys_sum_into_elements(ls, 1, 8, ent, rhs.data(), lhs.data(), nullptr, nullptr) < 0);
    EXPECT(nalu_linsys_sum_into_elements_noalloc(ls, 1, 8, ent, rhs.data(), lhs.data()) < 0);
    EXPECT(nalu_linsys_sum_into_elements_noalloc(nullptr, 1, 8, ent, rhs.data(), lhs.data()) < 0);
  }
  EXPECT(nalu_linsys_write_to_file(ls, "unused") < 0);
  EXPECT(nalu_linsys_write_to_file(ls, nullptr) < 0);
  EXPECT(nalu_linsys_set_write_matrix_files(ls, "unused") == NALU_OK);
  EXPECT(nalu_linsys_set_write_matrix_files(ls, nullptr) == NALU_OK);  // off again
  nalu_elem_algorithm* alg = nullptr;
  EXPECT(nalu_elem_algorithm_create(ls, nullptr, &alg) == NALU_ERR_ARG && alg == nullptr);
  EXPECT(nalu_elem_algorithm_add_kernel(nullptr, nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_elem_algorithm_execute(nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_linsys_destroy(ls) == NALU_OK);

  // null handles
  EXPECT(nalu_linsys_destroy(nullptr) == NALU_OK || nalu_linsys_destroy(nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_linsys_zero(nullptr) < 0);
  EXPECT(nalu_linsys_finalize(nullptr) < 0);
  EXPECT(nalu_linsys_solve_rhs(nullptr, nullptr, nullptr, nullptr, nullptr) < 0);
  EXPECT(nalu_linsys_iterations(nullptr) == -1);
  EXPECT(std::isnan(nalu_linsys_linear_residual(nullptr)));
  EXPECT(nalu_mesh_create(0, 0, nullptr, nullptr, nullptr, nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_mesh_exposed_faces(nullptr, nullptr, nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_compute_nodal_gradient(nullptr, nullptr, nullptr, 0, nullptr, nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_gnn_create(0, 8, 1u, nullptr) == NALU_ERR_ARG);
  EXPECT(nalu_gnn_num_params(nullptr) < 0);

  if (!g_fail) std::printf("host_abi: ok\n");
  return g_fail;
}
