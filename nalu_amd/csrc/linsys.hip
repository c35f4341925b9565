// Handle management and the LinearSystem half of the C ABI
// (TpetraLinearSystem replacement, src/TpetraLinearSystem.C).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "internal.hpp"

namespace nalu {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

PhaseTimer::PhaseTimer(nalu_linsys* l, int p) : ls(l), ph(p) {
  if (!ls || !ls->timing) return;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    a = b = nullptr;
    return;
  }
  (void)hipEventRecord(a, ls->stream);
}

PhaseTimer::~PhaseTimer() {
  if (!a) return;
  (void)hipEventRecord(b, ls->stream);
  ls->pending_timers.push_back({ph, a, b});
  if (ls->pending_timers.size() > 256) flush_timers(ls);
}

void flush_timers(nalu_linsys* ls) {
  for (auto& t : ls->pending_timers) {
    float ms = 0.f;
    if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess)
      ls->phase[t.phase] += 1e-3 * ms;
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  ls->pending_timers.clear();
}

}  // namespace nalu

using namespace nalu;

extern "C" const char* nalu_last_error(void) { return g_last_error.c_str(); }

extern "C" int nalu_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

// ------------------------------------------------------------------ mesh
extern "C" int nalu_mesh_create(int64_t nn, int64_t ne, const int32_t* conn, const double* coords,
                                const int64_t* ids, nalu_mesh** out) {
  return guarded([&]() -> int {
    NALU_REQUIRE(out && conn && coords && nn > 0 && ne > 0, NALU_ERR_ARG, "mesh_create: bad arguments");
    *out = nullptr;
    for (int64_t i = 0; i < 8 * ne; ++i)
      NALU_REQUIRE(conn[i] >= 0 && conn[i] < nn, NALU_ERR_ARG, "mesh_create: connectivity index out of range");
    if (ids)
      for (int64_t i = 0; i < nn; ++i)
        NALU_REQUIRE(ids[i] > 0 && ids[i] < ((int64_t)1 << 32), NALU_ERR_ARG, "mesh_create: ids must be in [1, 2^32)");
    auto* m = new nalu_mesh();
    try {
      m->num_nodes = nn;
      m->num_elems = ne;
      m->coords.alloc(3 * nn);
      m->conn.alloc(8 * ne);
      NALU_HIP_CHECK(hipMemcpy(m->coords.p, coords, sizeof(double) * 3 * nn, hipMemcpyHostToDevice));
      NALU_HIP_CHECK(hipMemcpy(m->conn.p, conn, sizeof(int32_t) * 8 * ne, hipMemcpyHostToDevice));
      if (ids) {
        m->ids.alloc(nn);
        NALU_HIP_CHECK(hipMemcpy(m->ids.p, ids, sizeof(int64_t) * nn, hipMemcpyHostToDevice));
      }
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
    return NALU_OK;
  });
}

extern "C" int nalu_mesh_destroy(nalu_mesh* m) {
  delete m;
  return NALU_OK;
}

extern "C" int nalu_mesh_device_ptrs(const nalu_mesh* m, const double** coords, const int32_t** conn) {
  return guarded([&]() -> int {
    NALU_REQUIRE(m, NALU_ERR_ARG, "mesh_device_ptrs: null mesh");
    if (coords) *coords = m->coords.p;
    if (conn) *conn = m->conn.p;
    return NALU_OK;
  });
}

// --------------------------------------------------------- linear system
extern "C" int nalu_linsys_create(int num_dof, const nalu_solver_config* cfg, nalu_linsys** out) {
  return guarded([&]() -> int {
    NALU_REQUIRE(out && cfg, NALU_ERR_ARG, "linsys_create: null argument");
    NALU_REQUIRE(num_dof >= 1 && num_dof <= 8, NALU_ERR_ARG, "linsys_create: numDof out of range");
    NALU_REQUIRE(cfg->method >= NALU_METHOD_GMRES && cfg->method <= NALU_METHOD_FGMRES, NALU_ERR_ARG,
                 "linsys_create: unknown method");
    NALU_REQUIRE(cfg->precond >= NALU_PREC_NONE && cfg->precond <= NALU_PREC_MUELU, NALU_ERR_ARG,
                 "linsys_create: unknown preconditioner");
    NALU_REQUIRE(cfg->precision == NALU_PRECISION_FP64 || cfg->precision == NALU_PRECISION_MIXED, NALU_ERR_ARG,
                 "linsys_create: unknown precision");
    NALU_REQUIRE(cfg->kspace >= 1 && cfg->kspace <= 1000, NALU_ERR_ARG, "linsys_create: kspace must be in [1, 1000]");
    NALU_REQUIRE(cfg->max_iterations >= 1, NALU_ERR_ARG, "linsys_create: max_iterations must be at least 1");
    NALU_REQUIRE(cfg->tolerance >= 0.0 && cfg->final_tolerance >= 0.0 && std::isfinite(cfg->tolerance) &&
                     std::isfinite(cfg->final_tolerance),
                 NALU_ERR_ARG, "linsys_create: tolerances must be finite and non-negative");
    // the network is a nonlinear operator: only flexible GMRES admits a
    // preconditioner that changes from one application to the next
    NALU_REQUIRE(cfg->precond != NALU_PREC_LEARNED || cfg->method == NALU_METHOD_FGMRES, NALU_ERR_ARG,
                 "linsys_create: the learned preconditioner needs method fgmres");
    auto* ls = new nalu_linsys();
    ls->num_dof = num_dof;
    ls->cfg = *cfg;
    *out = ls;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_destroy(nalu_linsys* ls) {
  if (!ls) return NALU_OK;
  if (ls->stream) (void)hipStreamSynchronize(ls->stream);
  else (void)hipDeviceSynchronize();
  flush_timers(ls);
  learned_release(ls);
  amg_release(ls);
  etile_release(ls);
  release_sum_into_queue(ls);
  delete ls;
  return NALU_OK;
}

extern "C" int nalu_linsys_set_stream(nalu_linsys* ls, void* stream) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls, NALU_ERR_ARG, "set_stream: null handle");
    ls->stream = (hipStream_t)stream;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_build_elem_to_node_graph(nalu_linsys* ls, const nalu_mesh* m) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && m, NALU_ERR_ARG, "buildElemToNodeGraph: null argument");
    NALU_REQUIRE(!ls->finalized, NALU_ERR_STATE, "buildElemToNodeGraph: system already finalized");
    NALU_REQUIRE(!ls->mesh || ls->mesh == m, NALU_ERR_UNSUPPORTED,
                 "buildElemToNodeGraph: one Hex8 block per linear system is supported");
    ls->mesh = m;
    ls->in_construction = true;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_finalize(nalu_linsys* ls) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls, NALU_ERR_ARG, "finalizeLinearSystem: null handle");
    NALU_REQUIRE(ls->in_construction && ls->mesh, NALU_ERR_STATE, "finalizeLinearSystem: no graph was built");
    PhaseTimer t(ls, PH_INIT);
    build_graph(ls, ls->stream);
    ls->in_construction = false;
    ls->finalized = true;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_zero(nalu_linsys* ls) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "zeroSystem: system not finalized");
    PhaseTimer t(ls, PH_ZERO);
    // values/rhs are zeroed lazily: the first owner-computes assembly writes
    // its rows instead of accumulating; anything else materializes the zero.
    discard_sum_into_queue(ls);
    ls->values_pending_zero = true;
    ls->rhs_pending_zero = true;
    ls->precond_ready = false;
    NALU_HIP_CHECK(hipMemsetAsync(ls->sln.p, 0, ls->sln.bytes(), ls->stream));
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_load_complete(nalu_linsys* ls) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "loadComplete: system not finalized");
    PhaseTimer t(ls, PH_LOADCOMPLETE);
    // single domain: no shared-not-owned rows to export
    materialize_pending_zero(ls);
    return NALU_OK;
  });
}

namespace nalu {
namespace {

// Global row index of every row, 1-based, as Tpetra numbers them:
// GID = numDof (nodeId - 1) + dof + 1 (src/TpetraLinearSystem.C:72).
std::vector<int64_t> row_gids(nalu_linsys* ls) {
  const int64_t N = ls->num_nodes, nd = ls->num_dof;
  std::vector<int64_t> id(N);
  if (ls->mesh->ids.p) {
    std::vector<int64_t> ids(N);
    std::vector<int32_t> rn(N);
    NALU_HIP_CHECK(hipMemcpy(ids.data(), ls->mesh->ids.p, sizeof(int64_t) * N, hipMemcpyDeviceToHost));
    NALU_HIP_CHECK(hipMemcpy(rn.data(), ls->row_node.p, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    for (int64_t l = 0; l < N; ++l) id[l] = ids[rn[l]];
  } else {
    for (int64_t l = 0; l < N; ++l) id[l] = l + 1;
  }
  std::vector<int64_t> g(N * nd);
  for (int64_t l = 0; l < N; ++l)
    for (int64_t d = 0; d < nd; ++d) g[l * nd + d] = nd * (id[l] - 1) + d + 1;
  return g;
}

std::string file_name(const nalu_linsys* ls, const char* kind) {
  return ls->write_base + "-O-" + std::to_string(ls->write_counter) + "." + kind + ".1";
}

void write_dense(const std::string& path, const std::vector<double>& v) {
  FILE* f = std::fopen(path.c_str(), "w");
  NALU_REQUIRE(f, NALU_ERR_ARG, "cannot open " + path);
  std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%lld 1\n", (long long)v.size());
  for (double x : v) std::fprintf(f, "%.17g\n", x);
  std::fclose(f);
}

// TpetraLinearSystem::writeToFile (src/TpetraLinearSystem.C:1969-1994): the owned
// matrix as a MatrixMarket coordinate file and the rhs as a dense array file.
void write_matrix_files(nalu_linsys* ls) {
  materialize_pending_zero(ls);
  NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
  const int64_t n = ls->nrows, nnz = ls->nnz;
  std::vector<int32_t> rp(n + 1), ci(nnz);
  std::vector<double> v(nnz), r(n);
  NALU_HIP_CHECK(hipMemcpy(rp.data(), ls->row_ptr.p, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  NALU_HIP_CHECK(hipMemcpy(ci.data(), ls->col_idx.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
  NALU_HIP_CHECK(hipMemcpy(v.data(), ls->values.p, sizeof(double) * nnz, hipMemcpyDeviceToHost));
  NALU_HIP_CHECK(hipMemcpy(r.data(), ls->rhs.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  const std::vector<int64_t> gid = row_gids(ls);
  const std::string path = file_name(ls, "mm");
  FILE* f = std::fopen(path.c_str(), "w");
  NALU_REQUIRE(f, NALU_ERR_ARG, "cannot open " + path);
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%% %s\n%% matrix for: %s\n",
               ls->write_base.c_str(), ls->write_base.c_str());
  std::fprintf(f, "%lld %lld %lld\n", (long long)n, (long long)n, (long long)nnz);
  for (int64_t i = 0; i < n; ++i)
    for (int32_t q = rp[i]; q < rp[i + 1]; ++q)
      std::fprintf(f, "%lld %lld %.17g\n", (long long)gid[i], (long long)gid[ci[q]], v[q]);
  std::fclose(f);
  write_dense(file_name(ls, "rhs"), r);
}

// TpetraLinearSystem::writeSolutionToFile (src/TpetraLinearSystem.C:2070-2090)
void write_solution_file(nalu_linsys* ls) {
  NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
  std::vector<double> x(ls->nrows);
  NALU_HIP_CHECK(hipMemcpy(x.data(), ls->sln.p, sizeof(double) * ls->nrows, hipMemcpyDeviceToHost));
  write_dense(file_name(ls, "sln"), x);
}

}  // namespace
}  // namespace nalu

// writeToFile / writeSolutionToFile on demand (include/LinearSystem.h:126-127):
// the same files as the solve-time output, named from base_name.
extern "C" int nalu_linsys_write_to_file(nalu_linsys* ls, const char* base_name) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && base_name && *base_name, NALU_ERR_ARG, "writeToFile: bad arguments");
    const std::string keep = ls->write_base;
    ls->write_base = base_name;
    try {
      write_matrix_files(ls);
    } catch (...) {
      ls->write_base = keep;
      throw;
    }
    ls->write_base = keep;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_write_solution_to_file(nalu_linsys* ls, const char* base_name) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && base_name && *base_name, NALU_ERR_ARG, "writeSolutionToFile: bad arguments");
    const std::string keep = ls->write_base;
    ls->write_base = base_name;
    try {
      write_solution_file(ls);
    } catch (...) {
      ls->write_base = keep;
      throw;
    }
    ls->write_base = keep;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_set_write_matrix_files(nalu_linsys* ls, const char* base_name) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls, NALU_ERR_ARG, "set_write_matrix_files: null handle");
    ls->write_base = base_name ? std::string(base_name) : std::string();
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_solve(nalu_linsys* ls, double* delta_out) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "solve: system not finalized");
    materialize_pending_zero(ls);
    if (!ls->write_base.empty()) write_matrix_files(ls);
    {
      PhaseTimer t(ls, PH_PRECOND);
      precond_setup(ls);
    }
    int status = 0;
    {
      PhaseTimer t(ls, PH_SOLVE);
      NALU_HIP_CHECK(hipMemsetAsync(ls->sln.p, 0, ls->sln.bytes(), ls->stream));
      int it = 0;
      double rel = 0.0;
      status = krylov_solve(ls, ls->rhs.p, ls->sln.p, &it, &rel);
      ls->iters = it;
    }
    if (!ls->write_base.empty()) {
      write_solution_file(ls);
      ++ls->write_counter;
    }
    if (delta_out)
      NALU_HIP_CHECK(hipMemcpyAsync(delta_out, ls->sln.p, ls->sln.bytes(), hipMemcpyDeviceToDevice, ls->stream));
    // nonlinear residual = l2Scaling * ||rhs|| (src/TpetraLinearSystem.C:1812-1820)
    const double nr = norm2(ls, ls->rhs.p, ls->nrows);
    ls->nonlinear_residual = ls->l2_scaling * nr;
    if (ls->first_nonlinear_residual < 0) ls->first_nonlinear_residual = ls->nonlinear_residual;
    ls->scaled_nonlinear_residual =
        ls->nonlinear_residual / std::max(std::numeric_limits<double>::epsilon(), ls->first_nonlinear_residual);
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    return status;
  });
}

extern "C" int nalu_linsys_solve_rhs(nalu_linsys* ls, const double* b, double* x, int* iters, double* rel) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && b && x, NALU_ERR_ARG, "solve_rhs: bad arguments");
    if (!ls->precond_ready) {
      PhaseTimer t(ls, PH_PRECOND);
      precond_setup(ls);
    }
    PhaseTimer t(ls, PH_SOLVE);
    int it = 0;
    double r = 0.0;
    const int st = krylov_solve(ls, b, x, &it, &r);
    ls->iters = it;
    if (iters) *iters = it;
    if (rel) *rel = r;
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    return st;
  });
}

// k right-hand sides sharing the system's matrix and preconditioner set-up
// (B, X: [k][nrows] device arrays; iters, rel: [k] host arrays or NULL).
extern "C" int nalu_linsys_solve_multi(nalu_linsys* ls, int k, const double* B, double* X, int* iters, double* rel) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && B && X && k >= 1, NALU_ERR_ARG, "solve_multi: bad arguments");
    if (!ls->precond_ready) {
      PhaseTimer t(ls, PH_PRECOND);
      precond_setup(ls);
    }
    PhaseTimer t(ls, PH_SOLVE);
    int status = 0, total = 0;
    for (int i = 0; i < k; ++i) {
      int it = 0;
      double r = 0.0;
      status = std::max(status, krylov_solve(ls, B + (size_t)i * ls->nrows, X + (size_t)i * ls->nrows, &it, &r));
      total += it;
      if (iters) iters[i] = it;
      if (rel) rel[i] = r;
    }
    ls->iters = total;
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    return status;
  });
}

extern "C" int nalu_linsys_solve_stacked(nalu_linsys* ls, int k, const double* B, double* X, int* iters,
                                         double* rel) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && B && X && k >= 1, NALU_ERR_ARG, "solve_stacked: bad arguments");
    if (!ls->precond_ready) {
      PhaseTimer t(ls, PH_PRECOND);
      precond_setup(ls);
    }
    PhaseTimer t(ls, PH_SOLVE);
    int it = 0;
    double r = 0.0;
    const int status = krylov_solve_k(ls, k, B, X, &it, &r, rel);
    ls->iters = it;
    if (iters) *iters = it;
    return status;
  });
}

extern "C" int nalu_linsys_iterations(const nalu_linsys* ls) { return ls ? ls->iters : -1; }
extern "C" double nalu_linsys_linear_residual(const nalu_linsys* ls) { return ls ? ls->linear_residual : NAN; }
extern "C" double nalu_linsys_nonlinear_residual(const nalu_linsys* ls) { return ls ? ls->nonlinear_residual : NAN; }
extern "C" double nalu_linsys_scaled_nonlinear_residual(const nalu_linsys* ls) {
  return ls ? ls->scaled_nonlinear_residual : NAN;
}
extern "C" int nalu_linsys_set_l2_scaling(nalu_linsys* ls, double s) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && s > 0, NALU_ERR_ARG, "set_l2_scaling: bad arguments");
    ls->l2_scaling = s;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_sizes(const nalu_linsys* ls, int64_t* nrows, int64_t* nnz) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "sizes: system not finalized");
    if (nrows) *nrows = ls->nrows;
    if (nnz) *nnz = ls->nnz;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_copy_csr(nalu_linsys* ls, int64_t* rp, int32_t* ci, double* vals, double* rhs) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "copy_csr: system not finalized");
    materialize_pending_zero(ls);
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    if (rp) {
      std::vector<int32_t> t(ls->nrows + 1);
      NALU_HIP_CHECK(hipMemcpy(t.data(), ls->row_ptr.p, sizeof(int32_t) * t.size(), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < t.size(); ++i) rp[i] = t[i];
    }
    if (ci) NALU_HIP_CHECK(hipMemcpy(ci, ls->col_idx.p, ls->col_idx.bytes(), hipMemcpyDeviceToHost));
    if (vals) NALU_HIP_CHECK(hipMemcpy(vals, ls->values.p, ls->values.bytes(), hipMemcpyDeviceToHost));
    if (rhs) NALU_HIP_CHECK(hipMemcpy(rhs, ls->rhs.p, ls->rhs.bytes(), hipMemcpyDeviceToHost));
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_device_ptrs(nalu_linsys* ls, const int32_t** rp, const int32_t** ci, double** vals,
                                       double** rhs) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "device_ptrs: system not finalized");
    materialize_pending_zero(ls);
    if (rp) *rp = ls->row_ptr.p;
    if (ci) *ci = ls->col_idx.p;
    if (vals) *vals = ls->values.p;
    if (rhs) *rhs = ls->rhs.p;
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_coloring(nalu_linsys* ls, int* nc, int64_t* cp, int32_t* cr) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "coloring: system not finalized");
    if (ls->num_colors == 0) build_coloring(ls, ls->stream);
    if (nc) *nc = ls->num_colors;
    if (cp)
      for (int c = 0; c <= ls->num_colors; ++c) cp[c] = ls->color_ptr_host[c];
    if (cr) NALU_HIP_CHECK(hipMemcpy(cr, ls->color_rows.p, ls->color_rows.bytes(), hipMemcpyDeviceToHost));
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_phase_times(const nalu_linsys* ls, double* t7) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && t7, NALU_ERR_ARG, "phase_times: null argument");
    flush_timers(const_cast<nalu_linsys*>(ls));
    for (int i = 0; i < PH_COUNT; ++i) t7[i] = ls->phase[i];
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_spmv(nalu_linsys* ls, const double* x, double* y) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && x && y, NALU_ERR_ARG, "spmv: bad arguments");
    materialize_pending_zero(ls);
    spmv(ls, x, y);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_precond_setup(nalu_linsys* ls) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "precond_setup: system not finalized");
    PhaseTimer t(ls, PH_PRECOND);
    precond_setup(ls);
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_precond_apply(nalu_linsys* ls, const double* r, double* z) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && r && z, NALU_ERR_ARG, "precond_apply: bad arguments");
    precond_apply(ls, r, z);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}
