// Internal definitions shared by the translation units of libnalu_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nalu_hip.h"

namespace nalu {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define NALU_HIP_CHECK(expr)                                                              \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      throw ::nalu::Error(NALU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e) + \
                                            " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define NALU_REQUIRE(cond, code, msg)                       \
  do {                                                      \
    if (!(cond)) throw ::nalu::Error((code), (msg));        \
  } while (0)

// Runs f() and converts exceptions into status codes + last-error text.
template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return NALU_ERR_ARG;
  }
}

// -------------------------------------------------------- device buffers
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count) NALU_HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Sub-buffers carved from one device workspace at 256-byte boundaries.  To size
// a workspace, carve the sequence from a null base first (off is then the byte
// count), then carve the same sequence from the real base.
struct Carve {
  unsigned char* base = nullptr;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) / 256 * 256;
    T* q = reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(base) + off);
    off += count * sizeof(T);
    return q;
  }
};

// Grows a byte workspace owned by a handle to at least `bytes` (never shrinks;
// growing frees the old buffer, which synchronizes the device).
inline unsigned char* ensure_ws(DevBuf<unsigned char>& ws, size_t bytes) {
  if (ws.n < bytes) ws.alloc(bytes);
  return ws.p;
}

// Page-locked host buffer (released with its owner).
struct HostPinned {
  void* p = nullptr;
  size_t n = 0;
  HostPinned() = default;
  HostPinned(const HostPinned&) = delete;
  HostPinned& operator=(const HostPinned&) = delete;
  ~HostPinned() {
    if (p) (void)hipHostFree(p);
  }
  void alloc(size_t bytes) {
    if (bytes <= n) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    NALU_HIP_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    n = bytes;
  }
};

// ---------------------------------------------------------------- timers
enum Phase { PH_INIT = 0, PH_ZERO, PH_ASSEMBLE, PH_DIRICHLET, PH_LOADCOMPLETE, PH_PRECOND, PH_SOLVE, PH_COUNT };

struct PendingTimer {
  int phase;
  hipEvent_t a, b;
};

}  // namespace nalu

// ------------------------------------------------------------------- mesh
struct nalu_mesh {
  int64_t num_nodes = 0, num_elems = 0;
  nalu::DevBuf<double> coords;   // [N][3]
  nalu::DevBuf<int32_t> conn;    // [E][8]
  nalu::DevBuf<int64_t> ids;     // [N] global ids (empty: id = index+1)
  // node -> element incidence (built on demand by the graph build)
  nalu::DevBuf<int32_t> inc_ptr;   // [N+1]
  nalu::DevBuf<int32_t> inc_elem;  // [8E] element*8 + local node index, sorted per node
  int max_incidence = 0;
  // exposed sides (built on demand): [num_exposed][4] nodes in HexSCS side order
  bool exposed_ready = false;
  int64_t num_exposed = 0;
  nalu::DevBuf<int32_t> exposed_faces;
};

// ---------------------------------------------------------- linear system
struct nalu_linsys {
  int num_dof = 1;
  nalu_solver_config cfg{};
  hipStream_t stream = nullptr;

  // construction state
  const nalu_mesh* mesh = nullptr;  // the mesh whose elements define the graph
  bool in_construction = false;
  bool finalized = false;

  int64_t num_nodes = 0;
  int64_t nrows = 0;
  int64_t nnz = 0;
  int max_node_row_len = 0;  // longest node-graph row (neighbors incl. self)

  // node graph (numDof = 1): rows ordered by global id
  nalu::DevBuf<int32_t> node_lid;  // [N] rank of node id (row of dof 0 = num_dof*lid)
  nalu::DevBuf<int32_t> nrp;       // node graph row_ptr [N+1]
  nalu::DevBuf<int32_t> nci;       // node graph col idx (node lids) [P]
  // element scatter table: for element e, local row node a: 8 bytes, byte c =
  // position of column node c in node-row of node a.
  nalu::DevBuf<uint64_t> elem_pos; // [E][8]

  // expanded CSR (numDof rows interleaved, src/TpetraLinearSystem.C:1227-1246)
  nalu::DevBuf<int32_t> row_ptr;   // [nrows+1]
  nalu::DevBuf<int32_t> col_idx;   // [nnz]
  nalu::DevBuf<int32_t> diag_pos;  // [nrows]
  nalu::DevBuf<double> values;     // [nnz]
  nalu::DevBuf<double> rhs;        // [nrows]
  nalu::DevBuf<double> sln;        // [nrows]
  bool values_pending_zero = false;  // zeroSystem deferred into the first gather assembly
  bool rhs_pending_zero = false;

  // multicolor ordering for SGS
  int num_colors = 0;
  std::vector<int64_t> color_ptr_host;
  nalu::DevBuf<int32_t> color_rows;  // rows sorted by color

  // solver workspace
  nalu::DevBuf<double> red;        // reduction partials
  nalu::DevBuf<double> red_out;    // reduction results
  std::vector<double> red_host;
  bool precond_ready = false;

  // color-permuted copy of the matrix used inside the solvers: rows grouped by
  // color, each row split into L (columns of earlier colors) and U (later
  // colors) parts plus the diagonal; the sources index the CSR values.
  bool perm_ready = false;
  nalu::DevBuf<int32_t> perm, iperm, cpd;        // new->old row, old->new row, color ranges
  nalu::DevBuf<int32_t> lrp, lci, lsrc, urp, uci, usrc, dsrc;
  nalu::DevBuf<double> lval, uval, dval, pinvd;
  nalu::DevBuf<float> lval32, uval32, dval32, pinvd32;
  bool split32_ready = false;
  int max_l_len = 0, max_u_len = 0;
  // the same split system in SELL-64 form: chunks of <= 64 consecutive rows of
  // one color; entry k of row row0 + r of a chunk sits at off + 64 k + r
  // (column-major in the chunk, padded with zero entries to the chunk's width)
  bool sell_ready = false;
  int64_t sell_nchunks = 0;
  nalu::DevBuf<unsigned char> sell_chunks;  // [nchunks] chunk descriptors (krylov.hip)
  std::vector<int64_t> sell_color_ptr;      // chunks of color c: [ptr[c], ptr[c+1])
  // all-color kernels visit the chunks in this order: fine spatial slabs, each
  // with its chunks of every color, so that neighbouring rows are touched close
  // together in time; the 8 XCDs take contiguous eighths of the order
  nalu::DevBuf<int32_t> sell_sched;
  nalu::DevBuf<int32_t> sl_col, su_col, sl_src, su_src;
  nalu::DevBuf<double> sl_val, su_val;
  nalu::DevBuf<float> sl_val32, su_val32;
  nalu::DevBuf<double> pwork;  // permuted in/out vectors of precond_apply
  // workspace of the fused CG (vectors, partial sums, device state), kept
  // across solves, and its page-locked host mirror of the state
  nalu::DevBuf<unsigned char> cg_work;
  nalu::HostPinned cg_host;
  // workspaces kept across solves: the permuted outer vectors of a solve
  // (right-hand side, solution, residual, refinement correction), and the
  // inner solver's (Krylov basis, residual, Hessenberg buffers; or the
  // vectors of the kernel-per-step CG) with its page-locked host buffer
  nalu::DevBuf<unsigned char> solve_ws;
  nalu::DevBuf<unsigned char> krylov_ws;
  nalu::HostPinned krylov_host;
  // layout of the vectors of a stacked multi-rhs solve while it runs: false =
  // [nrhs][rows] slices, true = interleaved [rows][nrhs] (batched kernels)
  bool stack_interleaved = false;

  // learned preconditioner (optional)
  void* learned = nullptr;
  // its base preconditioner: the input feature, the network's initial output
  // and the fallback (NALU_PREC_SGS or NALU_PREC_MUELU)
  int learned_base = NALU_PREC_SGS;
  // element-visit tile plans of the assembly (etile.hip)
  void* etile_plans = nullptr;
  // queued calls of the host (std::vector) sumInto overload (suminto.hip)
  void* sum_into_queue = nullptr;
  // AMG preconditioner: parameters and hierarchy (amg.hip)
  nalu_muelu_params muelu{1000, 10, 4, 20.0, 1.1, 30, 0.0, 4.0 / 3.0};
  void* amg = nullptr;

  // matrix-file output around solve() (LinearSolverConfig "write_matrix_files")
  std::string write_base;  // empty: off
  int write_counter = 0;

  // results
  int iters = 0;
  double linear_residual = 0.0;
  double nonlinear_residual = 0.0;
  double first_nonlinear_residual = -1.0;
  double scaled_nonlinear_residual = 0.0;
  double l2_scaling = 1.0;
  double phase[nalu::PH_COUNT] = {0};
  bool timing = true;
  std::vector<nalu::PendingTimer> pending_timers;
  nalu::DevBuf<int32_t> row_node;  // inverse of node_lid when global ids are explicit
};

namespace nalu {
// hipEvent pair around one API phase; accumulated lazily (no synchronization
// on the hot path) -- the reference's per-equation timers, src/EquationSystem.C:270-292.
struct PhaseTimer {
  nalu_linsys* ls;
  int ph;
  hipEvent_t a = nullptr, b = nullptr;
  PhaseTimer(nalu_linsys* l, int p);
  ~PhaseTimer();
};
void flush_timers(nalu_linsys* ls);

// One chunk of the SELL-64 solver copy: rows [row0, row0 + nrows) of one color;
// L entry k of row row0 + r at loff + 64 k + r (lw entries), U likewise.
struct SellChunk {
  int32_t row0, nrows, lw, uw;
  int64_t loff, uoff;
};

// XCD-aware block mapping: hardware dispatches block b to XCD b % 8; this
// returns the position of block b in a work order split into 8 contiguous
// ranges of nper blocks (one per XCD).  Blocks past nb return -1.
constexpr int kNumXcd = 8;
__host__ __device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t nper = (nb + kNumXcd - 1) / kNumXcd;
  const int64_t pos = (b % kNumXcd) * nper + b / kNumXcd;
  return pos < nb ? pos : -1;
}
inline int64_t xcd_grid(int64_t nb) { return (nb + kNumXcd - 1) / kNumXcd * kNumXcd; }

// graph.hip
void build_node_incidence(nalu_mesh* m, hipStream_t s);
void build_graph(nalu_linsys* ls, hipStream_t s);
void build_coloring(nalu_linsys* ls, hipStream_t s);
// assemble.hip
void materialize_pending_zero(nalu_linsys* ls);
void set_lds_limit(const void* kern, size_t bytes);
// etile.hip: element-visit tile assembly
void etile_scalar_diff(nalu_linsys* ls, const nalu_mesh* m, const double* q, const double* k, bool shifted, bool fp32);
void etile_continuity(nalu_linsys* ls, const nalu_mesh* m, const double* rho, const double* p, const double* u,
                      const double* gp, double tau, int key, bool fp32);
void etile_momentum(nalu_linsys* ls, const nalu_mesh* m, const double* u, const double* mu, const double* mdot,
                    double include_divu, bool skew, bool shifted, bool fp32);
void etile_release(nalu_linsys* ls);
// nodalgrad.hip: exposed sides and the boundary closure of the nodal gradient
void build_exposed_faces(nalu_mesh* m, hipStream_t s);
void nodal_grad_boundary(nalu_mesh* m, const double* q, const double* vol, bool shifted, double* dqdx, hipStream_t s);
// suminto.hip: queue of host sumInto calls, applied before anything else touches the system
void flush_sum_into_queue(nalu_linsys* ls);
void discard_sum_into_queue(nalu_linsys* ls);
void release_sum_into_queue(nalu_linsys* ls);
// krylov.hip
void precond_setup(nalu_linsys* ls);
void precond_apply(nalu_linsys* ls, const double* r, double* z);
void spmv(nalu_linsys* ls, const double* x, double* y);
int krylov_solve(nalu_linsys* ls, const double* b, double* x, int* iters, double* final_rel_resid);
int krylov_solve_k(nalu_linsys* ls, int nrhs, const double* b, double* x, int* iters, double* final_rel_resid,
                   double* rel_each);
double norm2(nalu_linsys* ls, const double* x, int64_t n);
// solver-order (color-permuted) building blocks, krylov.hip
void perm_sgs32(nalu_linsys* ls, const float* r, float* z);
void perm_sgs64(nalu_linsys* ls, const double* r, double* z);
void perm_spmv64(nalu_linsys* ls, const double* x, double* y);
void ensure_split32(nalu_linsys* ls);
// learned.hip
void learned_setup(nalu_linsys* ls);
bool learned_active(const nalu_linsys* ls);
bool learned_fallback(const nalu_linsys* ls);
void learned_set_fallback(nalu_linsys* ls, bool on);
void learned_apply_f32(nalu_linsys* ls, const float* r, float* z);
void learned_apply_f64(nalu_linsys* ls, const double* r, double* z);
void learned_release(nalu_linsys* ls);
// amg.hip
void amg_setup(nalu_linsys* ls);
template <typename T>
void amg_apply(nalu_linsys* ls, const T* r, T* z);
void amg_release(nalu_linsys* ls);
// krylov.hip: r = b - A x in solver order (SELL-64 copy)
template <typename T>
void perm_residual(nalu_linsys* ls, const T* x, T* r, const T* b);
}  // namespace nalu
