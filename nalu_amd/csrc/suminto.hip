// The two LinearSystem::sumInto overloads of the reference, for callers that
// produce dense element matrices themselves (the fused assembly kernels never
// materialize them).
//
//  * Kokkos overload (TpetraLinearSystem::sumInto, src/TpetraLinearSystem.C:1457-1522,
//    reached from SolverAlgorithm::apply_coeff, src/SolverAlgorithm.C:58-68):
//    per element, localIds = column LIDs of the entities' dofs and
//    sortPermutation = identity, both sorted by LID; then every row walks its
//    sorted CSR row once (sum_into_row / sum_into_row_vec_3, :1376-1453) and
//    adds atomically (the reference forces atomics off the Serial space).  The
//    device entry takes a batch of elements with caller-owned scratch, which it
//    overwrites exactly as the reference's team scratch views are overwritten.
//  * std::vector overload (:1524-1577): one element per call from host arrays.
//    Calls are copied into a host queue and applied on the device in one batch
//    when anything else touches the system, so a per-element caller costs a
//    host memcpy per element instead of device round trips.
#include <algorithm>
#include <mutex>

#include "internal.hpp"

namespace nalu {
namespace {

constexpr int kMaxRows = 64;  // numEntities * numDof of one element (Hex8, numDof <= 8)

// Phase 1, one thread per element: localIds and sortPermutation (keys sorted
// ascending with their permutation; ids within an element are distinct).
__global__ void k_elem_ids(const int32_t* __restrict__ ents, int64_t nelem, int nent, int nd,
                           const int32_t* __restrict__ lid, int64_t nnodes, int32_t* __restrict__ ids,
                           int32_t* __restrict__ perm, int* __restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int R = nent * nd;
  int32_t* id = ids + e * R;
  int32_t* pm = perm + e * R;
  for (int i = 0; i < nent; ++i) {
    const int32_t node = ents[e * nent + i];
    const bool ok = node >= 0 && node < nnodes;
    if (!ok) atomicOr(bad, 1);
    const int32_t base = ok ? nd * lid[node] : -1;
    for (int d = 0; d < nd; ++d) {
      id[i * nd + d] = base < 0 ? -1 : base + d;
      pm[i * nd + d] = i * nd + d;
    }
  }
  // insertion sort of (id, perm) by id
  for (int i = 1; i < R; ++i) {
    const int32_t k = id[i], v = pm[i];
    int j = i - 1;
    while (j >= 0 && id[j] > k) {
      id[j + 1] = id[j];
      pm[j + 1] = pm[j];
      --j;
    }
    id[j + 1] = k;
    pm[j + 1] = v;
  }
}

// Phase 2, one thread per (element, sorted row r): sum_into_row into the CSR.
__global__ void k_elem_rows(const int32_t* __restrict__ ents, int64_t nelem, int nent, int nd,
                            const int32_t* __restrict__ lid, const int32_t* __restrict__ ids,
                            const int32_t* __restrict__ perm, const double* __restrict__ lhs,
                            const double* __restrict__ rhs_e, const int32_t* __restrict__ rp,
                            const int32_t* __restrict__ ci, double* __restrict__ vals, double* __restrict__ rhs) {
  const int R = nent * nd;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nelem * R) return;
  const int64_t e = t / R;
  const int r = (int)(t - e * R);
  const int32_t* id = ids + e * R;
  const int32_t* pm = perm + e * R;
  if (id[0] < 0) return;  // an entity outside the mesh (reported by phase 1)
  const int32_t pr = pm[r];
  const int64_t row = (int64_t)nd * lid[ents[e * nent + pr / nd]] + pr % nd;
  const double* cur = lhs + (e * R + pr) * (int64_t)R;
  const int32_t b0 = rp[row], len = rp[row + 1] - b0;
  if (nd == 3) {
    // sum_into_row_vec_3: dof triples of a column node are contiguous
    int32_t off = 0;
    for (int j = 0; j < nent; ++j) {
      const int32_t col = id[3 * j];
      while (ci[b0 + off] != col) {
        off += 3;
        if (off >= len) goto rhs_part;
      }
      const int32_t eo = pm[3 * j];
      unsafeAtomicAdd(vals + b0 + off + 0, cur[eo + 0]);
      unsafeAtomicAdd(vals + b0 + off + 1, cur[eo + 1]);
      unsafeAtomicAdd(vals + b0 + off + 2, cur[eo + 2]);
      off += 3;
    }
  } else {
    int32_t off = 0;
    for (int j = 0; j < R; ++j) {
      const int32_t col = id[j];
      while (off < len && ci[b0 + off] != col) ++off;
      if (off < len) unsafeAtomicAdd(vals + b0 + off, cur[pm[j]]);
    }
  }
rhs_part:
  unsafeAtomicAdd(rhs + row, rhs_e[e * R + pr]);
}

inline unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

void launch_elements(nalu_linsys* ls, int64_t nelem, int nent, const int32_t* ents, const double* rhs_e,
                     const double* lhs, int32_t* ids, int32_t* perm, int* bad) {
  const int nd = ls->num_dof;
  hipLaunchKernelGGL(k_elem_ids, dim3(blocks(nelem)), dim3(256), 0, ls->stream, ents, nelem, nent, nd,
                     ls->node_lid.p, ls->num_nodes, ids, perm, bad);
  hipLaunchKernelGGL(k_elem_rows, dim3(blocks(nelem * nent * nd)), dim3(256), 0, ls->stream, ents, nelem, nent, nd,
                     ls->node_lid.p, ids, perm, lhs, rhs_e, ls->row_ptr.p, ls->col_idx.p, ls->values.p, ls->rhs.p);
  NALU_HIP_CHECK(hipGetLastError());
}

// Host queue of std::vector-overload calls, all with the same entity count.
struct HostQueue {
  int nent = 0;
  std::vector<int32_t> ents;
  std::vector<double> rhs, lhs;
  int64_t count() const { return nent ? (int64_t)ents.size() / nent : 0; }
};

HostQueue& queue_of(nalu_linsys* ls) {
  if (!ls->sum_into_queue) ls->sum_into_queue = new HostQueue();
  return *static_cast<HostQueue*>(ls->sum_into_queue);
}

constexpr int64_t kQueueFlush = 1 << 16;  // elements buffered before a flush

}  // namespace

void flush_sum_into_queue(nalu_linsys* ls) {
  if (!ls->sum_into_queue) return;
  HostQueue q;
  std::swap(q, queue_of(ls));  // the queue is empty from here on (no recursion below)
  const int64_t ne = q.count();
  if (ne == 0) return;
  const int R = q.nent * ls->num_dof;
  materialize_pending_zero(ls);
  DevBuf<int32_t> ents(ne * q.nent), ids(ne * R), perm(ne * R);
  DevBuf<double> rhs(ne * R), lhs(ne * R * R);
  DevBuf<int> bad(1);
  NALU_HIP_CHECK(hipMemcpyAsync(ents.p, q.ents.data(), ents.bytes(), hipMemcpyHostToDevice, ls->stream));
  NALU_HIP_CHECK(hipMemcpyAsync(rhs.p, q.rhs.data(), rhs.bytes(), hipMemcpyHostToDevice, ls->stream));
  NALU_HIP_CHECK(hipMemcpyAsync(lhs.p, q.lhs.data(), lhs.bytes(), hipMemcpyHostToDevice, ls->stream));
  NALU_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), ls->stream));
  launch_elements(ls, ne, q.nent, ents.p, rhs.p, lhs.p, ids.p, perm.p, bad.p);
  // the host vectors and device staging die here: wait for the copies and kernels
  NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
}

void discard_sum_into_queue(nalu_linsys* ls) {
  if (!ls->sum_into_queue) return;
  HostQueue& q = queue_of(ls);
  q.ents.clear();
  q.rhs.clear();
  q.lhs.clear();
  q.nent = 0;
}

void release_sum_into_queue(nalu_linsys* ls) {
  delete static_cast<HostQueue*>(ls->sum_into_queue);
  ls->sum_into_queue = nullptr;
}

}  // namespace nalu

using namespace nalu;

extern "C" int nalu_linsys_sum_into(nalu_linsys* ls, int ne, const int32_t* ents, const double* rhs_h,
                                    const double* lhs_h) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "sumInto: system not finalized");
    NALU_REQUIRE(ne > 0 && ne * ls->num_dof <= kMaxRows && ents && rhs_h && lhs_h, NALU_ERR_ARG,
                 "sumInto: bad arguments");
    for (int i = 0; i < ne; ++i)
      NALU_REQUIRE(ents[i] >= 0 && ents[i] < ls->num_nodes, NALU_ERR_ARG, "sumInto: entity out of range");
    PhaseTimer t(ls, PH_ASSEMBLE);
    HostQueue& q = queue_of(ls);
    if (q.count() > 0 && q.nent != ne) flush_sum_into_queue(ls);
    q.nent = ne;
    const int R = ne * ls->num_dof;
    q.ents.insert(q.ents.end(), ents, ents + ne);
    q.rhs.insert(q.rhs.end(), rhs_h, rhs_h + R);
    q.lhs.insert(q.lhs.end(), lhs_h, lhs_h + (size_t)R * R);
    if (q.count() >= kQueueFlush) flush_sum_into_queue(ls);
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_sum_into_elements(nalu_linsys* ls, int64_t num_elems, int num_entities,
                                             const int32_t* entities, const double* rhs, const double* lhs,
                                             int32_t* local_ids, int32_t* sort_permutation) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "sumInto: system not finalized");
    NALU_REQUIRE(num_elems >= 0 && num_entities > 0 && num_entities * ls->num_dof <= kMaxRows, NALU_ERR_ARG,
                 "sumInto: numEntities * numDof must be in [1, 64]");
    NALU_REQUIRE(num_elems == 0 || (entities && rhs && lhs && local_ids && sort_permutation), NALU_ERR_ARG,
                 "sumInto: null array");
    PhaseTimer t(ls, PH_ASSEMBLE);
    flush_sum_into_queue(ls);
    materialize_pending_zero(ls);
    if (num_elems == 0) return NALU_OK;
    DevBuf<int> bad(1);
    NALU_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), ls->stream));
    launch_elements(ls, num_elems, num_entities, entities, rhs, lhs, local_ids, sort_permutation, bad.p);
    int h = 0;
    NALU_HIP_CHECK(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, ls->stream));
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    NALU_REQUIRE(h == 0, NALU_ERR_ARG, "sumInto: entity index outside the mesh");
    return NALU_OK;
  });
}

// ---------------------------------------------------------------- no-allocation variant
// The Kokkos overload above for callers that cannot allocate device memory
// (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc): the localIds / sortPermutation tables and the error flag live in
// this file-scope device array instead of caller or library allocations.  Each
// GPU holds its own copy, at its own address, until the process ends; every
// call on a GPU shares it, so calls that overlap on one GPU race.
namespace nalu {
namespace {
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr size_t kScratchFlag = 256;  // the flag's slot, keeping the tables 256-byte aligned
constexpr int kMaxDevices = 64;
}  // namespace
}  // namespace nalu

__device__ __align__(256) unsigned char scratch[nalu::kScratchBytes];

namespace nalu {
namespace {
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// device address of this GPU's scratch (looked up once per device)
unsigned char* scratch_ptr() {
  int dev = 0;
  NALU_HIP_CHECK(hipGetDevice(&dev));
  NALU_REQUIRE(dev >= 0 && dev < kMaxDevices, NALU_ERR_UNSUPPORTED, "scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    NALU_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NALU_REQUIRE(p != nullptr, NALU_ERR_HIP, "scratch: symbol address not found");
    g_scratch_dev[dev] = p;
  }
  return static_cast<unsigned char*>(g_scratch_dev[dev]);
}

__global__ void k_zero_flag(int* f) { *f = 0; }
}  // namespace

bool sum_into_queue_empty(const nalu_linsys* ls) {
  return !ls->sum_into_queue || static_cast<const HostQueue*>(ls->sum_into_queue)->count() == 0;
}
}  // namespace nalu

extern "C" int nalu_linsys_sum_into_elements_noalloc(nalu_linsys* ls, int64_t num_elems, int num_entities,
                                                     const int32_t* entities, const double* rhs, const double* lhs) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "sumInto (no alloc): system not finalized");
    NALU_REQUIRE(num_elems >= 0 && num_entities > 0 && num_entities * ls->num_dof <= kMaxRows, NALU_ERR_ARG,
                 "sumInto (no alloc): numEntities * numDof must be in [1, 64]");
    NALU_REQUIRE(num_elems == 0 || (entities && rhs && lhs), NALU_ERR_ARG, "sumInto (no alloc): null array");
    const int64_t R = (int64_t)num_entities * ls->num_dof;
    const int64_t max_elems = (int64_t)((kScratchBytes - kScratchFlag) / (2 * sizeof(int32_t))) / R;
    NALU_REQUIRE(num_elems <= max_elems, NALU_ERR_ARG,
                 "sumInto (no alloc): " + std::to_string(num_elems) + " elements exceed the scratch limit of " +
                     std::to_string(max_elems) + " for " + std::to_string(R) + " rows per element");
    // applying queued host calls would allocate staging memory
    NALU_REQUIRE(sum_into_queue_empty(ls), NALU_ERR_STATE,
                 "sumInto (no alloc): host sumInto calls are queued; apply them first (e.g. loadComplete)");
    unsigned char* base = scratch_ptr();
    materialize_pending_zero(ls);
    if (num_elems == 0) return NALU_OK;
    int* bad = reinterpret_cast<int*>(base);
    int32_t* ids = reinterpret_cast<int32_t*>(base + kScratchFlag);
    int32_t* perm = ids + num_elems * R;
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, ls->stream, bad);
    launch_elements(ls, num_elems, num_entities, entities, rhs, lhs, ids, perm, bad);
    int h = 0;
    NALU_HIP_CHECK(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, ls->stream));
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    NALU_REQUIRE(h == 0, NALU_ERR_ARG, "sumInto (no alloc): entity index outside the mesh");
    return NALU_OK;
  });
}
