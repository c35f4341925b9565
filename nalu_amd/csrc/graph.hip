// GPU construction of the Tpetra-identical CSR graph, the node->element
// incidence, the element scatter table and the multicolor row ordering.
//
// Reference semantics (serial): rows are the nodes sorted by global id with
// LID = numDof * rank (src/TpetraLinearSystem.C:189-327); a row holds every
// node that shares an element with it, including itself (addConnections,
// :351-367); columns are sorted ascending (LocalGraphArrays::insertIndices,
// include/LinearSolver.h:77-89); each node's numDof rows are identical and
// each column node expands to numDof consecutive columns
// (fill_in_extra_dof_rows_per_node, :1227-1246).  The same graph is obtained
// here by sorting and de-duplicating the (row, col) keys of every element's
// 8x8 node pairs on the device.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace nalu {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

__global__ void k_iota(int32_t* lid, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lid[i] = (int32_t)i;
}

// ids are required to be < 2^32 so that (id, index) packs into one key
__global__ void k_id_keys(const int64_t* __restrict__ ids, int64_t n, uint64_t* __restrict__ k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) k[i] = ((uint64_t)ids[i] << 32) | (uint64_t)i;
}

__global__ void k_scatter_rank(const int32_t* sorted_idx, int32_t* lid, int64_t n) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) lid[sorted_idx[r]] = (int32_t)r;
}

// 64 (row, col) node-lid keys per element.
__global__ void k_pair_keys(const int32_t* __restrict__ conn, const int32_t* __restrict__ lid,
                            int64_t nelem, uint64_t* __restrict__ keys) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nelem * 8) return;
  const int64_t e = t >> 3;
  const int a = (int)(t & 7);
  const uint64_t ra = (uint64_t)(uint32_t)lid[conn[8 * e + a]];
  uint64_t* out = keys + 8 * t;
#pragma unroll
  for (int b = 0; b < 8; ++b) out[b] = (ra << 32) | (uint64_t)(uint32_t)lid[conn[8 * e + b]];
}

// Unique sorted keys -> node graph col idx; row starts found at key boundaries.
__global__ void k_split_keys(const uint64_t* __restrict__ ukeys, int64_t p, int32_t* __restrict__ nrp,
                             int32_t* __restrict__ nci) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p) return;
  const uint64_t k = ukeys[q];
  const uint32_t row = (uint32_t)(k >> 32);
  nci[q] = (int32_t)(uint32_t)(k & 0xffffffffu);
  if (q == 0 || (uint32_t)(ukeys[q - 1] >> 32) != row) nrp[row] = (int32_t)q;
}

__global__ void k_set_i32(int32_t* p, int32_t v) { *p = v; }

__global__ void k_row_len(const int32_t* __restrict__ rp, int64_t n, int32_t* __restrict__ len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) len[i] = rp[i + 1] - rp[i];
}

__device__ inline int find_in_row(const int32_t* __restrict__ cols, int len, int32_t key) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Element scatter table: for (e, a), byte c = position of node c in node-row a.
__global__ void k_elem_pos(const int32_t* __restrict__ conn, const int32_t* __restrict__ lid,
                           const int32_t* __restrict__ nrp, const int32_t* __restrict__ nci,
                           int64_t nelem, uint64_t* __restrict__ pos, int* __restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nelem * 8) return;
  const int64_t e = t >> 3;
  const int a = (int)(t & 7);
  const int32_t row = lid[conn[8 * e + a]];
  const int32_t b0 = nrp[row];
  const int len = nrp[row + 1] - b0;
  uint64_t packed = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int32_t key = lid[conn[8 * e + c]];
    const int p = find_in_row(nci + b0, len, key);
    if (p >= len || p > 255 || nci[b0 + p] != key) atomicOr(bad, 1);
    packed |= (uint64_t)(p & 0xff) << (8 * c);
  }
  pos[t] = packed;
}

// Expanded CSR for numDof > 1: row (a*nd+d) = node-row a with each column
// node b expanded to columns nd*b .. nd*b+nd-1.
__global__ void k_expand_rowptr(const int32_t* __restrict__ nrp, int64_t nn, int nd,
                                int32_t* __restrict__ rp) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a > nn) return;
  if (a == nn) {
    rp[nn * nd] = nd * nd * nrp[nn];
    return;
  }
  const int32_t L = nrp[a + 1] - nrp[a];
  for (int d = 0; d < nd; ++d) rp[a * nd + d] = nd * nd * nrp[a] + d * nd * L;
}

__global__ void k_expand_cols(const int32_t* __restrict__ nrp, const int32_t* __restrict__ nci,
                              int64_t nn, int nd, const int32_t* __restrict__ rp,
                              int32_t* __restrict__ ci) {
  // one thread per (node row a, position p)
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)nrp[nn]) return;
  // locate the row by binary search on nrp
  int64_t lo = 0, hi = nn;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (nrp[mid] <= q) lo = mid;
    else hi = mid;
  }
  const int64_t a = lo;
  const int p = (int)(q - nrp[a]);
  const int32_t b = nci[q];
  for (int d = 0; d < nd; ++d) {
    int32_t* row = ci + rp[a * nd + d];
    for (int e2 = 0; e2 < nd; ++e2) row[p * nd + e2] = nd * b + e2;
  }
}

__global__ void k_diag_pos(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, int64_t nrows,
                           int32_t* __restrict__ dpos, int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int32_t b = rp[i];
  const int len = rp[i + 1] - b;
  const int p = find_in_row(ci + b, len, (int32_t)i);
  if (p >= len || ci[b + p] != (int32_t)i) {
    atomicOr(bad, 2);
    dpos[i] = b;
  } else {
    dpos[i] = b + p;
  }
}

// node -> (element*8 + local) incidence keys
__global__ void k_inc_keys(const int32_t* __restrict__ conn, int64_t nelem, uint64_t* __restrict__ keys) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nelem * 8) return;
  keys[t] = ((uint64_t)(uint32_t)conn[t] << 32) | (uint64_t)t;
}

__global__ void k_inc_split(const uint64_t* __restrict__ skeys, int64_t m, int64_t nn,
                            int32_t* __restrict__ inc_ptr, int32_t* __restrict__ inc_elem) {
  // one thread per node: lower_bound of node in the sorted keys
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) inc_elem[t] = (int32_t)(uint32_t)(skeys[t] & 0xffffffffu);
  if (t > nn) return;
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(skeys[mid] >> 32) < t) lo = mid + 1;
    else hi = mid;
  }
  inc_ptr[t] = (int32_t)lo;
}

__global__ void k_inc_max(const int32_t* __restrict__ inc_ptr, int64_t nn, int* __restrict__ mx) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a < nn) atomicMax(mx, inc_ptr[a + 1] - inc_ptr[a]);
}

// ---------------------------------------------------------------- coloring
// Deterministic speculative greedy coloring (Jones-Plassmann style rounds):
// every uncolored row proposes the smallest color unused by its colored
// neighbors; a proposal is kept only if no uncolored neighbor of higher
// priority proposed the same color in the same round.
__device__ inline uint32_t prio(uint32_t i) {
  uint32_t x = i * 2654435761u;
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  return x;
}

constexpr int kMaxColors = 128;

__global__ void k_color_propose(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, int64_t n,
                                const int32_t* __restrict__ color, int32_t* __restrict__ prop,
                                int* __restrict__ overflow) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (color[i] >= 0) {
    prop[i] = -1;
    return;
  }
  uint64_t used0 = 0, used1 = 0;
  for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
    const int32_t j = ci[q];
    if (j == i) continue;
    const int32_t c = color[j];
    if (c >= 0) {
      if (c < 64) used0 |= 1ull << c;
      else if (c < 128) used1 |= 1ull << (c - 64);
    }
  }
  int c = -1;
  if (~used0) c = __ffsll((long long)~used0) - 1;
  else if (~used1) c = 64 + __ffsll((long long)~used1) - 1;
  if (c < 0) atomicOr(overflow, 1);
  prop[i] = c;
}

__global__ void k_color_resolve(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, int64_t n,
                                int32_t* __restrict__ color, const int32_t* __restrict__ prop,
                                int* __restrict__ remaining) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c = prop[i];
  if (c < 0) return;
  const uint32_t pi = prio((uint32_t)i);
  bool win = true;
  for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
    const int32_t j = ci[q];
    if (j == i || prop[j] != c) continue;
    const uint32_t pj = prio((uint32_t)j);
    if (pj > pi || (pj == pi && j > i)) {
      win = false;
      break;
    }
  }
  if (win) color[i] = c;
  else atomicAdd(remaining, 1);
}

// Node-parity coloring: on a hex mesh whose elements share consistent local
// axes (any structured-topology mesh), every node carries a 3-bit parity
// (i mod 2, j mod 2, k mod 2) that differs between any two nodes of an
// element, i.e. an 8-coloring of the node graph.  It is found by propagating
// parity across elements from a seed node; an inconsistent element means the
// mesh has no such labelling and the speculative coloring is used instead.
__global__ void k_parity_prop(const int32_t* __restrict__ conn, int64_t nelem, int32_t* par,
                              int* __restrict__ changed) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int32_t* nd = conn + 8 * e;
  int p0 = -1;
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int32_t pn = par[nd[n]];
    if (pn >= 0 && p0 < 0) p0 = pn ^ ((n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0) ^
                                    ((n == 2 || n == 3 || n == 6 || n == 7) ? 2 : 0) ^ (n >= 4 ? 4 : 0);
  }
  if (p0 < 0) return;
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int32_t want = p0 ^ ((n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0) ^
                         ((n == 2 || n == 3 || n == 6 || n == 7) ? 2 : 0) ^ (n >= 4 ? 4 : 0);
    if (atomicCAS(par + nd[n], -1, want) == -1) atomicOr(changed, 1);
  }
}

__global__ void k_parity_check(const int32_t* __restrict__ conn, int64_t nelem, const int32_t* __restrict__ par,
                               int* __restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int32_t* nd = conn + 8 * e;
  const int p0 = par[nd[0]];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int want = p0 ^ ((n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0) ^
                     ((n == 2 || n == 3 || n == 6 || n == 7) ? 2 : 0) ^ (n >= 4 ? 4 : 0);
    if (p0 < 0 || par[nd[n]] != want) atomicOr(bad, 1);
  }
}

// row color = numDof * parity(node of row) + dof; rows indexed by lid
__global__ void k_parity_rows(const int32_t* __restrict__ par, const int32_t* __restrict__ lid, int64_t nnodes,
                              int nd, int32_t* __restrict__ color) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nnodes) return;
  for (int d = 0; d < nd; ++d) color[(int64_t)nd * lid[a] + d] = nd * par[a] + d;
}

__global__ void k_first_unlabeled(const int32_t* __restrict__ par, int64_t n, int* __restrict__ first) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && par[i] < 0) atomicMin(first, (int)i);
}

__global__ void k_color_keys(const int32_t* __restrict__ color, int64_t n, uint64_t* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = ((uint64_t)(uint32_t)color[i] << 32) | (uint64_t)i;
}

// start of each color in the color-sorted keys (lower bound of c << 32)
__global__ void k_color_bounds(const uint64_t* __restrict__ skeys, int64_t n, int32_t* __restrict__ start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > kMaxColors) return;
  int64_t lo = 0, hi = n;
  const uint64_t key = (uint64_t)c << 32;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (skeys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  start[c] = (int32_t)lo;
}

__global__ void k_low32(const uint64_t* __restrict__ keys, int64_t n, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int32_t)(uint32_t)(keys[i] & 0xffffffffu);
}

int bits_for(uint64_t v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

void sort_u64(DevBuf<uint64_t>& keys, DevBuf<uint64_t>& out, int64_t n, int end_bit, hipStream_t s) {
  NALU_REQUIRE(n < (int64_t)1 << 31, NALU_ERR_UNSUPPORTED, "graph: too many keys for one sort");
  size_t tmp = 0;
  NALU_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, keys.p, out.p, (int)n, 0, end_bit, s));
  DevBuf<char> t(tmp ? tmp : 1);
  NALU_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(t.p, tmp, keys.p, out.p, (int)n, 0, end_bit, s));
}

}  // namespace

void build_node_incidence(nalu_mesh* m, hipStream_t s) {
  if (m->inc_ptr.p) return;
  const int64_t E = m->num_elems, N = m->num_nodes, M = 8 * E;
  DevBuf<uint64_t> keys(M), sorted(M);
  k_inc_keys<<<grid_for(M), kBlock, 0, s>>>(m->conn.p, E, keys.p);
  sort_u64(keys, sorted, M, 32 + bits_for((uint64_t)N), s);
  m->inc_ptr.alloc(N + 1);
  m->inc_elem.alloc(M);
  k_inc_split<<<grid_for(std::max<int64_t>(M, N + 1)), kBlock, 0, s>>>(sorted.p, M, N, m->inc_ptr.p,
                                                                        m->inc_elem.p);
  DevBuf<int> mx(1);
  NALU_HIP_CHECK(hipMemsetAsync(mx.p, 0, sizeof(int), s));
  k_inc_max<<<grid_for(N), kBlock, 0, s>>>(m->inc_ptr.p, N, mx.p);
  int h = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&h, mx.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  m->max_incidence = h;
}

void build_graph(nalu_linsys* ls, hipStream_t s) {
  const nalu_mesh* m = ls->mesh;
  const int64_t N = m->num_nodes, E = m->num_elems;
  const int nd = ls->num_dof;
  NALU_REQUIRE(N < ((int64_t)1 << 31) && E * 8 < ((int64_t)1 << 31), NALU_ERR_UNSUPPORTED,
               "graph: mesh too large for int32 indices");
  ls->num_nodes = N;

  // 1. LIDs: rank of the global id (identity when ids are implicit)
  ls->node_lid.alloc(N);
  if (m->ids.p) {
    DevBuf<uint64_t> keys(N), sorted(N);
    // pack (id, index); ids are assumed < 2^32
    k_id_keys<<<grid_for(N), kBlock, 0, s>>>(m->ids.p, N, keys.p);
    sort_u64(keys, sorted, N, 64, s);
    ls->row_node.alloc(N);  // row (= rank) -> node index
    k_low32<<<grid_for(N), kBlock, 0, s>>>(sorted.p, N, ls->row_node.p);
    k_scatter_rank<<<grid_for(N), kBlock, 0, s>>>(ls->row_node.p, ls->node_lid.p, N);
  } else {
    k_iota<<<grid_for(N), kBlock, 0, s>>>(ls->node_lid.p, N);
  }

  // 2. node graph from sorted unique element pair keys
  const int64_t K = E * 64;
  int64_t P = 0;
  {
    DevBuf<uint64_t> keys(K), sorted(K);
    k_pair_keys<<<grid_for(E * 8), kBlock, 0, s>>>(m->conn.p, ls->node_lid.p, E, keys.p);
    sort_u64(keys, sorted, K, 32 + bits_for((uint64_t)N), s);
    // unique into `keys`
    DevBuf<int> nsel(1);
    size_t tmp = 0;
    NALU_HIP_CHECK(hipcub::DeviceSelect::Unique(nullptr, tmp, sorted.p, keys.p, nsel.p, (int)K, s));
    {
      DevBuf<char> t(tmp ? tmp : 1);
      NALU_HIP_CHECK(hipcub::DeviceSelect::Unique(t.p, tmp, sorted.p, keys.p, nsel.p, (int)K, s));
    }
    int hP = 0;
    NALU_HIP_CHECK(hipMemcpyAsync(&hP, nsel.p, sizeof(int), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    P = hP;
    ls->nrp.alloc(N + 1);
    ls->nci.alloc(P);
    NALU_HIP_CHECK(hipMemsetAsync(ls->nrp.p, 0xff, sizeof(int32_t) * (N + 1), s));
    k_split_keys<<<grid_for(P), kBlock, 0, s>>>(keys.p, P, ls->nrp.p, ls->nci.p);
    k_set_i32<<<1, 1, 0, s>>>(ls->nrp.p + N, (int32_t)P);
  }
  // every node must appear in some element (Tpetra rejects empty connection
  // rows, verify_no_empty_connections :1248-1254)
  {
    DevBuf<int32_t> mn(1), mx(1), len(N);
    {
      size_t tmp = 0;
      NALU_HIP_CHECK(hipcub::DeviceReduce::Min(nullptr, tmp, ls->nrp.p, mn.p, (int)N, s));
      DevBuf<char> t(tmp ? tmp : 1);
      NALU_HIP_CHECK(hipcub::DeviceReduce::Min(t.p, tmp, ls->nrp.p, mn.p, (int)N, s));
    }
    k_row_len<<<grid_for(N), kBlock, 0, s>>>(ls->nrp.p, N, len.p);
    {
      size_t tmp = 0;
      NALU_HIP_CHECK(hipcub::DeviceReduce::Max(nullptr, tmp, len.p, mx.p, (int)N, s));
      DevBuf<char> t(tmp ? tmp : 1);
      NALU_HIP_CHECK(hipcub::DeviceReduce::Max(t.p, tmp, len.p, mx.p, (int)N, s));
    }
    int32_t hmx = 0, hmn = 0;
    NALU_HIP_CHECK(hipMemcpyAsync(&hmx, mx.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipMemcpyAsync(&hmn, mn.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    NALU_REQUIRE(hmn >= 0, NALU_ERR_ARG, "graph: a node belongs to no element (empty connections row)");
    ls->max_node_row_len = hmx;
  }

  // 3. element scatter table
  ls->elem_pos.alloc(E * 8);
  DevBuf<int> bad(1);
  NALU_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  k_elem_pos<<<grid_for(E * 8), kBlock, 0, s>>>(m->conn.p, ls->node_lid.p, ls->nrp.p, ls->nci.p, E,
                                                 ls->elem_pos.p, bad.p);

  // 4. expanded CSR
  const int64_t nrows = N * nd;
  const int64_t nnz = P * nd * nd;
  NALU_REQUIRE(nnz < ((int64_t)1 << 31), NALU_ERR_UNSUPPORTED, "graph: nnz exceeds int32");
  ls->nrows = nrows;
  ls->nnz = nnz;
  ls->row_ptr.alloc(nrows + 1);
  ls->col_idx.alloc(nnz);
  if (nd == 1) {
    NALU_HIP_CHECK(hipMemcpyAsync(ls->row_ptr.p, ls->nrp.p, sizeof(int32_t) * (N + 1), hipMemcpyDeviceToDevice, s));
    NALU_HIP_CHECK(hipMemcpyAsync(ls->col_idx.p, ls->nci.p, sizeof(int32_t) * P, hipMemcpyDeviceToDevice, s));
  } else {
    k_expand_rowptr<<<grid_for(N + 1), kBlock, 0, s>>>(ls->nrp.p, N, nd, ls->row_ptr.p);
    k_expand_cols<<<grid_for(P), kBlock, 0, s>>>(ls->nrp.p, ls->nci.p, N, nd, ls->row_ptr.p, ls->col_idx.p);
  }
  ls->diag_pos.alloc(nrows);
  k_diag_pos<<<grid_for(nrows), kBlock, 0, s>>>(ls->row_ptr.p, ls->col_idx.p, nrows, ls->diag_pos.p, bad.p);
  int hbad = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  NALU_REQUIRE(hbad == 0, NALU_ERR_NUMERIC,
               "graph: inconsistent element scatter table or missing diagonal (row longer than 255?)");

  ls->values.alloc(nnz);
  ls->rhs.alloc(nrows);
  ls->sln.alloc(nrows);
  NALU_HIP_CHECK(hipMemsetAsync(ls->values.p, 0, sizeof(double) * nnz, s));
  NALU_HIP_CHECK(hipMemsetAsync(ls->rhs.p, 0, sizeof(double) * nrows, s));
  NALU_HIP_CHECK(hipMemsetAsync(ls->sln.p, 0, sizeof(double) * nrows, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
}

// Parity labelling of the mesh nodes; false when the mesh admits none.
bool parity_coloring(nalu_linsys* ls, DevBuf<int32_t>& color, hipStream_t s) {
  const nalu_mesh* m = ls->mesh;
  const int64_t N = m->num_nodes, E = m->num_elems;
  DevBuf<int32_t> par(N);
  DevBuf<int> flag(2);
  NALU_HIP_CHECK(hipMemsetAsync(par.p, 0xff, sizeof(int32_t) * N, s));
  int seeds = 0;
  for (;;) {
    // seed the lowest unlabeled node (one seed per connected component)
    int first = 0x7fffffff;
    k_set_i32<<<1, 1, 0, s>>>(flag.p, 0x7fffffff);
    k_first_unlabeled<<<grid_for(N), kBlock, 0, s>>>(par.p, N, flag.p);
    NALU_HIP_CHECK(hipMemcpyAsync(&first, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    if (first == 0x7fffffff) break;
    if (++seeds > 1024) return false;  // too fragmented to be worth it
    k_set_i32<<<1, 1, 0, s>>>(par.p + first, 0);
    for (int round = 0;; ++round) {
      if (round > 100000) return false;
      NALU_HIP_CHECK(hipMemsetAsync(flag.p + 1, 0, sizeof(int), s));
      k_parity_prop<<<grid_for(E), kBlock, 0, s>>>(m->conn.p, E, par.p, flag.p + 1);
      int ch = 0;
      NALU_HIP_CHECK(hipMemcpyAsync(&ch, flag.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      if (!ch) break;
    }
  }
  NALU_HIP_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
  k_parity_check<<<grid_for(E), kBlock, 0, s>>>(m->conn.p, E, par.p, flag.p);
  int bad = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  if (bad) return false;
  k_parity_rows<<<grid_for(N), kBlock, 0, s>>>(par.p, ls->node_lid.p, N, ls->num_dof, color.p);
  return true;
}

void build_coloring(nalu_linsys* ls, hipStream_t s) {
  const int64_t n = ls->nrows;
  DevBuf<int32_t> color(n), prop(n);
  DevBuf<int> cnt(2);
  NALU_HIP_CHECK(hipMemsetAsync(color.p, 0xff, sizeof(int32_t) * n, s));
  int rounds = 0;
  const bool parity = parity_coloring(ls, color, s);
  for (; !parity;) {
    NALU_HIP_CHECK(hipMemsetAsync(cnt.p, 0, 2 * sizeof(int), s));
    k_color_propose<<<grid_for(n), kBlock, 0, s>>>(ls->row_ptr.p, ls->col_idx.p, n, color.p, prop.p, cnt.p + 1);
    k_color_resolve<<<grid_for(n), kBlock, 0, s>>>(ls->row_ptr.p, ls->col_idx.p, n, color.p, prop.p, cnt.p);
    int h[2] = {0, 0};
    NALU_HIP_CHECK(hipMemcpyAsync(h, cnt.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    NALU_REQUIRE(h[1] == 0, NALU_ERR_UNSUPPORTED, "coloring: more than 128 colors needed");
    if (h[0] == 0) break;
    NALU_REQUIRE(++rounds < 10000, NALU_ERR_NUMERIC, "coloring did not terminate");
  }
  DevBuf<int32_t> bounds(kMaxColors + 1);
  DevBuf<uint64_t> keys(n), sorted(n);
  k_color_keys<<<grid_for(n), kBlock, 0, s>>>(color.p, n, keys.p);
  sort_u64(keys, sorted, n, 32 + 8, s);
  ls->color_rows.alloc(n);
  k_low32<<<grid_for(n), kBlock, 0, s>>>(sorted.p, n, ls->color_rows.p);
  k_color_bounds<<<1, kMaxColors + 1, 0, s>>>(sorted.p, n, bounds.p);
  std::vector<int32_t> h(kMaxColors + 1);
  NALU_HIP_CHECK(hipMemcpyAsync(h.data(), bounds.p, sizeof(int32_t) * (kMaxColors + 1), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  int nc = 0;
  for (int c = 0; c < kMaxColors; ++c)
    if (h[c + 1] > h[c]) nc = c + 1;
  ls->num_colors = nc;
  ls->color_ptr_host.assign(nc + 1, 0);
  for (int c = 0; c <= nc; ++c) ls->color_ptr_host[c] = h[c];
}

}  // namespace nalu
