// Fused Hex8 CVFEM assembly: gather + scs geometry + element kernel + scatter
// in one pass over the mesh (the consolidated AssembleElemSolverAlgorithm,
// src/AssembleElemSolverAlgorithm.C:75-100 and include/AssembleElemSolverAlgorithm.h:52-105,
// whose per-element lhs/rhs go through SolverAlgorithm::apply_coeff ->
// TpetraLinearSystem::sumInto, src/TpetraLinearSystem.C:1457-1522).
//
// Two scatter strategies share the same per-(element, row-node) arithmetic
// (hex8_device.hpp):
//   * owner-computes row gather (mode 0): one thread per matrix row walks the
//     elements incident to its node, accumulates its row in LDS and the block
//     writes its contiguous CSR segment once, coalesced.  No atomics, fully
//     deterministic, and zeroSystem's memset of the values can be folded in.
//   * element-centric atomic scatter (mode 1): eight waves per block, wave w
//     computing row w of 64 elements, FP64 atomic adds into the CSR.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hex8_device.hpp"
#include "internal.hpp"

namespace nalu {
namespace {

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

__device__ __forceinline__ void load_nodes(const int32_t* __restrict__ conn, int64_t e, int32_t (&nd)[8]) {
  const int4* c4 = reinterpret_cast<const int4*>(conn + 8 * e);
  const int4 a = c4[0], b = c4[1];
  nd[0] = a.x; nd[1] = a.y; nd[2] = a.z; nd[3] = a.w;
  nd[4] = b.x; nd[5] = b.y; nd[6] = b.z; nd[7] = b.w;
}

template <typename R>
__device__ __forceinline__ void load_vec3(const double* __restrict__ f, const int32_t (&nd)[8], R (&v)[8][3]) {
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int d = 0; d < 3; ++d) v[n][d] = R(f[3 * (int64_t)nd[n] + d]);
}

// Element coordinates in the arithmetic type R.  The FP32 path works on
// coordinates relative to the element's first node, the differences formed in
// FP64, so that short edges of stretched elements keep their relative accuracy.
template <typename R>
__device__ __forceinline__ void load_coords(const double* __restrict__ f, const int32_t (&nd)[8], R (&x)[8][3]) {
  if constexpr (std::is_same<R, double>::value) {
    load_vec3(f, nd, x);
  } else {
    const double o[3] = {f[3 * (int64_t)nd[0]], f[3 * (int64_t)nd[0] + 1], f[3 * (int64_t)nd[0] + 2]};
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[n][d] = R(f[3 * (int64_t)nd[n] + d] - o[d]);
  }
}

template <typename R>
__device__ __forceinline__ void load_scalar(const double* __restrict__ f, const int32_t (&nd)[8], R (&v)[8]) {
#pragma unroll
  for (int n = 0; n < 8; ++n) v[n] = R(f[nd[n]]);
}

// ---------------------------------------------------------------- physics ops
// Each op evaluates row LA of its element matrix, calling emit(c, value) once
// per column node c, and returns the row's rhs contribution.  R is the
// arithmetic type (double, or float for the mixed-precision path); Jacobians
// and their determinants are always formed in FP64.
template <bool SHIFT, typename R = double>
struct ScalarDiffOp {
  using real = R;
  const double* q;
  const double* k;
  template <int LA, class Emit>
  __device__ __forceinline__ R row(const int32_t (&nd)[8], const R (&x)[8][3], Emit&& emit) const {
    const hex8::Gather gt{nd};
    return hex8::row_scalar_diff<LA, SHIFT, double>(x, gt, q, k, emit);
  }
};

template <bool SHIFT_MDOT, bool SHIFT_POISSON, bool REDUCED, typename R = double>
struct ContinuityOp {
  using real = R;
  const double* rho;
  const double* p;
  const double* u;
  const double* gp;
  double tau;
  double* mdot_out = nullptr;  // [E][12]: the mdot of each scs, written by the row of its left node
  template <int LA, class Emit>
  __device__ __forceinline__ R row(const int32_t (&nd)[8], const R (&x)[8][3], Emit&& emit) const {
    const hex8::Gather gt{nd};
    return hex8::row_continuity<LA, SHIFT_MDOT, SHIFT_POISSON, REDUCED, double>(x, gt, rho, p, u, gp, R(tau), emit);
  }
  // the same row, also delivering the mdot of the scs it owns (mdot_out != nullptr)
  template <int LA, class Emit>
  __device__ __forceinline__ R row_e(const int32_t (&nd)[8], const R (&x)[8][3], int64_t e, Emit&& emit) const {
    const hex8::Gather gt{nd};
    double* mo = mdot_out + 12 * e;
    return hex8::row_continuity<LA, SHIFT_MDOT, SHIFT_POISSON, REDUCED, double>(
        x, gt, rho, p, u, gp, R(tau), emit, [&](int ip, R v) { mo[ip] = (double)v; });
  }
};

// ops whose rows can also write per-element outputs (row_e)
template <class Op, class = void>
struct HasRowE : std::false_type {};
template <class Op>
struct HasRowE<Op, std::void_t<decltype(&Op::mdot_out)>> : std::true_type {};

// dispatch a wave-uniform (or mostly uniform) runtime row index to the
// compile-time specialisation
template <class Op, class Emit>
__device__ __forceinline__ typename Op::real dispatch_row(const Op& op, int la, const int32_t (&nd)[8],
                                                          const typename Op::real (&x)[8][3], Emit&& emit) {
  switch (la) {
    case 0: return op.template row<0>(nd, x, emit);
    case 1: return op.template row<1>(nd, x, emit);
    case 2: return op.template row<2>(nd, x, emit);
    case 3: return op.template row<3>(nd, x, emit);
    case 4: return op.template row<4>(nd, x, emit);
    case 5: return op.template row<5>(nd, x, emit);
    case 6: return op.template row<6>(nd, x, emit);
    default: return op.template row<7>(nd, x, emit);
  }
}

// dispatch_row with the element index, for ops with per-element outputs when
// they are requested
template <class Op, class Emit>
__device__ __forceinline__ typename Op::real dispatch_row_e(const Op& op, int la, const int32_t (&nd)[8],
                                                            const typename Op::real (&x)[8][3], int64_t e,
                                                            Emit&& emit) {
  if constexpr (HasRowE<Op>::value) {
    if (op.mdot_out) {
      switch (la) {
        case 0: return op.template row_e<0>(nd, x, e, emit);
        case 1: return op.template row_e<1>(nd, x, e, emit);
        case 2: return op.template row_e<2>(nd, x, e, emit);
        case 3: return op.template row_e<3>(nd, x, e, emit);
        case 4: return op.template row_e<4>(nd, x, e, emit);
        case 5: return op.template row_e<5>(nd, x, e, emit);
        case 6: return op.template row_e<6>(nd, x, e, emit);
        default: return op.template row_e<7>(nd, x, e, emit);
      }
    }
  }
  return dispatch_row(op, la, nd, x, emit);
}

__device__ __forceinline__ int32_t pick(const int32_t (&nd)[8], int la) {
  int32_t node = nd[0];
#pragma unroll
  for (int c = 1; c < 8; ++c) node = (la == c) ? nd[c] : node;
  return node;
}

// ---------------------------------------------------------------- scalar systems
// Element-centric kernels: a block handles one local row node LA of a chunk of
// kChunk elements.  The eight LA-blocks of a chunk are consecutive in the
// dispatch order of one XCD (block b runs on XCD b % 8), so the chunk's nodal
// data is shared through that XCD's L2.
constexpr int kChunk = 256;

struct ChunkMap {
  int64_t e;   // element of this thread
  int la;      // block-uniform local row node (0..NL-1)
  bool valid;
};
template <int NL>
__device__ __forceinline__ ChunkMap chunk_map(int64_t nelem) {
  const unsigned b = blockIdx.x;
  const unsigned xcd = b & 7, la = (b >> 3) % NL, grp = (b >> 3) / NL;
  const int64_t chunk = (int64_t)grp * 8 + xcd;
  const int64_t e = chunk * kChunk + threadIdx.x;
  return {e, (int)la, e < nelem};
}
template <int NL>
inline unsigned chunk_blocks(int64_t nelem) {
  const int64_t nchunk = (nelem + kChunk - 1) / kChunk;
  return (unsigned)(((nchunk + 7) / 8) * 8 * NL);
}

template <class Op, int LA>
__device__ __forceinline__ void scalar_atomic_body(const Op& op, const int32_t* __restrict__ conn,
                                                   const double* __restrict__ coords, const int32_t* __restrict__ lid,
                                                   const int32_t* __restrict__ nrp,
                                                   const uint64_t* __restrict__ elem_pos, int64_t e,
                                                   double* __restrict__ vals, double* __restrict__ rhs) {
  int32_t nd[8];
  load_nodes(conn, e, nd);
  typename Op::real x[8][3];
  load_coords(coords, nd, x);
  const int32_t row = lid[nd[LA]];
  double* rowp = vals + nrp[row];
  const uint64_t pk = elem_pos[8 * e + LA];
  const double rr = (double)op.template row<LA>(nd, x, [&](int c, auto v) {
    unsafeAtomicAdd(rowp + (int)((pk >> (8 * c)) & 0xff), (double)v);
  });
  unsafeAtomicAdd(rhs + row, rr);
}

template <class Op>
__global__ __launch_bounds__(kChunk) void k_scalar_atomic(Op op, const int32_t* __restrict__ conn,
                                                          const double* __restrict__ coords,
                                                          const int32_t* __restrict__ lid,
                                                          const int32_t* __restrict__ nrp,
                                                          const uint64_t* __restrict__ elem_pos, int64_t nelem,
                                                          double* __restrict__ vals, double* __restrict__ rhs) {
  const ChunkMap cm = chunk_map<8>(nelem);
  if (!cm.valid) return;
  switch (cm.la) {
#define NALU_SA(L) \
  case L: scalar_atomic_body<Op, L>(op, conn, coords, lid, nrp, elem_pos, cm.e, vals, rhs); break;
    NALU_SA(0) NALU_SA(1) NALU_SA(2) NALU_SA(3) NALU_SA(4) NALU_SA(5) NALU_SA(6) NALU_SA(7)
#undef NALU_SA
  }
}

// Owner-computes row gather.  Block b owns rows [b*B, b*B+B); its CSR segment
// lives in LDS while the rows accumulate, then is streamed out once.
template <class Op>
__global__ __launch_bounds__(256) void k_scalar_gather(Op op, const int32_t* __restrict__ conn,
                                                       const double* __restrict__ coords,
                                                       const int32_t* __restrict__ row_node,
                                                       const int32_t* __restrict__ inc_ptr,
                                                       const int32_t* __restrict__ inc_elem,
                                                       const int32_t* __restrict__ nrp,
                                                       const uint64_t* __restrict__ elem_pos, int64_t nrows,
                                                       double* __restrict__ vals, double* __restrict__ rhs,
                                                       int accumulate) {
  // the row segments accumulate in FP64 (FP32 LDS accumulation for the mixed
  // path halved the LDS per row but made the pass slower)
  using A = double;
  extern __shared__ __align__(16) unsigned char seg_raw[];
  A* seg = reinterpret_cast<A*>(seg_raw);
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t rend = min(r0 + (int64_t)blockDim.x, nrows);
  const int64_t r = r0 + threadIdx.x;
  const int32_t s0 = nrp[r0], s1 = nrp[rend];
  const int nseg = s1 - s0;
  for (int i = threadIdx.x; i < nseg; i += blockDim.x) seg[i] = A(0);
  __syncthreads();
  double rr = 0.0;
  if (r < nrows) {
    const int32_t a = row_node ? row_node[r] : (int32_t)r;
    A* myrow = seg + (nrp[r] - s0);
    const int32_t q1 = inc_ptr[a + 1];
    for (int32_t q = inc_ptr[a]; q < q1; ++q) {
      const int32_t t = inc_elem[q];
      const int64_t e = t >> 3;
      const int la = t & 7;
      int32_t nd[8];
      load_nodes(conn, e, nd);
      typename Op::real x[8][3];
      load_coords(coords, nd, x);
      const uint64_t pk = elem_pos[t];
      rr += (double)dispatch_row_e(op, la, nd, x, e,
                                   [&](int c, auto v) { myrow[(int)((pk >> (8 * c)) & 0xff)] += (A)v; });
    }
  }
  __syncthreads();
  if (accumulate) {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) vals[s0 + i] += (double)seg[i];
    if (r < nrows) rhs[r] += rr;
  } else {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) __builtin_nontemporal_store((double)seg[i], &vals[s0 + i]);
    if (r < nrows) rhs[r] = rr;
  }
}

// ---------------------------------------------------------------- mass (time derivative)
// MomentumMassElemKernel (src/kernel/MomentumMassElemKernel.C:95-170) and
// ContinuityMassElemKernel (src/kernel/ContinuityMassElemKernel.C:84-118).  The
// scv ip of an element belongs to one node (HexSCV::ipNodeMap_), so a node's
// rows receive exactly the contributions of the scvs at that node: k_mass_gather
// gives each node one thread that walks its incident elements (ip = the node's
// local index) and adds into its rows and rhs in place (no atomics, no LDS);
// the numDof 3 forms (k_mass_tile3, and the mass option of k_momentum_tile)
// accumulate node tiles in LDS.
struct MassArgs {
  const double* rho[3];  // Nm1, N, Np1 (nodal)
  const double* u[3];    // Nm1, N, Np1 (nodal, 3 components; momentum only)
  const double* gjp;     // projected pressure gradient (momentum only)
  double dt, g1, g2, g3;
  int balanced;          // densFac = rhoNp1 (balanced force) or 1
};

// ATOMIC: the row is shared with other threads (LDS row segment of a node tile)
template <int LA, bool MOM, bool LUMPED, bool SEG, bool ATOMIC = false>
__device__ __forceinline__ void mass_row(const MassArgs& A, const int32_t (&nd)[8], const double (&x)[8][3],
                                         uint64_t pk, double* __restrict__ row, int32_t L, double (&rh)[3]) {
  const double scV = hex8::scv_volume<LA>(x);
  double rs[3] = {0.0, 0.0, 0.0}, us[3][3] = {}, g[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const double r = hex8::scv_shape<LA, LUMPED>(c);
    if (LUMPED && c != LA) continue;
#pragma unroll
    for (int st = 0; st < 3; ++st) rs[st] += r * A.rho[st][nd[c]];
    if (MOM) {
#pragma unroll
      for (int st = 0; st < 3; ++st)
#pragma unroll
        for (int j = 0; j < 3; ++j) us[st][j] += r * A.u[st][3 * (int64_t)nd[c] + j];
#pragma unroll
      for (int j = 0; j < 3; ++j) g[j] += r * A.gjp[3 * (int64_t)nd[c] + j];
    }
  }
  if (MOM) {
    const double densFac = A.balanced ? rs[2] : 1.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
      rh[j] += -(A.g1 * rs[2] * us[2][j] + A.g2 * rs[1] * us[1][j] + A.g3 * rs[0] * us[0][j]) * scV / A.dt -
               g[j] * scV * densFac;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (LUMPED && c != LA) continue;
      const double lf = hex8::scv_shape<LA, LUMPED>(c) * A.g1 * rs[2] * scV / A.dt;
      const int p = (int)((pk >> (8 * c)) & 0xff);
      if (SEG) {
        row[p] += lf;
      } else if (ATOMIC) {
#pragma unroll
        for (int j = 0; j < 3; ++j) unsafeAtomicAdd(row + 3 * j * L + 3 * p + j, lf);
      } else {
#pragma unroll
        for (int j = 0; j < 3; ++j) row[3 * j * L + 3 * p + j] += lf;
      }
    }
  } else {
    const double projTimeScale = A.dt / A.g1;
    rh[0] += -(A.g1 * rs[2] + A.g2 * rs[1] + A.g3 * rs[0]) * scV / A.dt / projTimeScale;
  }
}

// ---------------------------------------------------------------- node-tile gather
// Owner-computes with one thread per (row node, local index): a block owns
// kTileNodes consecutive rows; thread (node j, slot s) evaluates row s of every
// element in which node j has local index s (one element on a structured
// lattice).  The slot is uniform across a wave, so the per-LA code path never
// diverges.  Contributions accumulate in the LDS copy of the block's CSR
// segment (LDS FP64 atomics: the 8 slots of a node share its row), which is
// then written to HBM once, coalesced -- no global atomics, no separate zero pass.
constexpr int kTileNodes = 64;
constexpr int kTileThreads = kTileNodes * 8;
constexpr size_t kMaxLds = 160 * 1024;  // LDS of one gfx950 CU, the most one block can hold

template <class Op>
__global__ __launch_bounds__(kTileThreads) void k_scalar_tile(Op op, const int32_t* __restrict__ conn,
                                                              const double* __restrict__ coords,
                                                              const int32_t* __restrict__ row_node,
                                                              const int32_t* __restrict__ inc_ptr,
                                                              const int32_t* __restrict__ inc_elem,
                                                              const int32_t* __restrict__ nrp,
                                                              const uint64_t* __restrict__ elem_pos, int64_t nrows,
                                                              double* __restrict__ vals, double* __restrict__ rhs,
                                                              int accumulate) {
  extern __shared__ double seg[];
  const int64_t r0 = (int64_t)blockIdx.x * kTileNodes;
  const int64_t r1 = min(r0 + (int64_t)kTileNodes, nrows);
  const int32_t s0 = nrp[r0], s1 = nrp[r1];
  const int nseg = s1 - s0;
  double* rseg = seg + nseg;
  for (int i = threadIdx.x; i < nseg + kTileNodes; i += blockDim.x) seg[i] = 0.0;
  __syncthreads();
  const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileNodes);  // wave-uniform local index
  const int j = threadIdx.x % kTileNodes;
  const int64_t r = r0 + j;
  if (r < r1) {
    const int32_t a = row_node ? row_node[r] : (int32_t)r;
    double* myrow = seg + (nrp[r] - s0);
    double acc = 0.0;
    for (int32_t q = inc_ptr[a]; q < inc_ptr[a + 1]; ++q) {
      const int32_t t = inc_elem[q];
      if ((t & 7) != slot) continue;
      const int64_t e = t >> 3;
      int32_t nd[8];
      load_nodes(conn, e, nd);
      typename Op::real x[8][3];
      load_coords(coords, nd, x);
      const uint64_t pk = elem_pos[t];
      acc += (double)dispatch_row(op, slot, nd, x, [&](int c, auto v) {
        unsafeAtomicAdd(myrow + (int)((pk >> (8 * c)) & 0xff), (double)v);
      });
    }
    if (acc != 0.0) unsafeAtomicAdd(rseg + j, acc);
  }
  __syncthreads();
  if (accumulate) {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) vals[s0 + i] += seg[i];
    for (int i = threadIdx.x; i < r1 - r0; i += blockDim.x) rhs[r0 + i] += rseg[i];
  } else {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) __builtin_nontemporal_store(seg[i], &vals[s0 + i]);
    for (int i = threadIdx.x; i < r1 - r0; i += blockDim.x) rhs[r0 + i] = rseg[i];
  }
}

template <int LA, bool SKEW, bool SHIFT, typename R>
__device__ __forceinline__ void momentum_row_lds(const int32_t (&nd)[8], const R (&x)[8][3],
                                                 const double* __restrict__ u, const double* __restrict__ mu,
                                                 const double* __restrict__ md, double incl_divu, double* myrow,
                                                 int32_t L, uint64_t pk, double (&rh)[3]) {
  const hex8::Gather gt{nd};
  const R m3[3] = {R(md[hex8::adj_ip(LA, 0)]), R(md[hex8::adj_ip(LA, 1)]), R(md[hex8::adj_ip(LA, 2)])};
  R rr[3] = {R(0), R(0), R(0)};
  hex8::row_momentum<LA, SKEW, SHIFT, double>(x, gt, u, mu, m3, R(incl_divu), rr, [&](int c, const R (&blk)[3][3]) {
    const int p = (int)((pk >> (8 * c)) & 0xff);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) unsafeAtomicAdd(myrow + 3 * i * L + 3 * p + jj, (double)blk[i][jj]);
  });
#pragma unroll
  for (int i = 0; i < 3; ++i) rh[i] += (double)rr[i];
}

// Momentum (numDof 3) node-tile gather: the block's segment holds the 3 rows of
// each of its nodes (rows 3a..3a+2 are contiguous in the CSR).  MASS (1
// consistent, 2 lumped) adds the momentum mass kernel of the same element and
// scv in the same pass, as one element algorithm runs its kernel list over an
// element: one LDS segment, one write of the matrix.
template <bool SKEW, bool SHIFT, typename R, int MASS = 0>
__global__ __launch_bounds__(kTileThreads) void k_momentum_tile(
    const int32_t* __restrict__ conn, const double* __restrict__ coords, const double* __restrict__ u,
    const double* __restrict__ mu, const double* __restrict__ mdot, double incl_divu,
    const int32_t* __restrict__ row_node, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_elem,
    const int32_t* __restrict__ nrp, const uint64_t* __restrict__ elem_pos, int64_t nnodes,
    double* __restrict__ vals, double* __restrict__ rhs, int accumulate, MassArgs MA) {
  extern __shared__ double seg[];
  const int64_t a0 = (int64_t)blockIdx.x * kTileNodes;
  const int64_t a1 = min(a0 + (int64_t)kTileNodes, nnodes);
  const int64_t s0 = 9 * (int64_t)nrp[a0], s1 = 9 * (int64_t)nrp[a1];
  const int nseg = (int)(s1 - s0);
  double* rseg = seg + nseg;
  for (int i = threadIdx.x; i < nseg + 3 * kTileNodes; i += blockDim.x) seg[i] = 0.0;
  __syncthreads();
  const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileNodes);
  const int j = threadIdx.x % kTileNodes;
  const int64_t r = a0 + j;
  if (r < a1) {
    const int32_t a = row_node ? row_node[r] : (int32_t)r;
    const int32_t L = nrp[r + 1] - nrp[r];
    double* myrow = seg + (9 * (int64_t)nrp[r] - s0);
    double rh[3] = {0, 0, 0};
    for (int32_t q = inc_ptr[a]; q < inc_ptr[a + 1]; ++q) {
      const int32_t t = inc_elem[q];
      if ((t & 7) != slot) continue;
      const int64_t e = t >> 3;
      int32_t nd[8];
      load_nodes(conn, e, nd);
      R x[8][3];
      load_coords(coords, nd, x);
      const uint64_t pk = elem_pos[t];
      const double* md = mdot + 12 * e;
      switch (slot) {
#define NALU_MT(Lx) \
  case Lx: momentum_row_lds<Lx, SKEW, SHIFT, R>(nd, x, u, mu, md, incl_divu, myrow, L, pk, rh); break;
        NALU_MT(0) NALU_MT(1) NALU_MT(2) NALU_MT(3) NALU_MT(4) NALU_MT(5) NALU_MT(6) NALU_MT(7)
#undef NALU_MT
      }
      if (MASS) {
        double xd[8][3];
        load_vec3(coords, nd, xd);
        switch (slot) {
#define NALU_MM(Lx) \
  case Lx: mass_row<Lx, true, MASS == 2, false, true>(MA, nd, xd, pk, myrow, L, rh); break;
          NALU_MM(0) NALU_MM(1) NALU_MM(2) NALU_MM(3) NALU_MM(4) NALU_MM(5) NALU_MM(6) NALU_MM(7)
#undef NALU_MM
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (rh[i] != 0.0) unsafeAtomicAdd(rseg + 3 * j + i, rh[i]);
  }
  __syncthreads();
  const int nr = (int)(3 * (a1 - a0));
  if (accumulate) {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) vals[s0 + i] += (double)seg[i];
    for (int i = threadIdx.x; i < nr; i += blockDim.x) rhs[3 * a0 + i] += (double)rseg[i];
  } else {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) __builtin_nontemporal_store((double)seg[i], &vals[s0 + i]);
    for (int i = threadIdx.x; i < nr; i += blockDim.x) rhs[3 * a0 + i] = (double)rseg[i];
  }
}

// ---------------------------------------------------------------- momentum
template <int LA, bool SKEW, bool SHIFT, typename R>
__device__ __forceinline__ void momentum_row_atomic(const int32_t (&nd)[8], const R (&x)[8][3],
                                                    const double* __restrict__ u, const double* __restrict__ mu,
                                                    const double* __restrict__ md, double incl_divu,
                                                    double* __restrict__ rowp, int32_t L, uint64_t pk,
                                                    double (&rh)[3]) {
  const hex8::Gather gt{nd};
  const R m3[3] = {R(md[hex8::adj_ip(LA, 0)]), R(md[hex8::adj_ip(LA, 1)]), R(md[hex8::adj_ip(LA, 2)])};
  R rr[3] = {R(0), R(0), R(0)};
  hex8::row_momentum<LA, SKEW, SHIFT, double>(x, gt, u, mu, m3, R(incl_divu), rr, [&](int c, const R (&blk)[3][3]) {
    const int p = (int)((pk >> (8 * c)) & 0xff);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) unsafeAtomicAdd(rowp + 3 * i * L + 3 * p + j, (double)blk[i][j]);
  });
#pragma unroll
  for (int i = 0; i < 3; ++i) rh[i] += (double)rr[i];
}

template <int LA, bool SKEW, bool SHIFT, typename R>
__device__ __forceinline__ void momentum_body(const int32_t* __restrict__ conn, const double* __restrict__ coords,
                                              const double* __restrict__ u, const double* __restrict__ mu,
                                              const double* __restrict__ mdot, double incl_divu,
                                              const int32_t* __restrict__ lid, const int32_t* __restrict__ nrp,
                                              const uint64_t* __restrict__ elem_pos, int64_t e,
                                              double* __restrict__ vals, double* __restrict__ rhs) {
  int32_t nd[8];
  load_nodes(conn, e, nd);
  R x[8][3];
  load_coords(coords, nd, x);
  const int32_t nr = lid[nd[LA]];
  const int32_t nb = nrp[nr];
  const int32_t L = nrp[nr + 1] - nb;
  double* rowp = vals + 9 * (int64_t)nb;
  const uint64_t pk = elem_pos[8 * e + LA];
  double rh[3] = {0, 0, 0};
  momentum_row_atomic<LA, SKEW, SHIFT, R>(nd, x, u, mu, mdot + 12 * e, incl_divu, rowp, L, pk, rh);
#pragma unroll
  for (int i = 0; i < 3; ++i) unsafeAtomicAdd(rhs + 3 * (int64_t)nr + i, rh[i]);
}

template <bool SKEW, bool SHIFT, typename R>
__global__ __launch_bounds__(kChunk) void k_momentum_atomic(
    const int32_t* __restrict__ conn, const double* __restrict__ coords, const double* __restrict__ u,
    const double* __restrict__ mu, const double* __restrict__ mdot, double incl_divu,
    const int32_t* __restrict__ lid, const int32_t* __restrict__ nrp, const uint64_t* __restrict__ elem_pos,
    int64_t nelem, double* __restrict__ vals, double* __restrict__ rhs) {
  const ChunkMap cm = chunk_map<8>(nelem);
  if (!cm.valid) return;
  switch (cm.la) {
#define NALU_MA(L)                                                                                          \
  case L:                                                                                                   \
    momentum_body<L, SKEW, SHIFT, R>(conn, coords, u, mu, mdot, incl_divu, lid, nrp, elem_pos, cm.e, vals, rhs); \
    break;
    NALU_MA(0) NALU_MA(1) NALU_MA(2) NALU_MA(3) NALU_MA(4) NALU_MA(5) NALU_MA(6) NALU_MA(7)
#undef NALU_MA
  }
}

// ---------------------------------------------------------------- mdot
// mdot, element-centric: one thread per element loads its connectivity,
// coordinates and nodal fields once (two dependent round trips) and evaluates
// all 12 scs from registers (a thread per (element, scs) gathers the same
// element data twelve times and measured 2.4x slower at 128^3).  The 12 values
// leave as six 16-byte stores (mdot must be 16-byte aligned).
template <bool SHIFT_MDOT, bool SHIFT_POISSON, typename R>
__global__ __launch_bounds__(256) void k_mdot_elem(const int32_t* __restrict__ conn,
                                                   const double* __restrict__ coords, const double* __restrict__ rho,
                                                   const double* __restrict__ p, const double* __restrict__ u,
                                                   const double* __restrict__ gp, double tau, int64_t nelem,
                                                   double* __restrict__ mdot) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  int32_t nd[8];
  load_nodes(conn, e, nd);
  R x[8][3];
  load_coords(coords, nd, x);
  double fr[8], fp[8], fu[8][3], fg[8][3];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int64_t a = nd[c];
    fr[c] = rho[a];
    fp[c] = p[a];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      fu[c][j] = u[3 * a + j];
      fg[c][j] = gp[3 * a + j];
    }
  }
  double out[12];
  hex8::static_for<12>([&](auto ii) {
    constexpr int IP = decltype(ii)::value;
    hex8::IpGeo<double, R> G;
    hex8::ip_geo<IP, SHIFT_POISSON>(x, G);
    out[IP] = (double)hex8::ip_mdot_regs<IP, SHIFT_MDOT, SHIFT_POISSON>(G, fr, fp, fu, fg, R(tau));
  });
  double2* o = reinterpret_cast<double2*>(mdot + 12 * e);
#pragma unroll
  for (int k = 0; k < 6; ++k) o[k] = make_double2(out[2 * k], out[2 * k + 1]);
}

// ---------------------------------------------------------------- generic sumInto
// Batched sumInto of caller-provided dense element matrices.
__global__ void k_sum_into_batch(const int32_t* __restrict__ conn, const double* __restrict__ lhs,
                                 const double* __restrict__ rhs_e, const int32_t* __restrict__ lid,
                                 const int32_t* __restrict__ nrp, const uint64_t* __restrict__ elem_pos,
                                 int64_t nelem, int nd, double* __restrict__ vals, double* __restrict__ rhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (element, row node, dof)
  if (t >= nelem * 8 * nd) return;
  const int64_t e = t / (8 * nd);
  const int rem = (int)(t - e * 8 * nd);
  const int a = rem / nd, d = rem % nd;
  const int R = 8 * nd;
  const int32_t nr = lid[conn[8 * e + a]];
  const int32_t nb = nrp[nr];
  const int32_t L = nrp[nr + 1] - nb;
  const uint64_t pk = elem_pos[8 * e + a];
  const double* lrow = lhs + (e * R + (int64_t)(a * nd + d)) * R;
  double* rowp = vals + (int64_t)nd * nd * nb + (int64_t)d * nd * L;
  for (int c = 0; c < 8; ++c) {
    const int p = (int)((pk >> (8 * c)) & 0xff);
    for (int j = 0; j < nd; ++j) unsafeAtomicAdd(rowp + nd * p + j, lrow[c * nd + j]);
  }
  unsafeAtomicAdd(rhs + (int64_t)nd * nr + d, rhs_e[e * R + a * nd + d]);
}

// ---------------------------------------------------------------- Dirichlet / resetRows
__global__ void k_dirichlet(const int32_t* __restrict__ nodes, int64_t nbc, int64_t nnodes, int nd, int begin,
                            int end, const int32_t* __restrict__ lid, const int32_t* __restrict__ rp,
                            const int32_t* __restrict__ ci, const double* __restrict__ sol,
                            const double* __restrict__ bcv, double* __restrict__ vals, double* __restrict__ rhs,
                            int reset_mode, double diag_value, double rhs_residual, int* __restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nsub = end - begin;
  if (t >= nbc * nsub) return;
  const int64_t k = t / nsub;
  const int d = begin + (int)(t % nsub);
  const int32_t node = nodes[k];
  if (node < 0 || node >= nnodes) {  // caller-provided list: never index outside the mesh
    atomicOr(bad, 1);
    return;
  }
  const int64_t row = (int64_t)nd * lid[node] + d;
  const double dv = reset_mode ? diag_value : 1.0;
  for (int32_t q = rp[row]; q < rp[row + 1]; ++q) vals[q] = (ci[q] == row) ? dv : 0.0;
  rhs[row] = reset_mode ? rhs_residual : bcv[(int64_t)nd * node + d] - sol[(int64_t)nd * node + d];
}

// rows (threads) per block of the row gather (128^3 continuity: 1.31 ms at
// 256, 1.34 ms at 64, 1.40 ms at 128)
int gather_rows() { return 256; }

template <class Op>
void launch_scalar(nalu_linsys* ls, const nalu_mesh* m, const Op& op, int mode) {
  hipStream_t s = ls->stream;
  const int64_t E = m->num_elems;
  if (mode == 0) {
    build_node_incidence(const_cast<nalu_mesh*>(m), s);
    const int accumulate = ls->values_pending_zero ? 0 : 1;
    NALU_REQUIRE(ls->values_pending_zero == ls->rhs_pending_zero, NALU_ERR_STATE,
                 "gather assembly: values/rhs zero state diverged");
    const size_t esz = sizeof(double);
    int B = gather_rows();
    while (B > 32 && (size_t)B * ls->max_node_row_len * esz > 64 * 1024) B >>= 1;
    const size_t lds = (size_t)B * ls->max_node_row_len * esz;
    hipLaunchKernelGGL((k_scalar_gather<Op>), dim3(blocks_for(ls->nrows, B)), dim3(B), lds, s, op, m->conn.p,
                       m->coords.p, (const int32_t*)ls->row_node.p, m->inc_ptr.p, m->inc_elem.p, ls->nrp.p,
                       ls->elem_pos.p, ls->nrows, ls->values.p, ls->rhs.p, accumulate);
    ls->values_pending_zero = ls->rhs_pending_zero = false;
  } else if (mode == 2) {
    build_node_incidence(const_cast<nalu_mesh*>(m), s);
    NALU_REQUIRE(ls->values_pending_zero == ls->rhs_pending_zero, NALU_ERR_STATE,
                 "tile assembly: values/rhs zero state diverged");
    const int accumulate = ls->values_pending_zero ? 0 : 1;
    const size_t lds = ((size_t)kTileNodes * ls->max_node_row_len + kTileNodes) * sizeof(double);
    set_lds_limit((const void*)k_scalar_tile<Op>, lds);
    hipLaunchKernelGGL((k_scalar_tile<Op>), dim3(blocks_for(ls->nrows, kTileNodes)), dim3(kTileThreads), lds, s, op,
                       m->conn.p, m->coords.p, (const int32_t*)ls->row_node.p, m->inc_ptr.p, m->inc_elem.p,
                       ls->nrp.p, ls->elem_pos.p, ls->nrows, ls->values.p, ls->rhs.p, accumulate);
    ls->values_pending_zero = ls->rhs_pending_zero = false;
  } else {
    materialize_pending_zero(ls);
    hipLaunchKernelGGL((k_scalar_atomic<Op>), dim3(chunk_blocks<8>(E)), dim3(kChunk), 0, s, op, m->conn.p,
                       m->coords.p, ls->node_lid.p, ls->nrp.p, ls->elem_pos.p, E, ls->values.p, ls->rhs.p);
  }
  NALU_HIP_CHECK(hipGetLastError());
}

// Systems created for the mixed-precision path assemble with FP32 arithmetic.
bool fp32_assembly(const nalu_linsys* ls) { return ls->cfg.precision == NALU_PRECISION_MIXED; }

template <typename R>
void launch_mdot(const nalu_mesh* m, const double* rho, const double* p, const double* u, const double* gp,
                 double tau, int shift_mdot, int shift_poisson, double* mdot, hipStream_t s) {
  const int64_t E = m->num_elems;
  NALU_REQUIRE(((uintptr_t)mdot & 15) == 0, NALU_ERR_ARG, "compute_mdot: mdot must be 16-byte aligned");
  const dim3 g(blocks_for(E, 256)), b(256);
  if (shift_mdot && shift_poisson)
    hipLaunchKernelGGL((k_mdot_elem<true, true, R>), g, b, 0, s, m->conn.p, m->coords.p, rho, p, u, gp, tau, E, mdot);
  else if (shift_mdot)
    hipLaunchKernelGGL((k_mdot_elem<true, false, R>), g, b, 0, s, m->conn.p, m->coords.p, rho, p, u, gp, tau, E, mdot);
  else if (shift_poisson)
    hipLaunchKernelGGL((k_mdot_elem<false, true, R>), g, b, 0, s, m->conn.p, m->coords.p, rho, p, u, gp, tau, E, mdot);
  else
    hipLaunchKernelGGL((k_mdot_elem<false, false, R>), g, b, 0, s, m->conn.p, m->coords.p, rho, p, u, gp, tau, E, mdot);
  NALU_HIP_CHECK(hipGetLastError());
}

void check_flag(DevBuf<int>& flag, hipStream_t s, const char* msg) {
  int h = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  NALU_REQUIRE(h == 0, NALU_ERR_ARG, msg);
}

// ---------------------------------------------------------------- kernel-input producers
// Dual nodal volume (ComputeGeometryInteriorAlgorithm, src/ComputeGeometryInteriorAlgorithm.C:77-120)
// and the interior Green-Gauss nodal gradient (AssembleNodalGradElemAlgorithm,
// src/AssembleNodalGradElemAlgorithm.C:85-140), owner-computes: one thread per
// node walks its incident elements; the node's scv (ip = its local index) and
// its three adjacent scs (sign +1 where it is the left node) are all that
// touch it, so every result is written once, with no atomics.
__global__ __launch_bounds__(256) void k_dual_volume(const int32_t* __restrict__ conn,
                                                     const double* __restrict__ coords,
                                                     const int32_t* __restrict__ inc_ptr,
                                                     const int32_t* __restrict__ inc_elem, int64_t nnodes,
                                                     double* __restrict__ vol) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nnodes) return;
  double v = 0.0;
  for (int32_t q = inc_ptr[a]; q < inc_ptr[a + 1]; ++q) {
    const int32_t t = inc_elem[q];
    int32_t nd[8];
    load_nodes(conn, t >> 3, nd);
    double x[8][3];
    load_vec3(coords, nd, x);
    switch (t & 7) {
#define NALU_DV(Lx) \
  case Lx: v += hex8::scv_volume<Lx>(x); break;
      NALU_DV(0) NALU_DV(1) NALU_DV(2) NALU_DV(3) NALU_DV(4) NALU_DV(5) NALU_DV(6) NALU_DV(7)
#undef NALU_DV
    }
  }
  vol[a] = v;
}

template <int LA, bool SHIFT>
__device__ __forceinline__ void nodal_grad_part(const double (&x)[8][3], const int32_t (&nd)[8],
                                                const double* __restrict__ qf, double (&g)[3]) {
  hex8::static_for<3>([&](auto kk) {
    constexpr int IP = hex8::adj_ip(LA, decltype(kk)::value);
    constexpr double S = hex8::adj_sign(LA, decltype(kk)::value);
    double A[3];
    hex8::scs_area<IP>(x, A);
    double qIp = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) qIp += hex8::kScsTab.shape[SHIFT][IP][c] * qf[nd[c]];
#pragma unroll
    for (int j = 0; j < 3; ++j) g[j] += S * (qIp * A[j]);
  });
}

template <bool SHIFT>
__global__ __launch_bounds__(256) void k_nodal_grad(const int32_t* __restrict__ conn,
                                                    const double* __restrict__ coords, const double* __restrict__ qf,
                                                    const double* __restrict__ vol,
                                                    const int32_t* __restrict__ inc_ptr,
                                                    const int32_t* __restrict__ inc_elem, int64_t nnodes,
                                                    double* __restrict__ dqdx) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nnodes) return;
  double g[3] = {0.0, 0.0, 0.0};
  for (int32_t q = inc_ptr[a]; q < inc_ptr[a + 1]; ++q) {
    const int32_t t = inc_elem[q];
    int32_t nd[8];
    load_nodes(conn, t >> 3, nd);
    double x[8][3];
    load_vec3(coords, nd, x);
    switch (t & 7) {
#define NALU_NG(Lx) \
  case Lx: nodal_grad_part<Lx, SHIFT>(x, nd, qf, g); break;
      NALU_NG(0) NALU_NG(1) NALU_NG(2) NALU_NG(3) NALU_NG(4) NALU_NG(5) NALU_NG(6) NALU_NG(7)
#undef NALU_NG
    }
  }
  const double inv = 1.0 / vol[a];
#pragma unroll
  for (int j = 0; j < 3; ++j) dqdx[3 * a + j] = g[j] * inv;
}

// ---------------------------------------------------------------- segregated momentum
// Momentum as three right-hand sides sharing one scalar matrix: the coupled
// MomentumAdvDiffElemKernel (src/kernel/MomentumAdvDiffElemKernel.C:72-186) with
// the transpose-gradient stress coupling kept on the rhs only (the rhs is the
// full residual, so converged velocities are unchanged).  Owner-computes row
// gather on a numDof 1 graph: the row segment accumulates in LDS, the three
// residual components go to rhs3[i][row].
template <bool SKEW, bool SHIFT, typename R, int MASS = 0>
__global__ __launch_bounds__(256) void k_momentum_seg_gather(
    const int32_t* __restrict__ conn, const double* __restrict__ coords, const double* __restrict__ u,
    const double* __restrict__ mu, const double* __restrict__ mdot, double incl_divu,
    const int32_t* __restrict__ row_node, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_elem,
    const int32_t* __restrict__ nrp, const uint64_t* __restrict__ elem_pos, int64_t nrows,
    double* __restrict__ vals, double* __restrict__ rhs3, int accumulate, MassArgs MA) {
  extern __shared__ double seg[];
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t rend = min(r0 + (int64_t)blockDim.x, nrows);
  const int64_t r = r0 + threadIdx.x;
  const int32_t s0 = nrp[r0], s1 = nrp[rend];
  const int nseg = s1 - s0;
  for (int i = threadIdx.x; i < nseg; i += blockDim.x) seg[i] = 0.0;
  __syncthreads();
  double rr[3] = {0.0, 0.0, 0.0};
  if (r < nrows) {
    const int32_t a = row_node ? row_node[r] : (int32_t)r;
    double* myrow = seg + (nrp[r] - s0);
    for (int32_t q = inc_ptr[a]; q < inc_ptr[a + 1]; ++q) {
      const int32_t t = inc_elem[q];
      const int64_t e = t >> 3;
      int32_t nd[8];
      load_nodes(conn, e, nd);
      R x[8][3];
      load_coords(coords, nd, x);
      const uint64_t pk = elem_pos[t];
      const double* md = mdot + 12 * e;
      const hex8::Gather gt{nd};
      auto emit = [&](int c, const R (&blk)[3][3]) { myrow[(int)((pk >> (8 * c)) & 0xff)] += (double)blk[0][0]; };
      R rl[3] = {R(0), R(0), R(0)};
      switch (t & 7) {
#define NALU_SG(Lx)                                                                                           \
  case Lx: {                                                                                                  \
    const R m3[3] = {R(md[hex8::adj_ip(Lx, 0)]), R(md[hex8::adj_ip(Lx, 1)]), R(md[hex8::adj_ip(Lx, 2)])};    \
    hex8::row_momentum<Lx, SKEW, SHIFT, double, true>(x, gt, u, mu, m3, R(incl_divu), rl, emit);              \
  } break;
        NALU_SG(0) NALU_SG(1) NALU_SG(2) NALU_SG(3) NALU_SG(4) NALU_SG(5) NALU_SG(6) NALU_SG(7)
#undef NALU_SG
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) rr[i] += (double)rl[i];
      if (MASS) {  // the momentum mass kernel of the same element and scv (MASS 1 consistent, 2 lumped)
        double xd[8][3];
        load_vec3(coords, nd, xd);
        switch (t & 7) {
#define NALU_SM(Lx) \
  case Lx: mass_row<Lx, true, MASS == 2, true>(MA, nd, xd, pk, myrow, 0, rr); break;
          NALU_SM(0) NALU_SM(1) NALU_SM(2) NALU_SM(3) NALU_SM(4) NALU_SM(5) NALU_SM(6) NALU_SM(7)
#undef NALU_SM
        }
      }
    }
  }
  __syncthreads();
  if (accumulate) {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) vals[s0 + i] += seg[i];
  } else {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) __builtin_nontemporal_store(seg[i], &vals[s0 + i]);
  }
  if (r < nrows) {
#pragma unroll
    for (int i = 0; i < 3; ++i) rhs3[i * nrows + r] += rr[i];
  }
}

template <bool MOM, bool LUMPED, bool SEG>
__global__ __launch_bounds__(256) void k_mass_gather(MassArgs A, const int32_t* __restrict__ conn,
                                                     const double* __restrict__ coords,
                                                     const int32_t* __restrict__ row_node,
                                                     const int32_t* __restrict__ inc_ptr,
                                                     const int32_t* __restrict__ inc_elem,
                                                     const int32_t* __restrict__ nrp,
                                                     const uint64_t* __restrict__ elem_pos, int64_t nnodes,
                                                     double* __restrict__ vals, double* __restrict__ rhs) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nnodes) return;
  const int32_t a = row_node ? row_node[r] : (int32_t)r;
  const int32_t L = nrp[r + 1] - nrp[r];
  double* row = vals + (MOM && !SEG ? 9 : 1) * (int64_t)nrp[r];
  double rh[3] = {0.0, 0.0, 0.0};
  for (int32_t q = inc_ptr[a]; q < inc_ptr[a + 1]; ++q) {
    const int32_t t = inc_elem[q];
    const int64_t e = t >> 3;
    int32_t nd[8];
    load_nodes(conn, e, nd);
    double x[8][3];
    load_vec3(coords, nd, x);
    const uint64_t pk = elem_pos[t];
    switch (t & 7) {
#define NALU_MS(Lx) \
  case Lx: mass_row<Lx, MOM, LUMPED, SEG>(A, nd, x, pk, row, L, rh); break;
      NALU_MS(0) NALU_MS(1) NALU_MS(2) NALU_MS(3) NALU_MS(4) NALU_MS(5) NALU_MS(6) NALU_MS(7)
#undef NALU_MS
    }
  }
  if (MOM && SEG) {
#pragma unroll
    for (int j = 0; j < 3; ++j) rhs[j * nnodes + r] += rh[j];
  } else if (MOM) {
#pragma unroll
    for (int j = 0; j < 3; ++j) rhs[3 * r + j] += rh[j];
  } else {
    rhs[r] += rh[0];
  }
}

// Momentum mass (numDof 3) as a node tile, the structure of k_momentum_tile:
// a block owns kTileNodes consecutive rows; thread (node j, slot s) adds the
// sub-control volume of local node s of each element in which node j has local
// index s (the 3x3 blocks' diagonal entries, LDS atomics: the 8 slots share
// the row), and the block adds its LDS segment to the matrix in one coalesced
// pass -- or stores it, folding a pending zeroSystem.  Replaces the
// thread-per-row form, whose scattered read-modify-writes of 81 diagonal
// entries per row went to HBM one 8-byte value at a time.
template <bool LUMPED>
__global__ __launch_bounds__(kTileThreads) void k_mass_tile3(MassArgs A, const int32_t* __restrict__ conn,
                                                             const double* __restrict__ coords,
                                                             const int32_t* __restrict__ row_node,
                                                             const int32_t* __restrict__ inc_ptr,
                                                             const int32_t* __restrict__ inc_elem,
                                                             const int32_t* __restrict__ nrp,
                                                             const uint64_t* __restrict__ elem_pos, int64_t nnodes,
                                                             double* __restrict__ vals, double* __restrict__ rhs,
                                                             int accumulate) {
  extern __shared__ double seg[];
  const int64_t a0 = (int64_t)blockIdx.x * kTileNodes;
  const int64_t a1 = min(a0 + (int64_t)kTileNodes, nnodes);
  const int64_t s0 = 9 * (int64_t)nrp[a0], s1 = 9 * (int64_t)nrp[a1];
  const int nseg = (int)(s1 - s0);
  double* rseg = seg + nseg;
  for (int i = threadIdx.x; i < nseg + 3 * kTileNodes; i += blockDim.x) seg[i] = 0.0;
  __syncthreads();
  const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileNodes);
  const int j = threadIdx.x % kTileNodes;
  const int64_t r = a0 + j;
  if (r < a1) {
    const int32_t a = row_node ? row_node[r] : (int32_t)r;
    const int32_t L = nrp[r + 1] - nrp[r];
    double* myrow = seg + (9 * (int64_t)nrp[r] - s0);
    double rh[3] = {0.0, 0.0, 0.0};
    for (int32_t q = inc_ptr[a]; q < inc_ptr[a + 1]; ++q) {
      const int32_t t = inc_elem[q];
      if ((t & 7) != slot) continue;
      const int64_t e = t >> 3;
      int32_t nd[8];
      load_nodes(conn, e, nd);
      double x[8][3];
      load_vec3(coords, nd, x);
      const uint64_t pk = elem_pos[t];
      switch (slot) {
#define NALU_MS(Lx) \
  case Lx: mass_row<Lx, true, LUMPED, false, true>(A, nd, x, pk, myrow, L, rh); break;
        NALU_MS(0) NALU_MS(1) NALU_MS(2) NALU_MS(3) NALU_MS(4) NALU_MS(5) NALU_MS(6) NALU_MS(7)
#undef NALU_MS
      }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (rh[i] != 0.0) unsafeAtomicAdd(rseg + 3 * j + i, rh[i]);
  }
  __syncthreads();
  const int nr = (int)(3 * (a1 - a0));
  if (accumulate) {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) vals[s0 + i] += seg[i];
    for (int i = threadIdx.x; i < nr; i += blockDim.x) rhs[3 * a0 + i] += rseg[i];
  } else {
    for (int i = threadIdx.x; i < nseg; i += blockDim.x) __builtin_nontemporal_store(seg[i], &vals[s0 + i]);
    for (int i = threadIdx.x; i < nr; i += blockDim.x) rhs[3 * a0 + i] = rseg[i];
  }
}

void check_assembly_args(nalu_linsys* ls, const nalu_mesh* m, int nd) {
  NALU_REQUIRE(ls && m, NALU_ERR_ARG, "assembly: null handle");
  NALU_REQUIRE(ls->finalized, NALU_ERR_STATE, "assembly: finalizeLinearSystem not called");
  NALU_REQUIRE(ls->mesh == m, NALU_ERR_ARG, "assembly: mesh differs from the one the graph was built on");
  NALU_REQUIRE(ls->num_dof == nd, NALU_ERR_ARG, "assembly: numDof mismatch for this kernel");
  // queued host sumInto calls precede this assembly (and keep it from folding the zero)
  flush_sum_into_queue(ls);
}

}  // namespace

// Opt a kernel into more than the default 64 KiB of dynamic LDS when needed.
void set_lds_limit(const void* kern, size_t bytes) {
  int dev = 0;
  NALU_HIP_CHECK(hipGetDevice(&dev));
  int maxopt = 0;
  NALU_HIP_CHECK(hipDeviceGetAttribute(&maxopt, hipDeviceAttributeSharedMemPerBlockOptin, dev));
  if (maxopt <= 0) NALU_HIP_CHECK(hipDeviceGetAttribute(&maxopt, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  NALU_REQUIRE(bytes <= (size_t)maxopt, NALU_ERR_UNSUPPORTED,
               "assembly tile needs " + std::to_string(bytes) + " B of LDS, device allows " + std::to_string(maxopt));
  if (bytes > 64 * 1024) NALU_HIP_CHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

void materialize_pending_zero(nalu_linsys* ls) {
  if (ls->values_pending_zero)
    NALU_HIP_CHECK(hipMemsetAsync(ls->values.p, 0, ls->values.bytes(), ls->stream));
  if (ls->rhs_pending_zero) NALU_HIP_CHECK(hipMemsetAsync(ls->rhs.p, 0, ls->rhs.bytes(), ls->stream));
  ls->values_pending_zero = ls->rhs_pending_zero = false;
  // host sumInto calls queued since the zero come next
  flush_sum_into_queue(ls);
}

}  // namespace nalu

using namespace nalu;

extern "C" int nalu_assemble_scalar_diff(nalu_linsys* ls, const nalu_mesh* m, const double* q, const double* k,
                                         int shifted, int mode) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 1);
    NALU_REQUIRE(q && k, NALU_ERR_ARG, "scalar_diff: null field");
    NALU_REQUIRE(mode >= 0 && mode <= 3, NALU_ERR_ARG, "scalar_diff: mode must be 0, 1, 2 or 3");
    PhaseTimer t(ls, PH_ASSEMBLE);
    if (mode == 3) {
      etile_scalar_diff(ls, m, q, k, shifted != 0, fp32_assembly(ls));
    } else if (fp32_assembly(ls)) {
      if (shifted) launch_scalar(ls, m, ScalarDiffOp<true, float>{q, k}, mode);
      else launch_scalar(ls, m, ScalarDiffOp<false, float>{q, k}, mode);
    } else {
      if (shifted) launch_scalar(ls, m, ScalarDiffOp<true>{q, k}, mode);
      else launch_scalar(ls, m, ScalarDiffOp<false>{q, k}, mode);
    }
    return NALU_OK;
  });
}

extern "C" int nalu_assemble_continuity(nalu_linsys* ls, const nalu_mesh* m, const double* rho, const double* p,
                                        const double* u, const double* gp, double tau, int shift_mdot,
                                        int shift_poisson, int reduced_sens, int mode) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 1);
    NALU_REQUIRE(rho && p && u && gp, NALU_ERR_ARG, "continuity: null field");
    NALU_REQUIRE(tau > 0.0, NALU_ERR_ARG, "continuity: projTimeScale must be positive");
    NALU_REQUIRE(mode >= 0 && mode <= 3, NALU_ERR_ARG, "continuity: mode must be 0, 1, 2 or 3");
    PhaseTimer t(ls, PH_ASSEMBLE);
    const int key = (shift_mdot ? 4 : 0) | (shift_poisson ? 2 : 0) | (reduced_sens ? 1 : 0);
    if (mode == 3) {
      etile_continuity(ls, m, rho, p, u, gp, tau, key, fp32_assembly(ls));
      return NALU_OK;
    }
    switch (key) {
#define NALU_CT(K, A, B, C)                                                          \
  case K:                                                                            \
    if (fp32_assembly(ls)) launch_scalar(ls, m, ContinuityOp<A, B, C, float>{rho, p, u, gp, tau}, mode); \
    else launch_scalar(ls, m, ContinuityOp<A, B, C>{rho, p, u, gp, tau}, mode);       \
    break;
      NALU_CT(0, false, false, false)
      NALU_CT(1, false, false, true)
      NALU_CT(2, false, true, false)
      NALU_CT(3, false, true, true)
      NALU_CT(4, true, false, false)
      NALU_CT(5, true, false, true)
      NALU_CT(6, true, true, false)
      NALU_CT(7, true, true, true)
#undef NALU_CT
    }
    return NALU_OK;
  });
}

extern "C" int nalu_assemble_continuity_mdot(nalu_linsys* ls, const nalu_mesh* m, const double* rho,
                                             const double* p, const double* u, const double* gp, double tau,
                                             int shift_mdot, int shift_poisson, int reduced_sens, double* mdot) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 1);
    NALU_REQUIRE(rho && p && u && gp && mdot, NALU_ERR_ARG, "continuity + mdot: null field");
    NALU_REQUIRE(tau > 0.0, NALU_ERR_ARG, "continuity + mdot: projTimeScale must be positive");
    PhaseTimer t(ls, PH_ASSEMBLE);
    const int key = (shift_mdot ? 4 : 0) | (shift_poisson ? 2 : 0) | (reduced_sens ? 1 : 0);
    switch (key) {
#define NALU_CM(K, A, B, C)                                                                              \
  case K:                                                                                                \
    if (fp32_assembly(ls)) launch_scalar(ls, m, ContinuityOp<A, B, C, float>{rho, p, u, gp, tau, mdot}, 0); \
    else launch_scalar(ls, m, ContinuityOp<A, B, C>{rho, p, u, gp, tau, mdot}, 0);                        \
    break;
      NALU_CM(0, false, false, false)
      NALU_CM(1, false, false, true)
      NALU_CM(2, false, true, false)
      NALU_CM(3, false, true, true)
      NALU_CM(4, true, false, false)
      NALU_CM(5, true, false, true)
      NALU_CM(6, true, true, false)
      NALU_CM(7, true, true, true)
#undef NALU_CM
    }
    return NALU_OK;
  });
}

extern "C" int nalu_compute_mdot(const nalu_mesh* m, const double* rho, const double* p, const double* u,
                                 const double* gp, double tau, int shift_mdot, int shift_poisson, double* mdot,
                                 void* stream) {
  return guarded([&]() -> int {
    NALU_REQUIRE(m && rho && p && u && gp && mdot, NALU_ERR_ARG, "mdot: null argument");
    NALU_REQUIRE(tau > 0.0, NALU_ERR_ARG, "mdot: projTimeScale must be positive");
    launch_mdot<double>(m, rho, p, u, gp, tau, shift_mdot, shift_poisson, mdot, (hipStream_t)stream);
    return NALU_OK;
  });
}

extern "C" int nalu_compute_mdot_f32(const nalu_mesh* m, const double* rho, const double* p, const double* u,
                                     const double* gp, double tau, int shift_mdot, int shift_poisson, double* mdot,
                                     void* stream) {
  return guarded([&]() -> int {
    NALU_REQUIRE(m && rho && p && u && gp && mdot, NALU_ERR_ARG, "mdot: null argument");
    NALU_REQUIRE(tau > 0.0, NALU_ERR_ARG, "mdot: projTimeScale must be positive");
    launch_mdot<float>(m, rho, p, u, gp, tau, shift_mdot, shift_poisson, mdot, (hipStream_t)stream);
    return NALU_OK;
  });
}

namespace {
// The node-tile momentum pass; ma != nullptr adds the momentum mass kernel in
// the same pass.
void launch_momentum_tile(nalu_linsys* ls, const nalu_mesh* m, const double* u, const double* mu, const double* mdot,
                          double include_divu, bool skew, bool shifted, const MassArgs* ma, bool lumped) {
  build_node_incidence(const_cast<nalu_mesh*>(m), ls->stream);
  NALU_REQUIRE(ls->values_pending_zero == ls->rhs_pending_zero, NALU_ERR_STATE,
               "tile assembly: values/rhs zero state diverged");
  const int accumulate = ls->values_pending_zero ? 0 : 1;
  const size_t lds = ((size_t)kTileNodes * ls->max_node_row_len * 9 + 3 * kTileNodes) * sizeof(double);
  const MassArgs none{};
  const MassArgs& A = ma ? *ma : none;
  auto go = [&](auto kern) {
    set_lds_limit((const void*)kern, lds);
    hipLaunchKernelGGL(kern, dim3(blocks_for(ls->num_nodes, kTileNodes)), dim3(kTileThreads), lds, ls->stream,
                       m->conn.p, m->coords.p, u, mu, mdot, include_divu, (const int32_t*)ls->row_node.p,
                       m->inc_ptr.p, m->inc_elem.p, ls->nrp.p, ls->elem_pos.p, ls->num_nodes, ls->values.p,
                       ls->rhs.p, accumulate, A);
  };
  const bool f32 = fp32_assembly(ls);
  if (ma) {  // the mass part is FP64 arithmetic at either precision of the momentum part
    // one-pass forms for the default advection options only (compile time)
    NALU_REQUIRE(!skew && !shifted, NALU_ERR_UNSUPPORTED, "momentum tile: one pass with mass needs skew = shifted = 0");
    if (lumped) f32 ? go(k_momentum_tile<false, false, float, 2>) : go(k_momentum_tile<false, false, double, 2>);
    else f32 ? go(k_momentum_tile<false, false, float, 1>) : go(k_momentum_tile<false, false, double, 1>);
  } else {
    if (skew && shifted) f32 ? go(k_momentum_tile<true, true, float>) : go(k_momentum_tile<true, true, double>);
    else if (skew) f32 ? go(k_momentum_tile<true, false, float>) : go(k_momentum_tile<true, false, double>);
    else if (shifted) f32 ? go(k_momentum_tile<false, true, float>) : go(k_momentum_tile<false, true, double>);
    else f32 ? go(k_momentum_tile<false, false, float>) : go(k_momentum_tile<false, false, double>);
  }
  ls->values_pending_zero = ls->rhs_pending_zero = false;
}
}  // namespace

extern "C" int nalu_assemble_momentum(nalu_linsys* ls, const nalu_mesh* m, const double* u, const double* mu,
                                      const double* mdot, double include_divu, int skew, int shifted, int mode) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 3);
    NALU_REQUIRE(u && mu && mdot, NALU_ERR_ARG, "momentum: null field");
    NALU_REQUIRE(mode >= 1 && mode <= 3, NALU_ERR_UNSUPPORTED,
                 "momentum: mode 1 (atomic), 2 (node tile) or 3 (element tile); the per-thread row gather is "
                 "scalar-only");
    PhaseTimer t(ls, PH_ASSEMBLE);
    const int64_t E = m->num_elems;
    if (mode == 3) {
      build_node_incidence(const_cast<nalu_mesh*>(m), ls->stream);
      etile_momentum(ls, m, u, mu, mdot, include_divu, skew != 0, shifted != 0, fp32_assembly(ls));
    } else if (mode == 2) {
      launch_momentum_tile(ls, m, u, mu, mdot, include_divu, skew != 0, shifted != 0, nullptr, false);
    } else {
      materialize_pending_zero(ls);
      const dim3 g(chunk_blocks<8>(E)), b(kChunk);
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, g, b, 0, ls->stream, m->conn.p, m->coords.p, u, mu, mdot, include_divu,
                           ls->node_lid.p, ls->nrp.p, ls->elem_pos.p, E, ls->values.p, ls->rhs.p);
      };
      const bool f32 = fp32_assembly(ls);
      if (skew && shifted) f32 ? go(k_momentum_atomic<true, true, float>) : go(k_momentum_atomic<true, true, double>);
      else if (skew) f32 ? go(k_momentum_atomic<true, false, float>) : go(k_momentum_atomic<true, false, double>);
      else if (shifted) f32 ? go(k_momentum_atomic<false, true, float>) : go(k_momentum_atomic<false, true, double>);
      else f32 ? go(k_momentum_atomic<false, false, float>) : go(k_momentum_atomic<false, false, double>);
    }
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_sum_into_batch(nalu_linsys* ls, int64_t nelem, const int32_t* conn, const double* lhs,
                                          const double* rhs_e) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "sum_into_batch: system not finalized");
    NALU_REQUIRE(nelem >= 0 && (nelem == 0 || (conn && lhs && rhs_e)), NALU_ERR_ARG, "sum_into_batch: bad args");
    NALU_REQUIRE(ls->mesh && nelem <= ls->mesh->num_elems, NALU_ERR_ARG, "sum_into_batch: more elements than the graph");
    PhaseTimer t(ls, PH_ASSEMBLE);
    materialize_pending_zero(ls);
    if (nelem == 0) return NALU_OK;
    // the element scatter table indexes the graph's own elements
    NALU_REQUIRE(conn == ls->mesh->conn.p, NALU_ERR_ARG,
                 "sum_into_batch: conn must be the graph mesh's connectivity (nalu_mesh_device_ptrs)");
    const int64_t n = nelem * 8 * ls->num_dof;
    hipLaunchKernelGGL(k_sum_into_batch, dim3(blocks_for(n, 256)), dim3(256), 0, ls->stream, conn, lhs, rhs_e,
                       ls->node_lid.p, ls->nrp.p, ls->elem_pos.p, nelem, ls->num_dof, ls->values.p, ls->rhs.p);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_apply_dirichlet(nalu_linsys* ls, int64_t nbc, const int32_t* nodes, const double* sol,
                                           const double* bcv, int begin, int end) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "applyDirichletBCs: system not finalized");
    NALU_REQUIRE(begin >= 0 && end <= ls->num_dof && begin <= end, NALU_ERR_ARG, "applyDirichletBCs: bad dof range");
    NALU_REQUIRE(nbc == 0 || (nodes && sol && bcv), NALU_ERR_ARG, "applyDirichletBCs: null array");
    PhaseTimer t(ls, PH_DIRICHLET);
    materialize_pending_zero(ls);
    const int64_t n = nbc * (end - begin);
    if (n == 0) return NALU_OK;
    DevBuf<int> bad(1);
    NALU_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), ls->stream));
    hipLaunchKernelGGL(k_dirichlet, dim3(blocks_for(n, 256)), dim3(256), 0, ls->stream, nodes, nbc, ls->num_nodes,
                       ls->num_dof, begin, end, ls->node_lid.p, ls->row_ptr.p, ls->col_idx.p, sol, bcv,
                       ls->values.p, ls->rhs.p, 0, 0.0, 0.0, bad.p);
    NALU_HIP_CHECK(hipGetLastError());
    check_flag(bad, ls->stream, "applyDirichletBCs: node index outside the mesh");
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_reset_rows(nalu_linsys* ls, int64_t nn, const int32_t* nodes, int begin, int end,
                                      double diag_value, double rhs_residual) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "resetRows: system not finalized");
    NALU_REQUIRE(begin >= 0 && end <= ls->num_dof && begin <= end, NALU_ERR_ARG, "resetRows: bad dof range");
    NALU_REQUIRE(nn == 0 || nodes, NALU_ERR_ARG, "resetRows: null node list");
    PhaseTimer t(ls, PH_DIRICHLET);
    materialize_pending_zero(ls);
    const int64_t n = nn * (end - begin);
    if (n == 0) return NALU_OK;
    DevBuf<int> bad(1);
    NALU_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), ls->stream));
    hipLaunchKernelGGL(k_dirichlet, dim3(blocks_for(n, 256)), dim3(256), 0, ls->stream, nodes, nn, ls->num_nodes,
                       ls->num_dof, begin, end, ls->node_lid.p, ls->row_ptr.p, ls->col_idx.p, (const double*)nullptr,
                       (const double*)nullptr, ls->values.p, ls->rhs.p, 1, diag_value, rhs_residual, bad.p);
    NALU_HIP_CHECK(hipGetLastError());
    check_flag(bad, ls->stream, "resetRows: node index outside the mesh");
    return NALU_OK;
  });
}

namespace {
void launch_mass(nalu_linsys* ls, const nalu_mesh* m, const MassArgs& A, bool mom, bool lumped,
                 double* rhs3 = nullptr) {
  build_node_incidence(const_cast<nalu_mesh*>(m), ls->stream);
  const size_t lds3 = ((size_t)kTileNodes * ls->max_node_row_len * 9 + 3 * kTileNodes) * sizeof(double);
  if (mom && !rhs3 && lds3 <= kMaxLds && ls->values_pending_zero == ls->rhs_pending_zero) {
    // numDof 3: node tile with the row segments in LDS
    const int accumulate = ls->values_pending_zero ? 0 : 1;
    auto go3 = [&](auto kern) {
      set_lds_limit((const void*)kern, lds3);
      hipLaunchKernelGGL(kern, dim3(blocks_for(ls->num_nodes, kTileNodes)), dim3(kTileThreads), lds3, ls->stream, A,
                         m->conn.p, m->coords.p, (const int32_t*)ls->row_node.p, m->inc_ptr.p, m->inc_elem.p,
                         ls->nrp.p, ls->elem_pos.p, ls->num_nodes, ls->values.p, ls->rhs.p, accumulate);
    };
    lumped ? go3(k_mass_tile3<true>) : go3(k_mass_tile3<false>);
    ls->values_pending_zero = ls->rhs_pending_zero = false;
    NALU_HIP_CHECK(hipGetLastError());
    return;
  }
  materialize_pending_zero(ls);
  const dim3 g(blocks_for(ls->num_nodes, 256)), b(256);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, g, b, 0, ls->stream, A, m->conn.p, m->coords.p, (const int32_t*)ls->row_node.p,
                       m->inc_ptr.p, m->inc_elem.p, ls->nrp.p, ls->elem_pos.p, ls->num_nodes, ls->values.p,
                       rhs3 ? rhs3 : ls->rhs.p);
  };
  if (rhs3) lumped ? go(k_mass_gather<true, true, true>) : go(k_mass_gather<true, false, true>);
  else if (mom) lumped ? go(k_mass_gather<true, true, false>) : go(k_mass_gather<true, false, false>);
  else lumped ? go(k_mass_gather<false, true, false>) : go(k_mass_gather<false, false, false>);
  NALU_HIP_CHECK(hipGetLastError());
}

void check_mass_args(double dt, const double* gamma) {
  NALU_REQUIRE(gamma, NALU_ERR_ARG, "mass: null gamma");
  NALU_REQUIRE(std::isfinite(dt) && dt > 0.0, NALU_ERR_ARG, "mass: dt must be positive");
  NALU_REQUIRE(std::isfinite(gamma[0]) && gamma[0] != 0.0 && std::isfinite(gamma[1]) && std::isfinite(gamma[2]),
               NALU_ERR_ARG, "mass: gamma1 must be nonzero and all gammas finite");
}
}  // namespace

extern "C" int nalu_assemble_momentum_mass(nalu_linsys* ls, const nalu_mesh* m, const double* rho_nm1,
                                           const double* rho_n, const double* rho_np1, const double* u_nm1,
                                           const double* u_n, const double* u_np1, const double* gjp, double dt,
                                           const double* gamma, int lumped, int balanced_force) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 3);
    check_mass_args(dt, gamma);
    NALU_REQUIRE(rho_nm1 && rho_n && rho_np1 && u_nm1 && u_n && u_np1 && gjp, NALU_ERR_ARG,
                 "momentum_mass: null field");
    PhaseTimer t(ls, PH_ASSEMBLE);
    const MassArgs A{{rho_nm1, rho_n, rho_np1}, {u_nm1, u_n, u_np1}, gjp, dt, gamma[0], gamma[1], gamma[2],
                     balanced_force ? 1 : 0};
    launch_mass(ls, m, A, true, lumped != 0);
    return NALU_OK;
  });
}

extern "C" int nalu_assemble_momentum_with_mass(nalu_linsys* ls, const nalu_mesh* m, const double* u, const double* mu,
                                                const double* mdot, double include_divu, int skew, int shifted,
                                                const double* rho_nm1, const double* rho_n, const double* rho_np1,
                                                const double* u_nm1, const double* u_n, const double* u_np1,
                                                const double* gjp, double dt, const double* gamma, int lumped,
                                                int balanced_force) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 3);
    check_mass_args(dt, gamma);
    NALU_REQUIRE(u && mu && mdot, NALU_ERR_ARG, "momentum_with_mass: null advection-diffusion field");
    NALU_REQUIRE(rho_nm1 && rho_n && rho_np1 && u_nm1 && u_n && u_np1 && gjp, NALU_ERR_ARG,
                 "momentum_with_mass: null mass field");
    const size_t lds = ((size_t)kTileNodes * ls->max_node_row_len * 9 + 3 * kTileNodes) * sizeof(double);
    // the two passes one after the other: rows too long for a node tile, or the
    // skew-symmetric / shifted advection options (one-pass forms exist for the defaults)
    if (lds > kMaxLds || skew || shifted) {
      int rc = nalu_assemble_momentum(ls, m, u, mu, mdot, include_divu, skew, shifted, lds > kMaxLds ? 1 : 2);
      if (rc != NALU_OK) return rc;
      return nalu_assemble_momentum_mass(ls, m, rho_nm1, rho_n, rho_np1, u_nm1, u_n, u_np1, gjp, dt, gamma, lumped,
                                         balanced_force);
    }
    PhaseTimer t(ls, PH_ASSEMBLE);
    const MassArgs A{{rho_nm1, rho_n, rho_np1}, {u_nm1, u_n, u_np1}, gjp, dt, gamma[0], gamma[1], gamma[2],
                     balanced_force ? 1 : 0};
    launch_momentum_tile(ls, m, u, mu, mdot, include_divu, skew != 0, shifted != 0, &A, lumped != 0);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

extern "C" int nalu_assemble_continuity_mass(nalu_linsys* ls, const nalu_mesh* m, const double* rho_nm1,
                                             const double* rho_n, const double* rho_np1, double dt,
                                             const double* gamma, int lumped) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 1);
    check_mass_args(dt, gamma);
    NALU_REQUIRE(rho_nm1 && rho_n && rho_np1, NALU_ERR_ARG, "continuity_mass: null field");
    PhaseTimer t(ls, PH_ASSEMBLE);
    const MassArgs A{{rho_nm1, rho_n, rho_np1}, {nullptr, nullptr, nullptr}, nullptr, dt, gamma[0], gamma[1],
                     gamma[2], 0};
    launch_mass(ls, m, A, false, lumped != 0);
    return NALU_OK;
  });
}

namespace {
// The segregated momentum row gather; ma != nullptr adds the momentum mass
// kernel in the same pass.
void launch_momentum_seg(nalu_linsys* ls, const nalu_mesh* m, const double* u, const double* mu, const double* mdot,
                         double include_divu, bool skew, bool shifted, double* rhs3, const MassArgs* ma, bool lumped) {
  build_node_incidence(const_cast<nalu_mesh*>(m), ls->stream);
  if (ls->rhs_pending_zero) NALU_HIP_CHECK(hipMemsetAsync(ls->rhs.p, 0, ls->rhs.bytes(), ls->stream));
  ls->rhs_pending_zero = false;
  const int accumulate = ls->values_pending_zero ? 0 : 1;
  constexpr int B = 256;
  const size_t lds = (size_t)B * ls->max_node_row_len * sizeof(double);
  const MassArgs none{};
  const MassArgs& A = ma ? *ma : none;
  auto go = [&](auto kern) {
    set_lds_limit((const void*)kern, lds);
    hipLaunchKernelGGL(kern, dim3(blocks_for(ls->nrows, B)), dim3(B), lds, ls->stream, m->conn.p, m->coords.p, u, mu,
                       mdot, include_divu, (const int32_t*)ls->row_node.p, m->inc_ptr.p, m->inc_elem.p, ls->nrp.p,
                       ls->elem_pos.p, ls->nrows, ls->values.p, rhs3, accumulate, A);
  };
  const bool f32 = fp32_assembly(ls);
  if (ma) {  // the mass part is FP64 arithmetic at either precision of the momentum part
    // one-pass forms for the default advection options only (compile time)
    NALU_REQUIRE(!skew && !shifted, NALU_ERR_UNSUPPORTED,
                 "momentum gather: one pass with mass needs skew = shifted = 0");
    if (lumped) f32 ? go(k_momentum_seg_gather<false, false, float, 2>) : go(k_momentum_seg_gather<false, false, double, 2>);
    else f32 ? go(k_momentum_seg_gather<false, false, float, 1>) : go(k_momentum_seg_gather<false, false, double, 1>);
  } else {
    if (skew && shifted) f32 ? go(k_momentum_seg_gather<true, true, float>) : go(k_momentum_seg_gather<true, true, double>);
    else if (skew) f32 ? go(k_momentum_seg_gather<true, false, float>) : go(k_momentum_seg_gather<true, false, double>);
    else if (shifted) f32 ? go(k_momentum_seg_gather<false, true, float>) : go(k_momentum_seg_gather<false, true, double>);
    else f32 ? go(k_momentum_seg_gather<false, false, float>) : go(k_momentum_seg_gather<false, false, double>);
  }
  ls->values_pending_zero = false;
  NALU_HIP_CHECK(hipGetLastError());
}
}  // namespace

extern "C" int nalu_assemble_momentum_segregated(nalu_linsys* ls, const nalu_mesh* m, const double* u,
                                                 const double* mu, const double* mdot, double include_divu,
                                                 int skew, int shifted, double* rhs3) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 1);
    NALU_REQUIRE(u && mu && mdot && rhs3, NALU_ERR_ARG, "momentum_segregated: null field");
    PhaseTimer t(ls, PH_ASSEMBLE);
    launch_momentum_seg(ls, m, u, mu, mdot, include_divu, skew != 0, shifted != 0, rhs3, nullptr, false);
    return NALU_OK;
  });
}

extern "C" int nalu_assemble_momentum_mass_segregated(nalu_linsys* ls, const nalu_mesh* m, const double* rho_nm1,
                                                      const double* rho_n, const double* rho_np1,
                                                      const double* u_nm1, const double* u_n, const double* u_np1,
                                                      const double* gjp, double dt, const double* gamma, int lumped,
                                                      int balanced_force, double* rhs3) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 1);
    check_mass_args(dt, gamma);
    NALU_REQUIRE(rho_nm1 && rho_n && rho_np1 && u_nm1 && u_n && u_np1 && gjp && rhs3, NALU_ERR_ARG,
                 "momentum_mass_segregated: null field");
    PhaseTimer t(ls, PH_ASSEMBLE);
    const MassArgs A{{rho_nm1, rho_n, rho_np1}, {u_nm1, u_n, u_np1}, gjp, dt, gamma[0], gamma[1], gamma[2],
                     balanced_force ? 1 : 0};
    launch_mass(ls, m, A, true, lumped != 0, rhs3);
    return NALU_OK;
  });
}

extern "C" int nalu_assemble_momentum_with_mass_segregated(
    nalu_linsys* ls, const nalu_mesh* m, const double* u, const double* mu, const double* mdot, double include_divu,
    int skew, int shifted, const double* rho_nm1, const double* rho_n, const double* rho_np1, const double* u_nm1,
    const double* u_n, const double* u_np1, const double* gjp, double dt, const double* gamma, int lumped,
    int balanced_force, double* rhs3) {
  return guarded([&]() -> int {
    check_assembly_args(ls, m, 1);
    check_mass_args(dt, gamma);
    NALU_REQUIRE(u && mu && mdot && rhs3, NALU_ERR_ARG, "momentum_with_mass_segregated: null field");
    NALU_REQUIRE(rho_nm1 && rho_n && rho_np1 && u_nm1 && u_n && u_np1 && gjp, NALU_ERR_ARG,
                 "momentum_with_mass_segregated: null mass field");
    if (skew || shifted) {  // one-pass forms exist for the default advection options
      const int rc = nalu_assemble_momentum_segregated(ls, m, u, mu, mdot, include_divu, skew, shifted, rhs3);
      if (rc != NALU_OK) return rc;
      return nalu_assemble_momentum_mass_segregated(ls, m, rho_nm1, rho_n, rho_np1, u_nm1, u_n, u_np1, gjp, dt, gamma,
                                                    lumped, balanced_force, rhs3);
    }
    PhaseTimer t(ls, PH_ASSEMBLE);
    const MassArgs A{{rho_nm1, rho_n, rho_np1}, {u_nm1, u_n, u_np1}, gjp, dt, gamma[0], gamma[1], gamma[2],
                     balanced_force ? 1 : 0};
    launch_momentum_seg(ls, m, u, mu, mdot, include_divu, false, false, rhs3, &A, lumped != 0);
    return NALU_OK;
  });
}

extern "C" int nalu_compute_dual_volume(const nalu_mesh* m, double* dual_vol, void* stream) {
  return guarded([&]() -> int {
    NALU_REQUIRE(m && dual_vol, NALU_ERR_ARG, "dual_volume: null argument");
    hipStream_t s = (hipStream_t)stream;
    build_node_incidence(const_cast<nalu_mesh*>(m), s);
    hipLaunchKernelGGL(k_dual_volume, dim3(blocks_for(m->num_nodes, 256)), dim3(256), 0, s, m->conn.p, m->coords.p,
                       m->inc_ptr.p, m->inc_elem.p, m->num_nodes, dual_vol);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

namespace {
void nodal_grad_interior(const nalu_mesh* m, const double* q, const double* dual_vol, int shifted, double* dqdx,
                         hipStream_t s) {
  build_node_incidence(const_cast<nalu_mesh*>(m), s);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(blocks_for(m->num_nodes, 256)), dim3(256), 0, s, m->conn.p, m->coords.p, q,
                       dual_vol, m->inc_ptr.p, m->inc_elem.p, m->num_nodes, dqdx);
  };
  if (shifted) go(k_nodal_grad<true>);
  else go(k_nodal_grad<false>);
  NALU_HIP_CHECK(hipGetLastError());
}
}  // namespace

extern "C" int nalu_compute_nodal_gradient(const nalu_mesh* m, const double* q, const double* dual_vol, int shifted,
                                           double* dqdx, void* stream) {
  return guarded([&]() -> int {
    NALU_REQUIRE(m && q && dual_vol && dqdx, NALU_ERR_ARG, "nodal_gradient: null argument");
    hipStream_t s = (hipStream_t)stream;
    nodal_grad_interior(m, q, dual_vol, shifted, dqdx, s);
    nodal_grad_boundary(const_cast<nalu_mesh*>(m), q, dual_vol, shifted != 0, dqdx, s);
    return NALU_OK;
  });
}

extern "C" int nalu_compute_nodal_gradient_interior(const nalu_mesh* m, const double* q, const double* dual_vol,
                                                    int shifted, double* dqdx, void* stream) {
  return guarded([&]() -> int {
    NALU_REQUIRE(m && q && dual_vol && dqdx, NALU_ERR_ARG, "nodal_gradient: null argument");
    nodal_grad_interior(m, q, dual_vol, shifted, dqdx, (hipStream_t)stream);
    return NALU_OK;
  });
}

extern "C" int nalu_nodal_gradient_boundary(const nalu_mesh* m, const double* q, const double* dual_vol, int shifted,
                                            double* dqdx, void* stream) {
  return guarded([&]() -> int {
    NALU_REQUIRE(m && q && dual_vol && dqdx, NALU_ERR_ARG, "nodal_gradient_boundary: null argument");
    nodal_grad_boundary(const_cast<nalu_mesh*>(m), q, dual_vol, shifted != 0, dqdx, (hipStream_t)stream);
    return NALU_OK;
  });
}

extern "C" int nalu_mesh_exposed_faces(const nalu_mesh* m, int64_t* num_faces, int32_t* faces_host) {
  return guarded([&]() -> int {
    NALU_REQUIRE(m && num_faces, NALU_ERR_ARG, "mesh_exposed_faces: null argument");
    build_exposed_faces(const_cast<nalu_mesh*>(m), nullptr);
    *num_faces = m->num_exposed;
    if (faces_host && m->num_exposed)
      NALU_HIP_CHECK(hipMemcpy(faces_host, m->exposed_faces.p, sizeof(int32_t) * 4 * m->num_exposed,
                               hipMemcpyDeviceToHost));
    return NALU_OK;
  });
}
