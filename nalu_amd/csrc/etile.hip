// Element-visit tile assembly: the scatter strategy of the consolidated
// element loop (AssembleElemSolverAlgorithm::execute,
// src/AssembleElemSolverAlgorithm.C:75-100, whose element matrices reach the
// matrix through SolverAlgorithm::apply_coeff -> TpetraLinearSystem::sumInto,
// src/TpetraLinearSystem.C:1457-1522) laid out for the GPU.
//
// A block owns T consecutive node rows.  Its work list holds every element
// incident to one of those rows, once ("visits"), with the set of its nodes the
// block owns.  A thread evaluates one visit: it gathers the element once and
// computes only the sub-control surfaces (scs) adjacent to an owned node, each
// scs once: the contribution of scs ip to row il is the negative of its
// contribution to row ir (lhs(ir,:) = -lhs(il,:), rhs likewise), so an scs with
// both end nodes owned is added to both rows from one evaluation.  Rows
// accumulate in the block's LDS copy of its CSR segment and are written to HBM
// once, coalesced, with no global atomics and no separate zeroing pass.
//
// Wave uniformity: every visit is rotated (a proper rotation of the reference
// cube relabels the local nodes) so that its owned nodes come first in one of
// four canonical sets -- {0}, the edge {0,1}, the face {0,1,2,3}, or any set
// ("all 12 scs") -- and visits are sorted by that class inside a tile.  The
// code of a wave is chosen by the largest class among its lanes and each lane
// emits only to the rows it owns.  On a mesh whose row order follows lattice
// lines every visit is an edge or a node visit, so waves run the 5-scs (or
// 3-scs) code without divergence.  Rotations preserve orientation, so a
// canonical scs is a physical scs with the same area vector, traversed either
// left->right or reversed; element fields indexed by scs (mdot) are mapped
// through the rotation's scs table with that sign.
#include <hipcub/hipcub.hpp>

#include "hex8_device.hpp"
#include "internal.hpp"

namespace nalu {
namespace {

using hex8::node_sign;
using hex8::scs_left;
using hex8::scs_right;

// ---------------------------------------------------------------- cube symmetries
struct SymTables {
  uint32_t perm[24];   // nibble k: physical local node of canonical node k
  uint64_t ipmap[24];  // 5 bits per canonical scs: physical scs | reversed << 4
  uint8_t code[256];   // owned-node mask -> class << 5 | rotation
};

constexpr int node_of(int sx, int sy, int sz) {
  for (int n = 0; n < 8; ++n)
    if (node_sign(n, 0) == sx && node_sign(n, 1) == sy && node_sign(n, 2) == sz) return n;
  return -1;
}

constexpr int kClassA = 0, kClassB = 1, kClassC = 2, kClassD = 3;
// scs adjacent to the canonical owned set of each class: {0}; {0,1}; {0,1,2,3}; all
constexpr uint32_t kClassScs[4] = {0x109u, 0x30bu, 0xf0fu, 0xfffu};
constexpr uint32_t kClassNodes[4] = {0x01u, 0x03u, 0x0fu, 0xffu};

constexpr SymTables make_sym() {
  SymTables t{};
  const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  const int Ps[6] = {1, -1, -1, 1, 1, -1};
  int nr = 0;
  for (int p = 0; p < 6; ++p)
    for (int f = 0; f < 8; ++f) {
      const int fs[3] = {(f & 1) ? -1 : 1, (f & 2) ? -1 : 1, (f & 4) ? -1 : 1};
      if (Ps[p] * fs[0] * fs[1] * fs[2] != 1) continue;
      int g[8] = {};
      uint32_t pm = 0;
      for (int k = 0; k < 8; ++k) {
        const int s[3] = {node_sign(k, 0), node_sign(k, 1), node_sign(k, 2)};
        g[k] = node_of(fs[0] * s[P[p][0]], fs[1] * s[P[p][1]], fs[2] * s[P[p][2]]);
        pm |= (uint32_t)g[k] << (4 * k);
      }
      uint64_t im = 0;
      for (int ip = 0; ip < 12; ++ip) {
        const int gl = g[scs_left(ip)], gr = g[scs_right(ip)];
        for (int q = 0; q < 12; ++q) {
          if (scs_left(q) == gl && scs_right(q) == gr) im |= (uint64_t)q << (5 * ip);
          if (scs_left(q) == gr && scs_right(q) == gl) im |= (uint64_t)(q | 16) << (5 * ip);
        }
      }
      t.perm[nr] = pm;
      t.ipmap[nr] = im;
      ++nr;
    }
  // rotation 0 is the identity (first permutation, no sign flips)
  for (int m = 0; m < 256; ++m) {
    int pc = 0;
    for (int i = 0; i < 8; ++i) pc += (m >> i) & 1;
    int cls = kClassD, rot = 0;
    for (int r = 0; r < 24 && cls == kClassD; ++r) {
      uint32_t img[4] = {};
      for (int k = 0; k < 4; ++k) img[k] = 1u << ((t.perm[r] >> (4 * k)) & 15);
      if (pc == 1 && (uint32_t)m == img[0]) cls = kClassA, rot = r;
      else if (pc == 2 && (uint32_t)m == (img[0] | img[1])) cls = kClassB, rot = r;
      else if (pc == 4 && (uint32_t)m == (img[0] | img[1] | img[2] | img[3])) cls = kClassC, rot = r;
    }
    t.code[m] = (uint8_t)((cls << 5) | rot);
  }
  return t;
}

__constant__ const SymTables kSym = make_sym();

// compile-time checks of the tables
constexpr SymTables kSymHost = make_sym();
static_assert(kSymHost.perm[0] == 0x76543210u, "rotation 0 must be the identity");
static_assert((kSymHost.code[0x03] >> 5) == kClassB && (kSymHost.code[0xc0] >> 5) == kClassB, "edge masks");
static_assert((kSymHost.code[0x01] >> 5) == kClassA && (kSymHost.code[0x80] >> 5) == kClassA, "node masks");
static_assert((kSymHost.code[0xf0] >> 5) == kClassC && (kSymHost.code[0x99] >> 5) == kClassC, "face masks");
static_assert((kSymHost.code[0x05] >> 5) == kClassD && (kSymHost.code[0xff] >> 5) == kClassD, "other masks");

// ---------------------------------------------------------------- per-visit context
// Element gathered in canonical order plus the owned rows' LDS addressing.
struct Visit {
  int64_t e;
  uint32_t perm;
  int32_t nd[8];     // node indices, canonical order
  int32_t row[8];    // node-graph rows (valid where owned)
  bool own[8];
  uint64_t pk[8];    // elem_pos of the owned rows (valid where owned)
};

__device__ __forceinline__ int phys(uint32_t perm, int k) { return (int)((perm >> (4 * k)) & 7u); }

// position of canonical column node c in the row of canonical node k
__device__ __forceinline__ int col_pos(const Visit& v, int k, int c) {
  return (int)((v.pk[k] >> (8 * phys(v.perm, c))) & 0xffu);
}

template <int CLS>
__device__ __forceinline__ void load_visit(Visit& v, int64_t e, uint32_t perm, const int32_t* __restrict__ conn,
                                           const int32_t* __restrict__ lid, const uint64_t* __restrict__ elem_pos,
                                           int32_t a0, int32_t a1) {
  v.e = e;
  v.perm = perm;
#pragma unroll
  for (int k = 0; k < 8; ++k) v.nd[k] = conn[8 * e + phys(perm, k)];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v.own[k] = false;
    v.row[k] = 0;
    v.pk[k] = 0;
    if ((kClassNodes[CLS] >> k) & 1u) {
      v.row[k] = lid[v.nd[k]];
      v.own[k] = v.row[k] >= a0 && v.row[k] < a1;
      if (v.own[k]) v.pk[k] = elem_pos[8 * e + phys(perm, k)];
    }
  }
}

template <typename R>
__device__ __forceinline__ void gather_coords(const double* __restrict__ f, const int32_t (&nd)[8], R (&x)[8][3]) {
  if constexpr (std::is_same<R, double>::value) {
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[n][d] = f[3 * (int64_t)nd[n] + d];
  } else {
    // FP32 arithmetic on coordinates relative to one node, differences formed in FP64
    const double o[3] = {f[3 * (int64_t)nd[0]], f[3 * (int64_t)nd[0] + 1], f[3 * (int64_t)nd[0] + 2]};
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[n][d] = R(f[3 * (int64_t)nd[n] + d] - o[d]);
  }
}

// ---------------------------------------------------------------- scalar physics, one scs
// Each returns the rhs contribution to the scs's left row and calls emit(c, v)
// with the left row's entry for column c; the right row receives the negatives.

// ScalarDiffElemKernel (src/kernel/ScalarDiffElemKernel.C:66-111)
template <bool SHIFT, typename R>
struct ScalarDiffEt {
  using real = R;
  const double* q;
  const double* k;
  struct Elem {
    R x[8][3];
    R q[8], k[8];
  };
  __device__ __forceinline__ void load(const Visit& v, const double* __restrict__ coords, Elem& el) const {
    gather_coords(coords, v.nd, el.x);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      el.q[c] = R(q[v.nd[c]]);
      el.k[c] = R(k[v.nd[c]]);
    }
  }
  template <int IP, class Emit>
  __device__ __forceinline__ R scs(const Elem& el, Emit&& emit) const {
    hex8::IpGeo<double, R> G;
    hex8::ip_geo<IP, SHIFT>(el.x, G);
    R kip = R(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) kip += R(hex8::shape(IP, c, false)) * el.k[c];
    R rhs = R(0);
    hex8::static_for<8>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      const R v = -kip * hex8::dsdot<IP, c, SHIFT>(G.g);
      rhs -= v * el.q[c];
      emit(c, v);
    });
    return rhs;
  }
};

// ContinuityAdvElemKernel (src/kernel/ContinuityAdvElemKernel.C:88-148).  The
// nodal fields enter mdot only as rho u + tau Gp and p, so those are gathered.
template <bool SHIFT_MDOT, bool SHIFT_POISSON, bool REDUCED, typename R>
struct ContinuityEt {
  using real = R;
  const double* rho;
  const double* p;
  const double* u;
  const double* gp;
  double tau;
  struct Elem {
    R x[8][3];
    R w[8][3];  // rho u + tau Gp
    R p[8];
  };
  __device__ __forceinline__ void load(const Visit& v, const double* __restrict__ coords, Elem& el) const {
    gather_coords(coords, v.nd, el.x);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int64_t a = v.nd[c];
      const double r = rho[a];
#pragma unroll
      for (int j = 0; j < 3; ++j) el.w[c][j] = R(r * u[3 * a + j] + tau * gp[3 * a + j]);
      el.p[c] = R(p[a]);
    }
  }
  template <int IP, class Emit>
  __device__ __forceinline__ R scs(const Elem& el, Emit&& emit) const {
    constexpr bool LHS_SHIFT = SHIFT_POISSON || REDUCED;
    R A[3], g[3], gl[3];
    {
      double C[3][3], inv;
      hex8::scs_area<IP>(el.x, A);
      hex8::jacobian<IP, SHIFT_POISSON>(el.x, C, inv);
      hex8::contract_area(C, inv, A, g);
      if constexpr (LHS_SHIFT != SHIFT_POISSON) {
        hex8::jacobian<IP, LHS_SHIFT>(el.x, C, inv);
        hex8::contract_area(C, inv, A, gl);
      } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) gl[i] = g[i];
      }
    }
    R wi[3] = {R(0), R(0), R(0)}, dp = R(0);
    hex8::static_for<8>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      const R r = R(hex8::shape(IP, c, SHIFT_MDOT));
#pragma unroll
      for (int j = 0; j < 3; ++j) wi[j] += r * el.w[c][j];
      dp += el.p[c] * hex8::dsdot<IP, c, SHIFT_POISSON>(g);
    });
    const R tau_r = R(tau);
    const R mdot = wi[0] * A[0] + wi[1] * A[1] + wi[2] * A[2] - tau_r * dp;
    hex8::static_for<8>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      emit(c, -hex8::dsdot<IP, c, LHS_SHIFT>(gl));
    });
    return -(mdot / tau_r);
  }
};

// Wave-wide maximum of a small non-negative int.
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

template <int CLS, class Op>
__device__ __forceinline__ void scalar_visit(const Op& op, int64_t e, uint32_t perm, const int32_t* __restrict__ conn,
                                             const double* __restrict__ coords, const int32_t* __restrict__ lid,
                                             const int32_t* __restrict__ nrp, const uint64_t* __restrict__ elem_pos,
                                             int32_t a0, int32_t a1, int32_t s0, double* seg, double* rseg) {
  Visit v;
  load_visit<CLS>(v, e, perm, conn, lid, elem_pos, a0, a1);
  typename Op::Elem el;
  op.load(v, coords, el);
  int32_t base[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) base[k] = ((kClassNodes[CLS] >> k) & 1u) && v.own[k] ? nrp[v.row[k]] - s0 : 0;
  hex8::static_for<12>([&](auto ii) {
    constexpr int IP = decltype(ii)::value;
    if constexpr ((kClassScs[CLS] >> IP) & 1u) {
      constexpr int il = scs_left(IP), ir = scs_right(IP);
      constexpr bool Lc = (kClassNodes[CLS] >> il) & 1u, Rc = (kClassNodes[CLS] >> ir) & 1u;
      const bool lo = Lc && v.own[il], ro = Rc && v.own[ir];
      if (lo || ro) {
        const auto r = op.template scs<IP>(el, [&](int c, auto val) {
          const double d = (double)val;
          if (lo) unsafeAtomicAdd(seg + base[il] + col_pos(v, il, c), d);
          if (ro) unsafeAtomicAdd(seg + base[ir] + col_pos(v, ir, c), -d);
        });
        if (lo) unsafeAtomicAdd(rseg + (v.row[il] - a0), (double)r);
        if (ro) unsafeAtomicAdd(rseg + (v.row[ir] - a0), -(double)r);
      }
    }
  });
}

// Tile kernel for numDof-1 systems.  Block b handles tile xcd_block(b): the
// tiles of one XCD are contiguous, so neighbouring tiles share nodal data
// through that XCD's L2.
// zero n doubles of LDS, four independent stores per thread and pass
__device__ __forceinline__ void lds_zero(double* seg, int n) {
  const int B = blockDim.x;
  for (int i = threadIdx.x; i < n; i += 4 * B) {
    seg[i] = 0.0;
    if (i + B < n) seg[i + B] = 0.0;
    if (i + 2 * B < n) seg[i + 2 * B] = 0.0;
    if (i + 3 * B < n) seg[i + 3 * B] = 0.0;
  }
}

// write (or add) n LDS doubles to dst, four LDS reads in flight per thread
__device__ __forceinline__ void lds_flush(const double* seg, int n, double* __restrict__ dst, bool accumulate) {
  const int B = blockDim.x;
  for (int i = threadIdx.x; i < n; i += 4 * B) {
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = i + k * B < n ? seg[i + k * B] : 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i + k * B < n) {
        if (accumulate) dst[i + k * B] += v[k];
        else __builtin_nontemporal_store(v[k], dst + i + k * B);
      }
  }
}

// Tile kernel for numDof-1 systems over a list of tiles.  Block b handles
// tiles[xcd_block(b)]: the tiles of one XCD are contiguous, so neighbouring
// tiles share nodal data through that XCD's L2.  FAST: every visit of the
// listed tiles is a node or edge visit, so only those code paths are built.
template <class Op, bool FAST>
__global__ __launch_bounds__(256) void k_scalar_etile(Op op, const int32_t* __restrict__ conn,
                                                      const double* __restrict__ coords,
                                                      const int32_t* __restrict__ lid,
                                                      const int32_t* __restrict__ nrp,
                                                      const uint64_t* __restrict__ elem_pos,
                                                      const int32_t* __restrict__ vptr,
                                                      const int32_t* __restrict__ velem,
                                                      const uint8_t* __restrict__ vcode,
                                                      const int32_t* __restrict__ tiles, int64_t ntl, int T,
                                                      int64_t nrows, double* __restrict__ vals,
                                                      double* __restrict__ rhs, int accumulate) {
  extern __shared__ double seg[];
  const int64_t pos = xcd_block(blockIdx.x, ntl);
  if (pos < 0) return;
  const int64_t tile = tiles[pos];
  const int32_t a0 = (int32_t)(tile * T);
  const int32_t a1 = (int32_t)min((int64_t)a0 + T, nrows);
  const int32_t s0 = nrp[a0], s1 = nrp[a1];
  const int nseg = s1 - s0, nr = a1 - a0;
  double* rseg = seg + nseg;
  lds_zero(seg, nseg + nr);
  __syncthreads();
  const int32_t v0 = vptr[tile], v1 = vptr[tile + 1];
  for (int32_t vb = v0; vb < v1; vb += blockDim.x) {
    const int32_t vi = vb + threadIdx.x;
    const bool act = vi < v1;
    const int code = act ? vcode[vi] : 0;
    const int cls = wave_max(act ? code >> 5 : 0);
    if (act) {
      const int64_t e = velem[vi];
      const uint32_t perm = kSym.perm[code & 31];
      if constexpr (FAST) {
        if (cls == kClassA) scalar_visit<kClassA>(op, e, perm, conn, coords, lid, nrp, elem_pos, a0, a1, s0, seg, rseg);
        else scalar_visit<kClassB>(op, e, perm, conn, coords, lid, nrp, elem_pos, a0, a1, s0, seg, rseg);
      } else {
        if (cls == kClassC) scalar_visit<kClassC>(op, e, perm, conn, coords, lid, nrp, elem_pos, a0, a1, s0, seg, rseg);
        else scalar_visit<kClassD>(op, e, perm, conn, coords, lid, nrp, elem_pos, a0, a1, s0, seg, rseg);
      }
    }
  }
  __syncthreads();
  lds_flush(seg, nseg, vals + s0, accumulate != 0);
  lds_flush(rseg, nr, rhs + a0, accumulate != 0);
}

// ---------------------------------------------------------------- momentum, one scs
// MomentumAdvDiffElemKernel (src/kernel/MomentumAdvDiffElemKernel.C:72-186) for
// one scs, left-row view (S = +1):
//   blk_c[i][j] = delta_ij (N_adv(ip,c) mdot - mu_ip dN_c/dx . A) - mu_ip dN_c/dx_i A_j
//   rhs[i]      = -mdot u_ip,i - 2/3 mu_ip divU A_i incl
//                 + mu_ip sum_c [dN_c/dx_i (A . u_c) + (dN_c/dx . A) u_c,i]
// SEG drops the transpose-gradient term from the lhs (it stays in the rhs).
template <int IP, bool SKEW, bool SHIFT, bool SEG, typename R, class Emit>
__device__ __forceinline__ void momentum_scs(const R (&x)[8][3], const R (&u)[8][3], const R (&mu)[8], R mdot,
                                             R incl_divu, R (&rhs)[3], Emit&& emit) {
  R Cs[3][3], A[3];
  {
    double C[3][3], inv;
    hex8::scs_area<IP>(x, A);
    hex8::jacobian<IP, SHIFT>(x, C, inv);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Cs[i][j] = R(C[i][j] * inv);
  }
  R muIp = R(0), uIp[3] = {R(0), R(0), R(0)}, divU = R(0);
  hex8::static_for<8>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    muIp += R(hex8::shape(IP, c, false)) * mu[c];
#pragma unroll
    for (int j = 0; j < 3; ++j) uIp[j] += R(hex8::shape(IP, c, SKEW)) * u[c][j];
    R dn[3];
    hex8::dndx_scaled<IP, c, SHIFT>(Cs, dn);
    divU += u[c][0] * dn[0] + u[c][1] * dn[1] + u[c][2] * dn[2];
  });
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    rhs[i] = -(mdot * uIp[i] + R(2.0 / 3.0) * muIp * divU * A[i] * incl_divu);
    A[i] *= muIp;
  }
  hex8::static_for<8>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    R dn[3];
    hex8::dndx_scaled<IP, c, SHIFT>(Cs, dn);
    const R h = dn[0] * A[0] + dn[1] * A[1] + dn[2] * A[2];
    const R adv = R(hex8::shape(IP, c, SKEW)) * mdot;
    const R Au = u[c][0] * A[0] + u[c][1] * A[1] + u[c][2] * A[2];
    R blk[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) blk[i][j] = SEG ? R(0) : -(dn[i] * A[j]);
      blk[i][i] += adv - h;
      rhs[i] += dn[i] * Au + h * u[c][i];
    }
    emit(c, blk);
  });
}

template <int CLS, bool SKEW, bool SHIFT, typename R>
__device__ __forceinline__ void momentum_visit(int64_t e, int code, const int32_t* __restrict__ conn,
                                               const double* __restrict__ coords, const double* __restrict__ u,
                                               const double* __restrict__ mu, const double* __restrict__ mdot,
                                               double incl_divu, const int32_t* __restrict__ lid,
                                               const int32_t* __restrict__ nrp,
                                               const uint64_t* __restrict__ elem_pos, int32_t a0, int32_t a1,
                                               int32_t s0, double* seg, double* rseg) {
  Visit v;
  const uint32_t perm = kSym.perm[code & 31];
  const uint64_t ipm = kSym.ipmap[code & 31];
  load_visit<CLS>(v, e, perm, conn, lid, elem_pos, a0, a1);
  R x[8][3], uu[8][3], mm[8];
  gather_coords(coords, v.nd, x);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int64_t a = v.nd[c];
#pragma unroll
    for (int j = 0; j < 3; ++j) uu[c][j] = R(u[3 * a + j]);
    mm[c] = R(mu[a]);
  }
  int32_t base[8], L[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    base[k] = 0;
    L[k] = 0;
    if (((kClassNodes[CLS] >> k) & 1u) && v.own[k]) {
      const int32_t b = nrp[v.row[k]];
      base[k] = 9 * (b - s0);
      L[k] = nrp[v.row[k] + 1] - b;
    }
  }
  const double* md = mdot + 12 * e;
  hex8::static_for<12>([&](auto ii) {
    constexpr int IP = decltype(ii)::value;
    if constexpr ((kClassScs[CLS] >> IP) & 1u) {
      constexpr int il = scs_left(IP), ir = scs_right(IP);
      constexpr bool Lc = (kClassNodes[CLS] >> il) & 1u, Rc = (kClassNodes[CLS] >> ir) & 1u;
      const bool lo = Lc && v.own[il], ro = Rc && v.own[ir];
      if (lo || ro) {
        const int q = (int)((ipm >> (5 * IP)) & 31u);
        const double m = md[q & 15];
        const R mdc = R((q & 16) ? -m : m);
        R rr[3];
        momentum_scs<IP, SKEW, SHIFT, false>(x, uu, mm, mdc, R(incl_divu), rr, [&](int c, const R (&blk)[3][3]) {
          if (lo) {
            double* dst = seg + base[il] + 3 * col_pos(v, il, c);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) unsafeAtomicAdd(dst + 3 * i * L[il] + j, (double)blk[i][j]);
          }
          if (ro) {
            double* dst = seg + base[ir] + 3 * col_pos(v, ir, c);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) unsafeAtomicAdd(dst + 3 * i * L[ir] + j, -(double)blk[i][j]);
          }
        });
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (lo) unsafeAtomicAdd(rseg + 3 * (v.row[il] - a0) + i, (double)rr[i]);
          if (ro) unsafeAtomicAdd(rseg + 3 * (v.row[ir] - a0) + i, -(double)rr[i]);
        }
      }
    }
  });
}

// Momentum (numDof 3): the segment holds the 3 rows of each owned node
// (rows 3a..3a+2 are contiguous in the CSR, 3L entries each).
template <bool SKEW, bool SHIFT, typename R, bool FAST>
__global__ __launch_bounds__(128) void k_momentum_etile(const int32_t* __restrict__ conn,
                                                        const double* __restrict__ coords,
                                                        const double* __restrict__ u, const double* __restrict__ mu,
                                                        const double* __restrict__ mdot, double incl_divu,
                                                        const int32_t* __restrict__ lid,
                                                        const int32_t* __restrict__ nrp,
                                                        const uint64_t* __restrict__ elem_pos,
                                                        const int32_t* __restrict__ vptr,
                                                        const int32_t* __restrict__ velem,
                                                        const uint8_t* __restrict__ vcode,
                                                        const int32_t* __restrict__ tiles, int64_t ntl, int T,
                                                        int64_t nnodes, double* __restrict__ vals,
                                                        double* __restrict__ rhs, int accumulate) {
  extern __shared__ double seg[];
  const int64_t pos = xcd_block(blockIdx.x, ntl);
  if (pos < 0) return;
  const int64_t tile = tiles[pos];
  const int32_t a0 = (int32_t)(tile * T);
  const int32_t a1 = (int32_t)min((int64_t)a0 + T, nnodes);
  const int32_t s0 = nrp[a0];
  const int nseg = 9 * (nrp[a1] - s0), nr = 3 * (a1 - a0);
  double* rseg = seg + nseg;
  lds_zero(seg, nseg + nr);
  __syncthreads();
  const int32_t v0 = vptr[tile], v1 = vptr[tile + 1];
  for (int32_t vb = v0; vb < v1; vb += blockDim.x) {
    const int32_t vi = vb + threadIdx.x;
    const bool act = vi < v1;
    const int code = act ? vcode[vi] : 0;
    const int cls = wave_max(act ? code >> 5 : 0);
    if (act) {
      const int64_t e = velem[vi];
#define NALU_MV(K)                                                                                             \
  momentum_visit<K, SKEW, SHIFT, R>(e, code, conn, coords, u, mu, mdot, incl_divu, lid, nrp, elem_pos, a0, a1, s0, \
                                    seg, rseg)
      if constexpr (FAST) {
        if (cls == kClassA) NALU_MV(kClassA);
        else NALU_MV(kClassB);
      } else {
        if (cls == kClassC) NALU_MV(kClassC);
        else NALU_MV(kClassD);
      }
#undef NALU_MV
    }
  }
  __syncthreads();
  lds_flush(seg, nseg, vals + 9 * (int64_t)s0, accumulate != 0);
  lds_flush(rseg, nr, rhs + 3 * (int64_t)a0, accumulate != 0);
}

// ---------------------------------------------------------------- item kernels
// One thread per (visit, scs) for tiles of node and edge visits: wavefront w of
// a block takes scs w of the edge class's five ({0, 1, 3, 8, 9} in the
// canonical frame) for 64 visits at a time, so every wavefront runs one
// compile-time scs and a block keeps five times as many wavefronts in flight
// per tile as one thread per visit.  Node visits (class A) ride along: scs 1
// and 9 touch no node they own.
constexpr int kItemSlots = 5;
__host__ __device__ constexpr int item_ip(int slot) {
  return slot == 0 ? 0 : slot == 1 ? 1 : slot == 2 ? 3 : slot == 3 ? 8 : 9;
}

template <int IP, class Op>
__device__ __forceinline__ void scalar_item(const Op& op, int64_t e, uint32_t perm, const int32_t* __restrict__ conn,
                                            const double* __restrict__ coords, const int32_t* __restrict__ lid,
                                            const int32_t* __restrict__ nrp, const uint64_t* __restrict__ elem_pos,
                                            int32_t a0, int32_t a1, int32_t s0, double* seg, double* rseg) {
  constexpr int il = scs_left(IP), ir = scs_right(IP);
  constexpr bool Lc = (kClassNodes[kClassB] >> il) & 1u, Rc = (kClassNodes[kClassB] >> ir) & 1u;
  Visit v;
  load_visit<kClassB>(v, e, perm, conn, lid, elem_pos, a0, a1);
  const bool lo = Lc && v.own[il], ro = Rc && v.own[ir];
  if (!(lo || ro)) return;
  typename Op::Elem el;
  op.load(v, coords, el);
  const int32_t bl = lo ? nrp[v.row[il]] - s0 : 0, br = ro ? nrp[v.row[ir]] - s0 : 0;
  const auto r = op.template scs<IP>(el, [&](int c, auto val) {
    const double d = (double)val;
    if (lo) unsafeAtomicAdd(seg + bl + col_pos(v, il, c), d);
    if (ro) unsafeAtomicAdd(seg + br + col_pos(v, ir, c), -d);
  });
  if (lo) unsafeAtomicAdd(rseg + (v.row[il] - a0), (double)r);
  if (ro) unsafeAtomicAdd(rseg + (v.row[ir] - a0), -(double)r);
}

template <class Op>
__global__ __launch_bounds__(64 * kItemSlots) void k_scalar_items(
    Op op, const int32_t* __restrict__ conn, const double* __restrict__ coords, const int32_t* __restrict__ lid,
    const int32_t* __restrict__ nrp, const uint64_t* __restrict__ elem_pos, const int32_t* __restrict__ vptr,
    const int32_t* __restrict__ velem, const uint8_t* __restrict__ vcode, const int32_t* __restrict__ tiles,
    int64_t ntl, int T, int64_t nrows, double* __restrict__ vals, double* __restrict__ rhs, int accumulate) {
  extern __shared__ double seg[];
  const int64_t pos = xcd_block(blockIdx.x, ntl);
  if (pos < 0) return;
  const int64_t tile = tiles[pos];
  const int32_t a0 = (int32_t)(tile * T);
  const int32_t a1 = (int32_t)min((int64_t)a0 + T, nrows);
  const int32_t s0 = nrp[a0], s1 = nrp[a1];
  const int nseg = s1 - s0, nr = a1 - a0;
  double* rseg = seg + nseg;
  lds_zero(seg, nseg + nr);
  __syncthreads();
  const int slot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int32_t v0 = vptr[tile], v1 = vptr[tile + 1];
  for (int32_t g = v0; g < v1; g += 64) {
    const int32_t vi = g + lane;
    if (vi < v1) {
      const int64_t e = velem[vi];
      const uint32_t perm = kSym.perm[vcode[vi] & 31];
      switch (slot) {
#define NALU_SI(S)                                                                                            \
  case S:                                                                                                     \
    scalar_item<item_ip(S)>(op, e, perm, conn, coords, lid, nrp, elem_pos, a0, a1, s0, seg, rseg);            \
    break;
        NALU_SI(0) NALU_SI(1) NALU_SI(2) NALU_SI(3) NALU_SI(4)
#undef NALU_SI
      }
    }
  }
  __syncthreads();
  lds_flush(seg, nseg, vals + s0, accumulate != 0);
  lds_flush(rseg, nr, rhs + a0, accumulate != 0);
}

template <int IP, bool SKEW, bool SHIFT, typename R>
__device__ __forceinline__ void momentum_item(int64_t e, int code, const int32_t* __restrict__ conn,
                                              const double* __restrict__ coords, const double* __restrict__ u,
                                              const double* __restrict__ mu, const double* __restrict__ mdot,
                                              double incl_divu, const int32_t* __restrict__ lid,
                                              const int32_t* __restrict__ nrp,
                                              const uint64_t* __restrict__ elem_pos, int32_t a0, int32_t a1,
                                              int32_t s0, double* seg, double* rseg) {
  constexpr int il = scs_left(IP), ir = scs_right(IP);
  constexpr bool Lc = (kClassNodes[kClassB] >> il) & 1u, Rc = (kClassNodes[kClassB] >> ir) & 1u;
  Visit v;
  const uint32_t perm = kSym.perm[code & 31];
  load_visit<kClassB>(v, e, perm, conn, lid, elem_pos, a0, a1);
  const bool lo = Lc && v.own[il], ro = Rc && v.own[ir];
  if (!(lo || ro)) return;
  R x[8][3], uu[8][3], mm[8];
  gather_coords(coords, v.nd, x);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int64_t a = v.nd[c];
#pragma unroll
    for (int j = 0; j < 3; ++j) uu[c][j] = R(u[3 * a + j]);
    mm[c] = R(mu[a]);
  }
  int32_t bl = 0, br = 0, Ll = 0, Lr = 0;
  if (lo) {
    const int32_t b = nrp[v.row[il]];
    bl = 9 * (b - s0);
    Ll = nrp[v.row[il] + 1] - b;
  }
  if (ro) {
    const int32_t b = nrp[v.row[ir]];
    br = 9 * (b - s0);
    Lr = nrp[v.row[ir] + 1] - b;
  }
  const int q = (int)((kSym.ipmap[code & 31] >> (5 * IP)) & 31u);
  const double m = mdot[12 * e + (q & 15)];
  const R mdc = R((q & 16) ? -m : m);
  R rr[3];
  momentum_scs<IP, SKEW, SHIFT, false>(x, uu, mm, mdc, R(incl_divu), rr, [&](int c, const R (&blk)[3][3]) {
    if (lo) {
      double* dst = seg + bl + 3 * col_pos(v, il, c);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) unsafeAtomicAdd(dst + 3 * i * Ll + j, (double)blk[i][j]);
    }
    if (ro) {
      double* dst = seg + br + 3 * col_pos(v, ir, c);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) unsafeAtomicAdd(dst + 3 * i * Lr + j, -(double)blk[i][j]);
    }
  });
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (lo) unsafeAtomicAdd(rseg + 3 * (v.row[il] - a0) + i, (double)rr[i]);
    if (ro) unsafeAtomicAdd(rseg + 3 * (v.row[ir] - a0) + i, -(double)rr[i]);
  }
}

template <bool SKEW, bool SHIFT, typename R>
__global__ __launch_bounds__(64 * kItemSlots) void k_momentum_items(
    const int32_t* __restrict__ conn, const double* __restrict__ coords, const double* __restrict__ u,
    const double* __restrict__ mu, const double* __restrict__ mdot, double incl_divu, const int32_t* __restrict__ lid,
    const int32_t* __restrict__ nrp, const uint64_t* __restrict__ elem_pos, const int32_t* __restrict__ vptr,
    const int32_t* __restrict__ velem, const uint8_t* __restrict__ vcode, const int32_t* __restrict__ tiles,
    int64_t ntl, int T, int64_t nnodes, double* __restrict__ vals, double* __restrict__ rhs, int accumulate) {
  extern __shared__ double seg[];
  const int64_t pos = xcd_block(blockIdx.x, ntl);
  if (pos < 0) return;
  const int64_t tile = tiles[pos];
  const int32_t a0 = (int32_t)(tile * T);
  const int32_t a1 = (int32_t)min((int64_t)a0 + T, nnodes);
  const int32_t s0 = nrp[a0];
  const int nseg = 9 * (nrp[a1] - s0), nr = 3 * (a1 - a0);
  double* rseg = seg + nseg;
  lds_zero(seg, nseg + nr);
  __syncthreads();
  const int slot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int32_t v0 = vptr[tile], v1 = vptr[tile + 1];
  for (int32_t g = v0; g < v1; g += 64) {
    const int32_t vi = g + lane;
    if (vi < v1) {
      const int64_t e = velem[vi];
      const int code = vcode[vi];
      switch (slot) {
#define NALU_MI(S)                                                                                              \
  case S:                                                                                                       \
    momentum_item<item_ip(S), SKEW, SHIFT, R>(e, code, conn, coords, u, mu, mdot, incl_divu, lid, nrp, elem_pos, \
                                              a0, a1, s0, seg, rseg);                                           \
    break;
        NALU_MI(0) NALU_MI(1) NALU_MI(2) NALU_MI(3) NALU_MI(4)
#undef NALU_MI
      }
    }
  }
  __syncthreads();
  lds_flush(seg, nseg, vals + 9 * (int64_t)s0, accumulate != 0);
  lds_flush(rseg, nr, rhs + 3 * (int64_t)a0, accumulate != 0);
}

// ---------------------------------------------------------------- tile plan build
constexpr int kB = 256;
inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kB - 1) / kB); }

// one key per (element, local node): tile of the node's row | element | local
__global__ void k_visit_keys(const int32_t* __restrict__ conn, const int32_t* __restrict__ lid, int64_t n8, int T,
                             uint64_t* __restrict__ keys) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n8) return;
  const uint64_t tile = (uint64_t)(lid[conn[t]] / T);
  keys[t] = (tile << 35) | ((uint64_t)(t >> 3) << 3) | (uint64_t)(t & 7);
}

__global__ void k_visit_heads(const uint64_t* __restrict__ s, int64_t n, int32_t* __restrict__ head) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) head[q] = (q == 0 || (s[q] >> 3) != (s[q - 1] >> 3)) ? 1 : 0;
}

// at each (tile, element) run head: owned mask -> class / rotation -> final sort key
__global__ void k_visit_emit(const uint64_t* __restrict__ s, int64_t n, const int32_t* __restrict__ head,
                             const int32_t* __restrict__ idx, uint64_t* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n || !head[q]) return;
  uint32_t mask = 0;
  for (int64_t r = q; r < n && (s[r] >> 3) == (s[q] >> 3); ++r) mask |= 1u << (uint32_t)(s[r] & 7);
  const uint64_t tile = s[q] >> 35;
  const uint64_t e = (s[q] >> 3) & 0xffffffffull;
  const uint64_t code = kSym.code[mask];
  out[idx[q]] = (tile << 40) | (code << 32) | e;
}

__global__ void k_visit_split(const uint64_t* __restrict__ s, int64_t n, int64_t ntiles, int32_t* __restrict__ vptr,
                              int32_t* __restrict__ velem, uint8_t* __restrict__ vcode) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  velem[q] = (int32_t)(uint32_t)(s[q] & 0xffffffffull);
  vcode[q] = (uint8_t)((s[q] >> 32) & 0xff);
  const int64_t t = (int64_t)(s[q] >> 40);
  const int64_t tp = q == 0 ? -1 : (int64_t)(s[q - 1] >> 40);
  for (int64_t k = tp + 1; k <= t; ++k) vptr[k] = (int32_t)q;  // tiles (tp, t] start here
  if (q == n - 1)
    for (int64_t k = t + 1; k <= ntiles; ++k) vptr[k] = (int32_t)n;
}

__global__ void k_visit_max(const int32_t* __restrict__ vptr, int64_t ntiles, int* __restrict__ mx) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) atomicMax(mx, vptr[t + 1] - vptr[t]);
}

// largest visit class of each tile (visits are sorted by class inside a tile)
__global__ void k_tile_class(const int32_t* __restrict__ vptr, const uint8_t* __restrict__ vcode, int64_t ntiles,
                             uint8_t* __restrict__ cls) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) cls[t] = vptr[t + 1] > vptr[t] ? (uint8_t)(vcode[vptr[t + 1] - 1] >> 5) : (uint8_t)0;
}

int bitsof(uint64_t v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

struct TilePlan {
  int T = 0;
  int64_t ntiles = 0, nvisits = 0;
  int max_visits = 0;
  DevBuf<int32_t> vptr, velem;
  DevBuf<uint8_t> vcode;
  // tiles whose visits are all node or edge visits (the compact kernel) and the rest
  DevBuf<int32_t> fast_tiles, other_tiles;
  int64_t nfast = 0, nother = 0;
};

struct Plans {
  std::vector<TilePlan*> list;
  ~Plans() {
    for (auto* p : list) delete p;
  }
};

void build_plan(nalu_linsys* ls, TilePlan& P, int T) {
  const nalu_mesh* m = ls->mesh;
  hipStream_t s = ls->stream;
  const int64_t N = ls->num_nodes, E = m->num_elems, n8 = 8 * E;
  NALU_REQUIRE(E < ((int64_t)1 << 31) / 8, NALU_ERR_UNSUPPORTED, "tile plan: mesh too large");
  P.T = T;
  P.ntiles = (N + T - 1) / T;
  NALU_REQUIRE(P.ntiles < ((int64_t)1 << 24), NALU_ERR_UNSUPPORTED, "tile plan: too many tiles");
  DevBuf<uint64_t> keys(n8), sorted(n8);
  k_visit_keys<<<grid_of(n8), kB, 0, s>>>(m->conn.p, ls->node_lid.p, n8, T, keys.p);
  const int tb = bitsof((uint64_t)P.ntiles);
  {
    size_t tmp = 0;
    NALU_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, keys.p, sorted.p, (int)n8, 0, 35 + tb, s));
    DevBuf<char> t(tmp ? tmp : 1);
    NALU_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(t.p, tmp, keys.p, sorted.p, (int)n8, 0, 35 + tb, s));
  }
  DevBuf<int32_t> head(n8), idx(n8);
  k_visit_heads<<<grid_of(n8), kB, 0, s>>>(sorted.p, n8, head.p);
  {
    size_t tmp = 0;
    NALU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, head.p, idx.p, (int)n8, s));
    DevBuf<char> t(tmp ? tmp : 1);
    NALU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, head.p, idx.p, (int)n8, s));
  }
  int32_t last_idx = 0, last_head = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&last_idx, idx.p + n8 - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipMemcpyAsync(&last_head, head.p + n8 - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  const int64_t U = (int64_t)last_idx + last_head;
  k_visit_emit<<<grid_of(n8), kB, 0, s>>>(sorted.p, n8, head.p, idx.p, keys.p);
  {
    size_t tmp = 0;
    NALU_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, keys.p, sorted.p, (int)U, 0, 40 + tb, s));
    DevBuf<char> t(tmp ? tmp : 1);
    NALU_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(t.p, tmp, keys.p, sorted.p, (int)U, 0, 40 + tb, s));
  }
  P.nvisits = U;
  P.vptr.alloc(P.ntiles + 1);
  if (U == 0) NALU_HIP_CHECK(hipMemsetAsync(P.vptr.p, 0, sizeof(int32_t) * (P.ntiles + 1), s));
  P.velem.alloc(U);
  P.vcode.alloc(U);
  k_visit_split<<<grid_of(U), kB, 0, s>>>(sorted.p, U, P.ntiles, P.vptr.p, P.velem.p, P.vcode.p);
  DevBuf<int> mx(1);
  NALU_HIP_CHECK(hipMemsetAsync(mx.p, 0, sizeof(int), s));
  k_visit_max<<<grid_of(P.ntiles), kB, 0, s>>>(P.vptr.p, P.ntiles, mx.p);
  NALU_HIP_CHECK(hipMemcpyAsync(&P.max_visits, mx.p, sizeof(int), hipMemcpyDeviceToHost, s));
  DevBuf<uint8_t> tcls(P.ntiles);
  k_tile_class<<<grid_of(P.ntiles), kB, 0, s>>>(P.vptr.p, P.vcode.p, P.ntiles, tcls.p);
  std::vector<uint8_t> hc(P.ntiles);
  NALU_HIP_CHECK(hipMemcpyAsync(hc.data(), tcls.p, P.ntiles, hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  std::vector<int32_t> fast, other;
  for (int64_t t = 0; t < P.ntiles; ++t) (hc[t] <= kClassB ? fast : other).push_back((int32_t)t);
  P.nfast = (int64_t)fast.size();
  P.nother = (int64_t)other.size();
  P.fast_tiles.alloc(std::max<int64_t>(1, P.nfast));
  P.other_tiles.alloc(std::max<int64_t>(1, P.nother));
  if (P.nfast)
    NALU_HIP_CHECK(hipMemcpy(P.fast_tiles.p, fast.data(), sizeof(int32_t) * P.nfast, hipMemcpyHostToDevice));
  if (P.nother)
    NALU_HIP_CHECK(hipMemcpy(P.other_tiles.p, other.data(), sizeof(int32_t) * P.nother, hipMemcpyHostToDevice));
  NALU_HIP_CHECK(hipGetLastError());
}

Plans& plans_of(nalu_linsys* ls) {
  if (!ls->etile_plans) ls->etile_plans = new Plans();
  return *static_cast<Plans*>(ls->etile_plans);
}

// plan cache on the linear system (one per tile size; the graph is fixed once finalized)
const TilePlan& tile_plan(nalu_linsys* ls, int T) {
  Plans& ps = plans_of(ls);
  for (auto* p : ps.list)
    if (p->T == T) return *p;
  auto* p = new TilePlan();
  try {
    build_plan(ls, *p, T);
  } catch (...) {
    delete p;
    throw;
  }
  ps.list.push_back(p);
  return *p;
}

}  // namespace

void etile_release(nalu_linsys* ls) {
  delete static_cast<Plans*>(ls->etile_plans);
  ls->etile_plans = nullptr;
}

// Tile sizes.  With rows following lattice lines a tile of T rows has at most
// 4T + 8 visits: scalar systems take 62 rows (27-entry rows, 14 KB of LDS) for
// one pass of 256 threads; the momentum system 29 nodes (3 rows of 81 entries
// each, 57 KB of LDS, two blocks per CU) for one pass of 128 threads.
constexpr int kScalarT = 62;
constexpr int kMomentumT = 29;

template <class Op>
void etile_scalar(nalu_linsys* ls, const nalu_mesh* m, const Op& op) {
  const TilePlan& P = tile_plan(ls, kScalarT);
  NALU_REQUIRE(ls->values_pending_zero == ls->rhs_pending_zero, NALU_ERR_STATE,
               "tile assembly: values/rhs zero state diverged");
  const int accumulate = ls->values_pending_zero ? 0 : 1;
  const size_t lds = ((size_t)P.T * ls->max_node_row_len + P.T) * sizeof(double);
  auto go = [&](auto kern, const int32_t* tiles, int64_t ntl) {
    if (ntl == 0) return;
    set_lds_limit((const void*)kern, lds);
    hipLaunchKernelGGL(kern, dim3((unsigned)xcd_grid(ntl)), dim3(256), lds, ls->stream, op, m->conn.p, m->coords.p,
                       ls->node_lid.p, ls->nrp.p, ls->elem_pos.p, P.vptr.p, P.velem.p, P.vcode.p, tiles, ntl, P.T,
                       ls->num_nodes, ls->values.p, ls->rhs.p, accumulate);
  };
  if (P.nfast) {
    set_lds_limit((const void*)k_scalar_items<Op>, lds);
    hipLaunchKernelGGL(k_scalar_items<Op>, dim3((unsigned)xcd_grid(P.nfast)), dim3(64 * kItemSlots), lds, ls->stream,
                       op, m->conn.p, m->coords.p, ls->node_lid.p, ls->nrp.p, ls->elem_pos.p, P.vptr.p, P.velem.p,
                       P.vcode.p, P.fast_tiles.p, P.nfast, P.T, ls->num_nodes, ls->values.p, ls->rhs.p, accumulate);
  }
  go(k_scalar_etile<Op, false>, P.other_tiles.p, P.nother);
  NALU_HIP_CHECK(hipGetLastError());
  ls->values_pending_zero = ls->rhs_pending_zero = false;
}

void etile_scalar_diff(nalu_linsys* ls, const nalu_mesh* m, const double* q, const double* k, bool shifted,
                       bool fp32) {
  if (fp32) {
    if (shifted) etile_scalar(ls, m, ScalarDiffEt<true, float>{q, k});
    else etile_scalar(ls, m, ScalarDiffEt<false, float>{q, k});
  } else {
    if (shifted) etile_scalar(ls, m, ScalarDiffEt<true, double>{q, k});
    else etile_scalar(ls, m, ScalarDiffEt<false, double>{q, k});
  }
}

void etile_continuity(nalu_linsys* ls, const nalu_mesh* m, const double* rho, const double* p, const double* u,
                      const double* gp, double tau, int key, bool fp32) {
  switch (key) {
#define NALU_EC(K, A, B, C)                                                                        \
  case K:                                                                                          \
    if (fp32) etile_scalar(ls, m, ContinuityEt<A, B, C, float>{rho, p, u, gp, tau});               \
    else etile_scalar(ls, m, ContinuityEt<A, B, C, double>{rho, p, u, gp, tau});                   \
    break;
    NALU_EC(0, false, false, false)
    NALU_EC(1, false, false, true)
    NALU_EC(2, false, true, false)
    NALU_EC(3, false, true, true)
    NALU_EC(4, true, false, false)
    NALU_EC(5, true, false, true)
    NALU_EC(6, true, true, false)
    NALU_EC(7, true, true, true)
#undef NALU_EC
  }
}

void etile_momentum(nalu_linsys* ls, const nalu_mesh* m, const double* u, const double* mu, const double* mdot,
                    double include_divu, bool skew, bool shifted, bool fp32) {
  const TilePlan& P = tile_plan(ls, kMomentumT);
  NALU_REQUIRE(ls->values_pending_zero == ls->rhs_pending_zero, NALU_ERR_STATE,
               "tile assembly: values/rhs zero state diverged");
  const int accumulate = ls->values_pending_zero ? 0 : 1;
  const size_t lds = ((size_t)P.T * ls->max_node_row_len * 9 + 3 * P.T) * sizeof(double);
  auto go1 = [&](auto kern, const int32_t* tiles, int64_t ntl) {
    if (ntl == 0) return;
    set_lds_limit((const void*)kern, lds);
    hipLaunchKernelGGL(kern, dim3((unsigned)xcd_grid(ntl)), dim3(128), lds, ls->stream, m->conn.p, m->coords.p, u,
                       mu, mdot, include_divu, ls->node_lid.p, ls->nrp.p, ls->elem_pos.p, P.vptr.p, P.velem.p,
                       P.vcode.p, tiles, ntl, P.T, ls->num_nodes, ls->values.p, ls->rhs.p, accumulate);
  };
  auto go = [&](auto fast, auto other) {
    if (P.nfast) {
      set_lds_limit((const void*)fast, lds);
      hipLaunchKernelGGL(fast, dim3((unsigned)xcd_grid(P.nfast)), dim3(64 * kItemSlots), lds, ls->stream, m->conn.p,
                         m->coords.p, u, mu, mdot, include_divu, ls->node_lid.p, ls->nrp.p, ls->elem_pos.p, P.vptr.p,
                         P.velem.p, P.vcode.p, P.fast_tiles.p, P.nfast, P.T, ls->num_nodes, ls->values.p, ls->rhs.p,
                         accumulate);
    }
    go1(other, P.other_tiles.p, P.nother);
  };
#define NALU_MG(SK, SH, R) go(k_momentum_items<SK, SH, R>, k_momentum_etile<SK, SH, R, false>)
  if (skew && shifted) fp32 ? NALU_MG(true, true, float) : NALU_MG(true, true, double);
  else if (skew) fp32 ? NALU_MG(true, false, float) : NALU_MG(true, false, double);
  else if (shifted) fp32 ? NALU_MG(false, true, float) : NALU_MG(false, true, double);
  else fp32 ? NALU_MG(false, false, float) : NALU_MG(false, false, double);
#undef NALU_MG
  NALU_HIP_CHECK(hipGetLastError());
  ls->values_pending_zero = ls->rhs_pending_zero = false;
}

}  // namespace nalu
