// Device-side Hex8 CVFEM sub-control-surface geometry and element kernels,
// evaluated per (element, local row node).  Everything that depends on the
// integration point is a compile-time constant: the row node LA selects its
// three adjacent sub-control surfaces, and for each of them the isoparametric
// location, the shape-function derivative table and the four sub-division
// points of the surface are folded by the compiler.
//
// Geometry restated from the mathematics of the reference:
//   * scs area vector (HexSCS::determinant, src/master_element/Hex8CVFEM.C:547-581,
//     quad_area_by_triangulation, include/master_element/Hex8GeometryFunctions.h:30-76):
//     the fan of four triangles around the quad centroid sums to the shoelace
//     form 1/2 (p2 - p0) x (p3 - p1), evaluated here directly;
//   * grad_op (generic_grad_op, include/master_element/MasterElementFunctions.h:51-108):
//     J_ij = sum_n dN_n/ds_j x_n,i, dN/dx = cof(J) dN/ds / det J.  Contractions
//     with the area vector use dN/dx . A = dN/ds . (cof(J)^T A) / det J.
// Floating-point results agree with the reference order to rounding.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace nalu {
namespace hex8 {

// corner signs of the Hex8 reference element (Hex8CVFEM.C nodeLoc_ :420-436)
__host__ __device__ constexpr int node_sign(int n, int d) {
  return d == 0 ? ((n == 1 || n == 2 || n == 5 || n == 6) ? 1 : -1)
       : d == 1 ? ((n == 2 || n == 3 || n == 6 || n == 7) ? 1 : -1)
                : (n >= 4 ? 1 : -1);
}

// scs left/right nodes (HexSCS::lrscv_, Hex8CVFEM.C:264-276)
__host__ __device__ constexpr int scs_left(int ip) {
  return ip < 4 ? (ip == 3 ? 0 : ip) : ip < 8 ? (ip == 7 ? 4 : ip) : ip - 8;
}
__host__ __device__ constexpr int scs_right(int ip) {
  return ip < 4 ? (ip == 3 ? 3 : ip + 1) : ip < 8 ? (ip == 7 ? 7 : ip + 1) : ip - 4;
}

// scs integration point location (HexSCS::intgLoc_, :318-330): midpoint of the
// edge (il, ir) pulled halfway to the element centroid in the other two
// directions; shifted (:333-345): on the edge itself.
__host__ __device__ constexpr double scs_loc(int ip, int d, bool shifted) {
  // edge direction of ip: 0 for ip 0,2,4,6; 1 for 1,3,5,7; 2 for 8..11
  return (ip >= 8 ? 2 : (ip & 1)) == d
             ? 0.0
             : (shifted ? 0.5 : 0.25) *
                   (double)(node_sign(scs_left(ip), d) + node_sign(scs_right(ip), d)) * 0.5;
}

// dN_n/ds_d at the scs ip = sign_d * prod_{e != d} (1/2 + sign_e s_e)
__host__ __device__ constexpr double deriv(int ip, int n, int d, bool shifted) {
  double v = (double)node_sign(n, d);
  for (int e = 0; e < 3; ++e)
    if (e != d) v *= 0.5 + (double)node_sign(n, e) * scs_loc(ip, e, shifted);
  return v;
}

// N_n at the scs ip = prod_e (1/2 + sign_e s_e)   (HexSCS::hex8_shape_fcn, :487-518)
__host__ __device__ constexpr double shape(int ip, int n, bool shifted) {
  double v = 1.0;
  for (int e = 0; e < 3; ++e) v *= 0.5 + (double)node_sign(n, e) * scs_loc(ip, e, shifted);
  return v;
}

// The three scs adjacent to local node LA and the sign of LA on each
// (+1 when LA is the scs's left node).
__host__ __device__ constexpr int adj_ip(int la, int k) {
  // scs of edges incident to node la, ordered as in lrscv
  int found = 0;
  for (int ip = 0; ip < 12; ++ip)
    if (scs_left(ip) == la || scs_right(ip) == la) {
      if (found == k) return ip;
      ++found;
    }
  return -1;
}
__host__ __device__ constexpr int adj_sign(int la, int k) { return scs_left(adj_ip(la, k)) == la ? 1 : -1; }

// Sub-division points (subdivide_hex_8, Hex8GeometryFunctions.h:263-321) as
// corner masks: bit n set when corner n is averaged into the point.
__host__ __device__ constexpr uint32_t subdiv_mask(int p) {
  switch (p) {
    case 8: return 0x03;   // edge 0-1
    case 9: return 0x06;   // edge 1-2
    case 10: return 0x0c;  // edge 2-3
    case 11: return 0x09;  // edge 3-0
    case 12: return 0x0f;  // face 0123
    case 13: return 0x30;  // edge 4-5
    case 14: return 0x60;  // edge 5-6
    case 15: return 0xc0;  // edge 6-7
    case 16: return 0x90;  // edge 7-4
    case 17: return 0xf0;  // face 4567
    case 18: return 0x22;  // edge 1-5
    case 19: return 0x11;  // edge 0-4
    case 20: return 0x33;  // face 0145
    case 21: return 0x88;  // edge 3-7
    case 22: return 0x44;  // edge 2-6
    case 23: return 0xcc;  // face 2367
    case 24: return 0x66;  // face 1256
    case 25: return 0x99;  // face 0347
    case 26: return 0xff;  // centroid
    default: return 1u << p;
  }
}
__host__ __device__ constexpr int popc8(uint32_t m) {
  int c = 0;
  for (int i = 0; i < 8; ++i) c += (m >> i) & 1;
  return c;
}
// facet table of the 12 scs quads (Hex8CVFEM.C:551-564)
__host__ __device__ constexpr int facet(int ip, int k) {
  constexpr int t[12][4] = {{20, 8, 12, 26},  {24, 9, 12, 26},  {10, 12, 26, 23}, {11, 25, 26, 12},
                            {13, 20, 26, 17}, {17, 14, 24, 26}, {17, 15, 23, 26}, {16, 17, 26, 25},
                            {19, 20, 26, 25}, {20, 18, 24, 26}, {22, 23, 26, 24}, {21, 25, 26, 23}};
  return t[ip][k];
}
// weight of corner n in (point q - point p)
__host__ __device__ constexpr double diff_weight(int q, int p, int n) {
  const uint32_t mq = subdiv_mask(q), mp = subdiv_mask(p);
  return (((mq >> n) & 1) ? 1.0 / popc8(mq) : 0.0) - (((mp >> n) & 1) ? 1.0 / popc8(mp) : 0.0);
}

// Shape-function and derivative tables at the scs ips, for loops whose column
// index is a runtime value (kept in the constant cache; the ip stays compile-time).
struct ScsTables {
  double deriv[2][12][8][3];  // [shifted][ip][node][dir]
  double shape[2][12][8];     // [shifted][ip][node]
};
__host__ __device__ constexpr ScsTables make_scs_tables() {
  ScsTables t{};
  for (int sh = 0; sh < 2; ++sh)
    for (int ip = 0; ip < 12; ++ip)
      for (int n = 0; n < 8; ++n) {
        t.shape[sh][ip][n] = shape(ip, n, sh != 0);
        for (int d = 0; d < 3; ++d) t.deriv[sh][ip][n][d] = deriv(ip, n, d, sh != 0);
      }
  return t;
}
__constant__ const ScsTables kScsTab = make_scs_tables();

// area vector of scs IP: 1/2 (p2 - p0) x (p3 - p1)
template <int IP, typename R>
__device__ __forceinline__ void scs_area(const R (&x)[8][3], R (&A)[3]) {
  R d1[3], d2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R a = R(0), b = R(0);
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const R w1 = R(diff_weight(facet(IP, 2), facet(IP, 0), n));
      const R w2 = R(diff_weight(facet(IP, 3), facet(IP, 1), n));
      a += w1 * x[n][i];
      b += w2 * x[n][i];
    }
    d1[i] = a;
    d2[i] = b;
  }
  A[0] = R(0.5) * (d1[1] * d2[2] - d1[2] * d2[1]);
  A[1] = R(0.5) * (d1[2] * d2[0] - d1[0] * d2[2]);
  A[2] = R(0.5) * (d1[0] * d2[1] - d1[1] * d2[0]);
}

// Jacobian J_ij = sum_n dN_n/ds_j x_n,i, its cofactor matrix and determinant.
// JR is the arithmetic type of the Jacobian (FP64 keeps det J exact enough on
// highly stretched elements while the rest of the kernel runs in R).
template <int IP, bool SHIFT, typename JR, typename R>
__device__ __forceinline__ void jacobian(const R (&x)[8][3], JR (&C)[3][3], JR& inv_det) {
  JR J[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      JR s = JR(0);
#pragma unroll
      for (int n = 0; n < 8; ++n) s += JR(deriv(IP, n, j, SHIFT)) * JR(x[n][i]);
      J[i][j] = s;
    }
  C[0][0] = J[1][1] * J[2][2] - J[2][1] * J[1][2];
  C[0][1] = J[1][2] * J[2][0] - J[2][2] * J[1][0];
  C[0][2] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
  C[1][0] = J[0][2] * J[2][1] - J[2][2] * J[0][1];
  C[1][1] = J[0][0] * J[2][2] - J[2][0] * J[0][2];
  C[1][2] = J[0][1] * J[2][0] - J[2][1] * J[0][0];
  C[2][0] = J[0][1] * J[1][2] - J[1][1] * J[0][2];
  C[2][1] = J[0][2] * J[1][0] - J[1][2] * J[0][0];
  C[2][2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
  const JR det = J[0][0] * C[0][0] + J[1][0] * C[1][0] + J[2][0] * C[2][0];
  inv_det = JR(1) / det;
}

// g_k = sum_i C_ik A_i / det, so that dN_c/dx . A = dN_c/ds . g
template <typename JR, typename R>
__device__ __forceinline__ void contract_area(const JR (&C)[3][3], JR inv_det, const R (&A)[3], R (&g)[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k)
    g[k] = R((C[0][k] * JR(A[0]) + C[1][k] * JR(A[1]) + C[2][k] * JR(A[2])) * inv_det);
}

// ---------------------------------------------------------------- kernels

// Per-ip geometry kept while a row streams over its columns.
template <typename JR, typename R>
struct IpGeo {
  R A[3];         // area vector
  JR C[3][3];     // cofactor matrix of J
  JR inv;         // 1 / det J
  R g[3];         // C^T A / det J: dN_c/dx . A = dN_c/ds . g
};

template <int IP, bool SHIFT, typename JR, typename R>
__device__ __forceinline__ void ip_geo(const R (&x)[8][3], IpGeo<JR, R>& G) {
  scs_area<IP>(x, G.A);
  jacobian<IP, SHIFT>(x, G.C, G.inv);
  contract_area(G.C, G.inv, G.A, G.g);
}

// dN_c/ds . g
template <int IP, int C, bool SHIFT, typename R>
__device__ __forceinline__ R dsdot(const R (&g)[3]) {
  return R(deriv(IP, C, 0, SHIFT)) * g[0] + R(deriv(IP, C, 1, SHIFT)) * g[1] + R(deriv(IP, C, 2, SHIFT)) * g[2];
}

// full dN_c/dx
template <int IP, int C, bool SHIFT, typename JR, typename R>
__device__ __forceinline__ void dndx(const IpGeo<JR, R>& G, R (&dn)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
    dn[i] = R((G.C[i][0] * JR(deriv(IP, C, 0, SHIFT)) + G.C[i][1] * JR(deriv(IP, C, 1, SHIFT)) +
               G.C[i][2] * JR(deriv(IP, C, 2, SHIFT))) * G.inv);
}

// compile-time loop helpers
template <int N, int I = 0, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<N, I + 1>(f);
  }
}

// Nodal field access for one element: node indices + base pointers.
struct Gather {
  const int32_t (&nd)[8];
  __device__ __forceinline__ double s(const double* __restrict__ f, int c) const { return f[nd[c]]; }
  __device__ __forceinline__ double v(const double* __restrict__ f, int c, int d) const {
    return f[3 * (int64_t)nd[c] + d];
  }
};

// ScalarDiffElemKernel, row LA (src/kernel/ScalarDiffElemKernel.C:66-111).
// The row's entry for column c is v_c = sum_{ip adj LA} S (-k_ip dN_c/dx . A);
// the rhs is -sum_c v_c q_c.  emit(c, v_c) is called once per column.
template <int LA, bool SHIFT, typename JR, typename R, class Emit>
__device__ __forceinline__ R row_scalar_diff(const R (&x)[8][3], const Gather& gt, const double* __restrict__ q,
                                             const double* __restrict__ k, Emit&& emit) {
  IpGeo<JR, R> G[3];
  R kip[3];
  static_for<3>([&](auto kk) {
    constexpr int IP = adj_ip(LA, decltype(kk)::value);
    ip_geo<IP, SHIFT>(x, G[kk]);
    kip[kk] = R(0);
  });
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const R kc = R(gt.s(k, c));
    static_for<3>([&](auto kk) { kip[kk] += R(shape(adj_ip(LA, decltype(kk)::value), c, false)) * kc; });
  }
  R rhs = R(0);
  static_for<8>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    R v = R(0);
    static_for<3>([&](auto kk) {
      constexpr int IP = adj_ip(LA, decltype(kk)::value);
      constexpr R S = R(adj_sign(LA, decltype(kk)::value));
      v += S * (-kip[kk] * dsdot<IP, c, SHIFT>(G[kk].g));
    });
    rhs -= v * R(gt.s(q, c));
    emit(c, v);
  });
  return rhs;
}

// ComputeMdotElemAlgorithm for one scs (src/ComputeMdotElemAlgorithm.C:204-232):
// mdot = (rho u)_ip . A - tau (grad p_ip - Gp_ip) . A, from an element's nodal
// fields held in registers.
template <int IP, bool SHIFT_MDOT, bool SHIFT_POISSON, typename JR, typename R>
__device__ __forceinline__ R ip_mdot_regs(const IpGeo<JR, R>& G, const double (&rho)[8], const double (&p)[8],
                                          const double (&u)[8][3], const double (&gp)[8][3], R tau) {
  R ru[3] = {R(0), R(0), R(0)}, gpi[3] = {R(0), R(0), R(0)}, dpa = R(0);
  static_for<8>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    const R r = R(shape(IP, c, SHIFT_MDOT));
    const R rr = r * R(rho[c]);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      ru[j] += rr * R(u[c][j]);
      gpi[j] += r * R(gp[c][j]);
    }
    dpa += R(p[c]) * dsdot<IP, c, SHIFT_POISSON>(G.g);
  });
  const R rua = ru[0] * G.A[0] + ru[1] * G.A[1] + ru[2] * G.A[2];
  const R gpa = gpi[0] * G.A[0] + gpi[1] * G.A[1] + gpi[2] * G.A[2];
  return rua - tau * (dpa - gpa);
}

// no-op receiver of the per-ip mdot values of row_continuity
struct NoIpOut {
  template <typename R>
  __device__ __forceinline__ void operator()(int, R) const {}
};

// ContinuityAdvElemKernel, row LA (src/kernel/ContinuityAdvElemKernel.C:88-148):
// lhs(la,c) = sum_ip S (-dN_lhs_c/dx . A); rhs(la) = -sum_ip S mdot_ip / tau,
// with mdot_ip as in ComputeMdotElemAlgorithm.  Nodal fields are read once per
// node and folded into the three ip accumulators.  ipout(ip, mdot_ip) receives
// the mdot of each adjacent scs whose left node is LA (each scs of an element
// has exactly one left node, so a pass over the element's rows delivers all 12).
template <int LA, bool SHIFT_MDOT, bool SHIFT_POISSON, bool REDUCED, typename JR, typename R, class Emit,
          class IpOut = NoIpOut>
__device__ __forceinline__ R row_continuity(const R (&x)[8][3], const Gather& gt, const double* __restrict__ rho,
                                            const double* __restrict__ p, const double* __restrict__ u,
                                            const double* __restrict__ gp, R tau, Emit&& emit,
                                            IpOut&& ipout = IpOut{}) {
  constexpr bool LHS_SHIFT = SHIFT_POISSON || REDUCED;
  R A[3][3], g[3][3], gl[3][3];
  static_for<3>([&](auto kk) {
    constexpr int IP = adj_ip(LA, decltype(kk)::value);
    JR C[3][3], inv;
    scs_area<IP>(x, A[kk]);
    jacobian<IP, SHIFT_POISSON>(x, C, inv);
    contract_area(C, inv, A[kk], g[kk]);
    if constexpr (LHS_SHIFT != SHIFT_POISSON) {
      jacobian<IP, LHS_SHIFT>(x, C, inv);
      contract_area(C, inv, A[kk], gl[kk]);
    } else {
#pragma unroll
      for (int i = 0; i < 3; ++i) gl[kk][i] = g[kk][i];
    }
  });
  R fl[3] = {R(0), R(0), R(0)};  // (rho u - tau (-Gp)) . A accumulated per ip, minus tau grad p . A
  static_for<8>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    const R rc = R(gt.s(rho, c)), pc = R(gt.s(p, c));
    const R uc[3] = {R(gt.v(u, c, 0)), R(gt.v(u, c, 1)), R(gt.v(u, c, 2))};
    const R gc[3] = {R(gt.v(gp, c, 0)), R(gt.v(gp, c, 1)), R(gt.v(gp, c, 2))};
    static_for<3>([&](auto kk) {
      constexpr int IP = adj_ip(LA, decltype(kk)::value);
      const R r = R(shape(IP, c, SHIFT_MDOT));
      const R ua = uc[0] * A[kk][0] + uc[1] * A[kk][1] + uc[2] * A[kk][2];
      const R ga = gc[0] * A[kk][0] + gc[1] * A[kk][1] + gc[2] * A[kk][2];
      fl[kk] += r * (rc * ua + tau * ga) - tau * pc * dsdot<IP, c, SHIFT_POISSON>(g[kk]);
    });
  });
  R rhs = R(0);
  static_for<3>([&](auto kk) {
    constexpr R S = R(adj_sign(LA, decltype(kk)::value));
    rhs -= S * (fl[kk] / tau);
    if constexpr (adj_sign(LA, decltype(kk)::value) > 0) ipout(adj_ip(LA, decltype(kk)::value), fl[kk]);
  });
  static_for<8>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    R v = R(0);
    static_for<3>([&](auto kk) {
      constexpr int IP = adj_ip(LA, decltype(kk)::value);
      constexpr R S = R(adj_sign(LA, decltype(kk)::value));
      v += S * (-dsdot<IP, c, LHS_SHIFT>(gl[kk]));
    });
    emit(c, v);
  });
  return rhs;
}

// MomentumAdvDiffElemKernel, row block LA (src/kernel/MomentumAdvDiffElemKernel.C:72-186).
// Rows (la, i), columns (c, j); per ip (S = +1 on the left node):
//   lhs += S [ delta_ij (N_adv(ip,c) mdot - mu_ip dN_c/dx . A) - mu_ip dN_c/dx_i A_j ]
//   rhs += S [ -mdot u_ip,i - 2/3 mu_ip divU A_i incl + mu_ip sum_c dN_c/dx_i (A . u_c)
//              + mu_ip sum_c (dN_c/dx . A) u_c,i ]
// emit(c, blk) receives the 3x3 block blk[i][j] of column node c.  SEG leaves
// the transpose-gradient term -mu_ip dN_c/dx_i A_j out of the lhs (it stays in
// the rhs), so that blk = a_c I: the three components share one scalar matrix.
template <int IP, int C, bool SHIFT, typename R>
__device__ __forceinline__ void dndx_scaled(const R (&Cs)[3][3], R (&dn)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
    dn[i] = Cs[i][0] * R(deriv(IP, C, 0, SHIFT)) + Cs[i][1] * R(deriv(IP, C, 1, SHIFT)) +
            Cs[i][2] * R(deriv(IP, C, 2, SHIFT));
}

template <int LA, bool SKEW, bool SHIFT, typename JR, bool SEG = false, typename R, class Emit>
__device__ __forceinline__ void row_momentum(const R (&x)[8][3], const Gather& gt, const double* __restrict__ u,
                                             const double* __restrict__ mu, const R (&mdot3)[3], R incl_divu,
                                             R (&rhs)[3], Emit&& emit) {
  R Cs[3][3][3], A[3][3];  // per ip: cofactors / det J, area vector
  static_for<3>([&](auto kk) {
    constexpr int IP = adj_ip(LA, decltype(kk)::value);
    JR C[3][3], inv;
    scs_area<IP>(x, A[kk]);
    jacobian<IP, SHIFT>(x, C, inv);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Cs[kk][i][j] = R(C[i][j] * inv);
  });
  constexpr int IP0 = adj_ip(LA, 0), IP1 = adj_ip(LA, 1), IP2 = adj_ip(LA, 2);
  const double(*dv[3])[3] = {kScsTab.deriv[SHIFT][IP0], kScsTab.deriv[SHIFT][IP1], kScsTab.deriv[SHIFT][IP2]};
  const double* sh[3] = {kScsTab.shape[0][IP0], kScsTab.shape[0][IP1], kScsTab.shape[0][IP2]};
  const double* sa[3] = {kScsTab.shape[SKEW][IP0], kScsTab.shape[SKEW][IP1], kScsTab.shape[SKEW][IP2]};
  // pass 1: ip values of mu, u (advection shape functions) and divU
  R muIp[3] = {R(0), R(0), R(0)}, divU[3] = {R(0), R(0), R(0)}, uIp[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) uIp[a][b] = R(0);
#pragma unroll 1
  for (int c = 0; c < 8; ++c) {
    const R muc = R(gt.s(mu, c));
    const R uc[3] = {R(gt.v(u, c, 0)), R(gt.v(u, c, 1)), R(gt.v(u, c, 2))};
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      muIp[kk] += R(sh[kk][c]) * muc;
      R dn[3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        dn[i] = Cs[kk][i][0] * R(dv[kk][c][0]) + Cs[kk][i][1] * R(dv[kk][c][1]) + Cs[kk][i][2] * R(dv[kk][c][2]);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        uIp[kk][j] += R(sa[kk][c]) * uc[j];
        divU[kk] += uc[j] * dn[j];
      }
    }
  }
  static_for<3>([&](auto kk) {
    constexpr R S = R(adj_sign(LA, decltype(kk)::value));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      rhs[i] -= S * (mdot3[kk] * uIp[kk][i] + R(2.0 / 3.0) * muIp[kk] * divU[kk] * A[kk][i] * incl_divu);
      A[kk][i] *= muIp[kk];  // A becomes mu_ip A
    }
  });
  constexpr R S0 = R(adj_sign(LA, 0)), S1 = R(adj_sign(LA, 1)), S2 = R(adj_sign(LA, 2));
  const R Sg[3] = {S0, S1, S2};
  // pass 2: one 3x3 column block at a time
#pragma unroll 1
  for (int c = 0; c < 8; ++c) {
    const R uc[3] = {R(gt.v(u, c, 0)), R(gt.v(u, c, 1)), R(gt.v(u, c, 2))};
    R blk[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) blk[i][j] = R(0);
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      const R S = Sg[kk];
      R dn[3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        dn[i] = Cs[kk][i][0] * R(dv[kk][c][0]) + Cs[kk][i][1] * R(dv[kk][c][1]) + Cs[kk][i][2] * R(dv[kk][c][2]);
      const R h = dn[0] * A[kk][0] + dn[1] * A[kk][1] + dn[2] * A[kk][2];
      const R adv = R(sa[kk][c]) * mdot3[kk];
      const R Au = uc[0] * A[kk][0] + uc[1] * A[kk][1] + uc[2] * A[kk][2];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        blk[i][i] += S * (adv - h);
        if (!SEG) {
#pragma unroll
          for (int j = 0; j < 3; ++j) blk[i][j] -= S * (dn[i] * A[kk][j]);
        }
        rhs[i] += S * (dn[i] * Au + h * uc[i]);
      }
    }
    emit(c, blk);
  }
}


// ---------------------------------------------------------------- sub-control volumes
// Sub-hexahedron of node LA as subdivision point numbers (HexSCV::determinant,
// src/master_element/Hex8CVFEM.C:149-178) and its volume by Grandy's formula
// (hex_volume_grandy, include/master_element/Hex8GeometryFunctions.h:78-163):
// six face centroids, 24 triangles, vol = 1/18 sum (p + q + r) . ((q - p) x (r - p)).
constexpr int kScvSub[8][8] = {{0, 8, 12, 11, 19, 20, 26, 25},  {8, 1, 9, 12, 20, 18, 24, 26},
                               {12, 9, 2, 10, 26, 24, 22, 23},  {11, 12, 10, 3, 25, 26, 23, 21},
                               {19, 20, 26, 25, 4, 13, 17, 16}, {20, 18, 24, 26, 13, 5, 14, 17},
                               {26, 24, 22, 23, 17, 14, 6, 15}, {25, 26, 23, 21, 16, 17, 15, 7}};
constexpr int kGrandyFace[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 3, 7}};
constexpr int kGrandyTri[24][3] = {{0, 8, 1},   {8, 2, 1},   {3, 2, 8},   {3, 8, 0},   {6, 9, 5},   {7, 9, 6},
                                   {4, 9, 7},   {4, 5, 9},   {10, 0, 1},  {5, 10, 1},  {4, 10, 5},  {4, 0, 10},
                                   {7, 6, 11},  {6, 2, 11},  {2, 3, 11},  {3, 7, 11},  {6, 12, 2},  {5, 12, 6},
                                   {5, 1, 12},  {1, 2, 12},  {0, 4, 13},  {4, 7, 13},  {7, 3, 13},  {3, 0, 13}};

template <int LA>
__device__ __forceinline__ double scv_volume(const double (&x)[8][3]) {
  double c[14][3];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const uint32_t m = subdiv_mask(kScvSub[LA][n]);
    const int cnt = popc8(m);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      bool first = true;
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if ((m >> i) & 1u) {
          s = first ? x[i][d] : s + x[i][d];
          first = false;
        }
      c[n][d] = cnt == 1 ? s : s * (1.0 / cnt);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      c[8 + k][d] = 0.25 * (((c[kGrandyFace[k][0]][d] + c[kGrandyFace[k][1]][d]) + c[kGrandyFace[k][2]][d]) +
                            c[kGrandyFace[k][3]][d]);
  double vol = 0.0;
#pragma unroll
  for (int k = 0; k < 24; ++k) {
    const int p = kGrandyTri[k][0], q = kGrandyTri[k][1], r = kGrandyTri[k][2];
    const double m0 = (c[p][0] + c[q][0]) + c[r][0];
    const double m1 = (c[p][1] + c[q][1]) + c[r][1];
    const double m2 = (c[p][2] + c[q][2]) + c[r][2];
    const double d0 = (c[q][1] - c[p][1]) * (c[r][2] - c[p][2]) - (c[r][1] - c[p][1]) * (c[q][2] - c[p][2]);
    const double d1 = (c[r][0] - c[p][0]) * (c[q][2] - c[p][2]) - (c[q][0] - c[p][0]) * (c[r][2] - c[p][2]);
    const double d2 = (c[q][0] - c[p][0]) * (c[r][1] - c[p][1]) - (c[r][0] - c[p][0]) * (c[q][1] - c[p][1]);
    vol += (m0 * d0 + m1 * d1) + m2 * d2;
  }
  return vol / 18.0;
}

// Shape function of column node c at the scv ip of node LA: the scv centroid is
// at +-1/4 (HexSCV::intgLoc_, Hex8CVFEM.C:88-96), so N_c = prod (1/2 + sign/4):
// (3/4)^k (1/4)^(3-k), k = axes on which c and LA agree (27/64, 9/64, 3/64, 1/64);
// the lumped form (shifted locations, :99-107, at the nodes) is delta_{c,LA}.
template <int LA, bool LUMPED>
__host__ __device__ constexpr double scv_shape(int c) {
  if (LUMPED) return c == LA ? 1.0 : 0.0;
  int k = 0;
  for (int d = 0; d < 3; ++d) k += node_sign(c, d) == node_sign(LA, d) ? 1 : 0;
  return k == 3 ? 27.0 / 64 : k == 2 ? 9.0 / 64 : k == 1 ? 3.0 / 64 : 1.0 / 64;
}

}  // namespace hex8
}  // namespace nalu
