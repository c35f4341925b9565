// Exposed-face closure of the projected nodal gradient
// (AssembleNodalGradBoundaryAlgorithm::execute, src/AssembleNodalGradBoundaryAlgorithm.C:54-152).
//
// The exposed sides of the Hex8 block are found once per mesh from the node
// incidence: side s of element e is exposed when no other element holds all
// four of its nodes.  Sides are listed in HexSCS side order
// (sideNodeOrdinals_, src/master_element/Hex8CVFEM.C:415-422), whose
// right-hand normal points out of the element.  Each exposed side adds, for
// its four sub-faces (Quad3DSCS::determinant, quad3d_scs_det,
// src/master_element/MasterElementWork.F:1467-1584), q_ip A_i / V_i to the
// gradient of its node i, with q_ip from the quad shape functions at
// (+-1/4, +-1/4) (quad3d_shape_fcn, :3651-3695) or at the node when shifted.
// Boundary nodes are shared by up to four exposed sides, so the additions are
// atomic.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace nalu {
namespace {

__constant__ const int kSide[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                        {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};

inline unsigned grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// flag[6e + s] = 1 when side s of element e is exposed
__global__ void k_side_exposed(const int32_t* __restrict__ conn, int64_t nelem, const int32_t* __restrict__ inc_ptr,
                               const int32_t* __restrict__ inc_elem, int32_t* __restrict__ flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nelem) return;
  const int64_t e = t / 6;
  const int s = (int)(t - 6 * e);
  int32_t fn[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) fn[k] = conn[8 * e + kSide[s][k]];
  bool shared = false;
  for (int32_t q = inc_ptr[fn[0]]; q < inc_ptr[fn[0] + 1] && !shared; ++q) {
    const int64_t o = inc_elem[q] >> 3;
    if (o == e) continue;
    int hits = 0;
#pragma unroll
    for (int k = 1; k < 4; ++k) {
      bool in = false;
#pragma unroll
      for (int n = 0; n < 8; ++n) in = in || conn[8 * o + n] == fn[k];
      hits += in ? 1 : 0;
    }
    shared = hits == 3;
  }
  flag[t] = shared ? 0 : 1;
}

__global__ void k_side_list(const int32_t* __restrict__ conn, int64_t nelem, const int32_t* __restrict__ flag,
                            const int32_t* __restrict__ pos, int32_t* __restrict__ faces) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nelem || !flag[t]) return;
  const int64_t e = t / 6;
  const int s = (int)(t - 6 * e);
#pragma unroll
  for (int k = 0; k < 4; ++k) faces[4 * (int64_t)pos[t] + k] = conn[8 * e + kSide[s][k]];
}

template <bool SHIFT>
__global__ void k_nodal_grad_boundary(const int32_t* __restrict__ faces, int64_t nf, const double* __restrict__ x,
                                      const double* __restrict__ q, const double* __restrict__ vol,
                                      double* __restrict__ dqdx) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int32_t nd[4];
  double p[4][3], qf[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    nd[k] = faces[4 * f + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) p[k][d] = x[3 * (int64_t)nd[k] + d];
    qf[k] = q[nd[k]];
  }
  double e[4][3], c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int i = 0; i < 4; ++i) e[i][d] = (p[i][d] + p[(i + 1) & 3][d]) * 0.5;
    c[d] = (p[0][d] + p[1][d] + p[2][d] + p[3][d]) * 0.25;
  }
  // quad shape functions at the ips (s1, s2) = h (-1,-1), (1,-1), (1,1), (-1,1)
  constexpr double h = SHIFT ? 0.5 : 0.25;
  constexpr double sg[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double s1 = h * sg[i][0], s2 = h * sg[i][1];
    const double N[4] = {0.25 + 0.5 * (-s1 - s2) + s1 * s2, 0.25 + 0.5 * (s1 - s2) - s1 * s2,
                         0.25 + 0.5 * (s1 + s2) + s1 * s2, 0.25 + 0.5 * (-s1 + s2) - s1 * s2};
    const double qIp = N[0] * qf[0] + N[1] * qf[1] + N[2] * qf[2] + N[3] * qf[3];
    const int im = (i + 3) & 3;
    const double d13[3] = {c[0] - p[i][0], c[1] - p[i][1], c[2] - p[i][2]};
    const double d24[3] = {e[im][0] - e[i][0], e[im][1] - e[i][1], e[im][2] - e[i][2]};
    const double A[3] = {0.5 * (d24[2] * d13[1] - d13[2] * d24[1]), 0.5 * (d24[0] * d13[2] - d13[0] * d24[2]),
                         0.5 * (d24[1] * d13[0] - d13[1] * d24[0])};
    const double inv = 1.0 / vol[nd[i]];
#pragma unroll
    for (int j = 0; j < 3; ++j) unsafeAtomicAdd(dqdx + 3 * (int64_t)nd[i] + j, qIp * A[j] * inv);
  }
}

}  // namespace

void build_exposed_faces(nalu_mesh* m, hipStream_t s) {
  if (m->exposed_ready) return;
  build_node_incidence(m, s);
  const int64_t E = m->num_elems, S = 6 * E;
  NALU_REQUIRE(S < ((int64_t)1 << 31), NALU_ERR_UNSUPPORTED, "exposed faces: mesh too large");
  DevBuf<int32_t> flag(S), pos(S);
  k_side_exposed<<<grid(S), 256, 0, s>>>(m->conn.p, E, m->inc_ptr.p, m->inc_elem.p, flag.p);
  size_t tmp = 0;
  NALU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flag.p, pos.p, (int)S, s));
  {
    DevBuf<char> t(tmp ? tmp : 1);
    NALU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, flag.p, pos.p, (int)S, s));
  }
  int32_t last_pos = 0, last_flag = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&last_pos, pos.p + S - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipMemcpyAsync(&last_flag, flag.p + S - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  m->num_exposed = (int64_t)last_pos + last_flag;
  m->exposed_faces.alloc(std::max<int64_t>(1, 4 * m->num_exposed));
  k_side_list<<<grid(S), 256, 0, s>>>(m->conn.p, E, flag.p, pos.p, m->exposed_faces.p);
  NALU_HIP_CHECK(hipGetLastError());
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  m->exposed_ready = true;
}

void nodal_grad_boundary(nalu_mesh* m, const double* q, const double* vol, bool shifted, double* dqdx,
                         hipStream_t s) {
  build_exposed_faces(m, s);
  const int64_t nf = m->num_exposed;
  if (nf == 0) return;
  if (shifted)
    k_nodal_grad_boundary<true><<<grid(nf), 256, 0, s>>>(m->exposed_faces.p, nf, m->coords.p, q, vol, dqdx);
  else
    k_nodal_grad_boundary<false><<<grid(nf), 256, 0, s>>>(m->exposed_faces.p, nf, m->coords.p, q, vol, dqdx);
  NALU_HIP_CHECK(hipGetLastError());
}

}  // namespace nalu
