// Sparse kernels and Krylov solvers on the assembled CSR system.
//
// Replaces Belos (GMRES with ICGS, CG; set up at src/LinearSolver.C:106-111,
// parameters src/LinearSolverConfig.C:54-69) and Ifpack2 Relaxation
// (SGS / Jacobi, src/LinearSolverConfig.C:71-85).  The algorithms are the
// textbook ones (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.:
// GMRES with Givens rotations Alg. 6.9, right preconditioning Alg. 9.5, FGMRES
// Alg. 9.6, PCG Alg. 9.1; symmetric Gauss-Seidel M = (D+L) D^-1 (D+U)).
// The SGS sweep is the multicolor variant ("mt_sgs"): rows of one color are
// independent, colors are swept 0..C-1 then C-1..0.
//
// HBM layout: CSR values / col_idx are streamed once per SpMV through LDS
// (row-block "CSR-stream"), so every wave issues fully coalesced loads;
// Krylov bases are column-major [k][rows].
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

#include "internal.hpp"

// unroll depth of the SELL entry loops (loads in flight per lane)
#ifndef NALU_SPARSE_UNROLL
#define NALU_SPARSE_UNROLL 4
#endif

namespace nalu {
namespace {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 1024;

inline unsigned nblocks(int64_t n, int per = kThreads) {
  return (unsigned)std::max<int64_t>(1, (n + per - 1) / per);
}
inline unsigned red_blocks(int64_t n) { return (unsigned)std::min<int64_t>(kRedBlocks, nblocks(n)); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kThreads / 64];
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < 64) {
    t = (l < (int)(blockDim.x >> 6)) ? sh[l] : 0.0;
    t = wave_sum(t);
  }
  __syncthreads();
  return t;  // valid in thread 0
}

// ------------------------------------------------------------ reductions
template <typename TA, typename TB>
__global__ __launch_bounds__(kThreads) void k_dot_part(const TA* __restrict__ a, const TB* __restrict__ b, int64_t n,
                                                       double* __restrict__ part) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[i] * (double)b[i];
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[j] = sum of part[j][0..nb)   (one block per j)
__global__ __launch_bounds__(kThreads) void k_reduce_rows(const double* __restrict__ part, int nb,
                                                          double* __restrict__ out, int accumulate) {
  const double* p = part + (int64_t)blockIdx.x * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
  s = block_sum(s);
  if (threadIdx.x == 0) out[blockIdx.x] = accumulate ? out[blockIdx.x] + s : s;
}

// w -= V h  (h on device, length cnt)
template <typename TV, typename TW>
__global__ __launch_bounds__(kThreads) void k_maxpy(const TV* __restrict__ V, int64_t ldv, int cnt,
                                                    const double* __restrict__ h, TW* __restrict__ w, int64_t n,
                                                    double sign) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = 0; j < cnt; ++j) acc += h[j] * (double)V[j * ldv + i];
  w[i] = (TW)((double)w[i] + sign * acc);
}

// ---- vectorized Gram-Schmidt passes (basis columns padded to 64 entries, so
// every column starts 16-byte aligned; 16-byte loads per lane)
template <typename T>
struct V16;
template <>
struct V16<float> {
  static constexpr int N = 4;
  __device__ static void load(const float* p, double* o) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
  }
};
template <>
struct V16<double> {
  static constexpr int N = 2;
  __device__ static void load(const double* p, double* o) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    o[0] = a.x; o[1] = a.y;
  }
};

// Coefficients part[j * gridDim.x + blk] = V_j . w over the chunks blk, blk + G,
// ... of R rows (j < cnt), plus w . w as row cnt when with_self; w's chunk is
// staged in LDS, one wavefront per basis vector, per-block sums kept in LDS.
template <typename TV>
__global__ __launch_bounds__(kThreads) void k_gs_dots(const TV* __restrict__ V, int64_t ldv, int cnt, int with_self,
                                                      const TV* __restrict__ w, int64_t n, int R,
                                                      double* __restrict__ part) {
  extern __shared__ unsigned char smem_raw[];
  TV* ws = reinterpret_cast<TV*>(smem_raw);         // [R]
  double* acc = reinterpret_cast<double*>(ws + R);  // [cnt + 1]
  constexpr int VN = V16<TV>::N;
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int rows = cnt + (with_self ? 1 : 0);
  for (int j = threadIdx.x; j < rows; j += blockDim.x) acc[j] = 0.0;
  const int64_t nch = (n + R - 1) / R;
  for (int64_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const int64_t r0 = c * R;
    const int m = (int)min<int64_t>(R, n - r0);
    __syncthreads();
    for (int t = threadIdx.x; t < m; t += blockDim.x) ws[t] = w[r0 + t];
    __syncthreads();
    const int mv = m / VN * VN;
    for (int j = wv; j < rows; j += kThreads / 64) {
      const TV* v = (j < cnt ? V + (int64_t)j * ldv : w) + r0;
      double sj = 0.0;
      for (int t = l * VN; t < mv; t += 64 * VN) {
        double a[VN];
        V16<TV>::load(v + t, a);
#pragma unroll
        for (int u = 0; u < VN; ++u) sj += a[u] * (double)ws[t + u];
      }
      for (int t = mv + l; t < m; t += 64) sj += (double)v[t] * (double)ws[t];
      sj = wave_sum(sj);
      if (l == 0) acc[j] += sj;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < rows; j += blockDim.x) part[(int64_t)j * gridDim.x + blockIdx.x] = acc[j];
}

// w -= V h (FP64 accumulation); each thread owns VN consecutive rows
template <typename TV>
__global__ __launch_bounds__(kThreads) void k_gs_update(const TV* __restrict__ V, int64_t ldv, int cnt,
                                                        const double* __restrict__ h, TV* __restrict__ w, int64_t n) {
  constexpr int VN = V16<TV>::N;
  const int64_t ng = (n + VN - 1) / VN;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = g * VN;
    if (i0 + VN <= n) {
      double a[VN];
      V16<TV>::load(w + i0, a);
      for (int j = 0; j < cnt; ++j) {
        double v[VN];
        V16<TV>::load(V + (int64_t)j * ldv + i0, v);
        const double hj = h[j];
#pragma unroll
        for (int u = 0; u < VN; ++u) a[u] -= hj * v[u];
      }
#pragma unroll
      for (int u = 0; u < VN; ++u) w[i0 + u] = (TV)a[u];
    } else {
      for (int64_t i = i0; i < n; ++i) {
        double t = (double)w[i];
        for (int j = 0; j < cnt; ++j) t -= h[j] * (double)V[(int64_t)j * ldv + i];
        w[i] = (TV)t;
      }
    }
  }
}

// pass-2 update with the norm of the result: w -= V h, part[blk] = ||w||^2 (partial);
// each thread owns VN consecutive rows
template <typename TV>
__global__ __launch_bounds__(kThreads) void k_gs_update_norm(const TV* __restrict__ V, int64_t ldv, int cnt,
                                                             const double* __restrict__ h, TV* __restrict__ w,
                                                             int64_t n, double* __restrict__ part) {
  constexpr int VN = V16<TV>::N;
  double s2 = 0.0;
  const int64_t ng = (n + VN - 1) / VN;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = g * VN;
    double a[VN];
    if (i0 + VN <= n) {
      V16<TV>::load(w + i0, a);
      for (int j = 0; j < cnt; ++j) {
        double v[VN];
        V16<TV>::load(V + (int64_t)j * ldv + i0, v);
        const double hj = h[j];
#pragma unroll
        for (int u = 0; u < VN; ++u) a[u] -= hj * v[u];
      }
#pragma unroll
      for (int u = 0; u < VN; ++u) {
        const TV t = (TV)a[u];
        w[i0 + u] = t;
        s2 += (double)t * (double)t;
      }
    } else {
      for (int64_t i = i0; i < n; ++i) {
        double t = (double)w[i];
        for (int j = 0; j < cnt; ++j) t -= h[j] * (double)V[(int64_t)j * ldv + i];
        const TV tv = (TV)t;
        w[i] = tv;
        s2 += (double)tv * (double)tv;
      }
    }
  }
  s2 = block_sum(s2);
  if (threadIdx.x == 0) part[blockIdx.x] = s2;
}

// ------------------------------------------------------------ SpMV (CSR-stream)
// Block owns rows [r0, r0+R); products of its contiguous nnz segment are
// formed with coalesced loads into LDS, then each row sums its slice.
// mode 0: y = A x;  mode 1: y = b - A x.
template <typename T, typename TX, typename TY>
__global__ __launch_bounds__(kThreads) void k_spmv_stream(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                          const T* __restrict__ v, const TX* __restrict__ x,
                                                          TY* __restrict__ y, const TY* __restrict__ b, int64_t nrows,
                                                          int rows_per_block, int mode) {
  extern __shared__ unsigned char smem_raw[];
  T* prod = reinterpret_cast<T*>(smem_raw);
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = min(r0 + (int64_t)rows_per_block, nrows);
  const int32_t s0 = rp[r0], s1 = rp[r1];
  const int ns = s1 - s0;
  for (int k = threadIdx.x; k < ns; k += blockDim.x) prod[k] = v[s0 + k] * (T)x[ci[s0 + k]];
  __syncthreads();
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    const int a = rp[r] - s0, e = rp[r + 1] - s0;
    T s = T(0);
    for (int k = a; k < e; ++k) s += prod[k];
    y[r] = mode ? (TY)((T)b[r] - s) : (TY)s;
  }
}

// ------------------------------------------------------------ relaxation
template <typename T>
__global__ void k_inv_diag(const int32_t* __restrict__ dpos, const double* __restrict__ v, int64_t n,
                           T* __restrict__ invd, int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = v[dpos[i]];
  if (d == 0.0 || !isfinite(d)) atomicOr(bad, 1);
  invd[i] = (T)(1.0 / d);
}

template <typename T>
__global__ void k_jacobi(const T* __restrict__ invd, const T* __restrict__ r, T* __restrict__ z, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = r[i] * invd[i];
}

// one color of a Gauss-Seidel sweep; LANES lanes cooperate on a row
template <typename T, int LANES>
__global__ __launch_bounds__(kThreads) void k_gs_color(const int32_t* __restrict__ rows, int64_t nr,
                                                       const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                       const T* __restrict__ v, const T* __restrict__ invd,
                                                       const T* __restrict__ r, T* __restrict__ z) {
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
  const int lane = threadIdx.x % LANES;
  if (g >= nr) return;  // whole lane groups exit together
  const int32_t i = rows[g];
  T s = T(0);
  for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += LANES) {
    const int32_t j = ci[q];
    if (j != i) s += v[q] * z[j];
  }
#pragma unroll
  for (int o = LANES / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) z[i] = (r[i] - s) * invd[i];
}

// ------------------------------------------------------------ BLAS-1
template <typename T>
__global__ void k_fill(T* __restrict__ x, int64_t n, T val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = val;
}
template <typename TA, typename TB>
__global__ void k_copy(const TA* __restrict__ a, TB* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (TB)a[i];
}
// x *= 1/sqrt(nrm2[0])
template <typename T>
__global__ void k_scale_invnorm(T* __restrict__ x, int64_t n, const double* __restrict__ nrm2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double s = 1.0 / sqrt(nrm2[0]);
  if (i < n) x[i] = (T)((double)x[i] * s);
}
// y = a*x + y with a host scalar
template <typename TX, typename TY>
__global__ void k_axpy(double a, const TX* __restrict__ x, TY* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (TY)((double)y[i] + a * (double)x[i]);
}
// CG update with device scalars: alpha = rz/pq; x += alpha p; r -= alpha q; partial ||r||^2
__global__ __launch_bounds__(kThreads) void k_cg_update(double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        const double* __restrict__ rz, const double* __restrict__ pq,
                                                        int64_t n, double* __restrict__ part) {
  const double alpha = rz[0] / pq[0];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += ri * ri;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// p = z + (rz_new/rz_old) p
__global__ void k_cg_p(double* __restrict__ p, const double* __restrict__ z, const double* __restrict__ rz_new,
                       const double* __restrict__ rz_old, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + (rz_new[0] / rz_old[0]) * p[i];
}

// ------------------------------------------------------------ fused CG building blocks
// Device-side CG for one right-hand side on the SELL copy with the SGS
// preconditioner: the dot products ride on the kernels that produce their
// operands (SpMV: p.q, the update: r.r, the sweeps that finalize z: r.z) as
// per-block partial sums, which one single-block launch reduces in a fixed
// order (deterministic).  No grid-wide election inside the producing kernels:
// on a multi-XCD part every device-scope release writes back the XCD's L2, and
// thousands of blocks doing so per launch cost more than the launch saved.
// The convergence test is a device flag the host reads every few iterations,
// so an iteration costs no host round trip; kernels of an iteration after
// convergence return at once.
struct CgState {
  double r0;       // |r_0|
  double tol2r02;  // (tol |r_0|)^2
  double rz[2];    // r.z, current and next (alternating slots)
  double pq;       // p.q
  double rr;       // r.r
  int done;        // 1 converged, 2 non-finite residual
  int iters;       // iteration at which done was set
};

// One block: sum of part[0, nb) in a fixed order into the CgState field
// `which` (0 p.q, 1 r.r with the convergence test at iteration it, 2 r.z of
// slot `slot`).
constexpr int kReduceThreads = 1024;
__global__ __launch_bounds__(kReduceThreads) void k_cg_reduce(const double* __restrict__ part, int nb, int which,
                                                              int slot, int it, CgState* __restrict__ st) {
  __shared__ double sh[kReduceThreads / 64];
  if (st->done) return;
  // four independent loads in flight per thread, then a fixed-order tree
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * kReduceThreads < nb; i += 4 * kReduceThreads) {
    a0 += part[i];
    a1 += part[i + kReduceThreads];
    a2 += part[i + 2 * kReduceThreads];
    a3 += part[i + 3 * kReduceThreads];
  }
  for (; i < nb; i += kReduceThreads) a0 += part[i];
  double v = wave_sum((a0 + a1) + (a2 + a3));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x >= 64) return;
  v = wave_sum(threadIdx.x < kReduceThreads / 64 ? sh[threadIdx.x] : 0.0);
  const double s = v;
  if (threadIdx.x != 0) return;
  if (which == 0) {
    st->pq = s;
  } else if (which == 2) {
    st->rz[slot] = s;
  } else {
    st->rr = s;
    if (!(s == s) || s > 1e300) {
      st->done = 2;
      st->iters = it;
    } else if (s <= st->tol2r02) {
      st->done = 1;
      st->iters = it;
    }
  }
}

// x += alpha p, r -= alpha q (alpha = r.z / p.q); per-block partials of r.r
__global__ __launch_bounds__(kThreads) void k_cg_update_f(double* __restrict__ x, double* __restrict__ r,
                                                          const double* __restrict__ p,
                                                          const double* __restrict__ q, int64_t n, int cur,
                                                          double* __restrict__ part, const CgState* __restrict__ st) {
  if (st->done) return;
  const double alpha = st->rz[cur] / st->pq;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += ri * ri;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// p = z + (r.z_new / r.z_old) p
__global__ void k_cg_p_f(double* __restrict__ p, const double* __restrict__ z, int64_t n, int nxt, int cur,
                         const CgState* __restrict__ st) {
  if (st->done) return;
  const double beta = st->rz[nxt] / st->rz[cur];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + beta * p[i];
}

// ------------------------------------------------------------ host helpers
struct Red {
  nalu_linsys* ls;
  hipStream_t s;
  explicit Red(nalu_linsys* l) : ls(l), s(l->stream) {
    const size_t need = (size_t)std::max(kRedBlocks, 2048) * std::max(64, ls->cfg.kspace + 2);
    if (ls->red.n < need) ls->red.alloc(need);
    if (ls->red_out.n < 256) ls->red_out.alloc(256);
  }
  // dot into device slot out (not synchronized)
  template <typename TA, typename TB>
  void dot_dev(const TA* a, const TB* b, int64_t n, double* out) {
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL((k_dot_part<TA, TB>), dim3(nb), dim3(kThreads), 0, s, a, b, n, ls->red.p);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kThreads), 0, s, ls->red.p, (int)nb, out, 0);
  }
  template <typename TA, typename TB>
  double dot(const TA* a, const TB* b, int64_t n) {
    dot_dev(a, b, n, ls->red_out.p);
    double h = 0.0;
    NALU_HIP_CHECK(hipMemcpyAsync(&h, ls->red_out.p, sizeof(double), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    return h;
  }
};

int spmv_rows_per_block(const nalu_linsys* ls, size_t elem_bytes) {
  const int maxlen = std::max(1, ls->max_node_row_len * ls->num_dof);
  int R = 256;
  while (R > 8 && (size_t)R * maxlen * elem_bytes > 48 * 1024) R >>= 1;
  return R;
}

template <typename T, typename TX, typename TY>
void spmv_t(nalu_linsys* ls, const T* vals, const TX* x, TY* y, const TY* b, int mode) {
  const int R = spmv_rows_per_block(ls, sizeof(T));
  const size_t lds = (size_t)R * std::max(1, ls->max_node_row_len * ls->num_dof) * sizeof(T);
  hipLaunchKernelGGL((k_spmv_stream<T, TX, TY>), dim3(nblocks(ls->nrows, R)), dim3(kThreads), lds, ls->stream,
                     ls->row_ptr.p, ls->col_idx.p, vals, x, y, b, ls->nrows, R, mode);
}

// ------------------------------------------------------------ color-permuted split system
// Inside the solvers the matrix is used in a color-contiguous row order with
// each row split into its L part (columns of earlier colors), U part (later
// colors) and diagonal.  One SGS application is then a forward sweep over L
// and a backward sweep over U, y = (D+L)^-1 r, z = (D+U)^-1 D y (M^-1 for
// M = (D+L) D^-1 (D+U)), i.e. one pass over the off-diagonal entries, and every
// color's rows and entries are contiguous in HBM.

__global__ void k_iperm(const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ iperm) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) iperm[perm[k]] = (int32_t)k;
}

__device__ __forceinline__ int color_of(const int32_t* __restrict__ cpd, int nc, int64_t k) {
  int lo = 0, hi = nc;  // cpd[lo] <= k < cpd[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cpd[mid] <= k) lo = mid;
    else hi = mid;
  }
  return lo;
}

// pass 0: count L/U entries per permuted row; pass 1: fill
__global__ void k_split(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                        const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm,
                        const int32_t* __restrict__ cpd, int nc, const int32_t* __restrict__ dpos, int64_t n,
                        int32_t* __restrict__ lcnt, int32_t* __restrict__ ucnt, const int32_t* __restrict__ lrp,
                        const int32_t* __restrict__ urp, int32_t* __restrict__ lci, int32_t* __restrict__ lsrc,
                        int32_t* __restrict__ uci, int32_t* __restrict__ usrc, int32_t* __restrict__ dsrc, int fill) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t i = perm[k];
  const int c = color_of(cpd, nc, k);
  const int32_t lo = cpd[c];
  int32_t nl = 0, nu = 0;
  for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
    const int32_t j = ci[q];
    if (j == i) continue;
    const int32_t kj = iperm[j];
    if (kj < lo) {
      if (fill) {
        lci[lrp[k] + nl] = kj;
        lsrc[lrp[k] + nl] = q;
      }
      ++nl;
    } else {
      if (fill) {
        uci[urp[k] + nu] = kj;
        usrc[urp[k] + nu] = q;
      }
      ++nu;
    }
  }
  if (!fill) {
    lcnt[k] = nl;
    ucnt[k] = nu;
  } else {
    dsrc[k] = dpos[i];
  }
}

template <typename T>
__global__ void k_take(const int32_t* __restrict__ src, const double* __restrict__ v, int64_t n, T* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = (T)v[src[q]];
}

template <typename T>
__global__ void k_take_diag(const int32_t* __restrict__ dsrc, const double* __restrict__ v, int64_t n,
                            T* __restrict__ d, T* __restrict__ invd, int* __restrict__ bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double dv = v[dsrc[k]];
  if (dv == 0.0 || !isfinite(dv)) atomicOr(bad, 1);
  d[k] = (T)dv;
  invd[k] = (T)(1.0 / dv);
}

template <typename TA, typename TB>
__global__ void k_gather(const TA* __restrict__ x, const int32_t* __restrict__ perm, int64_t n, TB* __restrict__ xp) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) xp[k] = (TB)x[perm[k]];
}

// interleaved forms: out[k * stride] = x[perm[k]], x[perm[k]] = in[k * stride]
__global__ void k_gather_strided(const double* __restrict__ x, const int32_t* __restrict__ perm, int64_t n,
                                 int stride, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k * stride] = x[perm[k]];
}
__global__ void k_scatter_strided(const double* __restrict__ in, const int32_t* __restrict__ perm, int64_t n,
                                  int stride, double* __restrict__ x) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[perm[k]] = in[k * stride];
}
__global__ void k_take_strided(const double* __restrict__ in, int64_t n, int stride, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[k * stride];
}

template <typename TA, typename TB>
__global__ void k_scatter(const TA* __restrict__ xp, const int32_t* __restrict__ perm, int64_t n, TB* __restrict__ x) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[perm[k]] = (TB)xp[k];
}

// y = (D + L + U) x  (mode 0)  or  y = b - (D + L + U) x  (mode 1), CSR-stream
// over the block's L and U segments.
template <typename T, typename TY>
__global__ __launch_bounds__(kThreads) void k_split_spmv(const int32_t* __restrict__ lrp, const int32_t* __restrict__ lci,
                                                         const T* __restrict__ lv, const int32_t* __restrict__ urp,
                                                         const int32_t* __restrict__ uci, const T* __restrict__ uv,
                                                         const T* __restrict__ dv, const T* __restrict__ x,
                                                         TY* __restrict__ y, const TY* __restrict__ b, int64_t n,
                                                         int rows, int lcap, int mode) {
  extern __shared__ unsigned char smem_raw[];
  T* pl = reinterpret_cast<T*>(smem_raw);
  T* pu = pl + lcap;
  const int64_t r0 = (int64_t)blockIdx.x * rows;
  const int64_t r1 = min(r0 + (int64_t)rows, n);
  const int32_t l0 = lrp[r0], l1 = lrp[r1], u0 = urp[r0], u1 = urp[r1];
  for (int k = threadIdx.x; k < l1 - l0; k += blockDim.x) pl[k] = lv[l0 + k] * x[lci[l0 + k]];
  for (int k = threadIdx.x; k < u1 - u0; k += blockDim.x) pu[k] = uv[u0 + k] * x[uci[u0 + k]];
  __syncthreads();
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    T s = dv[r] * x[r];
    for (int k = lrp[r] - l0; k < lrp[r + 1] - l0; ++k) s += pl[k];
    for (int k = urp[r] - u0; k < urp[r + 1] - u0; ++k) s += pu[k];
    y[r] = mode ? (TY)((T)b[r] - s) : (TY)s;
  }
}

// forward sweep of one color: z_k = (r_k - sum_L a z) / d_k
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fwd_color(const int32_t* __restrict__ lrp, const int32_t* __restrict__ lci,
                                                        const T* __restrict__ lv, const T* __restrict__ invd,
                                                        const T* __restrict__ r, T* __restrict__ z, int64_t c0,
                                                        int64_t c1, int rows) {
  extern __shared__ unsigned char smem_raw[];
  T* pl = reinterpret_cast<T*>(smem_raw);
  const int64_t r0 = c0 + (int64_t)blockIdx.x * rows;
  const int64_t r1 = min(r0 + (int64_t)rows, c1);
  const int32_t l0 = lrp[r0], l1 = lrp[r1];
  for (int k = threadIdx.x; k < l1 - l0; k += blockDim.x) pl[k] = lv[l0 + k] * z[lci[l0 + k]];
  __syncthreads();
  for (int64_t k = r0 + threadIdx.x; k < r1; k += blockDim.x) {
    T s = T(0);
    for (int q = lrp[k] - l0; q < lrp[k + 1] - l0; ++q) s += pl[q];
    z[k] = (r[k] - s) * invd[k];
  }
}

// backward sweep of one color, in place: z_k = y_k - (sum_U a z) / d_k
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_color(const int32_t* __restrict__ urp, const int32_t* __restrict__ uci,
                                                        const T* __restrict__ uv, const T* __restrict__ invd,
                                                        T* __restrict__ z, int64_t c0, int64_t c1, int rows) {
  extern __shared__ unsigned char smem_raw[];
  T* pu = reinterpret_cast<T*>(smem_raw);
  const int64_t r0 = c0 + (int64_t)blockIdx.x * rows;
  const int64_t r1 = min(r0 + (int64_t)rows, c1);
  const int32_t u0 = urp[r0], u1 = urp[r1];
  for (int k = threadIdx.x; k < u1 - u0; k += blockDim.x) pu[k] = uv[u0 + k] * z[uci[u0 + k]];
  __syncthreads();
  for (int64_t k = r0 + threadIdx.x; k < r1; k += blockDim.x) {
    T s = T(0);
    for (int q = urp[k] - u0; q < urp[k + 1] - u0; ++q) s += pu[q];
    z[k] = z[k] - s * invd[k];
  }
}

// "Vector" variants: G lanes cooperate on a row, no LDS staging and no barrier.
template <typename T, int G>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <typename T, typename TY, int G>
__global__ __launch_bounds__(kThreads) void k_split_spmv_vec(const int32_t* __restrict__ lrp, const int32_t* __restrict__ lci,
                                                             const T* __restrict__ lv, const int32_t* __restrict__ urp,
                                                             const int32_t* __restrict__ uci, const T* __restrict__ uv,
                                                             const T* __restrict__ dv, const T* __restrict__ x,
                                                             TY* __restrict__ y, const TY* __restrict__ b, int64_t n,
                                                             int mode) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x % G;
  if (r >= n) return;  // whole groups exit together
  T s = T(0);
  for (int32_t q = lrp[r] + lane; q < lrp[r + 1]; q += G) s += lv[q] * x[lci[q]];
  for (int32_t q = urp[r] + lane; q < urp[r + 1]; q += G) s += uv[q] * x[uci[q]];
  s = group_sum<T, G>(s);
  if (lane == 0) {
    s += dv[r] * x[r];
    y[r] = mode ? (TY)((T)b[r] - s) : (TY)s;
  }
}

template <typename T, int G>
__global__ __launch_bounds__(kThreads) void k_fwd_color_vec(const int32_t* __restrict__ lrp, const int32_t* __restrict__ lci,
                                                            const T* __restrict__ lv, const T* __restrict__ invd,
                                                            const T* __restrict__ r, T* __restrict__ z, int64_t c0,
                                                            int64_t c1) {
  const int64_t k = c0 + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x % G;
  if (k >= c1) return;
  T s = T(0);
  for (int32_t q = lrp[k] + lane; q < lrp[k + 1]; q += G) s += lv[q] * z[lci[q]];
  s = group_sum<T, G>(s);
  if (lane == 0) z[k] = (r[k] - s) * invd[k];
}

template <typename T, int G>
__global__ __launch_bounds__(kThreads) void k_bwd_color_vec(const int32_t* __restrict__ urp, const int32_t* __restrict__ uci,
                                                            const T* __restrict__ uv, const T* __restrict__ invd,
                                                            T* __restrict__ z, int64_t c0, int64_t c1) {
  const int64_t k = c0 + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x % G;
  if (k >= c1) return;
  T s = T(0);
  for (int32_t q = urp[k] + lane; q < urp[k + 1]; q += G) s += uv[q] * z[uci[q]];
  s = group_sum<T, G>(s);
  if (lane == 0) z[k] = z[k] - s * invd[k];
}

// ------------------------------------------------------------ SELL-64 kernels
// Each part (L or U) of the SELL copy stores, per chunk, int32 columns and
// values slot-major: entry q of the chunk's 64 rows at off + q * 64 + lane.

// One wavefront per chunk, one lane per row: every load of column indices and
// values is a fully coalesced 64-lane access, no reductions or LDS.
// chunk widths = longest L / U part of the chunk's rows
__global__ void k_sell_widths(SellChunk* __restrict__ ch, int64_t nch, const int32_t* __restrict__ lrp,
                              const int32_t* __restrict__ urp) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  int lw = 0, uw = 0;
  for (int r = ch[c].row0; r < ch[c].row0 + ch[c].nrows; ++r) {
    lw = max(lw, lrp[r + 1] - lrp[r]);
    uw = max(uw, urp[r + 1] - urp[r]);
  }
  ch[c].lw = lw;
  ch[c].uw = uw;
}

// column / source tables; padding entries carry no source (value 0).  Their
// column must hold a value that is already final when the sweep reads it, so
// that 0 * x stays 0: row 0 (the first color, written first by the forward
// sweep) for L, the row itself (holding the forward result) for U.
__global__ __launch_bounds__(kThreads) void k_sell_fill(const SellChunk* __restrict__ ch, int64_t nch,
                                                        const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                        const int32_t* __restrict__ src, int upper,
                                                        int32_t* __restrict__ col, int32_t* __restrict__ ssrc) {
  const int64_t c = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= nch) return;
  const SellChunk k = ch[c];
  const int w = upper ? k.uw : k.lw;
  const int64_t off = upper ? k.uoff : k.loff;
  const bool act = lane < k.nrows;
  const int row = k.row0 + (act ? lane : 0);
  const int len = act ? rp[row + 1] - rp[row] : 0;
  for (int q = 0; q < w; ++q) {
    const int64_t at = off + (int64_t)q * 64 + lane;
    if (q < len) {
      col[at] = ci[rp[row] + q];
      ssrc[at] = src[rp[row] + q];
    } else {
      col[at] = upper ? row : 0;
      ssrc[at] = -1;
    }
  }
}


template <typename T>
__global__ void k_take_sell(const int32_t* __restrict__ src, const double* __restrict__ v, int64_t n,
                            T* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = src[q] >= 0 ? (T)v[src[q]] : T(0);
}

// y = (D + L + U) x  (mode 0)  or  y = b - (D + L + U) x  (mode 1); the matrix
// entries stream through once (non-temporal loads), the gathered x stays cached
template <typename T, typename TY>
__global__ __launch_bounds__(kThreads) void k_sell_spmv(const SellChunk* __restrict__ ch, int64_t nch,
                                                        const int32_t* __restrict__ lcs, const T* __restrict__ lval,
                                                        const int32_t* __restrict__ ucs, const T* __restrict__ uval,
                                                        const T* __restrict__ dv, const T* __restrict__ x,
                                                        TY* __restrict__ y, const TY* __restrict__ b, int mode) {
  const int64_t c = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= nch) return;
  const SellChunk k = ch[c];
  const bool act = lane < k.nrows;
  const int row = k.row0 + (act ? lane : 0);
  T s = dv[row] * x[row];
#pragma unroll NALU_SPARSE_UNROLL
  for (int q = 0; q < k.lw; ++q) {
    s += __builtin_nontemporal_load(&lval[k.loff + lane + q * 64]) *
         x[__builtin_nontemporal_load(&lcs[k.loff + lane + q * 64])];
  }
#pragma unroll NALU_SPARSE_UNROLL
  for (int q = 0; q < k.uw; ++q) {
    s += __builtin_nontemporal_load(&uval[k.uoff + lane + q * 64]) *
         x[__builtin_nontemporal_load(&ucs[k.uoff + lane + q * 64])];
  }
  if (act) y[row] = mode ? (TY)((T)b[row] - s) : (TY)s;
}

// forward sweep over the chunks [c0, c1) of one color: z = (r - L z) / d.
// Two lanes per row (128 threads per 64-row chunk, even / odd entries, the
// halves combined through LDS) double the loads in flight per chunk; the matrix
// entries are read once per sweep, so they are loaded non-temporally and the
// gathered z entries keep the cache.  Each XCD sweeps a contiguous range of the
// color's chunks, so neighbouring rows gather z through the same L2.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sell_fwd(const SellChunk* __restrict__ ch, int64_t c0, int64_t c1,
                                                       const int32_t* __restrict__ lcs, const T* __restrict__ lval,
                                                       const T* __restrict__ invd, const T* __restrict__ r,
                                                       T* __restrict__ z) {
  __shared__ T part[kThreads / 2];
  const int64_t nb = (c1 - c0 + 1) / 2;  // blocks of two chunks, XCD-contiguous ranges
  const int64_t pos = xcd_block(blockIdx.x, nb);
  const int64_t c = pos < 0 ? c1 : c0 + pos * (kThreads / 128) + (threadIdx.x >> 7);
  const int lane = threadIdx.x & 63;
  const int half = (threadIdx.x >> 6) & 1;
  const int slot = (threadIdx.x >> 7) * 64 + lane;
  const bool ok = c < c1;
  SellChunk k{};
  if (ok) k = ch[c];
  const bool act = ok && lane < k.nrows;
  T s = T(0);
  if (act) {
#pragma unroll NALU_SPARSE_UNROLL
    for (int q = half; q < k.lw; q += 2)
      s += __builtin_nontemporal_load(&lval[k.loff + lane + q * 64]) *
           z[__builtin_nontemporal_load(&lcs[k.loff + lane + q * 64])];
  }
  if (half == 1) part[slot] = s;
  __syncthreads();
  if (!act || half == 1) return;
  const int row = k.row0 + lane;
  s += part[slot];
  z[row] = (r[row] - s) * invd[row];
}

// backward sweep over the chunks of one color, in place: z = y - (U z) / d
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sell_bwd(const SellChunk* __restrict__ ch, int64_t c0, int64_t c1,
                                                       const int32_t* __restrict__ ucs, const T* __restrict__ uval,
                                                       const T* __restrict__ invd, T* __restrict__ z) {
  __shared__ T part[kThreads / 2];
  const int64_t nb = (c1 - c0 + 1) / 2;  // blocks of two chunks, XCD-contiguous ranges
  const int64_t pos = xcd_block(blockIdx.x, nb);
  const int64_t c = pos < 0 ? c1 : c0 + pos * (kThreads / 128) + (threadIdx.x >> 7);
  const int lane = threadIdx.x & 63;
  const int half = (threadIdx.x >> 6) & 1;
  const int slot = (threadIdx.x >> 7) * 64 + lane;
  const bool ok = c < c1;
  SellChunk k{};
  if (ok) k = ch[c];
  const bool act = ok && lane < k.nrows;
  T s = T(0);
  if (act) {
#pragma unroll NALU_SPARSE_UNROLL
    for (int q = half; q < k.uw; q += 2)
      s += __builtin_nontemporal_load(&uval[k.uoff + lane + q * 64]) *
           z[__builtin_nontemporal_load(&ucs[k.uoff + lane + q * 64])];
  }
  if (half == 1) part[slot] = s;
  __syncthreads();
  if (!act || half == 1) return;
  const int row = k.row0 + lane;
  s += part[slot];
  z[row] = z[row] - s * invd[row];
}

// q = A p (SELL copy); per-block partials of p.q
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sell_spmv_dot(const SellChunk* __restrict__ ch, int64_t nch,
                                                            const int32_t* __restrict__ lcs, const T* __restrict__ lval,
                                                            const int32_t* __restrict__ ucs, const T* __restrict__ uval,
                                                            const T* __restrict__ dv, const T* __restrict__ x,
                                                            T* __restrict__ y, double* __restrict__ part,
                                                            const CgState* __restrict__ st) {
  if (st->done) return;
  const int64_t c = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double d = 0.0;
  if (c < nch) {
    const SellChunk k = ch[c];
    const bool act = lane < k.nrows;
    const int row = k.row0 + (act ? lane : 0);
    T s = dv[row] * x[row];
#pragma unroll NALU_SPARSE_UNROLL
    for (int q = 0; q < k.lw; ++q) {
      s += __builtin_nontemporal_load(&lval[k.loff + lane + q * 64]) *
           x[__builtin_nontemporal_load(&lcs[k.loff + lane + q * 64])];
    }
#pragma unroll NALU_SPARSE_UNROLL
    for (int q = 0; q < k.uw; ++q) {
      s += __builtin_nontemporal_load(&uval[k.uoff + lane + q * 64]) *
           x[__builtin_nontemporal_load(&ucs[k.uoff + lane + q * 64])];
    }
    if (act) {
      y[row] = s;
      d = (double)s * (double)x[row];
    }
  }
  d = block_sum(d);
  if (threadIdx.x == 0) part[blockIdx.x] = d;
}

// Sweeps of the fused CG: as k_sell_fwd / k_sell_bwd, plus the partial sums
// r.z of the rows they finalize (the forward sweep of the last color, every
// backward sweep) in part[off + block].  All return at once after convergence.
template <typename T, bool UPPER>
__global__ __launch_bounds__(kThreads) void k_sell_sweep_dot(const SellChunk* __restrict__ ch, int64_t c0,
                                                             int64_t c1, const int32_t* __restrict__ cs, const T* __restrict__ val,
                                                             const T* __restrict__ invd, const T* __restrict__ r,
                                                             T* __restrict__ z, int dot, double* __restrict__ part,
                                                             int off, const CgState* __restrict__ st) {
  __shared__ T hpart[kThreads / 2];
  if (st->done) return;
  const int64_t nb = (c1 - c0 + 1) / 2;
  const int64_t pos = xcd_block(blockIdx.x, nb);
  const int64_t c = pos < 0 ? c1 : c0 + pos * (kThreads / 128) + (threadIdx.x >> 7);
  const int lane = threadIdx.x & 63;
  const int half = (threadIdx.x >> 6) & 1;
  const int hslot = (threadIdx.x >> 7) * 64 + lane;
  const bool ok = c < c1;
  SellChunk k{};
  if (ok) k = ch[c];
  const bool act = ok && lane < k.nrows;
  T s = T(0);
  if (act) {
    const int w = UPPER ? k.uw : k.lw;
    const int64_t o = UPPER ? k.uoff : k.loff;
#pragma unroll NALU_SPARSE_UNROLL
    for (int q = half; q < w; q += 2)
      s += __builtin_nontemporal_load(&val[o + lane + q * 64]) * z[__builtin_nontemporal_load(&cs[o + lane + q * 64])];
  }
  if (half == 1) hpart[hslot] = s;
  __syncthreads();
  double d = 0.0;
  if (act && half == 0) {
    const int row = k.row0 + lane;
    s += hpart[hslot];
    const T zr = UPPER ? z[row] - s * invd[row] : (r[row] - s) * invd[row];
    z[row] = zr;
    if (dot) d = (double)r[row] * (double)zr;
  }
  if (!dot) return;  // grid-uniform
  d = block_sum(d);
  if (threadIdx.x == 0) part[off + blockIdx.x] = d;
}

// Batched forms for NR right-hand sides interleaved [n][NR] (one matrix read
// and one contiguous gather of NR values per entry): SpMV and the two sweeps,
// with the matrix layouts of the kernels above.
template <typename T, typename TY, int NR>
__global__ __launch_bounds__(kThreads) void k_sell_spmv_m(const SellChunk* __restrict__ ch, int64_t nch,
                                                          const int32_t* __restrict__ lcs, const T* __restrict__ lval,
                                                          const int32_t* __restrict__ ucs, const T* __restrict__ uval,
                                                          const T* __restrict__ dv, const T* __restrict__ x,
                                                          TY* __restrict__ y, const TY* __restrict__ b, int mode,
                                                          int64_t n) {
  const int64_t c = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= nch) return;
  const SellChunk k = ch[c];
  const bool act = lane < k.nrows;
  const int row = k.row0 + (act ? lane : 0);
  T s[NR];
  const T d = dv[row];
#pragma unroll
  for (int r = 0; r < NR; ++r) s[r] = d * x[row * NR + r];
#pragma unroll 2
  for (int q = 0; q < k.lw; ++q) {
    const int64_t at = k.loff + (int64_t)q * 64 + lane;
    const T a = __builtin_nontemporal_load(&lval[at]);
    const int64_t j = __builtin_nontemporal_load(&lcs[at]);
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] += a * x[j * NR + r];
  }
#pragma unroll 2
  for (int q = 0; q < k.uw; ++q) {
    const int64_t at = k.uoff + (int64_t)q * 64 + lane;
    const T a = __builtin_nontemporal_load(&uval[at]);
    const int64_t j = __builtin_nontemporal_load(&ucs[at]);
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] += a * x[j * NR + r];
  }
  if (!act) return;
#pragma unroll
  for (int r = 0; r < NR; ++r) y[row * NR + r] = mode ? (TY)((T)b[row * NR + r] - s[r]) : (TY)s[r];
}

// UPPER = false: forward sweep z = (r - L z) / d; true: backward z -= (U z) / d
template <typename T, int NR, bool UPPER>
__global__ __launch_bounds__(kThreads) void k_sell_sweep_m(const SellChunk* __restrict__ ch, int64_t c0, int64_t c1,
                                                           const int32_t* __restrict__ cs, const T* __restrict__ val,
                                                           const T* __restrict__ invd, const T* __restrict__ rr,
                                                           T* __restrict__ z, int64_t n) {
  __shared__ T part[NR][kThreads / 2];
  const int64_t nb = (c1 - c0 + 1) / 2;  // blocks of two chunks, XCD-contiguous ranges
  const int64_t pos = xcd_block(blockIdx.x, nb);
  const int64_t c = pos < 0 ? c1 : c0 + pos * (kThreads / 128) + (threadIdx.x >> 7);
  const int lane = threadIdx.x & 63;
  const int half = (threadIdx.x >> 6) & 1;
  const int slot = (threadIdx.x >> 7) * 64 + lane;
  const bool ok = c < c1;
  SellChunk k{};
  if (ok) k = ch[c];
  const bool act = ok && lane < k.nrows;
  T s[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) s[r] = T(0);
  if (act) {
    const int w = UPPER ? k.uw : k.lw;
    const int64_t off = UPPER ? k.uoff : k.loff;
#pragma unroll 2
    for (int q = half; q < w; q += 2) {
      const int64_t at = off + (int64_t)q * 64 + lane;
      const T a = __builtin_nontemporal_load(&val[at]);
      const int64_t j = __builtin_nontemporal_load(&cs[at]);
#pragma unroll
      for (int r = 0; r < NR; ++r) s[r] += a * z[j * NR + r];
    }
  }
  if (half == 1)
#pragma unroll
    for (int r = 0; r < NR; ++r) part[r][slot] = s[r];
  __syncthreads();
  if (!act || half == 1) return;
  const int row = k.row0 + lane;
  const T di = invd[row];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const T t = s[r] + part[r][slot];
    if (UPPER) z[row * NR + r] = z[row * NR + r] - t * di;
    else z[row * NR + r] = (rr[row * NR + r] - t) * di;
  }
}

// Sparse kernel family used by the solvers: 64 = SELL-64 (default), 0 =
// CSR-stream (LDS staged), 4 / 8 = G lanes per CSR row.
// NALU_SPARSE_KERNEL=sell|stream|vector4|vector8 overrides.
int sparse_kernel_choice() {
  static int choice = -1;
  if (choice < 0) {
    const char* e = std::getenv("NALU_SPARSE_KERNEL");
    choice = 64;
    if (e && std::string(e) == "stream") choice = 0;
    else if (e && std::string(e) == "vector4") choice = 4;
    else if (e && std::string(e) == "vector8") choice = 8;
  }
  return choice;
}

int stream_rows(int maxlen, size_t elem_bytes, int parts) {
  int R = 256;
  while (R > 8 && (size_t)R * std::max(1, maxlen) * elem_bytes * parts > 56 * 1024) R >>= 1;
  return R;
}

void build_perm(nalu_linsys* ls) {
  if (ls->perm_ready) return;
  hipStream_t s = ls->stream;
  if (ls->num_colors == 0) build_coloring(ls, s);
  const int64_t n = ls->nrows;
  const int nc = ls->num_colors;
  ls->perm.alloc(n);
  ls->iperm.alloc(n);
  ls->cpd.alloc(nc + 1);
  std::vector<int32_t> cph(nc + 1);
  for (int c = 0; c <= nc; ++c) cph[c] = (int32_t)ls->color_ptr_host[c];
  NALU_HIP_CHECK(hipMemcpy(ls->cpd.p, cph.data(), sizeof(int32_t) * (nc + 1), hipMemcpyHostToDevice));
  NALU_HIP_CHECK(hipMemcpyAsync(ls->perm.p, ls->color_rows.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_iperm, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->perm.p, n, ls->iperm.p);
  DevBuf<int32_t> lc(n + 1), uc(n + 1);
  NALU_HIP_CHECK(hipMemsetAsync(lc.p + n, 0, sizeof(int32_t), s));
  NALU_HIP_CHECK(hipMemsetAsync(uc.p + n, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(k_split, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->row_ptr.p, ls->col_idx.p, ls->perm.p,
                     ls->iperm.p, ls->cpd.p, nc, ls->diag_pos.p, n, lc.p, uc.p, (const int32_t*)nullptr,
                     (const int32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr,
                     (int32_t*)nullptr, (int32_t*)nullptr, 0);
  // max row lengths (LDS sizing) and exclusive scans
  int32_t mx[2] = {0, 0};
  for (int part = 0; part < 2; ++part) {
    DevBuf<int32_t>& cnt = part ? uc : lc;
    DevBuf<int32_t> out(1);
    size_t tmp = 0;
    NALU_HIP_CHECK(hipcub::DeviceReduce::Max(nullptr, tmp, cnt.p, out.p, (int)n, s));
    DevBuf<char> t(tmp ? tmp : 1);
    NALU_HIP_CHECK(hipcub::DeviceReduce::Max(t.p, tmp, cnt.p, out.p, (int)n, s));
    NALU_HIP_CHECK(hipMemcpyAsync(&mx[part], out.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    DevBuf<int32_t>& rp = part ? ls->urp : ls->lrp;
    rp.alloc(n + 1);
    tmp = 0;
    NALU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt.p, rp.p, (int)(n + 1), s));
    DevBuf<char> t2(tmp ? tmp : 1);
    NALU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(t2.p, tmp, cnt.p, rp.p, (int)(n + 1), s));
  }
  int32_t nl = 0, nu = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&nl, ls->lrp.p + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipMemcpyAsync(&nu, ls->urp.p + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  NALU_REQUIRE((int64_t)nl + nu + n == ls->nnz, NALU_ERR_NUMERIC, "split: L + U + D does not cover the matrix");
  ls->max_l_len = mx[0];
  ls->max_u_len = mx[1];
  ls->lci.alloc(std::max(nl, 1));
  ls->lsrc.alloc(std::max(nl, 1));
  ls->uci.alloc(std::max(nu, 1));
  ls->usrc.alloc(std::max(nu, 1));
  ls->dsrc.alloc(n);
  hipLaunchKernelGGL(k_split, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->row_ptr.p, ls->col_idx.p, ls->perm.p,
                     ls->iperm.p, ls->cpd.p, nc, ls->diag_pos.p, n, (int32_t*)nullptr, (int32_t*)nullptr,
                     (const int32_t*)ls->lrp.p, (const int32_t*)ls->urp.p, ls->lci.p, ls->lsrc.p, ls->uci.p,
                     ls->usrc.p, ls->dsrc.p, 1);
  NALU_HIP_CHECK(hipGetLastError());
  ls->perm_ready = true;
}

// SELL-64 tables of the split system (structure only; values are refreshed
// with the rest of the solver copy).
void build_sell(nalu_linsys* ls) {
  if (ls->sell_ready) return;
  hipStream_t s = ls->stream;
  const int nc = ls->num_colors;
  std::vector<SellChunk> ch;
  ls->sell_color_ptr.assign(nc + 1, 0);
  for (int c = 0; c < nc; ++c) {
    ls->sell_color_ptr[c] = (int64_t)ch.size();
    for (int64_t r = ls->color_ptr_host[c]; r < ls->color_ptr_host[c + 1]; r += 64)
      ch.push_back({(int32_t)r, (int32_t)std::min<int64_t>(64, ls->color_ptr_host[c + 1] - r), 0, 0, 0, 0});
  }
  ls->sell_color_ptr[nc] = (int64_t)ch.size();
  const int64_t nch = (int64_t)ch.size();
  ls->sell_nchunks = nch;
  {
    // schedule: S fine slabs; slab k holds the k-th S-th of every color's chunks
    const int64_t S = std::max<int64_t>(1, nch / (4 * std::max(1, nc)));
    std::vector<int32_t> sched;
    sched.reserve(nch);
    for (int64_t k = 0; k < S; ++k)
      for (int c = 0; c < nc; ++c) {
        const int64_t a = ls->sell_color_ptr[c], len = ls->sell_color_ptr[c + 1] - a;
        for (int64_t q = a + k * len / S; q < a + (k + 1) * len / S; ++q) sched.push_back((int32_t)q);
      }
    NALU_REQUIRE((int64_t)sched.size() == nch, NALU_ERR_STATE, "SELL schedule does not cover the chunks");
    ls->sell_sched.alloc(std::max<int64_t>(1, nch));
    NALU_HIP_CHECK(hipMemcpy(ls->sell_sched.p, sched.data(), sizeof(int32_t) * nch, hipMemcpyHostToDevice));
  }
  ls->sell_chunks.alloc(std::max<int64_t>(1, nch) * sizeof(SellChunk));
  SellChunk* dch = reinterpret_cast<SellChunk*>(ls->sell_chunks.p);
  NALU_HIP_CHECK(hipMemcpy(dch, ch.data(), nch * sizeof(SellChunk), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sell_widths, dim3(nblocks(nch)), dim3(kThreads), 0, s, dch, nch, ls->lrp.p, ls->urp.p);
  NALU_HIP_CHECK(hipMemcpyAsync(ch.data(), dch, nch * sizeof(SellChunk), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  int64_t lo = 0, uo = 0;
  for (auto& k : ch) {
    k.loff = lo;
    k.uoff = uo;
    lo += 64 * (int64_t)k.lw;
    uo += 64 * (int64_t)k.uw;
  }
  NALU_HIP_CHECK(hipMemcpy(dch, ch.data(), nch * sizeof(SellChunk), hipMemcpyHostToDevice));
  ls->sl_col.alloc(std::max<int64_t>(1, lo));
  ls->sl_src.alloc(std::max<int64_t>(1, lo));
  ls->su_col.alloc(std::max<int64_t>(1, uo));
  ls->su_src.alloc(std::max<int64_t>(1, uo));
  const unsigned nbw = (unsigned)std::max<int64_t>(1, (nch + 3) / 4);
  hipLaunchKernelGGL(k_sell_fill, dim3(nbw), dim3(kThreads), 0, s, dch, nch, ls->lrp.p, ls->lci.p, ls->lsrc.p, 0,
                     ls->sl_col.p, ls->sl_src.p);
  hipLaunchKernelGGL(k_sell_fill, dim3(nbw), dim3(kThreads), 0, s, dch, nch, ls->urp.p, ls->uci.p, ls->usrc.p, 1,
                     ls->su_col.p, ls->su_src.p);
  NALU_HIP_CHECK(hipGetLastError());
  ls->sell_ready = true;
}

template <typename T>
struct Split {
  const int32_t *lrp, *lci, *urp, *uci;
  const T *lv, *uv, *dv, *invd;
};

template <typename T>
Split<T> split_of(nalu_linsys* ls) {
  if constexpr (std::is_same<T, double>::value)
    return {ls->lrp.p, ls->lci.p, ls->urp.p, ls->uci.p, ls->lval.p, ls->uval.p, ls->dval.p, ls->pinvd.p};
  else
    return {ls->lrp.p, ls->lci.p, ls->urp.p, ls->uci.p, ls->lval32.p, ls->uval32.p, ls->dval32.p, ls->pinvd32.p};
}

const int32_t* sell_cols(const nalu_linsys* ls, bool upper) { return upper ? ls->su_col.p : ls->sl_col.p; }

template <typename T>
const T* sell_vals(const nalu_linsys* ls, bool upper) {
  if constexpr (std::is_same<T, double>::value) return upper ? ls->su_val.p : ls->sl_val.p;
  else return upper ? ls->su_val32.p : ls->sl_val32.p;
}

template <typename T, typename TY>
void pspmv(nalu_linsys* ls, const T* x, TY* y, const TY* b, int mode) {
  const Split<T> A = split_of<T>(ls);
  const int kc = sparse_kernel_choice();
  if (kc == 64) {
    const int64_t nch = ls->sell_nchunks;
    hipLaunchKernelGGL((k_sell_spmv<T, TY>), dim3((unsigned)std::max<int64_t>(1, (nch + 3) / 4)), dim3(kThreads), 0,
                       ls->stream, reinterpret_cast<const SellChunk*>(ls->sell_chunks.p), nch, sell_cols(ls, false),
                       sell_vals<T>(ls, false), sell_cols(ls, true), sell_vals<T>(ls, true), A.dv, x, y, b, mode);
    return;
  }
  if (kc == 8 || kc == 4) {
    const int64_t n = ls->nrows;
    if (kc == 8)
      hipLaunchKernelGGL((k_split_spmv_vec<T, TY, 8>), dim3(nblocks(n * 8)), dim3(kThreads), 0, ls->stream, A.lrp,
                         A.lci, A.lv, A.urp, A.uci, A.uv, A.dv, x, y, b, n, mode);
    else
      hipLaunchKernelGGL((k_split_spmv_vec<T, TY, 4>), dim3(nblocks(n * 4)), dim3(kThreads), 0, ls->stream, A.lrp,
                         A.lci, A.lv, A.urp, A.uci, A.uv, A.dv, x, y, b, n, mode);
    return;
  }
  const int R = stream_rows(std::max(ls->max_l_len, ls->max_u_len), sizeof(T), 2);
  const int lcap = R * std::max(1, ls->max_l_len);
  const size_t lds = (size_t)(lcap + R * std::max(1, ls->max_u_len)) * sizeof(T);
  hipLaunchKernelGGL((k_split_spmv<T, TY>), dim3(nblocks(ls->nrows, R)), dim3(kThreads), lds, ls->stream, A.lrp,
                     A.lci, A.lv, A.urp, A.uci, A.uv, A.dv, x, y, b, ls->nrows, R, lcap, mode);
}

template <typename T>
void psgs(nalu_linsys* ls, const T* r, T* z) {
  const Split<T> A = split_of<T>(ls);
  hipStream_t s = ls->stream;
  const int Rl = stream_rows(ls->max_l_len, sizeof(T), 1), Ru = stream_rows(ls->max_u_len, sizeof(T), 1);
  const size_t ldl = (size_t)Rl * std::max(1, ls->max_l_len) * sizeof(T);
  const size_t ldu = (size_t)Ru * std::max(1, ls->max_u_len) * sizeof(T);
  const auto& cp = ls->color_ptr_host;
  const int kc = sparse_kernel_choice();
  if (kc == 64) {
    const SellChunk* ch = reinterpret_cast<const SellChunk*>(ls->sell_chunks.p);
    const auto& ccp = ls->sell_color_ptr;
    for (int c = 0; c < ls->num_colors; ++c) {
      const int64_t c0 = ccp[c], c1 = ccp[c + 1];
      if (c1 > c0)
        hipLaunchKernelGGL(k_sell_fwd<T>, dim3((unsigned)xcd_grid((c1 - c0 + 1) / 2)), dim3(kThreads), 0, s, ch,
                           c0, c1, sell_cols(ls, false), sell_vals<T>(ls, false), A.invd, r, z);
    }
    for (int c = ls->num_colors - 2; c >= 0; --c) {  // the last color has no U entries
      const int64_t c0 = ccp[c], c1 = ccp[c + 1];
      if (c1 > c0)
        hipLaunchKernelGGL(k_sell_bwd<T>, dim3((unsigned)xcd_grid((c1 - c0 + 1) / 2)), dim3(kThreads), 0, s, ch,
                           c0, c1, sell_cols(ls, true), sell_vals<T>(ls, true), A.invd, z);
    }
    return;
  }
  if (kc == 8 || kc == 4) {
    for (int c = 0; c < ls->num_colors; ++c) {
      const int64_t c0 = cp[c], c1 = cp[c + 1];
      if (c1 <= c0) continue;
      if (kc == 8)
        hipLaunchKernelGGL((k_fwd_color_vec<T, 8>), dim3(nblocks((c1 - c0) * 8)), dim3(kThreads), 0, s, A.lrp, A.lci,
                           A.lv, A.invd, r, z, c0, c1);
      else
        hipLaunchKernelGGL((k_fwd_color_vec<T, 4>), dim3(nblocks((c1 - c0) * 4)), dim3(kThreads), 0, s, A.lrp, A.lci,
                           A.lv, A.invd, r, z, c0, c1);
    }
    for (int c = ls->num_colors - 2; c >= 0; --c) {
      const int64_t c0 = cp[c], c1 = cp[c + 1];
      if (c1 <= c0) continue;
      if (kc == 8)
        hipLaunchKernelGGL((k_bwd_color_vec<T, 8>), dim3(nblocks((c1 - c0) * 8)), dim3(kThreads), 0, s, A.urp, A.uci,
                           A.uv, A.invd, z, c0, c1);
      else
        hipLaunchKernelGGL((k_bwd_color_vec<T, 4>), dim3(nblocks((c1 - c0) * 4)), dim3(kThreads), 0, s, A.urp, A.uci,
                           A.uv, A.invd, z, c0, c1);
    }
    return;
  }
  for (int c = 0; c < ls->num_colors; ++c) {
    const int64_t c0 = cp[c], c1 = cp[c + 1];
    if (c1 > c0)
      hipLaunchKernelGGL((k_fwd_color<T>), dim3(nblocks(c1 - c0, Rl)), dim3(kThreads), ldl, s, A.lrp, A.lci, A.lv,
                         A.invd, r, z, c0, c1, Rl);
  }
  for (int c = ls->num_colors - 2; c >= 0; --c) {  // the last color has no U entries: z = y there
    const int64_t c0 = cp[c], c1 = cp[c + 1];
    if (c1 > c0)
      hipLaunchKernelGGL((k_bwd_color<T>), dim3(nblocks(c1 - c0, Ru)), dim3(kThreads), ldu, s, A.urp, A.uci, A.uv,
                         A.invd, z, c0, c1, Ru);
  }
}

// Operators on nrhs stacked vectors [nrhs][n] (solver order): the batched
// SELL kernels for up to 4 right-hand sides, else one slice at a time.
template <typename T, typename TY, int NR>
void spmv_nr(nalu_linsys* ls, const T* x, TY* y, const TY* b, int mode) {
  const Split<T> A = split_of<T>(ls);
  const int64_t nch = ls->sell_nchunks;
  hipLaunchKernelGGL((k_sell_spmv_m<T, TY, NR>), dim3((unsigned)std::max<int64_t>(1, (nch + 3) / 4)), dim3(kThreads),
                     0, ls->stream, reinterpret_cast<const SellChunk*>(ls->sell_chunks.p), nch, sell_cols(ls, false),
                     sell_vals<T>(ls, false), sell_cols(ls, true), sell_vals<T>(ls, true), A.dv, x, y, b, mode,
                     (int64_t)ls->nrows);
}

// The stacked vectors are interleaved [rows][nrhs] exactly when the batched
// kernels serve both operators (stacked_interleaved()), [nrhs][rows] otherwise.
bool stacked_interleaved(const nalu_linsys* ls, int nrhs) {
  return nrhs > 1 && nrhs <= 4 && sparse_kernel_choice() == 64 && ls->cfg.precond == NALU_PREC_SGS;
}

template <typename T, typename TY>
void spmv_m(nalu_linsys* ls, int nrhs, const T* x, TY* y, const TY* b, int mode) {
  const int64_t n = ls->nrows;
  if (nrhs > 1 && ls->stack_interleaved) {
    if (nrhs == 2) spmv_nr<T, TY, 2>(ls, x, y, b, mode);
    else if (nrhs == 3) spmv_nr<T, TY, 3>(ls, x, y, b, mode);
    else spmv_nr<T, TY, 4>(ls, x, y, b, mode);
    return;
  }
  for (int r = 0; r < nrhs; ++r) pspmv<T, TY>(ls, x + r * n, y + r * n, b ? b + r * n : b, mode);
}

template <typename T, int NR>
void sgs_nr(nalu_linsys* ls, const T* r, T* z) {
  const Split<T> A = split_of<T>(ls);
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  const SellChunk* ch = reinterpret_cast<const SellChunk*>(ls->sell_chunks.p);
  const auto& ccp = ls->sell_color_ptr;
  for (int c = 0; c < ls->num_colors; ++c) {
    const int64_t c0 = ccp[c], c1 = ccp[c + 1];
    if (c1 > c0)
      hipLaunchKernelGGL((k_sell_sweep_m<T, NR, false>), dim3((unsigned)xcd_grid((c1 - c0 + 1) / 2)), dim3(kThreads),
                         0, s, ch, c0, c1, sell_cols(ls, false), sell_vals<T>(ls, false), A.invd, r, z, n);
  }
  for (int c = ls->num_colors - 2; c >= 0; --c) {
    const int64_t c0 = ccp[c], c1 = ccp[c + 1];
    if (c1 > c0)
      hipLaunchKernelGGL((k_sell_sweep_m<T, NR, true>), dim3((unsigned)xcd_grid((c1 - c0 + 1) / 2)), dim3(kThreads),
                         0, s, ch, c0, c1, sell_cols(ls, true), sell_vals<T>(ls, true), A.invd, (const T*)nullptr, z, n);
  }
}

template <typename T>
void pprec(nalu_linsys* ls, const T* r, T* z);

template <typename T>
void prec_m(nalu_linsys* ls, int nrhs, const T* r, T* z) {
  const int64_t n = ls->nrows;
  if (nrhs > 1 && ls->stack_interleaved) {
    if (nrhs == 2) sgs_nr<T, 2>(ls, r, z);
    else if (nrhs == 3) sgs_nr<T, 3>(ls, r, z);
    else sgs_nr<T, 4>(ls, r, z);
    return;
  }
  for (int q = 0; q < nrhs; ++q) pprec<T>(ls, r + q * n, z + q * n);
}

bool mixed(const nalu_linsys* ls) { return ls->cfg.precision == NALU_PRECISION_MIXED; }

// preconditioner in the permuted space
template <typename T>
void pprec(nalu_linsys* ls, const T* r, T* z) {
  const int64_t n = ls->nrows;
  switch (ls->cfg.precond) {
    case NALU_PREC_JACOBI:
      hipLaunchKernelGGL(k_jacobi<T>, dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, split_of<T>(ls).invd, r, z, n);
      break;
    case NALU_PREC_LEARNED:
      if (learned_active(ls)) {
        if constexpr (std::is_same<T, float>::value) learned_apply_f32(ls, r, z);
        else learned_apply_f64(ls, r, z);
        break;
      }
      // no network attached, or fallen back: the base preconditioner
      if (ls->learned_base == NALU_PREC_MUELU) amg_apply<T>(ls, r, z);
      else psgs<T>(ls, r, z);
      break;
    case NALU_PREC_SGS:
      psgs<T>(ls, r, z);
      break;
    case NALU_PREC_MUELU:
      amg_apply<T>(ls, r, z);
      break;
    default:
      hipLaunchKernelGGL((k_copy<T, T>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, r, z, n);
  }
}

// ------------------------------------------------------------ Givens / Hessenberg (host, FP64)
struct Hessenberg {
  int m;
  std::vector<double> H, cs, sn, g;
  explicit Hessenberg(int mm) : m(mm), H((size_t)(mm + 1) * mm, 0.0), cs(mm), sn(mm), g(mm + 1) {}
  double& h(int i, int j) { return H[(size_t)i * m + j]; }
  void reset(double beta) {
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
  }
  // column k from hcol[0..k+1]; returns |g_{k+1}|
  double add_column(int k, const double* hcol) {
    for (int i = 0; i <= k + 1; ++i) h(i, k) = hcol[i];
    for (int j = 0; j < k; ++j) {
      const double a = h(j, k), c = h(j + 1, k);
      h(j, k) = cs[j] * a + sn[j] * c;
      h(j + 1, k) = -sn[j] * a + cs[j] * c;
    }
    const double a = h(k, k), c = h(k + 1, k);
    const double den = std::sqrt(a * a + c * c);
    cs[k] = den > 0 ? a / den : 1.0;
    sn[k] = den > 0 ? c / den : 0.0;
    h(k, k) = den;
    h(k + 1, k) = 0.0;
    g[k + 1] = -sn[k] * g[k];
    g[k] = cs[k] * g[k];
    return std::fabs(g[k + 1]);
  }
  void solve(int k, std::vector<double>& y) {
    y.assign(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
      double t = g[i];
      for (int j = i + 1; j < k; ++j) t -= h(i, j) * y[j];
      y[i] = t / h(i, i);
    }
  }
};

// Restarted (F)GMRES on A x = b in the permuted space (all vectors permuted).
// TV is the storage/arithmetic type of the Krylov basis, SpMV and
// preconditioner (double, or float in the mixed-precision inner solve); inner
// products, the Hessenberg matrix and the Givens least-squares are FP64.
// flexible: store Z_k = M^-1 v_k (FGMRES).  x (FP64) holds the initial guess.
template <typename TV>
int gmres_t(nalu_linsys* ls, const double* b, double* x, int m, int maxit, double tol, bool flexible,
            double* implicit_rel, int nrhs = 1) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows * nrhs;  // nrhs > 1: the block-diagonal stacked system
  Red red(ls);
  // basis columns padded to a multiple of 64 entries (16-byte aligned columns)
  const int64_t ldv = (n + 63) / 64 * 64;
  const size_t need = (size_t)ldv * ((m + 1) + (flexible ? m : 1) + 2);
  // basis, FP64 residual, Hessenberg column buffers: the handle's workspace
  TV* V = nullptr;
  double *r64 = nullptr, *hdev = nullptr, *hbuf = nullptr;
  auto layout = [&](unsigned char* base) {
    Carve c{base};
    V = c.take<TV>(need);
    r64 = c.take<double>(n);
    hdev = c.take<double>(m + 2);
    hbuf = c.take<double>(2 * (m + 2) + 2);
    return c.off;
  };
  layout(ensure_ws(ls->krylov_ws, layout(nullptr)));
  ls->krylov_host.alloc(sizeof(double) * (2 * (m + 2) + 2));
  double* hpin = static_cast<double*>(ls->krylov_host.p);
  TV* Z = V + (size_t)ldv * (m + 1);
  TV* tmp = Z + (size_t)ldv * (flexible ? m : 1);
  TV* xt = tmp + ldv;  // x in TV precision for the residual of the inner SpMV path
  // Gram-Schmidt launch shapes: chunks of R rows, at most kGsBlocks blocks.
  // The FP64 solve orthogonalizes with two classical passes (ICGS, as Belos is
  // configured in the reference); the FP32 inner solve of the mixed path does
  // one pass and reorthogonalizes only on severe cancellation, ||w'|| < ||w|| / 10
  // (a looser form of the "twice is enough" test of Parlett and Kahan: the inner
  // solve only computes a correction, the FP64 outer residual decides).
  constexpr int kGsBlocks = 2048;
  const bool always2 = std::is_same<TV, double>::value;
  const int R_dots = 2048;
  const unsigned nb_dots = (unsigned)std::min<int64_t>(kGsBlocks, (n + R_dots - 1) / R_dots);
  const size_t lds_dots = (size_t)R_dots * sizeof(TV) + (size_t)(m + 2) * 8;
  const unsigned nb_upd = red_blocks((n + 3) / 4);
  NALU_REQUIRE((size_t)(m + 2) * kGsBlocks <= ls->red.n, NALU_ERR_STATE, "gmres: reduction workspace too small");
  // hA: pass-1 coefficients (+ ||w||^2 at k+1), hB: pass-2 coefficients, nrm: ||w_new||^2
  double* hA = hbuf;
  double* hB = hbuf + (m + 2);
  double* nrm = hbuf + 2 * (m + 2);
  // learned preconditioner: FGMRES falls back to SGS for the rest of the solve
  // when a cycle stagnates or the network output is not finite
  const bool watch = flexible && ls->cfg.precond == NALU_PREC_LEARNED;
  Hessenberg H(m);
  std::vector<double> hcol(m + 2), y;
  (void)xt;

  int iters = 0;
  double r0 = -1.0, rel = 1.0;
  while (iters < maxit) {
    // r = b - A x in FP64, v0 = r / ||r||
    spmv_m<double, double>(ls, nrhs, x, r64, b, 1);
    const double beta = std::sqrt(red.dot(r64, r64, n));
    if (r0 < 0) r0 = beta;
    if (r0 == 0.0 || beta <= tol * r0) {
      rel = r0 > 0 ? beta / r0 : 0.0;
      break;
    }
    NALU_REQUIRE(std::isfinite(beta), NALU_ERR_NUMERIC, "gmres: non-finite residual");
    hipLaunchKernelGGL((k_fill<TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, V, n, TV(0));
    hipLaunchKernelGGL((k_axpy<double, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, 1.0 / beta, r64, V, n);
    H.reset(beta);
    int k = 0;
    bool done = false;
    const bool learned_cycle = watch && learned_active(ls);
    for (; k < m && iters < maxit; ++k) {
      ++iters;
      TV* vk = V + (size_t)ldv * k;
      TV* vk1 = V + (size_t)ldv * (k + 1);
      TV* zk = flexible ? Z + (size_t)ldv * k : Z;
      prec_m<TV>(ls, nrhs, vk, zk);
      spmv_m<TV, TV>(ls, nrhs, zk, vk1, (const TV*)nullptr, 0);
      const int cnt = k + 1;
      const size_t hbytes = sizeof(double) * (2 * (m + 2) + 1);
      hipLaunchKernelGGL(k_gs_dots<TV>, dim3(nb_dots), dim3(kThreads), lds_dots, s, V, ldv, cnt, always2 ? 0 : 1, vk1,
                         n, R_dots, ls->red.p);
      hipLaunchKernelGGL(k_reduce_rows, dim3(always2 ? cnt : cnt + 1), dim3(kThreads), 0, s, ls->red.p, (int)nb_dots,
                         hA, 0);
      auto second_pass = [&]() {
        hipLaunchKernelGGL(k_gs_dots<TV>, dim3(nb_dots), dim3(kThreads), lds_dots, s, V, ldv, cnt, 0, vk1, n, R_dots,
                           ls->red.p);
        hipLaunchKernelGGL(k_reduce_rows, dim3(cnt), dim3(kThreads), 0, s, ls->red.p, (int)nb_dots, hB, 0);
        hipLaunchKernelGGL(k_gs_update_norm<TV>, dim3(nb_upd), dim3(kThreads), 0, s, V, ldv, cnt, hB, vk1, n,
                           ls->red.p);
        hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kThreads), 0, s, ls->red.p, (int)nb_upd, nrm, 0);
      };
      if (always2) {
        hipLaunchKernelGGL(k_gs_update<TV>, dim3(nb_upd), dim3(kThreads), 0, s, V, ldv, cnt, hA, vk1, n);
        second_pass();
      } else {
        hipLaunchKernelGGL(k_gs_update_norm<TV>, dim3(nb_upd), dim3(kThreads), 0, s, V, ldv, cnt, hA, vk1, n,
                           ls->red.p);
        hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kThreads), 0, s, ls->red.p, (int)nb_upd, nrm, 0);
      }
      NALU_HIP_CHECK(hipMemcpyAsync(hpin, hbuf, hbytes, hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      bool two = always2;
      if (!always2) {
        const double w2 = hpin[cnt], n2 = hpin[2 * (m + 2)];
        if (!(n2 >= 0.01 * w2)) {  // severe cancellation (or NaN): reorthogonalize once
          second_pass();
          NALU_HIP_CHECK(hipMemcpyAsync(hpin, hbuf, hbytes, hipMemcpyDeviceToHost, s));
          NALU_HIP_CHECK(hipStreamSynchronize(s));
          two = true;
        }
      }
      for (int i = 0; i <= k; ++i) hcol[i] = hpin[i] + (two ? hpin[(m + 2) + i] : 0.0);
      hcol[k + 1] = std::sqrt(hpin[2 * (m + 2)]);
      hipLaunchKernelGGL(k_scale_invnorm<TV>, dim3(nblocks(n)), dim3(kThreads), 0, s, vk1, n, nrm);
      if (learned_cycle && !std::isfinite(hcol[k + 1])) {
        learned_set_fallback(ls, true);  // drop this direction, finish the cycle with k columns
        break;
      }
      NALU_REQUIRE(std::isfinite(hcol[k + 1]), NALU_ERR_NUMERIC, "gmres: non-finite Arnoldi vector");
      const double res = H.add_column(k, hcol.data());
      rel = res / r0;
      if (res <= tol * r0 || hcol[k + 1] == 0.0) {
        ++k;
        done = true;
        break;
      }
    }
    // x += M^-1 V y (right preconditioning) or Z y (flexible)
    H.solve(k, y);
    for (int i = 0; i < k; ++i) hpin[i] = y[i];
    NALU_HIP_CHECK(hipMemcpyAsync(hdev, hpin, sizeof(double) * k, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_fill<TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, tmp, n, TV(0));
    if (flexible) {
      hipLaunchKernelGGL((k_maxpy<TV, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, Z, ldv, k, hdev, tmp, n, 1.0);
      hipLaunchKernelGGL((k_axpy<TV, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, 1.0, tmp, x, n);
    } else {
      hipLaunchKernelGGL((k_maxpy<TV, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, V, ldv, k, hdev, tmp, n, 1.0);
      prec_m<TV>(ls, nrhs, tmp, Z);
      hipLaunchKernelGGL((k_axpy<TV, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, 1.0, Z, x, n);
    }
    NALU_HIP_CHECK(hipStreamSynchronize(s));  // hpin is reused by the next cycle
    if (done) break;
    // stagnation: a full cycle that did not reduce the residual by 2x
    if (learned_cycle && k >= std::min(m, 10) && rel * r0 > 0.5 * beta) learned_set_fallback(ls, true);
  }
  if (implicit_rel) *implicit_rel = rel;
  return iters;
}

// Preconditioned CG in the permuted space; scalars stay on the device, one
// host read per iteration for the convergence test.
// CG with the SGS preconditioner on the SELL copy for one right-hand side (b
// and x in the color-permuted row order, as every solver here takes them),
// every dot product fused into a producing kernel and the convergence test
// read from the device every kCheck iterations (see CgState).
constexpr int kCgCheck = 8;

int cg_fused(nalu_linsys* ls, const double* b, double* x, int maxit, double tol) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  const int64_t nch = ls->sell_nchunks;
  const int nc = ls->num_colors;
  const auto& ccp = ls->sell_color_ptr;
  const SellChunk* ch = reinterpret_cast<const SellChunk*>(ls->sell_chunks.p);
  const unsigned g_spmv = (unsigned)std::max<int64_t>(1, (nch + 3) / 4);
  const unsigned g_upd = red_blocks(n);
  // partial-sum slots of the sweeps that finalize z: forward of the last color, every backward
  std::vector<int> off(nc + 1, 0);
  int total = 0;
  auto sweep_grid = [&](int c) { return (unsigned)xcd_grid(std::max<int64_t>(1, (ccp[c + 1] - ccp[c] + 1) / 2)); };
  const int cl = nc - 1;
  off[cl] = total;
  total += (int)sweep_grid(cl);
  for (int c = nc - 2; c >= 0; --c) {
    off[c] = total;
    total += (int)sweep_grid(c);
  }
  const size_t nparts = (size_t)std::max<int64_t>({(int64_t)g_spmv, (int64_t)g_upd, (int64_t)total});
  const size_t need = 256 + (4 * (size_t)n + nparts) * sizeof(double);
  if (ls->cg_work.n < need) ls->cg_work.alloc(need);
  unsigned char* wb = ls->cg_work.p;
  CgState* st = reinterpret_cast<CgState*>(wb);
  double* r = reinterpret_cast<double*>(wb + 256);
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  double* part = q + n;
  ls->cg_host.alloc(sizeof(CgState));
  CgState* hs = static_cast<CgState*>(ls->cg_host.p);
  Red red(ls);
  const int32_t* lcs = sell_cols(ls, false);
  const int32_t* ucs = sell_cols(ls, true);
  const double* lval = sell_vals<double>(ls, false);
  const double* uval = sell_vals<double>(ls, true);
  const double* invd = split_of<double>(ls).invd;

  pspmv<double, double>(ls, x, r, b, 1);
  const double r0 = std::sqrt(red.dot(r, r, n));
  if (r0 == 0.0) return 0;
  NALU_REQUIRE(std::isfinite(r0), NALU_ERR_NUMERIC, "cg: non-finite residual");
  *hs = CgState{};
  hs->r0 = r0;
  hs->tol2r02 = (tol * r0) * (tol * r0);
  NALU_HIP_CHECK(hipMemcpyAsync(st, hs, sizeof(CgState), hipMemcpyHostToDevice, s));
  psgs<double>(ls, r, z);
  red.dot_dev(r, z, n, &st->rz[0]);
  hipLaunchKernelGGL((k_copy<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, z, p, n);
  int cur = 0, it = 0;
  bool stop = false;
  while (!stop && it < maxit) {
    ++it;
    hipLaunchKernelGGL(k_sell_spmv_dot<double>, dim3(g_spmv), dim3(kThreads), 0, s, ch, nch, lcs, lval, ucs, uval,
                       split_of<double>(ls).dv, (const double*)p, q, part, (const CgState*)st);
    hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(kReduceThreads), 0, s, (const double*)part, (int)g_spmv, 0, 0, it, st);
    hipLaunchKernelGGL(k_cg_update_f, dim3(g_upd), dim3(kThreads), 0, s, x, r, (const double*)p,
                       (const double*)q, n, cur, part, (const CgState*)st);
    hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(kReduceThreads), 0, s, (const double*)part, (int)g_upd, 1, 0, it, st);
    const int nxt = 1 - cur;
    for (int c = 0; c < nc; ++c)
      hipLaunchKernelGGL((k_sell_sweep_dot<double, false>), dim3(sweep_grid(c)), dim3(kThreads), 0, s, ch, ccp[c],
                         ccp[c + 1], lcs, lval, invd, (const double*)r, z, c == cl ? 1 : 0, part, off[c],
                         (const CgState*)st);
    for (int c = nc - 2; c >= 0; --c)
      hipLaunchKernelGGL((k_sell_sweep_dot<double, true>), dim3(sweep_grid(c)), dim3(kThreads), 0, s, ch, ccp[c],
                         ccp[c + 1], ucs, uval, invd, (const double*)r, z, 1, part, off[c], (const CgState*)st);
    hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(kReduceThreads), 0, s, (const double*)part, total, 2, nxt, it, st);
    hipLaunchKernelGGL(k_cg_p_f, dim3(nblocks(n)), dim3(kThreads), 0, s, p, (const double*)z, n, nxt, cur,
                       (const CgState*)st);
    cur = nxt;
    if (it % kCgCheck == 0 || it == maxit) {
      NALU_HIP_CHECK(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      stop = hs->done != 0;
    }
  }
  NALU_HIP_CHECK(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  NALU_REQUIRE(hs->done != 2, NALU_ERR_NUMERIC, "cg: non-finite residual");
  return hs->done ? hs->iters : it;
}

bool cg_fusable(const nalu_linsys* ls, int nrhs) {
  return nrhs == 1 && sparse_kernel_choice() == 64 && ls->cfg.precond == NALU_PREC_SGS && ls->num_colors >= 1 &&
         ls->sell_ready && ls->perm_ready && std::getenv("NALU_CG_UNFUSED") == nullptr;
}

int cg(nalu_linsys* ls, const double* b, double* x, int maxit, double tol, int nrhs = 1) {
  if (cg_fusable(ls, nrhs)) return cg_fused(ls, b, x, maxit, tol);
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows * nrhs;  // nrhs > 1: the block-diagonal stacked system
  Red red(ls);
  double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *sc = nullptr;  // the handle's workspace
  auto layout = [&](unsigned char* base) {
    Carve c{base};
    r = c.take<double>(n);
    z = c.take<double>(n);
    p = c.take<double>(n);
    q = c.take<double>(n);
    sc = c.take<double>(8);
    return c.off;
  };
  layout(ensure_ws(ls->krylov_ws, layout(nullptr)));
  ls->krylov_host.alloc(sizeof(double));
  double* pin = static_cast<double*>(ls->krylov_host.p);
  spmv_m<double, double>(ls, nrhs, x, r, b, 1);
  const double r0 = std::sqrt(red.dot(r, r, n));
  if (r0 == 0.0) return 0;
  NALU_REQUIRE(std::isfinite(r0), NALU_ERR_NUMERIC, "cg: non-finite residual");
  prec_m<double>(ls, nrhs, r, z);
  red.dot_dev(r, z, n, sc + 0);  // rz (slot alternates 0/3)
  hipLaunchKernelGGL((k_copy<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, z, p, n);
  int it = 0;
  int cur = 0;
  while (it < maxit) {
    ++it;
    spmv_m<double, double>(ls, nrhs, p, q, (const double*)nullptr, 0);
    red.dot_dev(p, q, n, sc + 1);
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(kThreads), 0, s, x, r, p, q, sc + cur, sc + 1, n,
                       ls->red.p);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kThreads), 0, s, ls->red.p, (int)nb, sc + 2, 0);
    NALU_HIP_CHECK(hipMemcpyAsync(pin, sc + 2, sizeof(double), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    const double rn = std::sqrt(*pin);
    NALU_REQUIRE(std::isfinite(rn), NALU_ERR_NUMERIC, "cg: non-finite residual");
    if (rn <= tol * r0) break;
    prec_m<double>(ls, nrhs, r, z);
    const int nxt = cur == 0 ? 3 : 0;
    red.dot_dev(r, z, n, sc + nxt);
    hipLaunchKernelGGL(k_cg_p, dim3(nblocks(n)), dim3(kThreads), 0, s, p, z, sc + nxt, sc + cur, n);
    cur = nxt;
  }
  return it;
}

}  // namespace

void spmv(nalu_linsys* ls, const double* x, double* y) {
  spmv_t<double, double, double>(ls, ls->values.p, x, y, (const double*)nullptr, 0);
}

double norm2(nalu_linsys* ls, const double* x, int64_t n) {
  Red red(ls);
  return std::sqrt(red.dot(x, x, n));
}

void perm_sgs32(nalu_linsys* ls, const float* r, float* z) { psgs<float>(ls, r, z); }
void perm_sgs64(nalu_linsys* ls, const double* r, double* z) { psgs<double>(ls, r, z); }
void perm_spmv64(nalu_linsys* ls, const double* x, double* y) {
  pspmv<double, double>(ls, x, y, (const double*)nullptr, 0);
}
template <typename T>
void perm_residual(nalu_linsys* ls, const T* x, T* r, const T* b) {
  pspmv<T, T>(ls, x, r, b, 1);
}
template void perm_residual<double>(nalu_linsys*, const double*, double*, const double*);
template void perm_residual<float>(nalu_linsys*, const float*, float*, const float*);

// FP32 copy of the permuted split values (mixed-precision inner solve, and the
// SGS input feature of the learned preconditioner)
void ensure_split32(nalu_linsys* ls) {
  NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "FP32 split copy: preconditioner not set up");
  if (ls->split32_ready) return;
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows, nl = ls->lci.n, nu = ls->uci.n;
  if (ls->lval32.n != (size_t)nl) ls->lval32.alloc(nl);
  if (ls->uval32.n != (size_t)nu) ls->uval32.alloc(nu);
  if (ls->dval32.n != (size_t)n) ls->dval32.alloc(n);
  if (ls->pinvd32.n != (size_t)n) ls->pinvd32.alloc(n);
  DevBuf<int> bad(1);
  NALU_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_take<float>, dim3(nblocks(nl)), dim3(kThreads), 0, s, ls->lsrc.p, ls->values.p, nl,
                     ls->lval32.p);
  hipLaunchKernelGGL(k_take<float>, dim3(nblocks(nu)), dim3(kThreads), 0, s, ls->usrc.p, ls->values.p, nu,
                     ls->uval32.p);
  hipLaunchKernelGGL(k_take_diag<float>, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->dsrc.p, ls->values.p, n,
                     ls->dval32.p, ls->pinvd32.p, bad.p);
  const int64_t sl = ls->sl_src.n, su = ls->su_src.n;
  if (ls->sl_val32.n != (size_t)sl) ls->sl_val32.alloc(sl);
  if (ls->su_val32.n != (size_t)su) ls->su_val32.alloc(su);
  hipLaunchKernelGGL(k_take_sell<float>, dim3(nblocks(sl)), dim3(kThreads), 0, s, ls->sl_src.p, ls->values.p, sl,
                     ls->sl_val32.p);
  hipLaunchKernelGGL(k_take_sell<float>, dim3(nblocks(su)), dim3(kThreads), 0, s, ls->su_src.p, ls->values.p, su,
                     ls->su_val32.p);
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  ls->split32_ready = true;
}

// Ifpack2 compute(): refresh the permuted split copy of the current values.
void precond_setup(nalu_linsys* ls) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  materialize_pending_zero(ls);
  build_perm(ls);
  const int64_t nl = ls->lci.n, nu = ls->uci.n;
  if (ls->lval.n != (size_t)nl) ls->lval.alloc(nl);
  if (ls->uval.n != (size_t)nu) ls->uval.alloc(nu);
  if (ls->dval.n != (size_t)n) ls->dval.alloc(n);
  if (ls->pinvd.n != (size_t)n) ls->pinvd.alloc(n);
  DevBuf<int> bad(1);
  NALU_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_take<double>, dim3(nblocks(nl)), dim3(kThreads), 0, s, ls->lsrc.p, ls->values.p, nl, ls->lval.p);
  hipLaunchKernelGGL(k_take<double>, dim3(nblocks(nu)), dim3(kThreads), 0, s, ls->usrc.p, ls->values.p, nu, ls->uval.p);
  hipLaunchKernelGGL(k_take_diag<double>, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->dsrc.p, ls->values.p, n,
                     ls->dval.p, ls->pinvd.p, bad.p);
  build_sell(ls);
  {
    const int64_t sl = ls->sl_src.n, su = ls->su_src.n;
    if (ls->sl_val.n != (size_t)sl) ls->sl_val.alloc(sl);
    if (ls->su_val.n != (size_t)su) ls->su_val.alloc(su);
    hipLaunchKernelGGL(k_take_sell<double>, dim3(nblocks(sl)), dim3(kThreads), 0, s, ls->sl_src.p, ls->values.p, sl,
                       ls->sl_val.p);
    hipLaunchKernelGGL(k_take_sell<double>, dim3(nblocks(su)), dim3(kThreads), 0, s, ls->su_src.p, ls->values.p, su,
                       ls->su_val.p);
  }
  int hb = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  NALU_REQUIRE(hb == 0, NALU_ERR_NUMERIC, "preconditioner: zero or non-finite diagonal");
  ls->precond_ready = true;
  ls->split32_ready = false;
  if (mixed(ls) || ls->learned) ensure_split32(ls);
  learned_setup(ls);
  if (ls->cfg.precond == NALU_PREC_MUELU ||
      (ls->cfg.precond == NALU_PREC_LEARNED && ls->learned_base == NALU_PREC_MUELU))
    amg_setup(ls);
  else
    amg_release(ls);
}

// M^-1 r in the original row order (permutes in and out through two
// persistent work vectors of the handle)
void precond_apply(nalu_linsys* ls, const double* r, double* z) {
  NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "preconditioner not set up");
  const int64_t n = ls->nrows;
  if (ls->pwork.n != (size_t)(2 * n)) ls->pwork.alloc(2 * n);
  double* rp = ls->pwork.p;
  double* zp = rp + n;
  hipLaunchKernelGGL((k_gather<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, r, ls->perm.p, n,
                     rp);
  pprec<double>(ls, rp, zp);
  hipLaunchKernelGGL((k_scatter<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, zp, ls->perm.p,
                     n, z);
}

// Solve with the configured method; x holds the initial guess (original order).
// nrhs > 1: B and X are [nrhs][rows] and the block-diagonal stacked system is
// solved in one Krylov space (the right-hand sides share every matrix read);
// convergence is measured on the stacked residual norm, as the reference
// measures a numDof 3 system over all of its components.  rel_each (optional,
// [nrhs]): each right-hand side's FP64 true relative residual.
int krylov_solve_k(nalu_linsys* ls, int nrhs, const double* b, double* x, int* iters_out, double* rel_out,
                   double* rel_each) {
  const auto& c = ls->cfg;
  const double tol = c.final_outer_iter ? c.final_tolerance : c.tolerance;
  const int maxit = std::max(1, c.max_iterations);
  const int m = std::max(1, c.kspace);
  const int64_t n1 = ls->nrows;
  const int64_t n = n1 * nrhs;
  hipStream_t s = ls->stream;
  Red red(ls);
  const bool inter = stacked_interleaved(ls, nrhs);
  // permuted right-hand side, solution and residual, the refinement correction
  // (mixed) and the per-component slices (interleaved stacks): the handle's workspace
  struct Vec {
    double* p = nullptr;
  } bp, xp, r, d, sl;
  auto layout = [&](unsigned char* base) {
    Carve c{base};
    bp.p = c.take<double>(n);
    xp.p = c.take<double>(n);
    r.p = c.take<double>(n);
    d.p = c.take<double>(mixed(ls) ? n : 0);
    sl.p = c.take<double>(rel_each && inter ? 2 * n1 : 0);
    return c.off;
  };
  layout(ensure_ws(ls->solve_ws, layout(nullptr)));
  struct LayoutReset {  // the flag describes this solve's vectors only
    nalu_linsys* ls;
    ~LayoutReset() { ls->stack_interleaved = false; }
  } reset{ls};
  ls->stack_interleaved = inter;
  for (int q = 0; q < nrhs; ++q) {
    if (inter) {
      hipLaunchKernelGGL(k_gather_strided, dim3(nblocks(n1)), dim3(kThreads), 0, s, b + q * n1, ls->perm.p, n1,
                         nrhs, bp.p + q);
      hipLaunchKernelGGL(k_gather_strided, dim3(nblocks(n1)), dim3(kThreads), 0, s, x + q * n1, ls->perm.p, n1,
                         nrhs, xp.p + q);
    } else {
      hipLaunchKernelGGL((k_gather<double, double>), dim3(nblocks(n1)), dim3(kThreads), 0, s, b + q * n1,
                         ls->perm.p, n1, bp.p + q * n1);
      hipLaunchKernelGGL((k_gather<double, double>), dim3(nblocks(n1)), dim3(kThreads), 0, s, x + q * n1,
                         ls->perm.p, n1, xp.p + q * n1);
    }
  }
  const double bn = std::sqrt(red.dot(bp.p, bp.p, n));
  NALU_REQUIRE(std::isfinite(bn), NALU_ERR_NUMERIC, "solve: non-finite right-hand side");
  learned_set_fallback(ls, false);
  int iters = 0;
  if (!mixed(ls)) {
    if (c.method == NALU_METHOD_CG) iters = cg(ls, bp.p, xp.p, maxit, tol, nrhs);
    else iters = gmres_t<double>(ls, bp.p, xp.p, m, maxit, tol, c.method == NALU_METHOD_FGMRES, nullptr, nrhs);
  } else {
    // FP64 iterative refinement around an FP32 inner (F)GMRES: r = b - A x in
    // FP64, d ~ A^-1 r in FP32 arithmetic with FP64 Arnoldi inner products and
    // least squares, x += d in FP64.
    spmv_m<double, double>(ls, nrhs, xp.p, r.p, bp.p, 1);
    double rn = std::sqrt(red.dot(r.p, r.p, n));
    const double target = tol * (bn > 0 ? bn : 1.0);
    int outer = 0;
    while (rn > target && iters < maxit && outer < 100) {
      ++outer;
      hipLaunchKernelGGL((k_fill<double>), dim3(nblocks(n)), dim3(kThreads), 0, s, d.p, n, 0.0);
      const double inner_tol = std::min(1e-1, std::max(1e-4, 0.5 * target / rn));
      iters += gmres_t<float>(ls, r.p, d.p, m, maxit - iters, inner_tol, c.method == NALU_METHOD_FGMRES, nullptr,
                              nrhs);
      hipLaunchKernelGGL((k_axpy<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, 1.0, d.p, xp.p, n);
      spmv_m<double, double>(ls, nrhs, xp.p, r.p, bp.p, 1);
      rn = std::sqrt(red.dot(r.p, r.p, n));
      NALU_REQUIRE(std::isfinite(rn), NALU_ERR_NUMERIC, "mixed precision: non-finite residual");
    }
  }
  // the FP64 true residual decides convergence
  spmv_m<double, double>(ls, nrhs, xp.p, r.p, bp.p, 1);
  const double rn = std::sqrt(red.dot(r.p, r.p, n));
  if (rel_each) {
    for (int q = 0; q < nrhs; ++q) {
      const double *bq_p = bp.p + q * n1, *rq_p = r.p + q * n1;
      if (inter) {  // contiguous copies of component q
        hipLaunchKernelGGL(k_take_strided, dim3(nblocks(n1)), dim3(kThreads), 0, s, bp.p + q, n1, nrhs, sl.p);
        hipLaunchKernelGGL(k_take_strided, dim3(nblocks(n1)), dim3(kThreads), 0, s, r.p + q, n1, nrhs, sl.p + n1);
        bq_p = sl.p;
        rq_p = sl.p + n1;
      }
      const double bq = std::sqrt(red.dot(bq_p, bq_p, n1));
      const double rq = std::sqrt(red.dot(rq_p, rq_p, n1));
      rel_each[q] = rq / (bq > 0 ? bq : 1.0);
    }
  }
  for (int q = 0; q < nrhs; ++q) {
    if (inter)
      hipLaunchKernelGGL(k_scatter_strided, dim3(nblocks(n1)), dim3(kThreads), 0, s, xp.p + q, ls->perm.p, n1, nrhs,
                         x + q * n1);
    else
      hipLaunchKernelGGL((k_scatter<double, double>), dim3(nblocks(n1)), dim3(kThreads), 0, s, xp.p + q * n1,
                         ls->perm.p, n1, x + q * n1);
  }
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  if (iters_out) *iters_out = iters;
  if (rel_out) *rel_out = rn / (bn > 0 ? bn : 1.0);
  ls->linear_residual = rn;
  return (rn <= tol * (bn > 0 ? bn : 1.0) * (1.0 + 1e-12) || bn == 0.0) ? NALU_OK : NALU_NOT_CONVERGED;
}

int krylov_solve(nalu_linsys* ls, const double* b, double* x, int* iters_out, double* rel_out) {
  return krylov_solve_k(ls, 1, b, x, iters_out, rel_out, nullptr);
}

}  // namespace nalu

extern "C" int nalu_linsys_spmv_f32(nalu_linsys* ls, const float* x, float* y) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && x && y, NALU_ERR_ARG, "spmv_f32: bad arguments");
    NALU_REQUIRE(ls->precond_ready && ls->lval32.p, NALU_ERR_STATE,
                 "spmv_f32: no FP32 matrix (set up a mixed-precision preconditioner first)");
    const int64_t n = ls->nrows;
    nalu::DevBuf<float> xp(n), yp(n);
    hipLaunchKernelGGL((nalu::k_gather<float, float>), dim3(nalu::nblocks(n)), dim3(nalu::kThreads), 0, ls->stream, x,
                       ls->perm.p, n, xp.p);
    nalu::pspmv<float, float>(ls, xp.p, yp.p, (const float*)nullptr, 0);
    hipLaunchKernelGGL((nalu::k_scatter<float, float>), dim3(nalu::nblocks(n)), dim3(nalu::kThreads), 0, ls->stream,
                       yp.p, ls->perm.p, n, y);
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    return NALU_OK;
  });
}

namespace nalu {
namespace {
__global__ void k_sub_mean(double* __restrict__ x, int64_t n, const double* __restrict__ sum) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] -= sum[0] / (double)n;
}
}  // namespace
}  // namespace nalu

extern "C" int nalu_linsys_remove_rhs_mean(nalu_linsys* ls) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized, NALU_ERR_STATE, "remove_rhs_mean: system not finalized");
    nalu::materialize_pending_zero(ls);
    nalu::DevBuf<double> one(ls->nrows);
    hipLaunchKernelGGL((nalu::k_fill<double>), dim3(nalu::nblocks(ls->nrows)), dim3(nalu::kThreads), 0, ls->stream,
                       one.p, ls->nrows, 1.0);
    nalu::Red red(ls);
    red.dot_dev(ls->rhs.p, one.p, ls->nrows, ls->red_out.p + 255);
    hipLaunchKernelGGL(nalu::k_sub_mean, dim3(nalu::nblocks(ls->nrows)), dim3(nalu::kThreads), 0, ls->stream,
                       ls->rhs.p, ls->nrows, ls->red_out.p + 255);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_precond_apply_permuted(nalu_linsys* ls, const double* r, double* z) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && r && z, NALU_ERR_ARG, "precond_apply_permuted: bad arguments");
    NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "precond_apply_permuted: preconditioner not set up");
    nalu::pprec<double>(ls, r, z);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_precond_apply_permuted_f32(nalu_linsys* ls, const float* r, float* z) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && r && z, NALU_ERR_ARG, "precond_apply_permuted_f32: bad arguments");
    NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "precond_apply_permuted_f32: preconditioner not set up");
    nalu::ensure_split32(ls);
    nalu::pprec<float>(ls, r, z);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_spmv_permuted(nalu_linsys* ls, const double* x, double* y) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && x && y, NALU_ERR_ARG, "spmv_permuted: bad arguments");
    NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "spmv_permuted: solver copy not set up");
    nalu::pspmv<double, double>(ls, x, y, (const double*)nullptr, 0);
    NALU_HIP_CHECK(hipGetLastError());
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_permutation(nalu_linsys* ls, int32_t* perm_host) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && perm_host, NALU_ERR_ARG, "permutation: bad arguments");
    nalu::build_perm(ls);
    NALU_HIP_CHECK(hipMemcpy(perm_host, ls->perm.p, sizeof(int32_t) * ls->nrows, hipMemcpyDeviceToHost));
    return NALU_OK;
  });
}
