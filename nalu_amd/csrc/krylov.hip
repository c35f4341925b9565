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

#include <cmath>
#include <vector>

#include "internal.hpp"

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

// multi-dot: part[j][blk] for j in [0, cnt), V column-major with leading dim ldv
template <typename TV, typename TW, int CHUNK>
__global__ __launch_bounds__(kThreads) void k_mdot_part(const TV* __restrict__ V, int64_t ldv, int cnt,
                                                        const TW* __restrict__ w, int64_t n, double* __restrict__ part) {
  double s[CHUNK];
#pragma unroll
  for (int j = 0; j < CHUNK; ++j) s[j] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < CHUNK; ++j)
      if (j < cnt) s[j] += (double)V[j * ldv + i] * wi;
  }
#pragma unroll
  for (int j = 0; j < CHUNK; ++j) {
    const double t = block_sum(s[j]);
    if (threadIdx.x == 0 && j < cnt) part[(int64_t)j * gridDim.x + blockIdx.x] = t;
  }
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

// ------------------------------------------------------------ host helpers
struct Red {
  nalu_linsys* ls;
  hipStream_t s;
  explicit Red(nalu_linsys* l) : ls(l), s(l->stream) {
    if (ls->red.n < (size_t)kRedBlocks * 64) ls->red.alloc((size_t)kRedBlocks * 64);
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
  // out[0..cnt) (+)= V[:, 0..cnt)^T w
  template <typename TV, typename TW>
  void mdot_dev(const TV* V, int64_t ldv, int cnt, const TW* w, int64_t n, double* out, int accumulate) {
    const unsigned nb = red_blocks(n);
    for (int j0 = 0; j0 < cnt; j0 += 8) {
      const int c = std::min(8, cnt - j0);
      hipLaunchKernelGGL((k_mdot_part<TV, TW, 8>), dim3(nb), dim3(kThreads), 0, s, V + j0 * ldv, ldv, c, w, n,
                         ls->red.p);
      hipLaunchKernelGGL(k_reduce_rows, dim3(c), dim3(kThreads), 0, s, ls->red.p, (int)nb, out + j0, accumulate);
    }
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

template <typename T>
void sgs_apply_t(nalu_linsys* ls, const T* vals, const T* invd, const T* r, T* z) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  hipLaunchKernelGGL((k_fill<T>), dim3(nblocks(n)), dim3(kThreads), 0, s, z, n, T(0));
  constexpr int L = 8;
  auto sweep = [&](int c) {
    const int64_t a = ls->color_ptr_host[c], b = ls->color_ptr_host[c + 1];
    if (b <= a) return;
    hipLaunchKernelGGL((k_gs_color<T, L>), dim3(nblocks((b - a) * L)), dim3(kThreads), 0, s, ls->color_rows.p + a,
                       b - a, ls->row_ptr.p, ls->col_idx.p, vals, invd, r, z);
  };
  for (int c = 0; c < ls->num_colors; ++c) sweep(c);
  for (int c = ls->num_colors - 1; c >= 0; --c) sweep(c);
}

bool mixed(const nalu_linsys* ls) { return ls->cfg.precision == NALU_PRECISION_MIXED; }

// FP32 preconditioner application (mixed-precision inner loop)
void precond_apply_f32(nalu_linsys* ls, const float* r, float* z) {
  const int64_t n = ls->nrows;
  switch (ls->cfg.precond) {
    case NALU_PREC_JACOBI:
      hipLaunchKernelGGL((k_jacobi<float>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, ls->inv_diag32.p, r, z, n);
      break;
    case NALU_PREC_SGS:
    case NALU_PREC_LEARNED:
      sgs_apply_t<float>(ls, ls->values32.p, ls->inv_diag32.p, r, z);
      break;
    default:
      hipLaunchKernelGGL((k_copy<float, float>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, r, z, n);
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

// Restarted (F)GMRES on A x = b.  TV is the storage/arithmetic type of the
// Krylov basis and of the SpMV / preconditioner (double, or float in the
// mixed-precision inner solve); inner products, the Hessenberg matrix and the
// Givens least-squares are FP64.  flexible: store Z_k = M^-1 v_k (FGMRES).
// x is FP64; returns iterations, *res_hist = implicit |g_k| / r0 at exit.
template <typename TV>
int gmres_t(nalu_linsys* ls, const TV* vals, const double* b, double* x, int m, int maxit, double tol,
            bool flexible, double r0_scale, double* implicit_rel) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  Red red(ls);
  // workspace: V (m+1) cols, Z (flexible: m cols, else 1), w temp
  const size_t need = (size_t)n * ((m + 1) + (flexible ? m : 1) + 1);
  DevBuf<TV> ws(need);
  TV* V = ws.p;
  TV* Z = V + (size_t)n * (m + 1);
  TV* tmp = Z + (size_t)n * (flexible ? m : 1);
  DevBuf<double> r64(n), hdev(m + 2), h2(m + 2);
  Hessenberg H(m);
  std::vector<double> hcol(m + 2), y;
  double* hpin = nullptr;
  NALU_HIP_CHECK(hipHostMalloc((void**)&hpin, sizeof(double) * (m + 2), hipHostMallocDefault));
  struct PinFree {
    double* p;
    ~PinFree() { if (p) (void)hipHostFree(p); }
  } pin_guard{hpin};

  auto precond = [&](const TV* in, TV* out) {
    if constexpr (std::is_same<TV, double>::value) precond_apply(ls, in, out);
    else precond_apply_f32(ls, in, out);
  };

  int iters = 0;
  double r0 = -1.0, rel = 1.0;
  while (iters < maxit) {
    // r = b - A x (FP64), v0 = r / ||r||
    spmv_t<double, double, double>(ls, ls->values.p, x, r64.p, b, 1);
    const double beta = std::sqrt(red.dot(r64.p, r64.p, n));
    if (r0 < 0) r0 = (r0_scale > 0 ? r0_scale : beta);
    if (r0 == 0.0 || beta <= tol * r0) {
      rel = r0 > 0 ? beta / r0 : 0.0;
      break;
    }
    hipLaunchKernelGGL((k_copy<double, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, r64.p, V, n);
    hipLaunchKernelGGL((k_axpy<double, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, 1.0 / beta - 1.0, r64.p, V, n);
    H.reset(beta);
    int k = 0;
    bool done = false;
    for (; k < m && iters < maxit; ++k) {
      ++iters;
      TV* vk = V + (size_t)n * k;
      TV* vk1 = V + (size_t)n * (k + 1);
      TV* zk = flexible ? Z + (size_t)n * k : Z;
      precond(vk, zk);
      spmv_t<TV, TV, TV>(ls, vals, zk, vk1, (const TV*)nullptr, 0);
      // ICGS: two classical Gram-Schmidt passes, coefficients summed on device
      red.mdot_dev(V, n, k + 1, vk1, n, hdev.p, 0);
      hipLaunchKernelGGL((k_maxpy<TV, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, V, n, k + 1, hdev.p, vk1, n, -1.0);
      red.mdot_dev(V, n, k + 1, vk1, n, h2.p, 0);
      hipLaunchKernelGGL((k_maxpy<TV, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, V, n, k + 1, h2.p, vk1, n, -1.0);
      hipLaunchKernelGGL((k_axpy<double, double>), dim3(nblocks(k + 1)), dim3(kThreads), 0, s, 1.0, h2.p, hdev.p,
                         (int64_t)(k + 1));
      red.dot_dev(vk1, vk1, n, hdev.p + k + 1);
      hipLaunchKernelGGL((k_scale_invnorm<TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, vk1, n, hdev.p + k + 1);
      NALU_HIP_CHECK(hipMemcpyAsync(hpin, hdev.p, sizeof(double) * (k + 2), hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      for (int i = 0; i <= k; ++i) hcol[i] = hpin[i];
      hcol[k + 1] = std::sqrt(hpin[k + 1]);
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
    NALU_HIP_CHECK(hipMemcpyAsync(hdev.p, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_fill<TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, tmp, n, TV(0));
    if (flexible) {
      hipLaunchKernelGGL((k_maxpy<TV, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, Z, n, k, hdev.p, tmp, n, 1.0);
      hipLaunchKernelGGL((k_axpy<TV, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, 1.0, tmp, x, n);
    } else {
      hipLaunchKernelGGL((k_maxpy<TV, TV>), dim3(nblocks(n)), dim3(kThreads), 0, s, V, n, k, hdev.p, tmp, n, 1.0);
      precond(tmp, Z);
      hipLaunchKernelGGL((k_axpy<TV, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, 1.0, Z, x, n);
    }
    NALU_HIP_CHECK(hipStreamSynchronize(s));  // y must outlive the copy
    if (done) break;
  }
  if (implicit_rel) *implicit_rel = rel;
  return iters;
}

int cg(nalu_linsys* ls, const double* b, double* x, int maxit, double tol) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  Red red(ls);
  DevBuf<double> r(n), z(n), p(n), q(n), sc(8);
  spmv_t<double, double, double>(ls, ls->values.p, x, r.p, b, 1);
  const double r0 = std::sqrt(red.dot(r.p, r.p, n));
  if (r0 == 0.0) return 0;
  precond_apply(ls, r.p, z.p);
  red.dot_dev(r.p, z.p, n, sc.p + 0);  // rz (slot alternates 0/3)
  hipLaunchKernelGGL((k_copy<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, z.p, p.p, n);
  double* pin = nullptr;
  NALU_HIP_CHECK(hipHostMalloc((void**)&pin, sizeof(double), hipHostMallocDefault));
  struct PinFree {
    double* p;
    ~PinFree() { if (p) (void)hipHostFree(p); }
  } guard{pin};
  int it = 0;
  int cur = 0;
  while (it < maxit) {
    ++it;
    spmv_t<double, double, double>(ls, ls->values.p, p.p, q.p, (const double*)nullptr, 0);
    red.dot_dev(p.p, q.p, n, sc.p + 1);
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(kThreads), 0, s, x, r.p, p.p, q.p, sc.p + cur, sc.p + 1, n,
                       ls->red.p);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kThreads), 0, s, ls->red.p, (int)nb, sc.p + 2, 0);
    NALU_HIP_CHECK(hipMemcpyAsync(pin, sc.p + 2, sizeof(double), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    const double rn = std::sqrt(*pin);
    NALU_REQUIRE(std::isfinite(rn), NALU_ERR_NUMERIC, "cg: non-finite residual");
    if (rn <= tol * r0) break;
    precond_apply(ls, r.p, z.p);
    const int nxt = cur == 0 ? 3 : 0;
    red.dot_dev(r.p, z.p, n, sc.p + nxt);
    hipLaunchKernelGGL(k_cg_p, dim3(nblocks(n)), dim3(kThreads), 0, s, p.p, z.p, sc.p + nxt, sc.p + cur, n);
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

void precond_setup(nalu_linsys* ls) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  materialize_pending_zero(ls);
  if (ls->cfg.precond == NALU_PREC_SGS || ls->cfg.precond == NALU_PREC_LEARNED) {
    if (ls->num_colors == 0) build_coloring(ls, s);
  }
  DevBuf<int> bad(1);
  NALU_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  if (ls->inv_diag.n != (size_t)n) ls->inv_diag.alloc(n);
  hipLaunchKernelGGL((k_inv_diag<double>), dim3(nblocks(n)), dim3(kThreads), 0, s, ls->diag_pos.p, ls->values.p, n,
                     ls->inv_diag.p, bad.p);
  if (mixed(ls)) {
    if (ls->values32.n != (size_t)ls->nnz) ls->values32.alloc(ls->nnz);
    if (ls->inv_diag32.n != (size_t)n) ls->inv_diag32.alloc(n);
    hipLaunchKernelGGL((k_copy<double, float>), dim3(nblocks(ls->nnz)), dim3(kThreads), 0, s, ls->values.p,
                       ls->values32.p, ls->nnz);
    hipLaunchKernelGGL((k_inv_diag<float>), dim3(nblocks(n)), dim3(kThreads), 0, s, ls->diag_pos.p, ls->values.p, n,
                       ls->inv_diag32.p, bad.p);
  }
  int hb = 0;
  NALU_HIP_CHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  NALU_REQUIRE(hb == 0, NALU_ERR_NUMERIC, "preconditioner: zero or non-finite diagonal");
  ls->precond_ready = true;
}

void precond_apply(nalu_linsys* ls, const double* r, double* z) {
  NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "preconditioner not set up");
  const int64_t n = ls->nrows;
  switch (ls->cfg.precond) {
    case NALU_PREC_JACOBI:
      hipLaunchKernelGGL((k_jacobi<double>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, ls->inv_diag.p, r, z, n);
      break;
    case NALU_PREC_SGS:
    case NALU_PREC_LEARNED:
      sgs_apply_t<double>(ls, ls->values.p, ls->inv_diag.p, r, z);
      break;
    default:
      hipLaunchKernelGGL((k_copy<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, r, z, n);
  }
}

// Solve with the configured method; x holds the initial guess.
int krylov_solve(nalu_linsys* ls, const double* b, double* x, int* iters_out, double* rel_out) {
  const auto& c = ls->cfg;
  const double tol = c.final_outer_iter ? c.final_tolerance : c.tolerance;
  const int maxit = std::max(1, c.max_iterations);
  const int m = std::max(1, c.kspace);
  const int64_t n = ls->nrows;
  int iters = 0;
  if (!mixed(ls)) {
    if (c.method == NALU_METHOD_CG) iters = cg(ls, b, x, maxit, tol);
    else iters = gmres_t<double>(ls, ls->values.p, b, x, m, maxit, tol, c.method == NALU_METHOD_FGMRES, -1.0, nullptr);
  } else {
    // FP64 iterative refinement around an FP32 inner (F)GMRES: r = b - A x in
    // FP64, d ~ A^-1 r in FP32 (inner tolerance 1e-3 relative to ||r||), x += d.
    DevBuf<double> r(n), d(n);
    Red red(ls);
    const double bn = std::sqrt(red.dot(b, b, n));
    double rn = bn;
    spmv_t<double, double, double>(ls, ls->values.p, x, r.p, b, 1);
    rn = std::sqrt(red.dot(r.p, r.p, n));
    const double target = tol * (bn > 0 ? bn : 1.0);
    int outer = 0;
    while (rn > target && iters < maxit && outer < 100) {
      ++outer;
      hipLaunchKernelGGL((k_fill<double>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, d.p, n, 0.0);
      const double inner_tol = std::max(1e-3, 0.5 * target / rn);
      iters += gmres_t<float>(ls, ls->values32.p, r.p, d.p, m, maxit - iters, std::min(inner_tol, 1e-1),
                              c.method == NALU_METHOD_FGMRES, -1.0, nullptr);
      hipLaunchKernelGGL((k_axpy<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, 1.0, d.p, x, n);
      spmv_t<double, double, double>(ls, ls->values.p, x, r.p, b, 1);
      rn = std::sqrt(red.dot(r.p, r.p, n));
      NALU_REQUIRE(std::isfinite(rn), NALU_ERR_NUMERIC, "mixed precision: non-finite residual");
    }
  }
  // true FP64 relative residual decides convergence
  DevBuf<double> r(n);
  spmv_t<double, double, double>(ls, ls->values.p, x, r.p, b, 1);
  Red red(ls);
  const double rn = std::sqrt(red.dot(r.p, r.p, n));
  const double bn = std::sqrt(red.dot(b, b, n));
  if (iters_out) *iters_out = iters;
  if (rel_out) *rel_out = rn / (bn > 0 ? bn : 1.0);
  ls->linear_residual = rn;
  return (rn <= tol * (bn > 0 ? bn : 1.0) * 1.0000001 || bn == 0.0) ? NALU_OK : NALU_NOT_CONVERGED;
}

}  // namespace nalu

extern "C" int nalu_linsys_spmv_f32(nalu_linsys* ls, const float* x, float* y) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(ls && ls->finalized && x && y, NALU_ERR_ARG, "spmv_f32: bad arguments");
    NALU_REQUIRE(ls->values32.n == (size_t)ls->nnz, NALU_ERR_STATE,
                 "spmv_f32: no FP32 matrix (mixed precision preconditioner not set up)");
    nalu::spmv_t<float, float, float>(ls, ls->values32.p, x, y, (const float*)nullptr, 0);
    NALU_HIP_CHECK(hipGetLastError());
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
