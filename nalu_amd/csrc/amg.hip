// Smoothed-aggregation algebraic multigrid preconditioner on the GPU: the role
// MueLu plays for the pressure system in the reference
// (TpetraLinearSolver::setMueLu, src/LinearSolver.C:124-152, configured by a
// MueLu parameter list such as reg_tests/xml/matches_ml_default.xml:
// uncoupled aggregation with a drop tolerance, smoothed-aggregation prolongator,
// Chebyshev smoothing with a power-iteration eigenvalue estimate, a direct
// coarse solve below "coarse: max size").
//
// Everything runs on the device in the solver's color-permuted row order:
//   setup, per level l with operator A_l (CSR):
//     1. inverse diagonal; lambda_max of D^-1 A by power iteration;
//     2. strength of connection: j strong for row i when a_ij != 0 and
//        a_ij^2 > tol^2 |a_ii a_jj| (every nonzero when tol == 0);
//     3. aggregation: distance-2 maximal independent set of the strength
//        graph (synchronous rounds, hash priorities) gives the roots; every
//        other row joins the root it is most strongly connected to, and rows
//        two steps away join their strongest aggregated neighbour;
//     4. tentative prolongator: per aggregate, the normalized restriction of
//        the near-null-space vector B (B = 1 on the finest level), B_c = its norm;
//     5. P = (I - omega D_f^-1 A_f) P_tent, omega = damping / lambda_max, where
//        A_f is A with the weak entries lumped into the diagonal (rows with no
//        strong entry keep their tentative row);
//     6. R = P^T (device radix sort), A_c = R (A P) with a hash SpGEMM
//        (one wave per row, hash table in LDS);
//   until the operator has at most max_coarse rows, whose dense inverse is
//   formed by Gauss-Jordan elimination with partial pivoting; pivots below
//   1e-10 of the scale max(max|a_ii|, max|a_ii of the level above| x
//   max_I |P(:, I)|^2) are skipped (the pseudo-solve of a pure Neumann
//   operator; the second term keeps the scale when the Galerkin product of a
//   singular operator cancels to round-off, e.g. a single coarse aggregate).
//   When the coarse near-null vector B_c satisfies |A_c B_c| <= 1e-5 scale |B_c|
//   (a pure Neumann operator whose row sums vanish only to its assembly
//   precision, e.g. FP32 assembly), the rank-one residual (A_c B_c) B_c^T / |B_c|^2
//   is removed first, so that B_c is an exact null vector and the coarse solve
//   adds no near-null component of size 1 / round-off.
//   apply: one V-cycle, Chebyshev pre- and post-smoothing of the configured
//   degree (Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1, on
//   D^-1 A with lambda in [lambda_max / ratio, lambda_max]), zero initial guess,
//   restriction R, coarse correction P.  The finest level's residuals use the
//   solver's SELL-64 copy of the matrix; templated on the arithmetic type (FP64
//   for the FP64 solvers, FP32 inside the mixed-precision inner solve).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>

#include "internal.hpp"

namespace nalu {

// Largest coarsest level solved densely.  The Gauss-Jordan step reads and
// writes the n x 2n augmented matrix once per pivot, 32 n^3 bytes in all:
// 275 GB (about 45 ms at HBM rate) at 2048 rows; MueLu's default "coarse: max
// size" is 1000.
constexpr int64_t kDenseCoarseMax = 2048;

namespace {

constexpr int kT = 256;
inline unsigned nbl(int64_t n, int t = kT) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

struct Csr {
  int64_t n = 0, m = 0, nnz = 0;  // rows, columns, entries
  int max_row = 0;
  DevBuf<int32_t> rp, ci;
  DevBuf<double> v;
  DevBuf<float> v32;
};

struct Level {
  Csr A;                      // level operator (level 0: solver-order copy, setup only)
  DevBuf<double> diag, invd;  // diagonal and its inverse
  DevBuf<float> invd32;
  Csr P, R;                   // prolongator n x nc, restriction nc x n
  DevBuf<int32_t> agg;        // aggregate of each row
  double lmax = 0.0;          // power-iteration estimate of lambda_max(D^-1 A)
  DevBuf<double> work;        // x, b, r, d: 4 n values (FP64 or FP32)
};

struct Amg {
  std::vector<Level> lev;
  int64_t nc = 0;             // rows of the dense coarse operator
  int nfree = 0;              // skipped (singular) pivots of the coarse inverse
  double null_rel = 0.0;      // |A_c B_c| / (scale |B_c|) of the coarsest operator
  bool deflated = false;      // the rank-one residual along B_c was removed
  DevBuf<double> cinv;        // dense (pseudo-)inverse of the coarsest operator, row-major
  DevBuf<float> cinv32;
  bool f32_ready = false;
  double setup_s = 0.0;
};

Amg* amg_of(const nalu_linsys* ls) { return reinterpret_cast<Amg*>(ls->amg); }

// ------------------------------------------------------------------ helpers
void scan_excl(const int32_t* in, int32_t* out, int64_t n, hipStream_t s) {
  NALU_REQUIRE(n < (int64_t)INT32_MAX, NALU_ERR_UNSUPPORTED, "amg: more than 2^31 items in a scan");
  size_t tmp = 0;
  NALU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, s));
  DevBuf<unsigned char> t(std::max<size_t>(tmp, 1));
  NALU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, (int)n, s));
}

template <typename T>
T to_host(const T* dev, hipStream_t s) {
  T h{};
  NALU_HIP_CHECK(hipMemcpyAsync(&h, dev, sizeof(T), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  return h;
}

__global__ void k_dot_part(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                           double* __restrict__ part) {
  __shared__ double sh[kT];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) acc += a[i] * b[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_part(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double sh[kT];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kT) acc += part[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// ------------------------------------------------------ finest-level operator
// Row kernels of the set-up run kL lanes per row (coalesced walks of the row's
// entries, shuffle reductions within the lane group).
constexpr int kL = 8;

// Solver-order CSR of the split system: row k = its L part, its diagonal, its
// U part (row_ptr[k] = lrp[k] + urp[k] + k because lrp/urp are prefix sums).
__global__ void k_fine_csr(const int32_t* __restrict__ lrp, const int32_t* __restrict__ lci,
                           const double* __restrict__ lv, const int32_t* __restrict__ urp,
                           const int32_t* __restrict__ uci, const double* __restrict__ uv,
                           const double* __restrict__ dv, int64_t n, int32_t* __restrict__ rp,
                           int32_t* __restrict__ ci, double* __restrict__ v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t k = t / kL;
  const int lane = (int)(t % kL);
  if (k > n) return;
  const int32_t o0 = lrp[k] + urp[k] + (int32_t)k;
  if (lane == 0) rp[k] = o0;
  if (k == n) return;
  const int32_t l0 = lrp[k], nl = lrp[k + 1] - l0;
  for (int32_t q = lane; q < nl; q += kL) {
    ci[o0 + q] = lci[l0 + q];
    v[o0 + q] = lv[l0 + q];
  }
  if (lane == 0) {
    ci[o0 + nl] = (int32_t)k;
    v[o0 + nl] = dv[k];
  }
  const int32_t u0 = urp[k], nu = urp[k + 1] - u0;
  for (int32_t q = lane; q < nu; q += kL) {
    ci[o0 + nl + 1 + q] = uci[u0 + q];
    v[o0 + nl + 1 + q] = uv[u0 + q];
  }
}

// diagonal of a CSR operator; bad = missing, zero or non-finite diagonal
__global__ void k_diag(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                       const double* __restrict__ v, int64_t n, double* __restrict__ d, double* __restrict__ invd,
                       int* __restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / kL;
  const int lane = (int)(t % kL);
  double a = 0.0;
  int found = 0;
  if (i < n)
    for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += kL)
      if (ci[q] == (int32_t)i) {
        a += v[q];
        found = 1;
      }
#pragma unroll
  for (int o = kL / 2; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, kL);
    found |= __shfl_xor(found, o, kL);
  }
  if (i >= n || lane != 0) return;
  if (!found || a == 0.0 || !isfinite(a)) {
    atomicOr(bad, 1);
    a = 1.0;
  }
  d[i] = a;
  invd[i] = 1.0 / a;
}

// ------------------------------------------------------------------ SpMV (CSR)
// L lanes per row.  mode 0: y = A x; 1: y = b - A x; 2: y += A x.
template <typename T, int L>
__global__ void k_csr_spmv(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ v,
                           const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ b, int64_t n,
                           int mode) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / L;
  const int lane = (int)(t % L);
  T acc = T(0);
  if (i < n)
    for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += L) acc += v[q] * x[ci[q]];
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, L);
  if (i < n && lane == 0) {
    if (mode == 0) y[i] = acc;
    else if (mode == 1) y[i] = b[i] - acc;
    else y[i] += acc;
  }
}

template <typename T>
void csr_spmv(const Csr& A, const T* vals, const T* x, T* y, const T* b, int mode, hipStream_t s) {
  if (A.n == 0) return;
  const double avg = A.nnz / (double)A.n;
  if (avg > 48)
    hipLaunchKernelGGL((k_csr_spmv<T, 16>), dim3(nbl(A.n * 16)), dim3(kT), 0, s, A.rp.p, A.ci.p, vals, x, y, b, A.n,
                       mode);
  else if (avg > 12)
    hipLaunchKernelGGL((k_csr_spmv<T, 8>), dim3(nbl(A.n * 8)), dim3(kT), 0, s, A.rp.p, A.ci.p, vals, x, y, b, A.n,
                       mode);
  else
    hipLaunchKernelGGL((k_csr_spmv<T, 2>), dim3(nbl(A.n * 2)), dim3(kT), 0, s, A.rp.p, A.ci.p, vals, x, y, b, A.n,
                       mode);
  NALU_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------- power iteration
__global__ void k_hash_vec(double* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t h = (uint32_t)i * 2654435761u + 0x9e3779b9u;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  x[i] = 0.5 + (double)(h & 0xffffff) / (double)0x1000000;  // (0.5, 1.5): no zero entries
}

// out = invd * y / sqrt(*nx2)
__global__ void k_scale_diag(const double* __restrict__ invd, const double* __restrict__ y,
                             const double* __restrict__ nx2, int64_t n, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = invd[i] * y[i] / sqrt(*nx2);
}

void dot_dev(const double* a, const double* b, int64_t n, double* part, double* out, hipStream_t s) {
  constexpr int kParts = 512;
  const unsigned np = std::min<unsigned>(kParts, nbl(n));
  hipLaunchKernelGGL(k_dot_part, dim3(np), dim3(kT), 0, s, a, b, n, part);
  hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(kT), 0, s, part, (int)np, out);
}

// lambda_max(D^-1 A) by the power method from a fixed pseudo-random start:
// x <- x / |x|, y = D^-1 A x, lambda = x . y (the Rayleigh-quotient form of the
// estimate Ifpack2's Chebyshev setup uses).  The scalars stay on the device;
// one host read at the end.  ls (finest level only): the product runs on the
// solver's SELL-64 copy of the same matrix.
double power_lmax(const Csr& A, const double* invd, int iters, hipStream_t s, nalu_linsys* ls = nullptr) {
  const int64_t n = A.n;
  const int it_n = std::max(1, iters);
  DevBuf<double> x(n), y(n), t(n), part(512), sc(2 * it_n);
  hipLaunchKernelGGL(k_hash_vec, dim3(nbl(n)), dim3(kT), 0, s, x.p, n);
  for (int it = 0; it < it_n; ++it) {
    double* nx2 = sc.p + 2 * it;
    dot_dev(x.p, x.p, n, part.p, nx2, s);
    if (ls) perm_spmv64(ls, x.p, t.p);
    else csr_spmv<double>(A, A.v.p, x.p, t.p, (const double*)nullptr, 0, s);
    hipLaunchKernelGGL(k_scale_diag, dim3(nbl(n)), dim3(kT), 0, s, invd, t.p, nx2, n, y.p);
    dot_dev(x.p, y.p, n, part.p, nx2 + 1, s);
    std::swap(x.p, y.p);
  }
  NALU_HIP_CHECK(hipGetLastError());
  std::vector<double> h(2 * it_n);
  NALU_HIP_CHECK(hipMemcpyAsync(h.data(), sc.p, sizeof(double) * 2 * it_n, hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  // lambda = (x . y) / |x| of the last iteration whose |x| was positive
  double lam = 0.0;
  for (int it = 0; it < it_n; ++it) {
    if (!(h[2 * it] > 0.0) || !std::isfinite(h[2 * it])) break;
    lam = h[2 * it + 1] / std::sqrt(h[2 * it]);
  }
  return lam;
}

// ------------------------------------------------------------ aggregation
// strong[q] = 1 when entry q is a strong off-diagonal connection, and the
// values of S = I - omega D_f^-1 A_f on A's pattern: weak entries 0, the
// filtered diagonal d_f = a_ii + the weak off-diagonal entries (lumping).  A row
// without strong entries, or whose d_f is below 1e-12 |a_ii|, keeps the
// identity row (its tentative prolongator row is not smoothed).
__global__ void k_strength(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                           const double* __restrict__ v, const double* __restrict__ d, int64_t n, double tol,
                           double omega, uint8_t* __restrict__ strong, double* __restrict__ sv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / kL;
  const int lane = (int)(t % kL);
  double l = 0.0;
  int has = 0;
  if (i < n) {
    const double di = d[i];
    for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += kL) {
      const int32_t j = ci[q];
      const double a = v[q];
      bool st = false;
      if (j != (int32_t)i) {
        st = (a != 0.0) && (tol == 0.0 || a * a > tol * tol * fabs(di * d[j]));
        if (!st) l += a;
      }
      strong[q] = st ? 1 : 0;
      has |= st ? 1 : 0;
    }
  }
#pragma unroll
  for (int o = kL / 2; o > 0; o >>= 1) {
    l += __shfl_xor(l, o, kL);
    has |= __shfl_xor(has, o, kL);
  }
  if (i >= n) return;
  const double di = d[i];
  const double df = di + l;
  const double w = (has && fabs(df) > 1e-12 * fabs(di)) ? omega / df : 0.0;
  for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += kL) {
    if (ci[q] == (int32_t)i) sv[q] = 1.0 - w * df;
    else sv[q] = strong[q] ? -w * v[q] : 0.0;
  }
}

__device__ __forceinline__ uint32_t prio_hash(uint32_t i) {
  uint32_t h = i * 0x9e3779b1u;
  h ^= h >> 15;
  h *= 0x2c1b3c6du;
  h ^= h >> 12;
  return h & 0x7fffffffu;
}

// MIS-2 key: class (root 3 > undecided 2 > covered 1) | priority | index
__device__ __forceinline__ uint64_t mis_key(int st, uint32_t i) {
  const uint64_t cls = st == 1 ? 3u : (st == 0 ? 2u : 1u);
  return (cls << 62) | ((uint64_t)prio_hash(i) << 31) | (uint64_t)i;
}

__global__ void k_mis_keys(const int8_t* __restrict__ state, int64_t n, uint64_t* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = mis_key(state[i], (uint32_t)i);
}

__global__ void k_mis_max(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                          const uint8_t* __restrict__ strong, const uint64_t* __restrict__ in, int64_t n,
                          uint64_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / kL;
  const int lane = (int)(t % kL);
  uint64_t m = 0;
  if (i < n) {
    if (lane == 0) m = in[i];
    for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += kL)
      if (strong[q]) m = max(m, in[ci[q]]);
  }
#pragma unroll
  for (int o = kL / 2; o > 0; o >>= 1) m = max(m, (uint64_t)__shfl_xor((unsigned long long)m, o, kL));
  if (i < n && lane == 0) out[i] = m;
}

// undecided row: root when it holds the largest key of its 2-neighbourhood,
// covered when a root lies within two steps; *undecided is set to 1 (one
// store per wave) while rows remain undecided
__global__ void k_mis_update(const uint64_t* __restrict__ key, const uint64_t* __restrict__ m2, int64_t n,
                             int8_t* __restrict__ state, int* __restrict__ undecided) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool left = false;
  if (i < n && state[i] == 0) {
    if (m2[i] == key[i]) state[i] = 1;
    else if ((m2[i] >> 62) == 3u) state[i] = 2;
    else left = true;
  }
  if (__ballot(left) != 0 && (threadIdx.x % 64) == 0) *undecided = 1;
}

__global__ void k_root_flags(const int8_t* __restrict__ state, int64_t n, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) flag[i] = (i < n && state[i] == 1) ? 1 : 0;
}

// best (|a|, key) candidate of a lane group; all lanes end with the group's best
__device__ __forceinline__ void group_best(double& best, uint64_t& bk, int32_t& a) {
#pragma unroll
  for (int o = kL / 2; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, kL);
    const uint64_t ok = (uint64_t)__shfl_xor((unsigned long long)bk, o, kL);
    const int32_t oa = __shfl_xor(a, o, kL);
    if (ob > best || (ob == best && ok > bk)) {
      best = ob;
      bk = ok;
      a = oa;
    }
  }
}

// pass 1: roots get their own aggregate; rows adjacent to roots join the root
// of the strongest connection (ties: the larger priority key)
__global__ void k_agg_pass1(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const double* __restrict__ v, const uint8_t* __restrict__ strong,
                            const int8_t* __restrict__ state, const int32_t* __restrict__ root_id, int64_t n,
                            int32_t* __restrict__ agg) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / kL;
  const int lane = (int)(t % kL);
  const bool live = i < n && state[i] != 1;
  double best = -1.0;
  uint64_t bk = 0;
  int32_t a = -1;
  if (live)
    for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += kL) {
      if (!strong[q]) continue;
      const int32_t j = ci[q];
      if (state[j] != 1) continue;
      const double w = fabs(v[q]);
      const uint64_t kj = mis_key(1, (uint32_t)j);
      if (w > best || (w == best && kj > bk)) {
        best = w;
        bk = kj;
        a = root_id[j];
      }
    }
  group_best(best, bk, a);
  if (i >= n || lane != 0) return;
  agg[i] = live ? a : root_id[i];
}

// pass 2: the remaining rows join the aggregate of their strongest neighbour
// aggregated in pass 1 (reads the pass-1 snapshot only)
__global__ void k_agg_pass2(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const double* __restrict__ v, const uint8_t* __restrict__ strong,
                            const int32_t* __restrict__ agg1, int64_t n, int32_t* __restrict__ agg,
                            int* __restrict__ orphan) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / kL;
  const int lane = (int)(t % kL);
  const bool live = i < n && agg1[i] < 0;
  double best = -1.0;
  uint64_t bk = 0;
  int32_t a = -1;
  if (live)
    for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += kL) {
      if (!strong[q]) continue;
      const int32_t j = ci[q];
      if (agg1[j] < 0) continue;
      const double w = fabs(v[q]);
      const uint64_t kj = mis_key(2, (uint32_t)j);
      if (w > best || (w == best && kj > bk)) {
        best = w;
        bk = kj;
        a = agg1[j];
      }
    }
  group_best(best, bk, a);
  if (i >= n || lane != 0) return;
  if (!live) {
    agg[i] = agg1[i];
    return;
  }
  if (a < 0) atomicOr(orphan, 1);
  agg[i] = a;
}

// ------------------------------------------------------ prolongator pieces
__global__ void k_agg_norm2(const int32_t* __restrict__ agg, const double* __restrict__ B, int64_t n,
                            double* __restrict__ s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(s + agg[i], B[i] * B[i]);
}

__global__ void k_sqrt(double* __restrict__ s, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = sqrt(s[i]);
}

// tentative prolongator (one entry per row) as a CSR
__global__ void k_tentative(const int32_t* __restrict__ agg, const double* __restrict__ B,
                            const double* __restrict__ Bc, int64_t n, int32_t* __restrict__ rp,
                            int32_t* __restrict__ ci, double* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  rp[i] = (int32_t)i;
  if (i == n) return;
  const int32_t a = agg[i];
  ci[i] = a;
  const double nb = Bc[a];
  v[i] = nb > 0.0 ? B[i] / nb : 0.0;
}

// ------------------------------------------------------------------ SpGEMM
// C = A B, one 64-lane block per row (grid-stride), open-addressing hash of H
// (power of two) slots in LDS.  Phase 0 counts each row's distinct columns;
// phase 1 writes them (unsorted) at C.rp[i].  Zero entries of A or B add
// nothing.  A row with more than H distinct columns sets *overflow.
__device__ __forceinline__ uint32_t slot_hash(int32_t c) { return (uint32_t)c * 0x9e3779b1u; }

template <int PHASE>
__global__ __launch_bounds__(64) void k_spgemm(const int32_t* __restrict__ arp, const int32_t* __restrict__ aci,
                                               const double* __restrict__ av, const int32_t* __restrict__ brp,
                                               const int32_t* __restrict__ bci, const double* __restrict__ bv,
                                               int64_t n, int H, const int32_t* __restrict__ crp,
                                               int32_t* __restrict__ cnt, int32_t* __restrict__ cci,
                                               double* __restrict__ cv, int* __restrict__ overflow) {
  extern __shared__ unsigned char smem[];
  double* hv = reinterpret_cast<double*>(smem);
  int32_t* hk = reinterpret_cast<int32_t*>(hv + H);
  const int lane = threadIdx.x;
  const uint32_t mask = (uint32_t)H - 1;
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    for (int s = lane; s < H; s += 64) {
      hk[s] = -1;
      hv[s] = 0.0;
    }
    __syncthreads();
    for (int32_t q = arp[i] + lane; q < arp[i + 1]; q += 64) {
      const double a = av[q];
      if (a == 0.0) continue;
      const int32_t k = aci[q];
      for (int32_t t = brp[k]; t < brp[k + 1]; ++t) {
        const double b = bv[t];
        if (b == 0.0) continue;
        const int32_t c = bci[t];
        uint32_t h = slot_hash(c) & mask;
        int probe = 0;
        for (; probe < H; ++probe) {
          const int32_t old = atomicCAS(&hk[h], -1, c);
          if (old == -1 || old == c) break;
          h = (h + 1) & mask;
        }
        if (probe == H) {
          atomicOr(overflow, 1);
          continue;
        }
        if (PHASE == 1) atomicAdd(&hv[h], a * b);
      }
    }
    __syncthreads();
    int base = 0;
    for (int s0 = 0; s0 < H; s0 += 64) {
      const int s = s0 + lane;
      const bool occ = s < H && hk[s] != -1;
      const uint64_t bal = __ballot(occ);
      if (PHASE == 1 && occ) {
        const int off = __popcll(bal & ((1ull << lane) - 1ull));
        const int64_t o = (int64_t)crp[i] + base + off;
        cci[o] = hk[s];
        cv[o] = hv[s];
      }
      base += __popcll(bal);
    }
    if (PHASE == 0 && lane == 0) cnt[i] = base;
    __syncthreads();
  }
}

// 64-bit total of int32 counts (block sums, one atomic per block)
__global__ void k_count_total(const int32_t* __restrict__ cnt, int64_t n, unsigned long long* __restrict__ total) {
  __shared__ unsigned long long sh[kT];
  unsigned long long acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    acc += (unsigned long long)cnt[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(total, sh[0]);
}

__global__ void k_row_bound(const int32_t* __restrict__ arp, const int32_t* __restrict__ aci,
                            const int32_t* __restrict__ brp, int64_t n, int32_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / kL;
  const int lane = (int)(t % kL);
  int64_t sum = 0;
  if (i < n)
    for (int32_t q = arp[i] + lane; q < arp[i + 1]; q += kL) sum += brp[aci[q] + 1] - brp[aci[q]];
#pragma unroll
  for (int o = kL / 2; o > 0; o >>= 1) sum += __shfl_xor((long long)sum, o, kL);
  if (i < n && lane == 0) out[i] = (int32_t)std::min<int64_t>(sum, 1 << 30);
}

int32_t dev_max(const int32_t* a, int64_t n, hipStream_t s) {
  NALU_REQUIRE(n > 0 && n < (int64_t)INT32_MAX, NALU_ERR_UNSUPPORTED, "amg: reduction size out of range");
  DevBuf<int32_t> mx(1);
  size_t tmp = 0;
  NALU_HIP_CHECK(hipcub::DeviceReduce::Max(nullptr, tmp, a, mx.p, (int)n, s));
  DevBuf<unsigned char> t(std::max<size_t>(tmp, 1));
  NALU_HIP_CHECK(hipcub::DeviceReduce::Max(t.p, tmp, a, mx.p, (int)n, s));
  return to_host(mx.p, s);
}

constexpr int kFirstHash = 1024;
constexpr int kMaxHash = 4096;

Csr spgemm(const Csr& A, const double* av, const Csr& B, hipStream_t s) {
  Csr C;
  C.n = A.n;
  C.m = B.m;
  C.rp.alloc(A.n + 1);
  if (A.n == 0) {
    NALU_HIP_CHECK(hipMemsetAsync(C.rp.p, 0, sizeof(int32_t), s));
    return C;
  }
  DevBuf<int32_t> cnt(A.n + 1);
  hipLaunchKernelGGL(k_row_bound, dim3(nbl(A.n * kL)), dim3(kT), 0, s, A.rp.p, A.ci.p, B.rp.p, A.n, cnt.p);
  const int64_t bound = std::min<int64_t>(dev_max(cnt.p, A.n, s), B.m);
  // the bound (sum of the B rows met) overestimates the distinct columns of a
  // product row several-fold: start with at most kFirstHash slots and grow to
  // kMaxHash only when a row overflows
  int H = 64;
  while (H < 2 * bound && H < kFirstHash) H <<= 1;
  DevBuf<int> ovf(1);
  const unsigned grid = (unsigned)std::min<int64_t>(A.n, 1 << 20);
  for (;;) {
    NALU_HIP_CHECK(hipMemsetAsync(ovf.p, 0, sizeof(int), s));
    NALU_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (A.n + 1), s));
    const size_t lds = (size_t)H * (sizeof(double) + sizeof(int32_t));
    hipLaunchKernelGGL(k_spgemm<0>, dim3(grid), dim3(64), lds, s, A.rp.p, A.ci.p, av, B.rp.p, B.ci.p, B.v.p, A.n,
                       H, (const int32_t*)nullptr, cnt.p, (int32_t*)nullptr, (double*)nullptr, ovf.p);
    NALU_HIP_CHECK(hipGetLastError());
    if (to_host(ovf.p, s) == 0) break;
    NALU_REQUIRE(H < kMaxHash && H < 2 * bound, NALU_ERR_UNSUPPORTED,
                 "amg: a product row has more than " + std::to_string(H) + " distinct columns");
    while (H < 2 * bound && H < kMaxHash) H <<= 1;
  }
  const size_t lds = (size_t)H * (sizeof(double) + sizeof(int32_t));
  {
    DevBuf<unsigned long long> tot(1);
    NALU_HIP_CHECK(hipMemsetAsync(tot.p, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_count_total, dim3(std::min<unsigned>(512, nbl(A.n))), dim3(kT), 0, s, cnt.p, A.n, tot.p);
    NALU_REQUIRE(to_host(tot.p, s) < (unsigned long long)INT32_MAX, NALU_ERR_UNSUPPORTED,
                 "amg: a product has more than 2^31 entries (32-bit row pointers)");
  }
  scan_excl(cnt.p, C.rp.p, A.n + 1, s);
  C.nnz = to_host(C.rp.p + A.n, s);
  C.max_row = dev_max(cnt.p, A.n, s);
  C.ci.alloc(std::max<int64_t>(C.nnz, 1));
  C.v.alloc(std::max<int64_t>(C.nnz, 1));
  hipLaunchKernelGGL(k_spgemm<1>, dim3(grid), dim3(64), lds, s, A.rp.p, A.ci.p, av, B.rp.p, B.ci.p, B.v.p, A.n, H,
                     C.rp.p, cnt.p, C.ci.p, C.v.p, ovf.p);
  NALU_HIP_CHECK(hipGetLastError());
  return C;
}

// --------------------------------------------------------------- transpose
__global__ void k_tr_keys(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, int64_t n,
                          uint64_t* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int32_t q = rp[i]; q < rp[i + 1]; ++q) key[q] = ((uint64_t)(uint32_t)ci[q] << 32) | (uint32_t)i;
}

// rp[c] = first position of column-major row c in the sorted keys
__global__ void k_tr_rows(const uint64_t* __restrict__ key, int64_t nnz, int64_t m, int32_t* __restrict__ rp,
                          int32_t* __restrict__ ci) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p > nnz) return;
  const int64_t c = p < nnz ? (int64_t)(key[p] >> 32) : m;
  const int64_t cprev = p > 0 ? (int64_t)(key[p - 1] >> 32) : -1;
  for (int64_t r = cprev + 1; r <= c; ++r) rp[r] = (int32_t)p;
  if (p < nnz) ci[p] = (int32_t)(key[p] & 0xffffffffu);
}

// R = P^T, rows sorted by column (deterministic): one radix sort of
// (column, row) keys carrying the values
Csr transpose(const Csr& P, hipStream_t s) {
  NALU_REQUIRE(P.nnz < (int64_t)INT32_MAX, NALU_ERR_UNSUPPORTED, "amg: prolongator has more than 2^31 entries");
  Csr R;
  R.n = P.m;
  R.m = P.n;
  R.nnz = P.nnz;
  R.rp.alloc(R.n + 1);
  R.ci.alloc(std::max<int64_t>(P.nnz, 1));
  R.v.alloc(std::max<int64_t>(P.nnz, 1));
  DevBuf<uint64_t> k0(std::max<int64_t>(P.nnz, 1)), k1(std::max<int64_t>(P.nnz, 1));
  hipLaunchKernelGGL(k_tr_keys, dim3(nbl(P.n)), dim3(kT), 0, s, P.rp.p, P.ci.p, P.n, k0.p);
  int bits = 1;
  while ((int64_t(1) << bits) < std::max<int64_t>(P.m, 2)) ++bits;
  size_t tmp = 0;
  NALU_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k0.p, k1.p, P.v.p, R.v.p, (int)P.nnz, 0, 32 + bits, s));
  DevBuf<unsigned char> t(std::max<size_t>(tmp, 1));
  NALU_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, k0.p, k1.p, P.v.p, R.v.p, (int)P.nnz, 0, 32 + bits, s));
  hipLaunchKernelGGL(k_tr_rows, dim3(nbl(P.nnz + 1)), dim3(kT), 0, s, k1.p, P.nnz, R.n, R.rp.p, R.ci.p);
  return R;
}

// ------------------------------------------------------ dense coarse solve
// sum of squares of each row (R rows = P columns)
__global__ void k_row_sumsq(const int32_t* __restrict__ rp, const double* __restrict__ v, int64_t n,
                            double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int32_t q = rp[i]; q < rp[i + 1]; ++q) s += v[q] * v[q];
  out[i] = s;
}

__global__ void k_dense_fill(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                             const double* __restrict__ v, int64_t n, double* __restrict__ M) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* row = M + i * 2 * n;
  for (int64_t j = 0; j < 2 * n; ++j) row[j] = (j == n + i) ? 1.0 : 0.0;
  for (int32_t q = rp[i]; q < rp[i + 1]; ++q) row[ci[q]] += v[q];
}

// M[:, :n] -= c y B^T (the residual rank-one part of a near-singular operator)
__global__ void k_rank1_sub(double* __restrict__ M, const double* __restrict__ y, const double* __restrict__ B,
                            int64_t n, double c) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int64_t i = t / n, j = t % n;
  M[i * 2 * n + j] -= c * y[i] * B[j];
}

// One Gauss-Jordan step as a single launch, out of place (Min -> Mout, no
// block reads what another writes): every block (one row each) finds the same
// pivot p of column k (largest |value| in rows >= k, ties to the smaller row),
// row k of Mout is row p / pivot, row p takes the old row k, and every other
// row eliminates column k with the normalized pivot row.  A pivot at or below
// tiny leaves Min unchanged in Mout and is recorded as skipped (piv[k] = 0).
__global__ void k_gj_step(const double* __restrict__ Min, double* __restrict__ Mout, int64_t n, int64_t k,
                          double tiny, int8_t* __restrict__ piv) {
  __shared__ double bv[kT];
  __shared__ int64_t bi[kT];
  const int64_t w = 2 * n;
  const int64_t i = blockIdx.x;
  double best = -1.0;
  int64_t arg = k;
  for (int64_t r = k + threadIdx.x; r < n; r += kT) {
    const double a = fabs(Min[r * w + k]);
    if (a > best) {
      best = a;
      arg = r;
    }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = arg;
  __syncthreads();
  for (int sft = kT / 2; sft > 0; sft >>= 1) {
    if ((int)threadIdx.x < sft) {
      const double o = bv[threadIdx.x + sft];
      const int64_t oi = bi[threadIdx.x + sft];
      if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
        bv[threadIdx.x] = o;
        bi[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  const int64_t p = bi[0];
  const bool ok = bv[0] > tiny;
  if (i == 0 && threadIdx.x == 0) piv[k] = ok ? 1 : 0;
  if (!ok) {
    for (int64_t j = threadIdx.x; j < w; j += kT) Mout[i * w + j] = Min[i * w + j];
    return;
  }
  const double pv = Min[p * w + k];
  if (i == k) {
    for (int64_t j = threadIdx.x; j < w; j += kT) Mout[k * w + j] = Min[p * w + j] / pv;
    return;
  }
  const int64_t src = (i == p) ? k : i;
  const double f = Min[src * w + k];
  for (int64_t j = threadIdx.x; j < w; j += kT) Mout[i * w + j] = Min[src * w + j] - f * (Min[p * w + j] / pv);
}

__global__ void k_gj_extract(const double* __restrict__ M, const int8_t* __restrict__ piv, int64_t n,
                             double* __restrict__ X, float* __restrict__ X32) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int64_t i = t / n, j = t % n;
  const double x = piv[i] ? M[i * 2 * n + n + j] : 0.0;
  X[t] = x;
  X32[t] = (float)x;
}

template <typename T>
__global__ void k_dense_apply(const T* __restrict__ X, const T* __restrict__ b, int64_t n, T* __restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / 64;
  const int lane = (int)(t % 64);
  T acc = T(0);
  if (i < n)
    for (int64_t j = lane; j < n; j += 64) acc += X[i * n + j] * b[j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (i < n && lane == 0) x[i] = acc;
}

// ------------------------------------------------------------- smoothing
template <typename T>
__global__ void k_cheb_first(const T* __restrict__ invd, const T* __restrict__ r, T c, int64_t n,
                             T* __restrict__ d, T* __restrict__ x, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T di = c * invd[i] * r[i];
  d[i] = di;
  x[i] = add ? x[i] + di : di;
}

template <typename T>
__global__ void k_cheb_step(const T* __restrict__ invd, const T* __restrict__ r, T c0, T c1, int64_t n,
                            T* __restrict__ d, T* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T di = c0 * d[i] + c1 * invd[i] * r[i];
  d[i] = di;
  x[i] += di;
}

template <typename T>
__global__ void k_fill(T* __restrict__ x, T v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

template <typename T>
__global__ void k_to32(const double* __restrict__ a, int64_t n, T* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}

}  // namespace

// ====================================================================== setup
void amg_release(nalu_linsys* ls) {
  delete amg_of(ls);
  ls->amg = nullptr;
}

void amg_setup(nalu_linsys* ls) {
  NALU_REQUIRE(ls->num_dof == 1, NALU_ERR_UNSUPPORTED, "amg preconditioner: scalar systems (numDof 1) only");
  NALU_REQUIRE(ls->nrows > 0, NALU_ERR_STATE, "amg preconditioner: empty system");
  const nalu_muelu_params& pr = ls->muelu;
  hipStream_t s = ls->stream;
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  const auto t0 = std::chrono::steady_clock::now();
  amg_release(ls);
  std::unique_ptr<Amg> holder(new Amg());  // installed in the handle once complete
  Amg* amg = holder.get();
  const int64_t n0 = ls->nrows;

  // finest operator in solver order
  Level L0;
  L0.A.n = L0.A.m = n0;
  L0.A.nnz = (int64_t)to_host(ls->lrp.p + n0, s) + (int64_t)to_host(ls->urp.p + n0, s) + n0;
  L0.A.rp.alloc(n0 + 1);
  L0.A.ci.alloc(L0.A.nnz);
  L0.A.v.alloc(L0.A.nnz);
  hipLaunchKernelGGL(k_fine_csr, dim3(nbl((n0 + 1) * kL)), dim3(kT), 0, s, ls->lrp.p, ls->lci.p, ls->lval.p, ls->urp.p,
                     ls->uci.p, ls->uval.p, ls->dval.p, n0, L0.A.rp.p, L0.A.ci.p, L0.A.v.p);
  NALU_HIP_CHECK(hipGetLastError());
  L0.A.max_row = ls->max_l_len + ls->max_u_len + 1;
  amg->lev.push_back(std::move(L0));

  DevBuf<double> B(n0);
  hipLaunchKernelGGL(k_fill<double>, dim3(nbl(n0)), dim3(kT), 0, s, B.p, 1.0, n0);
  DevBuf<int> flag(1);
  bool stalled = false;
  for (int l = 0;; ++l) {
    Level& L = amg->lev[l];
    const int64_t n = L.A.n;
    L.diag.alloc(n);
    L.invd.alloc(n);
    NALU_HIP_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_diag, dim3(nbl(n * kL)), dim3(kT), 0, s, L.A.rp.p, L.A.ci.p, L.A.v.p, n, L.diag.p, L.invd.p,
                       flag.p);
    NALU_REQUIRE(to_host(flag.p, s) == 0, NALU_ERR_NUMERIC,
                 "amg: level " + std::to_string(l) + " has a zero, missing or non-finite diagonal");
    L.lmax = power_lmax(L.A, L.invd.p, pr.cheb_eig_iters, s, l == 0 ? ls : nullptr);
    NALU_REQUIRE(L.lmax > 0.0 && std::isfinite(L.lmax), NALU_ERR_NUMERIC,
                 "amg: non-positive eigenvalue estimate on level " + std::to_string(l));
    L.work.alloc(4 * n);
    const bool last = stalled || n <= pr.max_coarse_size || l + 1 >= pr.max_levels;
    if (last) break;

    // strength, MIS-2 roots, aggregates
    DevBuf<uint8_t> strong(std::max<int64_t>(L.A.nnz, 1));
    DevBuf<double> sv(std::max<int64_t>(L.A.nnz, 1));  // values of I - omega D_f^-1 A_f
    const double omega = pr.sa_damping / L.lmax;
    hipLaunchKernelGGL(k_strength, dim3(nbl(n * kL)), dim3(kT), 0, s, L.A.rp.p, L.A.ci.p, L.A.v.p, L.diag.p, n,
                       pr.drop_tol, omega, strong.p, sv.p);
    DevBuf<int8_t> state(n);
    NALU_HIP_CHECK(hipMemsetAsync(state.p, 0, n, s));
    DevBuf<uint64_t> key(n), m1(n), m2(n);
    DevBuf<int> und(1);
    int rounds = 0;
    for (;; ++rounds) {
      NALU_REQUIRE(rounds <= 4096, NALU_ERR_NUMERIC, "amg: aggregation did not settle");
      hipLaunchKernelGGL(k_mis_keys, dim3(nbl(n)), dim3(kT), 0, s, state.p, n, key.p);
      hipLaunchKernelGGL(k_mis_max, dim3(nbl(n * kL)), dim3(kT), 0, s, L.A.rp.p, L.A.ci.p, strong.p, key.p, n, m1.p);
      hipLaunchKernelGGL(k_mis_max, dim3(nbl(n * kL)), dim3(kT), 0, s, L.A.rp.p, L.A.ci.p, strong.p, m1.p, n, m2.p);
      NALU_HIP_CHECK(hipMemsetAsync(und.p, 0, sizeof(int), s));
      hipLaunchKernelGGL(k_mis_update, dim3(nbl(n)), dim3(kT), 0, s, key.p, m2.p, n, state.p, und.p);
      if (to_host(und.p, s) == 0) break;
    }
    DevBuf<int32_t> rflag(n + 1), rid(n + 1);
    hipLaunchKernelGGL(k_root_flags, dim3(nbl(n + 1)), dim3(kT), 0, s, state.p, n, rflag.p);
    scan_excl(rflag.p, rid.p, n + 1, s);
    const int64_t nc = to_host(rid.p + n, s);
    NALU_REQUIRE(nc > 0, NALU_ERR_NUMERIC, "amg: no aggregates");
    DevBuf<int32_t> agg1(n);
    L.agg.alloc(n);
    hipLaunchKernelGGL(k_agg_pass1, dim3(nbl(n * kL)), dim3(kT), 0, s, L.A.rp.p, L.A.ci.p, L.A.v.p, strong.p, state.p,
                       rid.p, n, agg1.p);
    NALU_HIP_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_agg_pass2, dim3(nbl(n * kL)), dim3(kT), 0, s, L.A.rp.p, L.A.ci.p, L.A.v.p, strong.p, agg1.p, n,
                       L.agg.p, flag.p);
    NALU_REQUIRE(to_host(flag.p, s) == 0, NALU_ERR_NUMERIC, "amg: a row was left without an aggregate");

    // tentative and smoothed prolongators
    DevBuf<double> Bc(nc);
    NALU_HIP_CHECK(hipMemsetAsync(Bc.p, 0, sizeof(double) * nc, s));
    hipLaunchKernelGGL(k_agg_norm2, dim3(nbl(n)), dim3(kT), 0, s, L.agg.p, B.p, n, Bc.p);
    hipLaunchKernelGGL(k_sqrt, dim3(nbl(nc)), dim3(kT), 0, s, Bc.p, nc);
    Csr Pt;
    Pt.n = n;
    Pt.m = nc;
    Pt.nnz = n;
    Pt.max_row = 1;
    Pt.rp.alloc(n + 1);
    Pt.ci.alloc(n);
    Pt.v.alloc(n);
    hipLaunchKernelGGL(k_tentative, dim3(nbl(n + 1)), dim3(kT), 0, s, L.agg.p, B.p, Bc.p, n, Pt.rp.p, Pt.ci.p,
                       Pt.v.p);
    L.P = spgemm(L.A, sv.p, Pt, s);
    L.R = transpose(L.P, s);
    // A_c = R (A P); the other association, (R A) P, forms fewer partial
    // products but its ~500-column R A rows load the LDS hash tables to half
    // full and measured 2.2x slower
    Csr AP = spgemm(L.A, L.A.v.p, L.P, s);
    Level next;
    next.A = spgemm(L.R, L.R.v.p, AP, s);
    next.A.m = nc;
    amg->lev.push_back(std::move(next));
    B = std::move(Bc);
    stalled = nc >= n;  // no coarsening: the new level is the last one
  }

  // dense (pseudo-)inverse of the coarsest operator
  Level& Lc = amg->lev.back();
  const int64_t nc = Lc.A.n;
  NALU_REQUIRE(nc <= kDenseCoarseMax, NALU_ERR_UNSUPPORTED,
               "amg: coarsest level has " + std::to_string(nc) + " rows (dense solve limit " +
                   std::to_string(kDenseCoarseMax) + ")");
  amg->nc = nc;
  {
    DevBuf<double> M(2 * nc * nc);
    DevBuf<int8_t> piv(nc);
    hipLaunchKernelGGL(k_dense_fill, dim3(nbl(nc)), dim3(kT), 0, s, Lc.A.rp.p, Lc.A.ci.p, Lc.A.v.p, nc, M.p);
    auto host_absmax = [&](const double* d, int64_t m) {
      std::vector<double> h(m);
      NALU_HIP_CHECK(hipMemcpyAsync(h.data(), d, sizeof(double) * m, hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      double mx = 0.0;
      for (double x : h) mx = std::max(mx, std::fabs(x));
      return mx;
    };
    double scale = host_absmax(Lc.diag.p, nc);
    if (amg->lev.size() >= 2) {
      const Level& Lp = amg->lev[amg->lev.size() - 2];
      DevBuf<double> sq(nc);
      hipLaunchKernelGGL(k_row_sumsq, dim3(nbl(nc)), dim3(kT), 0, s, Lp.R.rp.p, Lp.R.v.p, nc, sq.p);
      scale = std::max(scale, host_absmax(Lp.diag.p, Lp.A.n) * host_absmax(sq.p, nc));
    }
    const double tiny = 1e-10 * scale;
    NALU_REQUIRE((int64_t)B.n == nc, NALU_ERR_STATE, "amg: near-null vector size mismatch");
    {
      DevBuf<double> y(nc);
      csr_spmv<double>(Lc.A, Lc.A.v.p, B.p, y.p, (const double*)nullptr, 0, s);
      std::vector<double> yh(nc), bh(nc);
      NALU_HIP_CHECK(hipMemcpyAsync(yh.data(), y.p, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipMemcpyAsync(bh.data(), B.p, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      double yy = 0.0, bb = 0.0;
      for (int64_t i = 0; i < nc; ++i) {
        yy += yh[i] * yh[i];
        bb += bh[i] * bh[i];
      }
      amg->null_rel = (scale > 0.0 && bb > 0.0) ? std::sqrt(yy) / (scale * std::sqrt(bb)) : 1.0;
      amg->deflated = amg->null_rel <= 1e-5;
      if (amg->deflated)
        hipLaunchKernelGGL(k_rank1_sub, dim3(nbl(nc * nc)), dim3(kT), 0, s, M.p, y.p, B.p, nc, 1.0 / bb);
    }
    DevBuf<double> M2(2 * nc * nc);
    double* cur = M.p;
    double* nxt = M2.p;
    for (int64_t k = 0; k < nc; ++k) {
      hipLaunchKernelGGL(k_gj_step, dim3((unsigned)nc), dim3(kT), 0, s, cur, nxt, nc, k, tiny, piv.p);
      std::swap(cur, nxt);
    }
    NALU_HIP_CHECK(hipGetLastError());
    amg->cinv.alloc(nc * nc);
    amg->cinv32.alloc(nc * nc);
    hipLaunchKernelGGL(k_gj_extract, dim3(nbl(nc * nc)), dim3(kT), 0, s, cur, piv.p, nc, amg->cinv.p,
                       amg->cinv32.p);
    std::vector<int8_t> ph(nc);
    NALU_HIP_CHECK(hipMemcpyAsync(ph.data(), piv.p, nc, hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    amg->nfree = (int)std::count(ph.begin(), ph.end(), (int8_t)0);
  }
  NALU_HIP_CHECK(hipGetLastError());
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  amg->setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  ls->amg = holder.release();
}

namespace {
// FP32 copies of the level operators (mixed-precision inner solve)
void ensure_amg32(nalu_linsys* ls) {
  Amg* amg = amg_of(ls);
  if (amg->f32_ready) return;
  hipStream_t s = ls->stream;
  auto conv = [&](Csr& C) {
    if (C.nnz == 0) return;
    C.v32.alloc(C.nnz);
    hipLaunchKernelGGL(k_to32<float>, dim3(nbl(C.nnz)), dim3(kT), 0, s, C.v.p, C.nnz, C.v32.p);
  };
  for (size_t l = 0; l < amg->lev.size(); ++l) {
    Level& L = amg->lev[l];
    if (l > 0) conv(L.A);
    conv(L.P);
    conv(L.R);
    L.invd32.alloc(L.A.n);
    hipLaunchKernelGGL(k_to32<float>, dim3(nbl(L.A.n)), dim3(kT), 0, s, L.invd.p, L.A.n, L.invd32.p);
  }
  NALU_HIP_CHECK(hipGetLastError());
  amg->f32_ready = true;
}

template <typename T>
const T* vals_of(const Csr& C) {
  if constexpr (std::is_same<T, double>::value) return C.v.p;
  else return C.v32.p;
}

template <typename T>
const T* invd_of(nalu_linsys* ls, const Level& L, int l) {
  if (l == 0) {
    if constexpr (std::is_same<T, double>::value) return ls->pinvd.p;
    else return ls->pinvd32.p;
  }
  if constexpr (std::is_same<T, double>::value) return L.invd.p;
  else return L.invd32.p;
}

// r = b - A_l x
template <typename T>
void residual(nalu_linsys* ls, const Level& L, int l, const T* x, T* r, const T* b) {
  if (l == 0) perm_residual<T>(ls, x, r, b);
  else csr_spmv<T>(L.A, vals_of<T>(L.A), x, r, b, 1, ls->stream);
}

// Chebyshev smoothing of A_l x = b (Saad Alg. 12.1 on D^-1 A): zero_start
// skips the first residual (x = 0 on entry)
template <typename T>
void chebyshev(nalu_linsys* ls, const Level& L, int l, const T* b, T* x, T* r, T* d, bool zero_start) {
  const nalu_muelu_params& pr = ls->muelu;
  hipStream_t s = ls->stream;
  const int64_t n = L.A.n;
  const double lmax = L.lmax * pr.cheb_boost, lmin = lmax / pr.cheb_ratio;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  const T* invd = invd_of<T>(ls, L, l);
  if (zero_start) {
    hipLaunchKernelGGL(k_cheb_first<T>, dim3(nbl(n)), dim3(kT), 0, s, invd, b, (T)(1.0 / theta), n, d, x, 0);
  } else {
    residual<T>(ls, L, l, x, r, b);
    hipLaunchKernelGGL(k_cheb_first<T>, dim3(nbl(n)), dim3(kT), 0, s, invd, r, (T)(1.0 / theta), n, d, x, 1);
  }
  for (int k = 1; k < pr.cheb_degree; ++k) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    residual<T>(ls, L, l, x, r, b);
    hipLaunchKernelGGL(k_cheb_step<T>, dim3(nbl(n)), dim3(kT), 0, s, invd, r, (T)(rn * rho), (T)(2.0 * rn / delta),
                       n, d, x);
    rho = rn;
  }
}

template <typename T>
void vcycle(nalu_linsys* ls, int l, const T* b, T* x) {
  Amg* amg = amg_of(ls);
  hipStream_t s = ls->stream;
  const Level& L = amg->lev[l];
  const int64_t n = L.A.n;
  if (l + 1 == (int)amg->lev.size()) {
    const T* X;
    if constexpr (std::is_same<T, double>::value) X = amg->cinv.p;
    else X = amg->cinv32.p;
    hipLaunchKernelGGL(k_dense_apply<T>, dim3(nbl(n * 64)), dim3(kT), 0, s, X, b, n, x);
    return;
  }
  T* w = reinterpret_cast<T*>(const_cast<double*>(L.work.p));
  T* r = w + 2 * n;
  T* d = w + 3 * n;
  chebyshev<T>(ls, L, l, b, x, r, d, true);
  residual<T>(ls, L, l, x, r, b);
  const Level& C = amg->lev[l + 1];
  T* wc = reinterpret_cast<T*>(const_cast<double*>(C.work.p));
  T* xc = wc;
  T* bc = wc + C.A.n;
  csr_spmv<T>(L.R, vals_of<T>(L.R), r, bc, (const T*)nullptr, 0, s);
  vcycle<T>(ls, l + 1, bc, xc);
  csr_spmv<T>(L.P, vals_of<T>(L.P), xc, x, (const T*)nullptr, 2, s);
  chebyshev<T>(ls, L, l, b, x, r, d, false);
}
}  // namespace

template <typename T>
void amg_apply(nalu_linsys* ls, const T* r, T* z) {
  NALU_REQUIRE(ls->amg, NALU_ERR_STATE, "amg preconditioner: not set up");
  if constexpr (std::is_same<T, float>::value) {
    ensure_split32(ls);
    ensure_amg32(ls);
  }
  vcycle<T>(ls, 0, r, z);
  NALU_HIP_CHECK(hipGetLastError());
}
template void amg_apply<double>(nalu_linsys*, const double*, double*);
template void amg_apply<float>(nalu_linsys*, const float*, float*);

}  // namespace nalu

// ===================================================================== C ABI
using namespace nalu;

extern "C" int nalu_linsys_set_muelu_params(nalu_linsys* ls, const nalu_muelu_params* p) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && p, NALU_ERR_ARG, "set_muelu_params: null argument");
    NALU_REQUIRE(p->max_coarse_size >= 1 && p->max_coarse_size <= kDenseCoarseMax, NALU_ERR_ARG,
                 "set_muelu_params: coarse: max size must be in [1, " + std::to_string(kDenseCoarseMax) + "]");
    NALU_REQUIRE(p->max_levels >= 1 && p->max_levels <= 32, NALU_ERR_ARG, "set_muelu_params: max levels in [1, 32]");
    NALU_REQUIRE(p->cheb_degree >= 1 && p->cheb_degree <= 16, NALU_ERR_ARG,
                 "set_muelu_params: chebyshev: degree in [1, 16]");
    NALU_REQUIRE(p->cheb_ratio > 1.0 && p->cheb_boost >= 1.0 && p->cheb_eig_iters >= 1 && p->drop_tol >= 0.0 &&
                     p->sa_damping >= 0.0,
                 NALU_ERR_ARG, "set_muelu_params: parameter out of range");
    ls->muelu = *p;
    ls->precond_ready = false;
    return NALU_OK;
  });
}

extern "C" int nalu_amg_info(const nalu_linsys* ls, int* num_levels, int64_t* rows, int64_t* nnz, double* lmax,
                             int max_levels, double* setup_s, int* coarse_free, double* coarse_null_rel) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && num_levels, NALU_ERR_ARG, "amg_info: null argument");
    const Amg* amg = amg_of(ls);
    NALU_REQUIRE(amg, NALU_ERR_STATE, "amg_info: no hierarchy (preconditioner not set up with muelu)");
    *num_levels = (int)amg->lev.size();
    for (int l = 0; l < (int)amg->lev.size() && l < max_levels; ++l) {
      if (rows) rows[l] = amg->lev[l].A.n;
      if (nnz) nnz[l] = amg->lev[l].A.nnz;
      if (lmax) lmax[l] = amg->lev[l].lmax;
    }
    if (setup_s) *setup_s = amg->setup_s;
    if (coarse_free) *coarse_free = amg->nfree;
    if (coarse_null_rel) *coarse_null_rel = amg->null_rel;
    return NALU_OK;
  });
}

extern "C" int nalu_amg_level_csr(const nalu_linsys* ls, int level, int which, int32_t* rp, int32_t* ci, double* v,
                                  int64_t* n, int64_t* m, int64_t* nnz) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls, NALU_ERR_ARG, "amg_level_csr: null handle");
    const Amg* amg = amg_of(ls);
    NALU_REQUIRE(amg, NALU_ERR_STATE, "amg_level_csr: no hierarchy");
    NALU_REQUIRE(level >= 0 && level < (int)amg->lev.size(), NALU_ERR_ARG, "amg_level_csr: level out of range");
    NALU_REQUIRE(which >= 0 && which <= 1, NALU_ERR_ARG, "amg_level_csr: which is 0 (A) or 1 (P)");
    const Level& L = amg->lev[level];
    NALU_REQUIRE(which == 0 || level + 1 < (int)amg->lev.size(), NALU_ERR_ARG,
                 "amg_level_csr: the coarsest level has no prolongator");
    const Csr& C = which == 0 ? L.A : L.P;
    if (n) *n = C.n;
    if (m) *m = C.m;
    if (nnz) *nnz = C.nnz;
    hipStream_t s = ls->stream;
    if (rp) NALU_HIP_CHECK(hipMemcpyAsync(rp, C.rp.p, sizeof(int32_t) * (C.n + 1), hipMemcpyDeviceToHost, s));
    if (ci && C.nnz) NALU_HIP_CHECK(hipMemcpyAsync(ci, C.ci.p, sizeof(int32_t) * C.nnz, hipMemcpyDeviceToHost, s));
    if (v && C.nnz) NALU_HIP_CHECK(hipMemcpyAsync(v, C.v.p, sizeof(double) * C.nnz, hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    return NALU_OK;
  });
}

extern "C" int nalu_amg_aggregates(const nalu_linsys* ls, int level, int32_t* agg) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls && agg, NALU_ERR_ARG, "amg_aggregates: null argument");
    const Amg* amg = amg_of(ls);
    NALU_REQUIRE(amg, NALU_ERR_STATE, "amg_aggregates: no hierarchy");
    NALU_REQUIRE(level >= 0 && level + 1 < (int)amg->lev.size(), NALU_ERR_ARG,
                 "amg_aggregates: level out of range (the coarsest level is not aggregated)");
    const Level& L = amg->lev[level];
    NALU_HIP_CHECK(hipMemcpyAsync(agg, L.agg.p, sizeof(int32_t) * L.A.n, hipMemcpyDeviceToHost, ls->stream));
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    return NALU_OK;
  });
}
