// Learned message-passing preconditioner for the pressure-Poisson system.
//
// No counterpart in the reference: this sits beside the Ifpack2 SGS
// preconditioner (src/LinearSolverConfig.C:71-85) as an option of the flexible
// GMRES solve, with SGS as the automatic fallback.
//
// Network (per node i of one right-hand side; d_i = a_ii, dbar = mean(d),
// delta_i = dbar / d_i, s = rms(r)):
//   features  f = (r_i delta_i / s, zs_i dbar / s, delta_i), zs = SGS(r)
//   encoder   h0 = W_in f + b_in                               3 -> 32
//   layer l   m_i = sum_{j != i} w_ij h_j,  w_ij = -a_ij / a_ii (BF16)
//             h' = h + relu(W_l [h ; m] + b_l)                 64 -> 32
//   output    out = w_out . h + c0 f0 + c1 f1,  z = out s / dbar, minus its
//             mean when every row of A sums to zero (pure Neumann: the loss
//             below cannot see the constant part of the output)
// Weights are BF16, hidden states are stored as BF16, every accumulation is
// FP32; the 64 -> 32 layer transforms run on the MFMA units
// (v_mfma_f32_32x32x16_bf16, one 32-node tile per wavefront).  With w_out = 0,
// c = (0, 1) the network returns SGS(r) exactly, which is the initial state.
//
// Training minimises, per harvested pair (r, e) with A e = r, the A-norm
// relative error of the optimally scaled output direction:
//   L = 1 - <out, r>^2 / (<out, A out> <e, r>),  A = D (I - W),
// with Adam on the FP32 master weights.  The pairs are residuals and errors of
// plain FP64 CG + SGS iterations on seeded Gaussian zero-mean right-hand sides.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <vector>

#include "internal.hpp"

namespace nalu {
namespace {

constexpr int kH = 32;        // hidden width
constexpr int kIn = 3;        // input features
constexpr int kThreads = 256;
constexpr int kWaves = 4;     // wavefronts per block in the tile kernels
constexpr int kTile = 32;     // nodes per wavefront tile (MFMA M dimension)
constexpr int kLdA = 2 * kH + 8;  // LDS row stride (bf16) of the [h ; m] tile
constexpr int kLdD = kH + 8;      // LDS row stride (bf16) of the delta tile
constexpr int kLayerParams = kH * 2 * kH + kH;
constexpr int kRedBlocks = 512;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

inline unsigned nblocks(int64_t n, int per = kThreads) {
  return (unsigned)std::max<int64_t>(1, (n + per - 1) / per);
}

// Parameter vector layout (FP32 master copy; the BF16 copy has the same layout).
struct Layout {
  int L;
  int64_t w_in() const { return 0; }
  int64_t b_in() const { return kH * kIn; }
  int64_t w(int l) const { return kH * kIn + kH + (int64_t)l * kLayerParams; }
  int64_t b(int l) const { return w(l) + kH * 2 * kH; }
  int64_t w_out() const { return w(L); }
  int64_t c() const { return w_out() + kH; }
  int64_t total() const { return c() + 2; }
};

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x16 mfma_bf16(u16x8 a, u16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}
// row of accumulator element r in the 32x32 MFMA output held by a lane of half `half`
__device__ __forceinline__ int acc_row(int r, int half) { return 8 * (r >> 2) + 4 * half + (r & 3); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// ------------------------------------------------------------------ small kernels
__global__ void k_fill_i32(int* __restrict__ a, int64_t n, int v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}

__global__ void k_to_bf16(const float* __restrict__ a, int64_t n, uint16_t* __restrict__ b) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = f2bf(a[k]);
}

template <typename TA, typename TB>
__global__ void k_convert(const TA* __restrict__ a, int64_t n, TB* __restrict__ b) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = (TB)a[k];
}

template <typename TA, typename TB>
__global__ void k_gather_perm(const TA* __restrict__ x, const int32_t* __restrict__ perm, int64_t n, int nb,
                              TB* __restrict__ y) {  // y[b][k] = x[b][perm[k]]
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nb) return;
  const int64_t b = k / n, i = k - b * n;
  y[k] = (TB)x[b * n + perm[i]];
}
template <typename TA, typename TB>
__global__ void k_scatter_perm(const TA* __restrict__ y, const int32_t* __restrict__ perm, int64_t n, int nb,
                               TB* __restrict__ x) {  // x[b][perm[k]] = y[b][k]
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nb) return;
  const int64_t b = k / n, i = k - b * n;
  x[b * n + perm[i]] = (TB)y[k];
}

// Batched FP64 dot products: grid (blocks, B); part[b * gridDim.x + blk]
template <typename TA, typename TB>
__global__ __launch_bounds__(kThreads) void k_bdot(const TA* __restrict__ a, const TB* __restrict__ b, int64_t n,
                                                   double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  const int64_t off = (int64_t)blockIdx.y * n;
  double s = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[off + k] * (double)b[off + k];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += sh[w];
    part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
  }
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bsum(const T* __restrict__ a, int64_t n, double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  const int64_t off = (int64_t)blockIdx.y * n;
  double s = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[off + k];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += sh[w];
    part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
  }
}
// out[b * stride + slot] = sum of nb partials of row b (one block per row)
__global__ __launch_bounds__(kThreads) void k_brow_sum(const double* __restrict__ part, int nb,
                                                       double* __restrict__ out, int stride, int slot) {
  __shared__ double sh[kThreads / 64];
  double s = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) s += part[(int64_t)blockIdx.x * nb + k];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += sh[w];
    out[(int64_t)blockIdx.x * stride + slot] = t;
  }
}

// Column sums of a [rows][cols] FP32 partial table (FP64 accumulation); grid.y
// splits the rows into chunks of `chunk`, out[blockIdx.y][col].
__global__ void k_col_sum(const float* __restrict__ part, int rows, int cols, int chunk, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  const int r0 = blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
  double s = 0.0;
  for (int r = r0; r < r1; ++r) s += part[(int64_t)r * cols + c];
  out[(int64_t)blockIdx.y * cols + c] = (float)s;
}

// ------------------------------------------------------------------ message graph
// Row i of the graph is row i of the permuted split system without its diagonal:
// its L entries followed by its U entries; weights w_ij = -a_ij / a_ii.
__global__ void k_graph_fill(const int32_t* __restrict__ lrp, const int32_t* __restrict__ lci,
                             const double* __restrict__ lval, const int32_t* __restrict__ urp,
                             const int32_t* __restrict__ uci, const double* __restrict__ uval,
                             const double* __restrict__ dval, int64_t n, int32_t* __restrict__ gptr,
                             int32_t* __restrict__ gcol, float* __restrict__ gw, float* __restrict__ dF) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int l0 = lrp[i], l1 = lrp[i + 1], u0 = urp[i], u1 = urp[i + 1];
  const double d = dval[i];
  int g = l0 + u0;
  gptr[i] = g;
  if (i == n - 1) gptr[n] = l1 + u1;
  for (int p = l0; p < l1; ++p, ++g) {
    gcol[g] = lci[p];
    gw[g] = (float)(-lval[p] / d);
  }
  for (int p = u0; p < u1; ++p, ++g) {
    gcol[g] = uci[p];
    gw[g] = (float)(-uval[p] / d);
  }
  dF[i] = (float)d;
}
__global__ void k_graph_delta(const double* __restrict__ dval, int64_t n, const double* __restrict__ dsum,
                              float* __restrict__ delta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) delta[i] = (float)((dsum[0] / (double)n) / dval[i]);
}
// transposed weights: entry (i -> j) receives w_ji (stored in row j at its entry -> i)
__global__ void k_graph_wt(const int32_t* __restrict__ gptr, const int32_t* __restrict__ gcol,
                           const float* __restrict__ gw, int64_t n, float* __restrict__ gwt, int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int p = gptr[i]; p < gptr[i + 1]; ++p) {
    const int j = gcol[p];
    float w = 0.f;
    bool found = false;
    for (int q = gptr[j]; q < gptr[j + 1]; ++q)
      if (gcol[q] == (int)i) {
        w = gw[q];
        found = true;
        break;
      }
    if (!found) atomicOr(bad, 1);
    gwt[p] = w;
  }
}

// ------------------------------------------------------------------ forward
// scal[2b] = 1/s_b, scal[2b+1] = s_b / dbar from sumsq[b] (FP64)
__global__ void k_scale(const double* __restrict__ sumsq, int nb, int64_t n, const double* __restrict__ dsum,
                        float* __restrict__ scal) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double s = std::sqrt(sumsq[b] / (double)n);
  const double dbar = dsum[0] / (double)n;
  scal[2 * b] = s > 0 ? (float)(1.0 / s) : 0.f;
  scal[2 * b + 1] = (float)(s / dbar);
}

// h0 = W_in f + b_in; one lane per (node, feature)
__global__ __launch_bounds__(kThreads) void k_encode(const float* __restrict__ r, const float* __restrict__ zs,
                                                     const float* __restrict__ delta, const double* __restrict__ dsum,
                                                     const float* __restrict__ scal, int64_t n, int64_t total,
                                                     const uint16_t* __restrict__ wbf, const float* __restrict__ th,
                                                     uint16_t* __restrict__ h0, float2* __restrict__ F) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t node = t >> 5;
  const int f = (int)(t & 31);
  if (node >= total) return;
  const int64_t b = node / n, i = node - b * n;
  const float inv_s = scal[2 * b];
  const float dbar = (float)(dsum[0] / (double)n);
  const float dl = delta[i];
  const float f0 = r[node] * inv_s * dl;
  const float f1 = zs[node] * inv_s * dbar;
  float h = th[kH * kIn + f];
  h = fmaf(bf2f(wbf[f * kIn + 0]), f0, h);
  h = fmaf(bf2f(wbf[f * kIn + 1]), f1, h);
  h = fmaf(bf2f(wbf[f * kIn + 2]), dl, h);
  h0[node * kH + f] = f2bf(h);
  if (f == 0) F[node] = make_float2(f0, f1);
}

// ------------------------------------------------------------------ inference on message tiles
// The inference layers run on spatial tiles of the matrix graph.  Rows are put
// in the order of a Morton curve through the node coordinates and cut greedily
// into tiles of at most kTileRows rows whose footprint -- the rows plus their
// neighbours outside the tile (the halo) -- fits kTileFoot hidden rows of LDS.
// Hidden states live in this tile order.  A block loads its tile's footprint
// once (64 B per row) and gathers the neighbour rows of each node from LDS.
constexpr int kTileRows = 256;  // rows per tile, one thread each
constexpr int kTileFoot = 640;  // footprint rows held in LDS (50 KB: three blocks per CU)
constexpr int kFootIter = kTileFoot * 4 / kTileRows;  // 16-byte footprint pieces per thread
static_assert(kTileFoot * 4 % kTileRows == 0, "footprint pieces per thread");

struct MsgTile {
  int32_t start, nint, nfoot, width;  // first row (tile order), rows, footprint rows, edges per row (even)
  int64_t eoff, foff;                 // entry k of row t at eoff + k nint + t; footprint rows at foff
};

// A footprint row occupies kFootStride 16-byte slots (64 B of data, 16 B of
// padding): with an odd stride any 8 (16) consecutive rows put a given quarter
// on distinct four-byte LDS banks, and the four quarters of a row are one base
// address plus immediate offsets.
constexpr int kFootStride = 5;
static_assert(kTileFoot * kFootStride * 16 < 65536, "footprint byte offsets fit 16 bits");

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

typedef float f32x2 __attribute__((ext_vector_type(2)));

// two floats -> packed BF16 pair, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pk_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

// a.lo b.lo + a.hi b.hi + c (BF16 pairs, FP32 accumulate)
__device__ __forceinline__ float dot2_bf16(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a), __builtin_bit_cast(bf16x2, b), c, false);
}

// Edge words: low 16 bits the LDS byte offset of the neighbour's footprint row, high 16 bits the
// BF16 message weight w_ij = -a_ij / a_ii (rounded from FP32), refreshed from
// the FP64 split system at every preconditioner set-up.  One block per tile.
__global__ __launch_bounds__(kTileRows) void k_tile_weights(const MsgTile* __restrict__ tiles,
                                                            const int32_t* __restrict__ esrc,
                                                            const int32_t* __restrict__ trow,
                                                            const double* __restrict__ lval,
                                                            const double* __restrict__ uval, int64_t nl,
                                                            const double* __restrict__ dval,
                                                            uint32_t* __restrict__ edge) {
  const MsgTile T = tiles[blockIdx.x];
  const int t = threadIdx.x;
  if (t >= T.nint) return;
  const double d = dval[trow[T.start + t]];
  for (int k = 0; k < T.width; ++k) {
    const int64_t e = T.eoff + (int64_t)k * T.nint + t;
    const int32_t src = esrc[e];
    uint32_t w = 0;
    if (src >= 0) {
      const double a = src < nl ? lval[src] : uval[src - nl];
      w = f2bf((float)(-a / d));
    }
    edge[e] = (edge[e] & 0xFFFFu) | (w << 16);
  }
}

constexpr int kFixedWidth = 26;  // the 27-point stencil: tiles padded to this width use the unrolled kernel

// acc[2p + h] += w_a h_a,(2p+h) + w_b h_b,(2p+h) for the two edges (ea, eb) (low
// 16 bits: byte offset of the neighbour's footprint row): the two weights and
// each pair of BF16 features go into one register each (v_perm_b32) and are
// summed with v_dot2_f32_bf16
__device__ __forceinline__ void pair_step(const uint4* __restrict__ foot, uint32_t ea, uint32_t eb, float* acc) {
  const uint4* pa = reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(foot) + (ea & 0xFFFFu));
  const uint4* pb = reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(foot) + (eb & 0xFFFFu));
  uint4 x[4], y[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    x[q] = pa[q];
    y[q] = pb[q];
  }
  const uint32_t wp = __builtin_amdgcn_perm(eb, ea, 0x07060302u);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t a4[4] = {x[q].x, x[q].y, x[q].z, x[q].w}, b4[4] = {y[q].x, y[q].y, y[q].z, y[q].w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = 4 * q + i;  // features 2p (low halves) and 2p + 1 (high halves)
      acc[2 * p] = dot2_bf16(__builtin_amdgcn_perm(b4[i], a4[i], 0x05040100u), wp, acc[2 * p]);
      acc[2 * p + 1] = dot2_bf16(__builtin_amdgcn_perm(b4[i], a4[i], 0x07060302u), wp, acc[2 * p + 1]);
    }
  }
}

// One inference layer on one tile per block (kTileRows threads, a thread per row).
//  1. the footprint's hidden rows are copied to LDS (interior rows, then halo);
//  2. each thread forms m_i = sum_j w_ij h_j over its edges two at a time: the
//     two weights and each pair of BF16 features (h_a,t, h_b,t) are packed into
//     one register and summed with v_dot2_f32_bf16 (FP32 accumulation);
//  3. each wavefront computes W_l [h ; m] for its 64 rows as the transposed
//     product (W_l [h ; m]^T, M = 32 features, N = 32 rows, K = 64) on two
//     32x32x16 BF16 MFMA tiles; the operand halves and the outputs move between
//     lanes l and l + 32 with one exchange each, so every lane ends up with the
//     32 outputs of its own row and writes h' = h + relu(. + b) as one 64-byte
//     row -- or, in the last layer, z = (w_out . h' + c0 f0 + c1 f1) s / dbar
//     at the row's place in the solver order.
// The first layer encodes its footprint itself: h0 = W_in f + b_in from the
// residual r, the SGS feature zs and delta of each footprint row (solver order,
// gathered through foot_srows), and writes f0, f1 of its own rows to F.
struct EncodeArgs {
  const int32_t* foot_srows;  // footprint rows in the solver order
  const float *r, *zs, *delta;
  const double* dsum;
  int64_t n;
  const uint16_t* win;  // W_in (BF16, [32][3])
  const float* bin;     // b_in (FP32)
  float2* F;            // f0, f1 per row of the tile order
};

template <bool FIRST, bool LAST, int KW>
__global__ __launch_bounds__(kTileRows) __attribute__((amdgpu_waves_per_eu(3))) void k_layer_tile(
    const MsgTile* __restrict__ tiles, int64_t ntiles, const int32_t* __restrict__ foot_rows,
    const uint32_t* __restrict__ edge, const uint16_t* __restrict__ hin, uint16_t* __restrict__ hout,
    const uint16_t* __restrict__ wl, const float* __restrict__ bl, const uint16_t* __restrict__ wo,
    const float2* __restrict__ F, const float* __restrict__ scal, const int32_t* __restrict__ trow,
    float* __restrict__ z, EncodeArgs enc) {
  __shared__ uint4 foot[kTileFoot * kFootStride];
  const int64_t tb = xcd_block(blockIdx.x, ntiles);
  if (tb < 0) return;  // the whole block
  const MsgTile T = tiles[tb];
  // Round trip 1: the footprint row indices of this thread's pieces and the
  // row's edge words (fixed width); round trip 2: the footprint rows.  Loads are
  // unconditional (clamped to valid addresses) so that each round trip is one
  // batch of requests with one wait.
  const int t = threadIdx.x;
  const bool act = t < T.nint;
  const int q = t & 3, fr0 = t >> 2;
  int32_t grow[kFootIter];
  const int32_t* frows = FIRST ? enc.foot_srows : foot_rows;
#pragma unroll
  for (int u = 0; u < kFootIter; ++u) grow[u] = frows[T.foff + min(fr0 + u * (kTileRows / 4), T.nfoot - 1)];
  const uint32_t* ep = edge + T.eoff + (act ? t : 0);
  const int64_t stride = T.nint;
  uint32_t ew[KW > 0 ? KW : 2];
  if (KW > 0) {
#pragma unroll
    for (int k = 0; k < KW; ++k) ew[k] = ep[k * stride];
  }
  if (FIRST) {
    float rv[kFootIter], zv[kFootIter], dv[kFootIter];
#pragma unroll
    for (int u = 0; u < kFootIter; ++u) {
      rv[u] = enc.r[grow[u]];
      zv[u] = enc.zs[grow[u]];
      dv[u] = enc.delta[grow[u]];
    }
    const float inv_s = scal[0];
    const float dbar = (float)(enc.dsum[0] / (double)enc.n);
    float w0[8], w1[8], w2[8], b[8];  // encoder rows of features 8q .. 8q + 7
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int fi = 8 * q + i;
      w0[i] = bf2f(enc.win[fi * kIn + 0]);
      w1[i] = bf2f(enc.win[fi * kIn + 1]);
      w2[i] = bf2f(enc.win[fi * kIn + 2]);
      b[i] = enc.bin[fi];
    }
#pragma unroll
    for (int u = 0; u < kFootIter; ++u) {
      const int fr = fr0 + u * (kTileRows / 4);
      const float f0 = rv[u] * inv_s * dv[u], f1 = zv[u] * inv_s * dbar;
      float h[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) h[i] = fmaf(w2[i], dv[u], fmaf(w1[i], f1, fmaf(w0[i], f0, b[i])));
      foot[fr * kFootStride + q] = make_uint4(pk_bf16(h[0], h[1]), pk_bf16(h[2], h[3]), pk_bf16(h[4], h[5]),
                                              pk_bf16(h[6], h[7]));
      if (q == 0 && fr < T.nint) enc.F[(int64_t)T.start + fr] = make_float2(f0, f1);
    }
  } else {
    const uint4* hq = reinterpret_cast<const uint4*>(hin);
    uint4 v[kFootIter];
#pragma unroll
    for (int u = 0; u < kFootIter; ++u) v[u] = hq[(int64_t)grow[u] * 4 + q];
    // rows past nfoot (< kTileFoot) get a copy of the last row: never read,
    // stored unconditionally so that no load is sunk into a branch
#pragma unroll
    for (int u = 0; u < kFootIter; ++u) foot[(fr0 + u * (kTileRows / 4)) * kFootStride + q] = v[u];
  }
  __syncthreads();
  const int lane = t & 63, half = lane >> 5, f = lane & 31;
  if ((t & ~63) >= T.nint) return;  // a wavefront without rows (no barrier follows)
  uint32_t own[16], mp[16];
  float acc[kH];
#pragma unroll
  for (int p = 0; p < 16; ++p) own[p] = 0u;
#pragma unroll
  for (int p = 0; p < kH; ++p) acc[p] = 0.f;
  if (act) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint4 v = foot[t * kFootStride + q];
      own[4 * q] = v.x, own[4 * q + 1] = v.y, own[4 * q + 2] = v.z, own[4 * q + 3] = v.w;
    }
    if (KW > 0) {
#pragma unroll
      for (int k = 0; k < KW; k += 2) pair_step(foot, ew[k], ew[k + 1], acc);
    } else {
      for (int k = 0; k < T.width; k += 2) pair_step(foot, ep[k * stride], ep[(k + 1) * stride], acc);
    }
  }
#pragma unroll
  for (int p = 0; p < 16; ++p) mp[p] = pk_bf16(acc[2 * p], acc[2 * p + 1]);
  // A operand: W_l[o = f][k = 16 ks + 8 half + 0..7].  B operand of tile 0: rows
  // 0..31 of the wavefront (lane l < 32 gives k = 16 ks + 0..7 of its own row,
  // lane l + 32 gives k + 8 of row l); tile 1: rows 32..63 likewise.  Each lane
  // holds both halves (P: k + 0..7, Q: k + 8..15) of its own row; one
  // v_permlane32_swap per register pair forms both operands.
  f32x16 d0 = {}, d1 = {};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const u16x8 a = *reinterpret_cast<const u16x8*>(wl + f * 2 * kH + ks * 16 + 8 * half);
    const uint32_t* src = ks < 2 ? own + 8 * ks : mp + 8 * (ks - 2);
    u32x4 b0, b1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      // lanes 32..63 of P <-> lanes 0..31 of Q: P becomes the tile-0 operand, Q the tile-1 one
      const auto pq = __builtin_amdgcn_permlane32_swap(src[i], src[4 + i], false, false);
      b0[i] = pq[0];
      b1[i] = pq[1];
    }
    d0 = mfma_bf16(a, __builtin_bit_cast(u16x8, b0), d0);
    d1 = mfma_bf16(a, __builtin_bit_cast(u16x8, b1), d1);
  }
  // lane l < 32 holds outputs 8g+4h'+s (h' = 0) of row l in d0, lane l + 32 the
  // h' = 1 outputs of the same row; tile 1 mirrored.  One swap per register pair
  // gives every lane lo[4g+s] = output 8g+s and hi[4g+s] = output 8g+4+s of its
  // own row.
  float lo[16], hi[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const auto x = __builtin_amdgcn_permlane32_swap(__float_as_uint(d0[r]), __float_as_uint(d1[r]), false, false);
    lo[r] = __uint_as_float(x[0]);
    hi[r] = __uint_as_float(x[1]);
  }
  if (!act) return;
  float hv[kH];  // h' = h + relu(W [h ; m] + b) of the lane's row
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int o0 = 8 * g + s, o1 = o0 + 4;
      const float h0v = __uint_as_float(((o0 & 1) ? own[o0 >> 1] : own[o0 >> 1] << 16) & 0xFFFF0000u);
      const float h1v = __uint_as_float(((o1 & 1) ? own[o1 >> 1] : own[o1 >> 1] << 16) & 0xFFFF0000u);
      hv[o0] = h0v + fmaxf(lo[4 * g + s] + bl[o0], 0.f);
      hv[o1] = h1v + fmaxf(hi[4 * g + s] + bl[o1], 0.f);
    }
  uint32_t hn[16];
#pragma unroll
  for (int p = 0; p < 16; ++p) hn[p] = pk_bf16(hv[2 * p], hv[2 * p + 1]);
  const int64_t row = (int64_t)T.start + t;
  if (!LAST) {
    uint4* o = reinterpret_cast<uint4*>(hout + row * kH);
#pragma unroll
    for (int p = 0; p < 4; ++p) o[p] = make_uint4(hn[4 * p], hn[4 * p + 1], hn[4 * p + 2], hn[4 * p + 3]);
  } else {
    float out = 0.f;
#pragma unroll
    for (int o = 0; o < kH; ++o)
      out = fmaf(bf2f(wo[o]), __uint_as_float((o & 1) ? hn[o >> 1] & 0xFFFF0000u : hn[o >> 1] << 16), out);
    const float2 ff = F[row];
    out = fmaf(bf2f(wo[kH]), ff.x, out);
    out = fmaf(bf2f(wo[kH + 1]), ff.y, out);
    z[trow[row]] = out * scal[1];
  }
}

// One message-passing layer.  A wavefront owns 32 consecutive nodes: it
// aggregates m_i for two nodes at a time (32 lanes = 32 features per node,
// neighbour lists read cooperatively and broadcast with shuffles), stages the
// [h ; m] tile as BF16 in LDS, and computes W_l [h ; m] with four
// 32x32x16 BF16 MFMAs (FP32 accumulation).
template <bool TRAIN>
__global__ __launch_bounds__(kThreads) void k_layer(const int32_t* __restrict__ gptr, const int32_t* __restrict__ gcol,
                                                    const float* __restrict__ gw, int64_t n, int64_t total,
                                                    const uint16_t* __restrict__ hin, uint16_t* __restrict__ hout,
                                                    const uint16_t* __restrict__ wl, const float* __restrict__ bl,
                                                    uint16_t* __restrict__ mst, uint32_t* __restrict__ mask) {
  __shared__ __align__(16) uint16_t As[kWaves][kTile][kLdA];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, f = lane & 31, half = lane >> 5;
  const int64_t tile0 = ((int64_t)blockIdx.x * kWaves + wv) * kTile;
  // B operand: B[k][o] = W_l[o][k]; lane holds o = f, k = 16 ks + 8 half + 0..7
  u16x8 wfrag[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) wfrag[ks] = *reinterpret_cast<const u16x8*>(wl + f * 2 * kH + ks * 16 + 8 * half);
  const float bias = bl[f];

  // aggregation: 8 lanes per node (4 features each), 8 nodes per pass; the
  // neighbour list is read 8 entries at a time and broadcast within the group,
  // so each lane has 8 independent 8-byte loads in flight
  {
    const int sub = lane >> 3, q = lane & 7;
    for (int p = 0; p < kTile / 8; ++p) {
      const int nl = p * 8 + sub;
      const int64_t node = tile0 + nl;
      uint2 hs = make_uint2(0u, 0u);
      float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
      if (node < total) {
        const int64_t b = node / n, i = node - b * n;
        const uint16_t* hb = hin + b * n * kH + 4 * q;
        hs = *reinterpret_cast<const uint2*>(hb + i * kH);
        const int s0 = gptr[i], s1 = gptr[i + 1];
        for (int c0 = s0; c0 < s1; c0 += 8) {
          int cj = 0;
          float cw = 0.f;
          if (c0 + q < s1) {
            cj = gcol[c0 + q];
            cw = bf2f(f2bf(gw[c0 + q]));  // message weights are BF16
          }
          uint2 hv[8];
          float ww[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int j = __shfl(cj, u, 8);
            ww[u] = __shfl(cw, u, 8);
            hv[u] = *reinterpret_cast<const uint2*>(hb + (int64_t)j * kH);
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            a0 = fmaf(ww[u], __uint_as_float(hv[u].x << 16), a0);
            a1 = fmaf(ww[u], __uint_as_float(hv[u].x & 0xFFFF0000u), a1);
            a2 = fmaf(ww[u], __uint_as_float(hv[u].y << 16), a2);
            a3 = fmaf(ww[u], __uint_as_float(hv[u].y & 0xFFFF0000u), a3);
          }
        }
      }
      const uint2 mb = make_uint2((uint32_t)f2bf(a0) | ((uint32_t)f2bf(a1) << 16),
                                  (uint32_t)f2bf(a2) | ((uint32_t)f2bf(a3) << 16));
      *reinterpret_cast<uint2*>(&As[wv][nl][4 * q]) = hs;
      *reinterpret_cast<uint2*>(&As[wv][nl][kH + 4 * q]) = mb;
      if (TRAIN && node < total) *reinterpret_cast<uint2*>(mst + node * kH + 4 * q) = mb;
    }
  }
  __syncthreads();

  f32x16 acc = {};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const u16x8 a = *reinterpret_cast<const u16x8*>(&As[wv][f][ks * 16 + 8 * half]);
    acc = mfma_bf16(a, wfrag[ks], acc);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = acc_row(r, half);
    const int64_t node = tile0 + row;
    const float u = acc[r] + bias;
    if (TRAIN) {
      const uint64_t bal = __ballot(u > 0.f);
      if (f == 0 && node < total) mask[node] = (uint32_t)(half ? (bal >> 32) : bal);
    }
    if (node < total) hout[node * kH + f] = f2bf(bf2f(As[wv][row][f]) + fmaxf(u, 0.f));
  }
}

// out = w_out . h + c0 f0 + c1 f1; z = out * scale (scale = s/dbar, or 1 for the raw output)
template <typename TZ>
__global__ __launch_bounds__(kThreads) void k_output(const uint16_t* __restrict__ h, const float2* __restrict__ F,
                                                     const uint16_t* __restrict__ wo, const float* __restrict__ scal,
                                                     int64_t n, int64_t total, int raw, TZ* __restrict__ z,
                                                     int* __restrict__ flag) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= total) return;
  const uint4* hv = reinterpret_cast<const uint4*>(h + node * kH);
  float out = 0.f;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 v = hv[q];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      out = fmaf(bf2f(wo[q * 8 + 2 * t]), __uint_as_float(w[t] << 16), out);
      out = fmaf(bf2f(wo[q * 8 + 2 * t + 1]), __uint_as_float(w[t] & 0xFFFF0000u), out);
    }
  }
  const float2 ff = F[node];
  out = fmaf(bf2f(wo[kH]), ff.x, out);
  out = fmaf(bf2f(wo[kH + 1]), ff.y, out);
  if (!raw) out *= scal[2 * (node / n) + 1];
  if (flag && !std::isfinite(out)) atomicOr(flag, 1);
  z[node] = (TZ)out;
}

// ------------------------------------------------------------------ loss
// Per node: Ax = d (x - W x), sym = (A + A^T) x = 2 d x - d W x - W^T (d x);
// per-sample partials of <x, r> and <x, A x>.  Grid (blocks, B).
__global__ __launch_bounds__(kThreads) void k_aop(const int32_t* __restrict__ gptr, const int32_t* __restrict__ gcol,
                                                  const float* __restrict__ gw, const float* __restrict__ gwt,
                                                  const float* __restrict__ dF, int64_t n, const float* __restrict__ x,
                                                  const float* __restrict__ r, float* __restrict__ sym,
                                                  double* __restrict__ part) {
  __shared__ double sh[2][kThreads / 64];
  const int64_t off = (int64_t)blockIdx.y * n;
  const float* xb = x + off;
  double sp = 0.0, sq = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float wx = 0.f, wtdx = 0.f;
    for (int p = gptr[i]; p < gptr[i + 1]; ++p) {
      const int j = gcol[p];
      const float xj = xb[j];
      wx = fmaf(gw[p], xj, wx);
      wtdx = fmaf(gwt[p], dF[j] * xj, wtdx);
    }
    const float d = dF[i], xi = xb[i];
    const float ax = d * (xi - wx);
    sym[off + i] = 2.f * d * xi - d * wx - wtdx;
    sp += (double)xi * (double)r[off + i];
    sq += (double)xi * (double)ax;
  }
  sp = wave_sum(sp);
  sq = wave_sum(sq);
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = sp;
    sh[1][threadIdx.x >> 6] = sq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) {
      a += sh[0][w];
      c += sh[1][w];
    }
    const int64_t k = ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 2;
    part[k] = a;
    part[k + 1] = c;
  }
}
// pq[b] = (P, Q) from partial pairs
__global__ void k_pq_sum(const double* __restrict__ part, int nb, double* __restrict__ pq) {
  __shared__ double sh[2][kThreads / 64];
  double a = 0.0, c = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    a += part[((int64_t)blockIdx.x * nb + k) * 2];
    c += part[((int64_t)blockIdx.x * nb + k) * 2 + 1];
  }
  a = wave_sum(a);
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0.0, tc = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) {
      ta += sh[0][w];
      tc += sh[1][w];
    }
    pq[2 * blockIdx.x] = ta;
    pq[2 * blockIdx.x + 1] = tc;
  }
}
// batch-mean loss into loss_acc[slot] (single thread)
__global__ void k_loss(const double* __restrict__ pq, const double* __restrict__ S, int nb,
                       double* __restrict__ loss_acc, int slot) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double L = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double P = pq[2 * b], Q = pq[2 * b + 1];
    L += 1.0 - P * P / (Q * S[b]);
  }
  loss_acc[slot] = L / nb;
}
// dL/dout = (1/B) (-2P/(QS) r + P^2/(Q^2 S) (A + A^T) out)
__global__ void k_gout(const float* __restrict__ r, const float* __restrict__ sym, const double* __restrict__ pq,
                       const double* __restrict__ S, int64_t n, int nb, float* __restrict__ g) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nb) return;
  const int64_t b = k / n;
  const double P = pq[2 * b], Q = pq[2 * b + 1], s = S[b];
  const double a = -2.0 * P / (Q * s) / nb, c = P * P / (Q * Q * s) / nb;
  g[k] = (float)(a * (double)r[k] + c * (double)sym[k]);
}

// ------------------------------------------------------------------ backward
// G^L = gout w_out; partials of dw_out (32) and dc (2): part[blk][34]
__global__ __launch_bounds__(kThreads) void k_output_bwd(const float* __restrict__ gout, const uint16_t* __restrict__ h,
                                                         const float2* __restrict__ F, const uint16_t* __restrict__ wo,
                                                         int64_t total, float* __restrict__ G, float* __restrict__ part) {
  __shared__ float sh[2][kThreads];
  const int f = threadIdx.x & 31;
  const float w = bf2f(wo[f]);
  float aw = 0.f, ac = 0.f;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total * kH;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = t >> 5;
    const float g = gout[node];
    G[t] = g * w;
    aw = fmaf(g, bf2f(h[t]), aw);
    if (f < 2) {
      const float2 ff = F[node];
      ac = fmaf(g, f == 0 ? ff.x : ff.y, ac);
    }
  }
  sh[0][threadIdx.x] = aw;
  sh[1][threadIdx.x] = ac;
  __syncthreads();
  if (threadIdx.x < 32) {
    float s0 = 0.f, s1 = 0.f;
    for (int k = threadIdx.x; k < kThreads; k += 32) {
      s0 += sh[0][k];
      s1 += sh[1][k];
    }
    part[(int64_t)blockIdx.x * 34 + threadIdx.x] = s0;
    if (threadIdx.x < 2) part[(int64_t)blockIdx.x * 34 + 32 + threadIdx.x] = s1;
  }
}

// Backward of one layer on 32-node tiles:
//   delta = G ⊙ [u > 0];  [dh | dm] = delta W_l;  Gp = G + dh;  DM = dm;
//   dW_l += delta^T [h ; m],  db_l += sum delta   (per-wavefront partials)
// Blocks stride over tile groups so every wavefront of a block runs the same
// number of iterations (the LDS tiles are per wavefront, the barriers per block).
__global__ __launch_bounds__(kThreads) void k_layer_bwd(int64_t total, const float* __restrict__ G,
                                                        const uint32_t* __restrict__ mask,
                                                        const uint16_t* __restrict__ h, const uint16_t* __restrict__ mst,
                                                        const uint16_t* __restrict__ wl, float* __restrict__ Gp,
                                                        float* __restrict__ DM, float* __restrict__ part) {
  __shared__ __align__(16) uint16_t As[kWaves][kTile][kLdA];
  __shared__ __align__(16) uint16_t Ds[kWaves][kTile][kLdD];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, f = lane & 31, half = lane >> 5;
  // B operand of [dh | dm] = delta W: B[o][k] = W[o][k]; lane holds k = f (+32), o = 16 ks + 8 half + t
  u16x8 wb[2][2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int t = 0; t < 8; ++t) wb[ks][nt][t] = wl[(ks * 16 + 8 * half + t) * 2 * kH + nt * kH + f];
  f32x16 dw0 = {}, dw1 = {};
  float db = 0.f;
  const int64_t ntiles = (total + kTile - 1) / kTile;
  for (int64_t grp = blockIdx.x; grp * kWaves < ntiles; grp += gridDim.x) {
    const int64_t tile0 = (grp * kWaves + wv) * kTile;
    for (int p = 0; p < kTile / 2; ++p) {
      const int nl = 2 * p + half;
      const int64_t node = tile0 + nl;
      float dl = 0.f;
      uint16_t hh = 0, mm = 0;
      if (node < total) {
        const float gv = G[node * kH + f];
        dl = ((mask[node] >> f) & 1u) ? gv : 0.f;
        hh = h[node * kH + f];
        mm = mst[node * kH + f];
      }
      db += dl;
      Ds[wv][nl][f] = f2bf(dl);
      As[wv][nl][f] = hh;
      As[wv][nl][kH + f] = mm;
    }
    __syncthreads();
    f32x16 ch = {}, cm = {};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const u16x8 a = *reinterpret_cast<const u16x8*>(&Ds[wv][f][ks * 16 + 8 * half]);
      ch = mfma_bf16(a, wb[ks][0], ch);
      cm = mfma_bf16(a, wb[ks][1], cm);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t node = tile0 + acc_row(r, half);
      if (node < total) {
        Gp[node * kH + f] = G[node * kH + f] + ch[r];
        DM[node * kH + f] = cm[r];
      }
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      u16x8 a, b0, b1;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int nd = ks * 16 + 8 * half + t;
        a[t] = Ds[wv][nd][f];
        b0[t] = As[wv][nd][f];
        b1[t] = As[wv][nd][kH + f];
      }
      dw0 = mfma_bf16(a, b0, dw0);
      dw1 = mfma_bf16(a, b1, dw1);
    }
    __syncthreads();
  }
  float* pw = part + ((int64_t)blockIdx.x * kWaves + wv) * kLayerParams;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int o = acc_row(r, half);
    pw[o * 2 * kH + f] = dw0[r];
    pw[o * 2 * kH + kH + f] = dw1[r];
  }
  db += __shfl_xor(db, 32);
  if (half == 0) pw[2 * kH * kH + f] = db;
}

// G_j = Gp_j + sum_{i in N(j)} w_ij DM_i  (transposed aggregation; FP32);
// 8 lanes per node, 4 features (16 bytes) per lane
__global__ __launch_bounds__(kThreads) void k_agg_t(const int32_t* __restrict__ gptr, const int32_t* __restrict__ gcol,
                                                    const float* __restrict__ gwt, int64_t n, int64_t total,
                                                    const float* __restrict__ Gp, const float* __restrict__ DM,
                                                    float* __restrict__ Gout) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t node = t >> 3;
  const int q = threadIdx.x & 7;
  if (node >= total) return;  // the 8 lanes of a node exit together
  const int64_t b = node / n, i = node - b * n;
  const float* db = DM + b * n * kH + 4 * q;
  float4 acc = *reinterpret_cast<const float4*>(Gp + node * kH + 4 * q);
  const int s0 = gptr[i], s1 = gptr[i + 1];
  for (int c0 = s0; c0 < s1; c0 += 8) {
    int cj = 0;
    float cw = 0.f;
    if (c0 + q < s1) {
      cj = gcol[c0 + q];
      cw = bf2f(f2bf(gwt[c0 + q]));  // the forward's BF16 message weights
    }
    float4 v[8];
    float w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = __shfl(cj, u, 8);
      w[u] = __shfl(cw, u, 8);
      v[u] = *reinterpret_cast<const float4*>(db + (int64_t)j * kH);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      acc.x = fmaf(w[u], v[u].x, acc.x);
      acc.y = fmaf(w[u], v[u].y, acc.y);
      acc.z = fmaf(w[u], v[u].z, acc.z);
      acc.w = fmaf(w[u], v[u].w, acc.w);
    }
  }
  *reinterpret_cast<float4*>(Gout + node * kH + 4 * q) = acc;
}

// dW_in += G0 f^T, db_in += G0: part[blk][128] = [w_in 96 | b_in 32]
__global__ __launch_bounds__(kThreads) void k_encode_bwd(const float* __restrict__ G0, const float2* __restrict__ F,
                                                         const float* __restrict__ delta, int64_t n, int64_t total,
                                                         float* __restrict__ part) {
  __shared__ float sh[4][kThreads];
  const int f = threadIdx.x & 31;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, ab = 0.f;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total * kH;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = t >> 5;
    const float g = G0[t];
    const float2 ff = F[node];
    a0 = fmaf(g, ff.x, a0);
    a1 = fmaf(g, ff.y, a1);
    a2 = fmaf(g, delta[node % n], a2);
    ab += g;
  }
  sh[0][threadIdx.x] = a0;
  sh[1][threadIdx.x] = a1;
  sh[2][threadIdx.x] = a2;
  sh[3][threadIdx.x] = ab;
  __syncthreads();
  if (threadIdx.x < 32) {
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = threadIdx.x; k < kThreads; k += 32)
      for (int q = 0; q < 4; ++q) s[q] += sh[q][k];
    float* pb = part + (int64_t)blockIdx.x * 128;
    pb[f * kIn + 0] = s[0];
    pb[f * kIn + 1] = s[1];
    pb[f * kIn + 2] = s[2];
    pb[kH * kIn + f] = s[3];
  }
}

__global__ void k_adam(float* __restrict__ th, float* __restrict__ m1, float* __restrict__ m2,
                       const float* __restrict__ g, int64_t n, float lr, float b1, float b2, float bc1, float bc2,
                       float eps, uint16_t* __restrict__ wbf) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float m = b1 * m1[k] + (1.f - b1) * gk;
  const float v = b2 * m2[k] + (1.f - b2) * gk * gk;
  m1[k] = m;
  m2[k] = v;
  const float t = th[k] - lr * (m / bc1) / (std::sqrt(v / bc2) + eps);
  th[k] = t;
  wbf[k] = f2bf(t);
}

// ------------------------------------------------------------------ harvest helpers
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// unit Gaussian (Box-Muller over two hashed uniforms) for entry k of stream `key`
__global__ void k_gauss(uint64_t key, int64_t n, double* __restrict__ x) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t a = mix64(key ^ mix64((uint64_t)k * 2 + 1)), c = mix64(a ^ 0xD1B54A32D192ED03ull);
  const double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53, u2 = (double)(c >> 11) * 0x1.0p-53;
  x[k] = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
__global__ void k_shift(double* __restrict__ x, int64_t n, const double* __restrict__ sum) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] -= sum[0] / (double)n;
}
__global__ void k_cg_step(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                          const double* __restrict__ q, const double* __restrict__ sc, int64_t n) {
  // sc[0] = rz, sc[1] = pq
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double a = sc[0] / sc[1];
  x[k] += a * p[k];
  r[k] -= a * q[k];
}
__global__ void k_cg_dir(double* __restrict__ p, const double* __restrict__ z, const double* __restrict__ rz_new,
                         const double* __restrict__ rz_old, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = z[k] + (rz_new[0] / rz_old[0]) * p[k];
}
__global__ void k_sub(const double* __restrict__ a, const double* __restrict__ b, int64_t n, double* __restrict__ c) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) c[k] = a[k] - b[k];
}
// R = r / nr, E = (e - mean e) / nr in FP32 (sum[0] = sum of e)
__global__ void k_pair(const double* __restrict__ r, const double* __restrict__ e, const double* __restrict__ sum,
                       const double* __restrict__ nr2, int64_t n, float* __restrict__ R, float* __restrict__ E) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double inv = 1.0 / std::sqrt(nr2[0]);
  R[k] = (float)(r[k] * inv);
  E[k] = (float)((e[k] - sum[0] / (double)n) * inv);
}
__global__ void k_gather_rows(const float* __restrict__ src, const int32_t* __restrict__ idx, int64_t n, int nb,
                              float* __restrict__ dst) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nb) return;
  const int64_t b = k / n, i = k - b * n;
  dst[k] = src[(int64_t)idx[b] * n + i];
}
__global__ void k_gather_s(const double* __restrict__ S, const int32_t* __restrict__ idx, int nb,
                           double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) out[b] = S[idx[b]];
}

}  // namespace
}  // namespace nalu

// =================================================================== host side
struct nalu_gnn {
  int L = 4;
  nalu::DevBuf<float> theta, m1, m2, grad;  // FP32 master parameters, Adam moments, last gradient
  nalu::DevBuf<uint16_t> wbf;                // BF16 copy read by the kernels
  int64_t adam_t = 0;
};

struct nalu_gnn_data {
  const nalu_linsys* owner = nullptr;
  int64_t n = 0;
  int npairs = 0;
  nalu::DevBuf<float> R, E, ZS;  // [npairs][n] in the owner's solver (color-permuted) order
  nalu::DevBuf<double> S;        // <E, R> per pair
  std::vector<int> cg_iters;     // iterations of each harvested CG solve
};

namespace nalu {

// Per-system state of the learned preconditioner.
struct Learned {
  nalu_gnn* net = nullptr;
  int64_t n = 0;
  // training graph (CSR of the solver order, FP32 weights) for the current matrix
  bool graph_ready = false, wt_ready = false;
  DevBuf<int32_t> gptr, gcol;
  DevBuf<float> gw, gwt, dF, delta;
  DevBuf<double> dsum;  // [1] sum of the diagonal
  // inference: message tiles (built once per system) and their weights
  bool infer_ready = false;
  int64_t tiles_n = 0, ntiles = 0, tnl = 0;
  bool tile_fixed = false;  // every tile has kFixedWidth edges per row
  DevBuf<unsigned char> tiles;  // MsgTile[ntiles]
  DevBuf<int32_t> tfoot, tfoot_s, tesrc, trow;  // footprint rows (tile order, solver order), edge sources,
                                                // tile order -> solver row
  DevBuf<uint32_t> tedge;
  // one-vector workspace
  DevBuf<uint16_t> h0, h1;
  DevBuf<float> zs, r32, z32, scal;
  DevBuf<float2> F;
  DevBuf<double> part, sumsq, zsum;
  // every row of A sums to zero (pure-Neumann pressure system): constants are
  // A's null space, and the network output is projected off them
  bool neumann = false;
  DevBuf<int> rs_flag;
  bool fallback = false;
};

namespace {

Learned* state(const nalu_linsys* ls) { return reinterpret_cast<Learned*>(ls->learned); }

Learned* ensure_state(nalu_linsys* ls) {
  if (!ls->learned) ls->learned = new Learned();
  return state(ls);
}

void check_net(const nalu_gnn* g) {
  NALU_REQUIRE(g && g->theta.p && g->wbf.p, NALU_ERR_ARG, "learned preconditioner: null or empty network");
}

// (Re)build the message graph from the permuted split system.
void build_msg_graph(nalu_linsys* ls, Learned* st, bool need_wt) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  NALU_REQUIRE(ls->num_dof == 1, NALU_ERR_UNSUPPORTED, "learned preconditioner: scalar systems only (numDof 1)");
  NALU_REQUIRE(ls->precond_ready && ls->perm_ready, NALU_ERR_STATE,
               "learned preconditioner: the preconditioner of the system is not set up");
  NALU_REQUIRE(ls->nnz < ((int64_t)1 << 31), NALU_ERR_UNSUPPORTED, "learned preconditioner: nnz >= 2^31");
  const int64_t nnz = (int64_t)ls->lci.n + (int64_t)ls->uci.n;
  if (st->n != n || st->gw.n != (size_t)std::max<int64_t>(1, nnz)) {
    st->n = n;
    st->gptr.alloc(n + 1);
    st->gcol.alloc(std::max<int64_t>(1, nnz));
    st->gw.alloc(std::max<int64_t>(1, nnz));
    st->gwt.release();
    st->dF.alloc(n);
    st->delta.alloc(n);
    st->dsum.alloc(1);
  }
  if (st->part.n < (size_t)kRedBlocks * 64) st->part.alloc((size_t)kRedBlocks * 64);
  hipLaunchKernelGGL(k_graph_fill, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->lrp.p, ls->lci.p, ls->lval.p,
                     ls->urp.p, ls->uci.p, ls->uval.p, ls->dval.p, n, st->gptr.p, st->gcol.p, st->gw.p, st->dF.p);
  const unsigned nb = std::min<unsigned>(kRedBlocks, nblocks(n));
  hipLaunchKernelGGL(k_bsum<double>, dim3(nb, 1), dim3(kThreads), 0, s, ls->dval.p, n, st->part.p);
  hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, st->part.p, (int)nb, st->dsum.p, 1, 0);
  hipLaunchKernelGGL(k_graph_delta, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->dval.p, n, st->dsum.p, st->delta.p);
  st->graph_ready = true;
  st->wt_ready = false;
  if (need_wt) {
    if (st->gwt.n != st->gw.n) st->gwt.alloc(st->gw.n);
    DevBuf<int> bad(1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, s, bad.p, 1, 0);
    hipLaunchKernelGGL(k_graph_wt, dim3(nblocks(n)), dim3(kThreads), 0, s, st->gptr.p, st->gcol.p, st->gw.p, n,
                       st->gwt.p, bad.p);
    int hb = 0;
    NALU_HIP_CHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    NALU_REQUIRE(hb == 0, NALU_ERR_UNSUPPORTED, "learned preconditioner: matrix graph is not structurally symmetric");
    st->wt_ready = true;
  }
}

void ensure_one(Learned* st, int64_t n) {
  if (st->h0.n != (size_t)(n * kH)) {
    st->h0.alloc(n * kH);
    st->h1.alloc(n * kH);
    st->zs.alloc(n);
    st->r32.alloc(n);
    st->z32.alloc(n);
    st->F.alloc(n);
  }
  if (st->scal.n < 2) st->scal.alloc(2);
  if (st->sumsq.n < 1) st->sumsq.alloc(1);
  if (st->part.n < (size_t)kRedBlocks * 64) st->part.alloc((size_t)kRedBlocks * 64);
}

// ---------------------------------------------------------------- message tiles
uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1FFFFFu;
  v = (v | v << 32) & 0x1F00000000FFFFull;
  v = (v | v << 16) & 0x1F0000FF0000FFull;
  v = (v | v << 8) & 0x100F00F00F00F00Full;
  v = (v | v << 4) & 0x10C30C30C30C30C3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

// Tile order and tiles of the solver-order graph (host, once per system): rows
// sorted along a Morton curve of their node coordinates (cells of the mean node
// spacing), then cut greedily -- a row joins the current tile while the tile
// keeps <= kTileRows rows and <= kTileFoot footprint rows.  Each tile lists its
// footprint rows (its own, then the halo in ascending tile-order position) and,
// per row, its off-diagonal
// entries as (footprint index, source in [lval ; uval]) in the column-major
// layout of the layer kernel, padded to an even width with zero-weight self
// entries.
void build_tiles(nalu_linsys* ls, Learned* st) {
  const int64_t n = ls->nrows;
  NALU_REQUIRE(ls->mesh, NALU_ERR_STATE, "learned preconditioner: the system has no mesh");
  NALU_REQUIRE(n >= 1 && n < ((int64_t)1 << 31), NALU_ERR_UNSUPPORTED, "learned preconditioner: bad row count");
  std::vector<int32_t> lrp(n + 1), urp(n + 1), perm(n);
  NALU_HIP_CHECK(hipMemcpy(lrp.data(), ls->lrp.p, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  NALU_HIP_CHECK(hipMemcpy(urp.data(), ls->urp.p, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  NALU_HIP_CHECK(hipMemcpy(perm.data(), ls->perm.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  const int64_t nl = lrp[n], nu = urp[n];
  NALU_REQUIRE(nl + nu < ((int64_t)1 << 31), NALU_ERR_UNSUPPORTED, "learned preconditioner: nnz >= 2^31");
  std::vector<int32_t> lci(std::max<int64_t>(nl, 1)), uci(std::max<int64_t>(nu, 1));
  if (nl) NALU_HIP_CHECK(hipMemcpy(lci.data(), ls->lci.p, sizeof(int32_t) * nl, hipMemcpyDeviceToHost));
  if (nu) NALU_HIP_CHECK(hipMemcpy(uci.data(), ls->uci.p, sizeof(int32_t) * nu, hipMemcpyDeviceToHost));
  const nalu_mesh* m = ls->mesh;
  std::vector<double> X(3 * m->num_nodes);
  NALU_HIP_CHECK(hipMemcpy(X.data(), m->coords.p, sizeof(double) * X.size(), hipMemcpyDeviceToHost));
  std::vector<int32_t> rn;
  if (m->ids.p) {
    rn.resize(n);
    NALU_HIP_CHECK(hipMemcpy(rn.data(), ls->row_node.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  }
  auto node_of = [&](int64_t p) -> int64_t {
    const int64_t old = perm[p];
    return rn.empty() ? old : rn[old];
  };
  // Morton keys on cells of the mean node spacing
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t p = 0; p < n; ++p)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], X[3 * node_of(p) + d]);
      hi[d] = std::max(hi[d], X[3 * node_of(p) + d]);
    }
  const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
  double vol = 1.0;
  int nd = 0;
  for (int d = 0; d < 3; ++d)
    if (hi[d] - lo[d] > 1e-12 * ext) vol *= hi[d] - lo[d], ++nd;
  const double h = std::max(nd ? std::pow(vol / (double)n, 1.0 / nd) : ext, ext / 2097151.0);
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  for (int64_t p = 0; p < n; ++p) {
    uint64_t k = 0;
    for (int d = 0; d < 3; ++d) {
      const double c = std::floor((X[3 * node_of(p) + d] - lo[d]) / h);
      k |= spread3((uint64_t)std::min(std::max(c, 0.0), 2097151.0)) << d;
    }
    key[p] = {k, (int32_t)p};
  }
  std::sort(key.begin(), key.end());
  std::vector<int32_t> order(n), pos(n);
  for (int64_t q = 0; q < n; ++q) {
    order[q] = key[q].second;
    pos[key[q].second] = (int32_t)q;
  }
  std::vector<std::pair<uint64_t, int32_t>>().swap(key);
  // greedy cut
  std::vector<int32_t> stamp(n, -1);
  std::vector<int64_t> starts{0};
  int cur = 0, nint = 0, nfoot = 0;
  auto fresh = [&](int64_t p, int tile) {
    int c = stamp[p] != tile;
    for (int32_t e = lrp[p]; e < lrp[p + 1]; ++e) c += stamp[lci[e]] != tile;
    for (int32_t e = urp[p]; e < urp[p + 1]; ++e) c += stamp[uci[e]] != tile;
    return c;
  };
  for (int64_t q = 0; q < n; ++q) {
    const int64_t p = order[q];
    int add = fresh(p, cur);
    if (nint > 0 && (nint + 1 > kTileRows || nfoot + add > kTileFoot)) {
      ++cur;
      starts.push_back(q);
      nint = nfoot = 0;
      add = fresh(p, cur);
    }
    NALU_REQUIRE(add <= kTileFoot, NALU_ERR_UNSUPPORTED,
                 "learned preconditioner: a row has more neighbours than a message tile holds");
    stamp[p] = cur;
    for (int32_t e = lrp[p]; e < lrp[p + 1]; ++e) stamp[lci[e]] = cur;
    for (int32_t e = urp[p]; e < urp[p + 1]; ++e) stamp[uci[e]] = cur;
    nfoot += add;
    ++nint;
  }
  starts.push_back(n);
  const int64_t nt = (int64_t)starts.size() - 1;
  // per tile: halo, local indices, edges (all tiles padded to kFixedWidth when
  // no row has more off-diagonal entries)
  int maxdeg = 0;
  for (int64_t p = 0; p < n; ++p) maxdeg = std::max(maxdeg, (lrp[p + 1] - lrp[p]) + (urp[p + 1] - urp[p]));
  // NALU_GNN_RUNTIME_WIDTH=1 selects the runtime-width kernel for any graph (tests)
  const char* rw = std::getenv("NALU_GNN_RUNTIME_WIDTH");
  const bool fixed = maxdeg <= kFixedWidth && !(rw && std::atoi(rw) != 0);
  std::vector<MsgTile> tiles(nt);
  std::vector<int32_t> flist, esrc, loc(n, -1), mark(n, -1);
  std::vector<uint32_t> edge;
  std::vector<int32_t> hl;
  std::vector<std::pair<int32_t, int32_t>> ent;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t s0 = starts[t], s1 = starts[t + 1];
    const int ni = (int)(s1 - s0);
    hl.clear();
    int width = 0;
    for (int64_t q = s0; q < s1; ++q) {
      const int64_t p = order[q];
      width = std::max(width, (lrp[p + 1] - lrp[p]) + (urp[p + 1] - urp[p]));
      auto visit = [&](int32_t j) {
        const int32_t g = pos[j];
        if ((g < s0 || g >= s1) && mark[g] != (int32_t)t) {
          mark[g] = (int32_t)t;
          hl.push_back(g);
        }
      };
      for (int32_t e = lrp[p]; e < lrp[p + 1]; ++e) visit(lci[e]);
      for (int32_t e = urp[p]; e < urp[p + 1]; ++e) visit(uci[e]);
    }
    std::sort(hl.begin(), hl.end());
    NALU_REQUIRE(ni + (int64_t)hl.size() <= kTileFoot, NALU_ERR_STATE, "learned preconditioner: tile footprint");
    for (size_t k = 0; k < hl.size(); ++k) loc[hl[k]] = ni + (int32_t)k;
    for (int64_t q = s0; q < s1; ++q) loc[q] = (int32_t)(q - s0);
    width = fixed ? kFixedWidth : width + (width & 1);
    MsgTile& T = tiles[t];
    T.start = (int32_t)s0;
    T.nint = ni;
    T.nfoot = ni + (int32_t)hl.size();
    T.width = width;
    T.eoff = (int64_t)edge.size();
    T.foff = (int64_t)flist.size();
    for (int64_t q = s0; q < s1; ++q) flist.push_back((int32_t)q);  // footprint: the rows, then the halo
    flist.insert(flist.end(), hl.begin(), hl.end());
    edge.resize(edge.size() + (size_t)ni * width);
    esrc.resize(edge.size());
    for (int64_t q = s0; q < s1; ++q) {
      const int64_t p = order[q];
      const int tl = (int)(q - s0);
      ent.clear();
      for (int32_t e = lrp[p]; e < lrp[p + 1]; ++e) ent.push_back({loc[pos[lci[e]]], e});
      for (int32_t e = urp[p]; e < urp[p + 1]; ++e) ent.push_back({loc[pos[uci[e]]], (int32_t)(nl + e)});
      std::sort(ent.begin(), ent.end());
      for (int k = 0; k < width; ++k) {
        const int64_t slot = T.eoff + (int64_t)k * ni + tl;
        const bool real = k < (int)ent.size();
        edge[slot] = (uint32_t)((real ? ent[k].first : tl) * kFootStride * 16);
        esrc[slot] = real ? ent[k].second : -1;
      }
    }
  }
  st->ntiles = nt;
  st->tiles.alloc(sizeof(MsgTile) * nt);
  st->tfoot.alloc(flist.size());
  st->tfoot_s.alloc(flist.size());
  st->tedge.alloc(edge.size());
  st->tesrc.alloc(esrc.size());
  st->trow.alloc(n);
  NALU_HIP_CHECK(hipMemcpy(st->tiles.p, tiles.data(), sizeof(MsgTile) * nt, hipMemcpyHostToDevice));
  NALU_HIP_CHECK(hipMemcpy(st->tfoot.p, flist.data(), sizeof(int32_t) * flist.size(), hipMemcpyHostToDevice));
  for (auto& g : flist) g = order[g];
  NALU_HIP_CHECK(hipMemcpy(st->tfoot_s.p, flist.data(), sizeof(int32_t) * flist.size(), hipMemcpyHostToDevice));
  NALU_HIP_CHECK(hipMemcpy(st->tedge.p, edge.data(), sizeof(uint32_t) * edge.size(), hipMemcpyHostToDevice));
  NALU_HIP_CHECK(hipMemcpy(st->tesrc.p, esrc.data(), sizeof(int32_t) * esrc.size(), hipMemcpyHostToDevice));
  NALU_HIP_CHECK(hipMemcpy(st->trow.p, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  st->tnl = nl;
  st->tile_fixed = fixed;
  st->tiles_n = n;
}

// flag = 1 when some row sum of A = D + L + U exceeds 1e-5 of its diagonal
// (a pure-Neumann operator assembled in FP32 arithmetic sums to zero only to
// that precision)
__global__ void k_rowsum_check(const int32_t* __restrict__ lrp, const double* __restrict__ lv,
                               const int32_t* __restrict__ urp, const double* __restrict__ uv,
                               const double* __restrict__ dv, int64_t n, int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t = dv[i];
  for (int32_t q = lrp[i]; q < lrp[i + 1]; ++q) t += lv[q];
  for (int32_t q = urp[i]; q < urp[i + 1]; ++q) t += uv[q];
  if (!(std::fabs(t) <= 1e-5 * std::fabs(dv[i]))) flag[0] = 1;
}

// z -= mean(z) (sum[0] = sum of z)
__global__ void k_sub_mean(float* __restrict__ z, int64_t n, const double* __restrict__ sum) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) z[k] -= (float)(sum[0] / (double)n);
}

// diagonal sum and delta_i = dbar / a_ii of the current matrix (solver order)
void refresh_diag(nalu_linsys* ls, Learned* st) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  if (st->delta.n != (size_t)n) st->delta.alloc(n);
  if (st->dsum.n < 1) st->dsum.alloc(1);
  if (st->part.n < (size_t)kRedBlocks * 64) st->part.alloc((size_t)kRedBlocks * 64);
  const unsigned nb = std::min<unsigned>(kRedBlocks, nblocks(n));
  hipLaunchKernelGGL(k_bsum<double>, dim3(nb, 1), dim3(kThreads), 0, s, ls->dval.p, n, st->part.p);
  hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, st->part.p, (int)nb, st->dsum.p, 1, 0);
  hipLaunchKernelGGL(k_graph_delta, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->dval.p, n, st->dsum.p, st->delta.p);
  // pure-Neumann test of the current matrix
  if (st->rs_flag.n < 1) st->rs_flag.alloc(1);
  if (st->zsum.n < 1) st->zsum.alloc(1);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, s, st->rs_flag.p, (int64_t)1, 0);
  hipLaunchKernelGGL(k_rowsum_check, dim3(nblocks(n)), dim3(kThreads), 0, s, ls->lrp.p, ls->lval.p, ls->urp.p,
                     ls->uval.p, ls->dval.p, n, st->rs_flag.p);
  int h = 1;
  NALU_HIP_CHECK(hipMemcpyAsync(&h, st->rs_flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NALU_HIP_CHECK(hipStreamSynchronize(s));
  st->neumann = h == 0;
}

// Inference state for the current matrix: tiles (built once per system), message
// weights and diagonal scaling (every set-up).
void refresh_inference(nalu_linsys* ls, Learned* st) {
  NALU_REQUIRE(ls->num_dof == 1, NALU_ERR_UNSUPPORTED, "learned preconditioner: scalar systems only (numDof 1)");
  NALU_REQUIRE(ls->precond_ready && ls->perm_ready, NALU_ERR_STATE,
               "learned preconditioner: the preconditioner of the system is not set up");
  if (st->tiles_n != ls->nrows) build_tiles(ls, st);
  refresh_diag(ls, st);
  hipLaunchKernelGGL(k_tile_weights, dim3((unsigned)st->ntiles), dim3(kTileRows), 0, ls->stream,
                     reinterpret_cast<const MsgTile*>(st->tiles.p), st->tesrc.p, st->trow.p, ls->lval.p, ls->uval.p,
                     st->tnl, ls->dval.p, st->tedge.p);
  st->infer_ready = true;
}

// the network's base preconditioner in the permuted order (FP32): SGS, or the
// AMG V-cycle when the system's learned base is NALU_PREC_MUELU
void base32(nalu_linsys* ls, const float* r, float* z) {
  if (ls->learned_base == NALU_PREC_MUELU) amg_apply<float>(ls, r, z);
  else perm_sgs32(ls, r, z);
}

// z = net(r) for one vector in the permuted order (FP32)
void apply_perm(nalu_linsys* ls, Learned* st, const nalu_gnn* g, const float* r, float* z) {
  hipStream_t s = ls->stream;
  const int64_t n = ls->nrows;
  const Layout lay{g->L};
  NALU_REQUIRE(st->infer_ready, NALU_ERR_STATE, "learned preconditioner: inference state not set up");
  ensure_one(st, n);
  base32(ls, r, st->zs.p);
  const unsigned nb = std::min<unsigned>(kRedBlocks, nblocks(n));
  hipLaunchKernelGGL((k_bdot<float, float>), dim3(nb, 1), dim3(kThreads), 0, s, r, r, n, st->part.p);
  hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, st->part.p, (int)nb, st->sumsq.p, 1, 0);
  hipLaunchKernelGGL(k_scale, dim3(1), dim3(64), 0, s, st->sumsq.p, 1, n, st->dsum.p, st->scal.p);
  const int32_t* trow = st->trow.p;
  const MsgTile* tl = reinterpret_cast<const MsgTile*>(st->tiles.p);
  const EncodeArgs enc{st->tfoot_s.p, r, st->zs.p, st->delta.p, st->dsum.p, n, g->wbf.p + lay.w_in(),
                       g->theta.p + lay.b_in(), st->F.p};
  uint16_t *hin = st->h0.p, *hout = st->h1.p;
  for (int l = 0; l < g->L; ++l) {
    const bool first = l == 0, last = l == g->L - 1;
    using K = decltype(&k_layer_tile<false, false, 0>);
    K kern;
    if (st->tile_fixed)
      kern = first ? (last ? k_layer_tile<true, true, kFixedWidth> : k_layer_tile<true, false, kFixedWidth>)
                   : (last ? k_layer_tile<false, true, kFixedWidth> : k_layer_tile<false, false, kFixedWidth>);
    else
      kern = first ? (last ? k_layer_tile<true, true, 0> : k_layer_tile<true, false, 0>)
                   : (last ? k_layer_tile<false, true, 0> : k_layer_tile<false, false, 0>);
    hipLaunchKernelGGL(kern, dim3((unsigned)xcd_grid(st->ntiles)), dim3(kTileRows), 0, s, tl, st->ntiles,
                       st->tfoot.p, st->tedge.p, hin, hout, g->wbf.p + lay.w(l), g->theta.p + lay.b(l),
                       g->wbf.p + lay.w_out(), st->F.p, st->scal.p, trow, z, enc);
    std::swap(hin, hout);
  }
  if (st->neumann) {
    // the training loss cannot see the output's constant part (A 1 = 0 and the
    // residuals are zero-mean), so it is left undetermined; drop it
    hipLaunchKernelGGL(k_bsum<float>, dim3(nb, 1), dim3(kThreads), 0, s, (const float*)z, n, st->part.p);
    hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, st->part.p, (int)nb, st->zsum.p, 1, 0);
    hipLaunchKernelGGL(k_sub_mean, dim3(nblocks(n)), dim3(kThreads), 0, s, z, n, (const double*)st->zsum.p);
  }
}

// ---------------------------------------------------------------- batch work
struct Work {
  int64_t n = 0;
  int B = 0, L = 0;
  DevBuf<float> R, E, ZS, out, sym, gout, G, Gp, DM, scal, fpart, ftmp;
  std::vector<DevBuf<uint16_t>> h, mst;
  std::vector<DevBuf<uint32_t>> mask;
  DevBuf<float2> F;
  DevBuf<double> S, sumsq, pq, dpart, loss;
  void alloc(int64_t nn, int bb, int ll) {
    if (n == nn && B == bb && L == ll) return;
    n = nn;
    B = bb;
    L = ll;
    const int64_t T = n * B;
    R.alloc(T);
    E.alloc(T);
    ZS.alloc(T);
    out.alloc(T);
    sym.alloc(T);
    gout.alloc(T);
    G.alloc(T * kH);
    Gp.alloc(T * kH);
    DM.alloc(T * kH);
    scal.alloc(2 * B);
    F.alloc(T);
    h.clear();
    mst.clear();
    mask.clear();
    for (int l = 0; l <= L; ++l) h.emplace_back(T * kH);
    for (int l = 0; l < L; ++l) {
      mst.emplace_back(T * kH);
      mask.emplace_back(T);
    }
    S.alloc(B);
    sumsq.alloc(B);
    pq.alloc(2 * B);
    dpart.alloc((size_t)kRedBlocks * B * 2);
    loss.alloc(1);
    fpart.alloc((size_t)256 * kWaves * kLayerParams);
    ftmp.alloc((size_t)64 * kLayerParams);
  }
};

constexpr int kBwdBlocks = 256;  // grid of the reducing backward kernels
constexpr int kColChunk = 64;

// g[0..cols) = column sums of part[rows][cols]
void col_sum(hipStream_t s, const float* part, int rows, int cols, float* g, Work& w) {
  const int nch = (rows + kColChunk - 1) / kColChunk;
  if (nch == 1) {
    hipLaunchKernelGGL(k_col_sum, dim3(nblocks(cols), 1), dim3(kThreads), 0, s, part, rows, cols, rows, g);
    return;
  }
  NALU_REQUIRE((size_t)nch * cols <= w.ftmp.n, NALU_ERR_STATE, "learned preconditioner: reduction workspace");
  hipLaunchKernelGGL(k_col_sum, dim3(nblocks(cols), nch), dim3(kThreads), 0, s, part, rows, cols, kColChunk,
                     w.ftmp.p);
  hipLaunchKernelGGL(k_col_sum, dim3(nblocks(cols), 1), dim3(kThreads), 0, s, w.ftmp.p, nch, cols, nch, g);
}

// Forward + loss + backward on the batch held in w.R / w.ZS / w.S; loss into
// loss_slot (device), gradient into g->grad.
void fwd_bwd(nalu_linsys* ls, Learned* st, nalu_gnn* g, Work& w, double* loss_slot) {
  hipStream_t s = ls->stream;
  const int64_t n = w.n;
  const int B = w.B;
  const int64_t T = n * B;
  const Layout lay{g->L};
  const unsigned nb = std::min<unsigned>(kRedBlocks, nblocks(n));
  // forward
  hipLaunchKernelGGL((k_bdot<float, float>), dim3(nb, B), dim3(kThreads), 0, s, w.R.p, w.R.p, n, w.dpart.p);
  hipLaunchKernelGGL(k_brow_sum, dim3(B), dim3(kThreads), 0, s, w.dpart.p, (int)nb, w.sumsq.p, 1, 0);
  hipLaunchKernelGGL(k_scale, dim3(1), dim3(64), 0, s, w.sumsq.p, B, n, st->dsum.p, w.scal.p);
  hipLaunchKernelGGL(k_encode, dim3(nblocks(T * kH)), dim3(kThreads), 0, s, w.R.p, w.ZS.p, st->delta.p, st->dsum.p,
                     w.scal.p, n, T, g->wbf.p, g->theta.p, w.h[0].p, w.F.p);
  for (int l = 0; l < g->L; ++l)
    hipLaunchKernelGGL(k_layer<true>, dim3(nblocks(T, kTile * kWaves)), dim3(kThreads), 0, s, st->gptr.p, st->gcol.p,
                       st->gw.p, n, T, w.h[l].p, w.h[l + 1].p, g->wbf.p + lay.w(l), g->theta.p + lay.b(l),
                       w.mst[l].p, w.mask[l].p);
  hipLaunchKernelGGL(k_output<float>, dim3(nblocks(T)), dim3(kThreads), 0, s, w.h[g->L].p, w.F.p,
                     g->wbf.p + lay.w_out(), w.scal.p, n, T, 1, w.out.p, (int*)nullptr);
  // loss
  hipLaunchKernelGGL(k_aop, dim3(nb, B), dim3(kThreads), 0, s, st->gptr.p, st->gcol.p, st->gw.p, st->gwt.p, st->dF.p,
                     n, w.out.p, w.R.p, w.sym.p, w.dpart.p);
  hipLaunchKernelGGL(k_pq_sum, dim3(B), dim3(kThreads), 0, s, w.dpart.p, (int)nb, w.pq.p);
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(64), 0, s, w.pq.p, w.S.p, B, loss_slot, 0);
  hipLaunchKernelGGL(k_gout, dim3(nblocks(T)), dim3(kThreads), 0, s, w.R.p, w.sym.p, w.pq.p, w.S.p, n, B, w.gout.p);
  // backward
  const unsigned nbo = std::min<unsigned>(kBwdBlocks, nblocks(T * kH));
  hipLaunchKernelGGL(k_output_bwd, dim3(nbo), dim3(kThreads), 0, s, w.gout.p, w.h[g->L].p, w.F.p,
                     g->wbf.p + lay.w_out(), T, w.G.p, w.fpart.p);
  col_sum(s, w.fpart.p, (int)nbo, 34, g->grad.p + lay.w_out(), w);
  const int64_t ntiles = (T + kTile - 1) / kTile;
  const unsigned nbl = (unsigned)std::min<int64_t>(kBwdBlocks, (ntiles + kWaves - 1) / kWaves);
  for (int l = g->L - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_layer_bwd, dim3(nbl), dim3(kThreads), 0, s, T, w.G.p, w.mask[l].p, w.h[l].p, w.mst[l].p,
                       g->wbf.p + lay.w(l), w.Gp.p, w.DM.p, w.fpart.p);
    col_sum(s, w.fpart.p, (int)nbl * kWaves, kLayerParams, g->grad.p + lay.w(l), w);
    hipLaunchKernelGGL(k_agg_t, dim3(nblocks(T * 8)), dim3(kThreads), 0, s, st->gptr.p, st->gcol.p, st->gwt.p, n, T,
                       w.Gp.p, w.DM.p, w.G.p);
  }
  const unsigned nbe = std::min<unsigned>(kBwdBlocks, nblocks(T * kH));
  hipLaunchKernelGGL(k_encode_bwd, dim3(nbe), dim3(kThreads), 0, s, w.G.p, w.F.p, st->delta.p, n, T, w.fpart.p);
  col_sum(s, w.fpart.p, (int)nbe, 128, g->grad.p, w);
}

// Seeded initial weights: N(0, 1/3) encoder, N(0, 0.25/64) layers, zero biases,
// zero output weights and c = (0, 1), so the untrained network is SGS.
std::vector<float> init_params(int L, uint64_t seed) {
  const Layout lay{L};
  std::vector<float> th((size_t)lay.total(), 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int64_t k = lay.w_in(); k < lay.b_in(); ++k) th[k] = (float)(nd(rng) / std::sqrt((double)kIn));
  for (int l = 0; l < L; ++l)
    for (int64_t k = lay.w(l); k < lay.b(l); ++k) th[k] = (float)(nd(rng) * 0.5 / std::sqrt(2.0 * kH));
  th[lay.c() + 1] = 1.f;
  return th;
}

void upload_params(nalu_gnn* g, const float* host, hipStream_t s) {
  const int64_t np = (int64_t)g->theta.n;
  NALU_HIP_CHECK(hipMemcpyAsync(g->theta.p, host, sizeof(float) * np, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_to_bf16, dim3(nblocks(np)), dim3(kThreads), 0, s, g->theta.p, np, g->wbf.p);
  NALU_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

// ------------------------------------------------------------ hooks for the solver
void learned_setup(nalu_linsys* ls) {
  Learned* st = state(ls);
  if (!st) return;
  st->graph_ready = st->wt_ready = st->infer_ready = false;  // the matrix changed
  if (st->net) refresh_inference(ls, st);
}

bool learned_active(const nalu_linsys* ls) {
  const Learned* st = state(ls);
  return st && st->net && st->infer_ready && ls->precond_ready && !st->fallback;
}

void learned_set_fallback(nalu_linsys* ls, bool on) {
  Learned* st = state(ls);
  if (st) st->fallback = on;
}

bool learned_fallback(const nalu_linsys* ls) {
  const Learned* st = state(ls);
  return st && st->fallback;
}

void learned_apply_f32(nalu_linsys* ls, const float* r, float* z) {
  Learned* st = state(ls);
  apply_perm(ls, st, st->net, r, z);
}

void learned_apply_f64(nalu_linsys* ls, const double* r, double* z) {
  Learned* st = state(ls);
  const int64_t n = ls->nrows;
  ensure_one(st, n);
  hipLaunchKernelGGL((k_convert<double, float>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, r, n, st->r32.p);
  apply_perm(ls, st, st->net, st->r32.p, st->z32.p);
  hipLaunchKernelGGL((k_convert<float, double>), dim3(nblocks(n)), dim3(kThreads), 0, ls->stream, st->z32.p, n, z);
}

void learned_release(nalu_linsys* ls) {
  delete state(ls);
  ls->learned = nullptr;
}

}  // namespace nalu

// =================================================================== C ABI
using nalu::DevBuf;
using nalu::guarded;

extern "C" int nalu_gnn_create(int num_layers, int hidden, uint64_t seed, nalu_gnn** out) {
  return guarded([&]() -> int {
    NALU_REQUIRE(out, NALU_ERR_ARG, "gnn_create: null output");
    NALU_REQUIRE(num_layers >= 1 && num_layers <= 16, NALU_ERR_ARG, "gnn_create: layers must be in [1, 16]");
    NALU_REQUIRE(hidden == nalu::kH, NALU_ERR_UNSUPPORTED, "gnn_create: hidden width must be 32");
    auto g = new nalu_gnn();
    g->L = num_layers;
    const int64_t np = nalu::Layout{num_layers}.total();
    g->theta.alloc(np);
    g->m1.alloc(np);
    g->m2.alloc(np);
    g->grad.alloc(np);
    g->wbf.alloc(np);
    NALU_HIP_CHECK(hipMemset(g->m1.p, 0, sizeof(float) * np));
    NALU_HIP_CHECK(hipMemset(g->m2.p, 0, sizeof(float) * np));
    NALU_HIP_CHECK(hipMemset(g->grad.p, 0, sizeof(float) * np));
    const std::vector<float> th = nalu::init_params(num_layers, seed);
    nalu::upload_params(g, th.data(), nullptr);
    *out = g;
    return NALU_OK;
  });
}

extern "C" int nalu_gnn_destroy(nalu_gnn* g) {
  if (!g) return NALU_OK;
  (void)hipDeviceSynchronize();
  delete g;
  return NALU_OK;
}

extern "C" int64_t nalu_gnn_num_params(const nalu_gnn* g) {
  return g ? nalu::Layout{g->L}.total() : (int64_t)NALU_ERR_ARG;
}

extern "C" int nalu_gnn_num_layers(const nalu_gnn* g) { return g ? g->L : NALU_ERR_ARG; }

extern "C" int nalu_gnn_get_params(const nalu_gnn* g, float* host, int64_t n) {
  return guarded([&]() -> int {
    nalu::check_net(g);
    NALU_REQUIRE(host && n == (int64_t)g->theta.n, NALU_ERR_ARG, "gnn_get_params: size mismatch");
    NALU_HIP_CHECK(hipDeviceSynchronize());
    NALU_HIP_CHECK(hipMemcpy(host, g->theta.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return NALU_OK;
  });
}

extern "C" int nalu_gnn_set_params(nalu_gnn* g, const float* host, int64_t n) {
  return guarded([&]() -> int {
    nalu::check_net(g);
    NALU_REQUIRE(host && n == (int64_t)g->theta.n, NALU_ERR_ARG, "gnn_set_params: size mismatch");
    NALU_HIP_CHECK(hipDeviceSynchronize());
    nalu::upload_params(g, host, nullptr);
    return NALU_OK;
  });
}

extern "C" int nalu_gnn_get_grad(const nalu_gnn* g, float* host, int64_t n) {
  return guarded([&]() -> int {
    nalu::check_net(g);
    NALU_REQUIRE(host && n == (int64_t)g->grad.n, NALU_ERR_ARG, "gnn_get_grad: size mismatch");
    NALU_HIP_CHECK(hipDeviceSynchronize());
    NALU_HIP_CHECK(hipMemcpy(host, g->grad.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_set_learned_base(nalu_linsys* ls, int precond) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls, NALU_ERR_ARG, "set_learned_base: null system");
    NALU_REQUIRE(precond == NALU_PREC_SGS || precond == NALU_PREC_MUELU, NALU_ERR_ARG,
                 "set_learned_base: the base is NALU_PREC_SGS or NALU_PREC_MUELU");
    if (precond != ls->learned_base) {
      ls->learned_base = precond;
      ls->precond_ready = false;  // the next set-up builds (or drops) the hierarchy
    }
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_set_learned_precond(nalu_linsys* ls, nalu_gnn* g) {
  return guarded([&]() -> int {
    NALU_REQUIRE(ls, NALU_ERR_ARG, "set_learned_precond: null system");
    if (!g) {
      if (ls->learned) nalu::state(ls)->net = nullptr;
      return NALU_OK;
    }
    nalu::check_net(g);
    NALU_REQUIRE(ls->num_dof == 1, NALU_ERR_UNSUPPORTED, "set_learned_precond: scalar systems only (numDof 1)");
    NALU_REQUIRE(ls->cfg.method == NALU_METHOD_FGMRES && ls->cfg.precond == NALU_PREC_LEARNED, NALU_ERR_ARG,
                 "set_learned_precond: the system must be configured with method fgmres and preconditioner learned");
    nalu::Learned* st = nalu::ensure_state(ls);
    st->net = g;
    st->fallback = false;
    if (ls->precond_ready) {
      nalu::ensure_split32(ls);
      nalu::refresh_inference(ls, st);
    }
    return NALU_OK;
  });
}

extern "C" int nalu_linsys_learned_fallback(const nalu_linsys* ls) {
  return ls ? (nalu::learned_fallback(ls) ? 1 : 0) : NALU_ERR_ARG;
}

// z = net(r) for one right-hand side in the original row order (FP32 device arrays)
extern "C" int nalu_gnn_apply(nalu_gnn* g, nalu_linsys* ls, const float* r_dev, float* z_dev) {
  return guarded([&]() -> int {
    nalu::check_net(g);
    NALU_REQUIRE(ls && r_dev && z_dev, NALU_ERR_ARG, "gnn_apply: bad arguments");
    NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "gnn_apply: set up the preconditioner first");
    nalu::ensure_split32(ls);
    nalu::Learned* st = nalu::ensure_state(ls);
    if (!st->infer_ready) nalu::refresh_inference(ls, st);
    const int64_t n = ls->nrows;
    hipStream_t s = ls->stream;
    nalu::ensure_one(st, n);
    hipLaunchKernelGGL((nalu::k_gather_perm<float, float>), dim3(nalu::nblocks(n)), dim3(nalu::kThreads), 0, s, r_dev,
                       ls->perm.p, n, 1, st->r32.p);
    nalu::apply_perm(ls, st, g, st->r32.p, st->z32.p);
    hipLaunchKernelGGL((nalu::k_scatter_perm<float, float>), dim3(nalu::nblocks(n)), dim3(nalu::kThreads), 0, s,
                       st->z32.p, ls->perm.p, n, 1, z_dev);
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    return NALU_OK;
  });
}

// Loss and gradient of one batch of B pairs given in the original row order.
extern "C" int nalu_gnn_loss_grad(nalu_gnn* g, nalu_linsys* ls, int batch, const float* r_dev, const float* e_dev,
                                  float* zs_dev, float* out_dev, double* loss_host, float* grad_host) {
  return guarded([&]() -> int {
    nalu::check_net(g);
    NALU_REQUIRE(ls && r_dev && e_dev && batch >= 1 && batch <= 64, NALU_ERR_ARG, "gnn_loss_grad: bad arguments");
    NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "gnn_loss_grad: set up the preconditioner first");
    nalu::ensure_split32(ls);
    nalu::Learned* st = nalu::ensure_state(ls);
    if (!st->wt_ready) nalu::build_msg_graph(ls, st, true);
    const int64_t n = ls->nrows, T = n * batch;
    hipStream_t s = ls->stream;
    nalu::Work w;
    w.alloc(n, batch, g->L);
    hipLaunchKernelGGL((nalu::k_gather_perm<float, float>), dim3(nalu::nblocks(T)), dim3(nalu::kThreads), 0, s, r_dev,
                       ls->perm.p, n, batch, w.R.p);
    hipLaunchKernelGGL((nalu::k_gather_perm<float, float>), dim3(nalu::nblocks(T)), dim3(nalu::kThreads), 0, s, e_dev,
                       ls->perm.p, n, batch, w.E.p);
    for (int b = 0; b < batch; ++b) nalu::base32(ls, w.R.p + b * n, w.ZS.p + b * n);
    const unsigned nb = std::min<unsigned>(nalu::kRedBlocks, nalu::nblocks(n));
    hipLaunchKernelGGL((nalu::k_bdot<float, float>), dim3(nb, batch), dim3(nalu::kThreads), 0, s, w.E.p, w.R.p, n,
                       w.dpart.p);
    hipLaunchKernelGGL(nalu::k_brow_sum, dim3(batch), dim3(nalu::kThreads), 0, s, w.dpart.p, (int)nb, w.S.p, 1, 0);
    nalu::fwd_bwd(ls, st, g, w, w.loss.p);
    if (zs_dev)
      hipLaunchKernelGGL((nalu::k_scatter_perm<float, float>), dim3(nalu::nblocks(T)), dim3(nalu::kThreads), 0, s,
                         w.ZS.p, ls->perm.p, n, batch, zs_dev);
    if (out_dev)
      hipLaunchKernelGGL((nalu::k_scatter_perm<float, float>), dim3(nalu::nblocks(T)), dim3(nalu::kThreads), 0, s,
                         w.out.p, ls->perm.p, n, batch, out_dev);
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    if (loss_host) NALU_HIP_CHECK(hipMemcpy(loss_host, w.loss.p, sizeof(double), hipMemcpyDeviceToHost));
    if (grad_host)
      NALU_HIP_CHECK(hipMemcpy(grad_host, g->grad.p, sizeof(float) * g->grad.n, hipMemcpyDeviceToHost));
    return NALU_OK;
  });
}

// Harvest (residual, error) pairs from plain FP64 CG + SGS solves of the
// system's matrix with seeded unit-variance Gaussian zero-mean right-hand sides.
extern "C" int nalu_gnn_harvest(nalu_linsys* ls, int num_rhs, int pairs_per_rhs, uint64_t seed, double tol,
                                int max_iterations, nalu_gnn_data** out) {
  return guarded([&]() -> int {
    using namespace nalu;
    NALU_REQUIRE(ls && out && num_rhs >= 1 && pairs_per_rhs >= 1 && tol > 0 && max_iterations >= 1, NALU_ERR_ARG,
                 "gnn_harvest: bad arguments");
    NALU_REQUIRE(ls->finalized && ls->num_dof == 1, NALU_ERR_STATE, "gnn_harvest: needs a finalized scalar system");
    if (!ls->precond_ready) precond_setup(ls);
    ensure_split32(ls);
    hipStream_t s = ls->stream;
    const int64_t n = ls->nrows;
    const int64_t cap = std::min<int64_t>(max_iterations, std::max<int64_t>(pairs_per_rhs, ((int64_t)1 << 31) / (16 * n)));
    NALU_REQUIRE(cap >= pairs_per_rhs, NALU_ERR_ARG, "gnn_harvest: too many pairs per right-hand side for this size");
    auto d = new nalu_gnn_data();
    std::unique_ptr<nalu_gnn_data> guard(d);
    d->owner = ls;
    d->n = n;
    d->npairs = num_rhs * pairs_per_rhs;
    d->R.alloc((size_t)d->npairs * n);
    d->E.alloc((size_t)d->npairs * n);
    d->ZS.alloc((size_t)d->npairs * n);
    d->S.alloc(d->npairs);
    DevBuf<double> X((size_t)cap * n), Rr((size_t)cap * n), b(n), x(n), r(n), z(n), p(n), q(n), sc(8), part((size_t)kRedBlocks);
    const unsigned nb = std::min<unsigned>(kRedBlocks, nblocks(n));
    auto dot_into = [&](const double* a, const double* c, double* slot) {
      hipLaunchKernelGGL((k_bdot<double, double>), dim3(nb, 1), dim3(kThreads), 0, s, a, c, n, part.p);
      hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, part.p, (int)nb, slot, 1, 0);
    };
    double host[2];
    int pair = 0;
    for (int rhs = 0; rhs < num_rhs; ++rhs) {
      hipLaunchKernelGGL(k_gauss, dim3(nblocks(n)), dim3(kThreads), 0, s, seed * 0x100000001B3ull + (uint64_t)rhs, n, b.p);
      hipLaunchKernelGGL(k_bsum<double>, dim3(nb, 1), dim3(kThreads), 0, s, b.p, n, part.p);
      hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, part.p, (int)nb, sc.p + 4, 1, 0);
      hipLaunchKernelGGL(k_shift, dim3(nblocks(n)), dim3(kThreads), 0, s, b.p, n, sc.p + 4);
      NALU_HIP_CHECK(hipMemsetAsync(x.p, 0, sizeof(double) * n, s));
      NALU_HIP_CHECK(hipMemcpyAsync(r.p, b.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      dot_into(r.p, r.p, sc.p + 2);
      NALU_HIP_CHECK(hipMemcpyAsync(host, sc.p + 2, sizeof(double), hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      const double r0 = std::sqrt(host[0]);
      NALU_REQUIRE(r0 > 0 && std::isfinite(r0), NALU_ERR_NUMERIC, "gnn_harvest: bad right-hand side");
      perm_sgs64(ls, r.p, z.p);
      NALU_HIP_CHECK(hipMemcpyAsync(p.p, z.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      dot_into(r.p, z.p, sc.p + 0);
      int64_t rec = 0;
      int it = 0;
      bool conv = false;
      double cur = r0, fin = r0;
      std::vector<double> rec_norm;
      while (it < max_iterations) {
        if (rec < cap) {
          NALU_HIP_CHECK(hipMemcpyAsync(X.p + rec * n, x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
          NALU_HIP_CHECK(hipMemcpyAsync(Rr.p + rec * n, r.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
          rec_norm.push_back(cur);
          ++rec;
        }
        ++it;
        perm_spmv64(ls, p.p, q.p);
        dot_into(p.p, q.p, sc.p + 1);
        hipLaunchKernelGGL(k_cg_step, dim3(nblocks(n)), dim3(kThreads), 0, s, x.p, r.p, p.p, q.p, sc.p, n);
        dot_into(r.p, r.p, sc.p + 2);
        NALU_HIP_CHECK(hipMemcpyAsync(host, sc.p + 2, sizeof(double), hipMemcpyDeviceToHost, s));
        NALU_HIP_CHECK(hipStreamSynchronize(s));
        NALU_REQUIRE(std::isfinite(host[0]), NALU_ERR_NUMERIC, "gnn_harvest: non-finite CG residual");
        cur = std::sqrt(host[0]);
        if (cur <= tol * r0) {
          conv = true;
          fin = cur;
          break;
        }
        perm_sgs64(ls, r.p, z.p);
        dot_into(r.p, z.p, sc.p + 3);
        hipLaunchKernelGGL(k_cg_dir, dim3(nblocks(n)), dim3(kThreads), 0, s, p.p, z.p, sc.p + 3, sc.p + 0, n);
        NALU_HIP_CHECK(hipMemcpyAsync(sc.p + 0, sc.p + 3, sizeof(double), hipMemcpyDeviceToDevice, s));
      }
      NALU_REQUIRE(conv, NALU_ERR_NUMERIC, "gnn_harvest: CG did not reach the harvest tolerance");
      // the error x* - x_k is known to |r_final| / |r_k|: keep the iterates whose
      // residual is at least 1e3 times the final one (errors accurate to 1e-3)
      int64_t usable = 0;
      while (usable < rec && rec_norm[usable] >= 1e3 * fin) ++usable;
      NALU_REQUIRE(usable >= pairs_per_rhs, NALU_ERR_ARG,
                   "gnn_harvest: too few accurate CG iterates for the pairs per right-hand side (lower tol)");
      d->cg_iters.push_back(it);
      for (int k = 0; k < pairs_per_rhs; ++k, ++pair) {
        const int64_t j = (int64_t)k * usable / pairs_per_rhs;  // evenly spread over the usable iterates
        dot_into(Rr.p + j * n, Rr.p + j * n, sc.p + 5);
        // e = x* - x_j with its constant part removed: the right-hand sides are
        // zero-mean (pure-Neumann pressure systems), whose solutions are fixed
        // up to a constant, and a drifting constant would swamp e in FP32
        hipLaunchKernelGGL(k_sub, dim3(nblocks(n)), dim3(kThreads), 0, s, x.p, X.p + j * n, n, q.p);
        hipLaunchKernelGGL(k_bsum<double>, dim3(nb, 1), dim3(kThreads), 0, s, q.p, n, part.p);
        hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, part.p, (int)nb, sc.p + 6, 1, 0);
        hipLaunchKernelGGL(k_pair, dim3(nblocks(n)), dim3(kThreads), 0, s, Rr.p + j * n, q.p, sc.p + 6, sc.p + 5, n,
                           d->R.p + (int64_t)pair * n, d->E.p + (int64_t)pair * n);
        base32(ls, d->R.p + (int64_t)pair * n, d->ZS.p + (int64_t)pair * n);
        hipLaunchKernelGGL((k_bdot<float, float>), dim3(nb, 1), dim3(kThreads), 0, s, d->E.p + (int64_t)pair * n,
                           d->R.p + (int64_t)pair * n, n, part.p);
        hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, part.p, (int)nb, d->S.p + pair, 1, 0);
      }
    }
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    *out = guard.release();
    return NALU_OK;
  });
}

extern "C" int nalu_gnn_data_destroy(nalu_gnn_data* d) {
  if (!d) return NALU_OK;
  (void)hipDeviceSynchronize();
  delete d;
  return NALU_OK;
}

extern "C" int nalu_gnn_data_size(const nalu_gnn_data* d) { return d ? d->npairs : NALU_ERR_ARG; }

extern "C" int nalu_gnn_data_cg_iterations(const nalu_gnn_data* d, int* iters_host, int count) {
  return guarded([&]() -> int {
    NALU_REQUIRE(d && iters_host && count == (int)d->cg_iters.size(), NALU_ERR_ARG, "gnn_data_cg_iterations: size");
    std::copy(d->cg_iters.begin(), d->cg_iters.end(), iters_host);
    return NALU_OK;
  });
}

// Pair idx in the original row order (FP32 device arrays of the owner's size).
extern "C" int nalu_gnn_data_pair(const nalu_gnn_data* d, int idx, float* r_dev, float* e_dev) {
  return guarded([&]() -> int {
    NALU_REQUIRE(d && r_dev && e_dev && idx >= 0 && idx < d->npairs, NALU_ERR_ARG, "gnn_data_pair: bad arguments");
    const nalu_linsys* ls = d->owner;
    const int64_t n = d->n;
    hipLaunchKernelGGL((nalu::k_scatter_perm<float, float>), dim3(nalu::nblocks(n)), dim3(nalu::kThreads), 0,
                       ls->stream, d->R.p + (int64_t)idx * n, ls->perm.p, n, 1, r_dev);
    hipLaunchKernelGGL((nalu::k_scatter_perm<float, float>), dim3(nalu::nblocks(n)), dim3(nalu::kThreads), 0,
                       ls->stream, d->E.p + (int64_t)idx * n, ls->perm.p, n, 1, e_dev);
    NALU_HIP_CHECK(hipStreamSynchronize(ls->stream));
    return NALU_OK;
  });
}

// Overwrites pair idx with a caller's (r, e), FP64 device vectors in the row
// order with A e = r (for instance the residual and error of an iterate of a
// solve whose solution is known), stored like a harvested pair: solver order,
// both scaled by 1/|r|, e off its mean, the base preconditioner's output of r
// as the stored feature.
extern "C" int nalu_gnn_data_set_pair(nalu_gnn_data* d, nalu_linsys* ls, int idx, const double* r_dev,
                                      const double* e_dev) {
  return guarded([&]() -> int {
    using namespace nalu;
    NALU_REQUIRE(d && ls && r_dev && e_dev && idx >= 0 && idx < d->npairs, NALU_ERR_ARG,
                 "gnn_data_set_pair: bad arguments");
    NALU_REQUIRE(d->owner == ls && d->n == ls->nrows, NALU_ERR_ARG, "gnn_data_set_pair: data of another system");
    hipStream_t s = ls->stream;
    const int64_t n = d->n;
    const unsigned nb = std::min<unsigned>(kRedBlocks, nblocks(n));
    DevBuf<double> rp(n), ep(n), part(kRedBlocks), sc(2);
    hipLaunchKernelGGL((k_gather_perm<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, r_dev, ls->perm.p, n,
                       1, rp.p);
    hipLaunchKernelGGL((k_gather_perm<double, double>), dim3(nblocks(n)), dim3(kThreads), 0, s, e_dev, ls->perm.p, n,
                       1, ep.p);
    hipLaunchKernelGGL((k_bdot<double, double>), dim3(nb, 1), dim3(kThreads), 0, s, rp.p, rp.p, n, part.p);
    hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, part.p, (int)nb, sc.p + 0, 1, 0);
    hipLaunchKernelGGL(k_bsum<double>, dim3(nb, 1), dim3(kThreads), 0, s, ep.p, n, part.p);
    hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, part.p, (int)nb, sc.p + 1, 1, 0);
    double nr2 = 0.0;
    NALU_HIP_CHECK(hipMemcpyAsync(&nr2, sc.p, sizeof(double), hipMemcpyDeviceToHost, s));
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    NALU_REQUIRE(nr2 > 0 && std::isfinite(nr2), NALU_ERR_NUMERIC, "gnn_data_set_pair: zero or non-finite residual");
    float* R = d->R.p + (int64_t)idx * n;
    float* E = d->E.p + (int64_t)idx * n;
    hipLaunchKernelGGL(k_pair, dim3(nblocks(n)), dim3(kThreads), 0, s, rp.p, ep.p, sc.p + 1, sc.p + 0, n, R, E);
    if (!ls->precond_ready) precond_setup(ls);
    ensure_split32(ls);
    base32(ls, R, d->ZS.p + (int64_t)idx * n);
    hipLaunchKernelGGL((k_bdot<float, float>), dim3(nb, 1), dim3(kThreads), 0, s, E, R, n, part.p);
    hipLaunchKernelGGL(k_brow_sum, dim3(1), dim3(kThreads), 0, s, part.p, (int)nb, d->S.p + idx, 1, 0);
    NALU_HIP_CHECK(hipStreamSynchronize(s));
    return NALU_OK;
  });
}

// Adam training on the harvested pairs of the same system; epoch_loss_host[e]
// receives the mean batch loss of epoch e.  The learning rate decays from lr to
// lr/10 on a cosine over all steps.
// Epochs [first, first + count) of a run of total_epochs: the learning-rate
// schedule spans the whole run and the pair order of epoch k is the k-th shuffle
// of one generator seeded with seed, so that a run split into ranges (with
// validation in between) follows the same sequence as one call.
extern "C" int nalu_gnn_train_epochs(nalu_gnn* g, nalu_linsys* ls, const nalu_gnn_data* d, int first_epoch,
                                     int num_epochs, int total_epochs, int batch, double lr, uint64_t seed,
                                     double* epoch_loss_host) {
  return guarded([&]() -> int {
    using namespace nalu;
    check_net(g);
    NALU_REQUIRE(ls && d && num_epochs >= 1 && first_epoch >= 0 && first_epoch + num_epochs <= total_epochs &&
                     batch >= 1 && batch <= 64 && lr > 0,
                 NALU_ERR_ARG, "gnn_train: bad arguments");
    NALU_REQUIRE(d->owner == ls && d->n == ls->nrows, NALU_ERR_ARG, "gnn_train: data harvested from another system");
    NALU_REQUIRE(ls->precond_ready, NALU_ERR_STATE, "gnn_train: set up the preconditioner first");
    ensure_split32(ls);
    Learned* st = ensure_state(ls);
    if (!st->wt_ready) build_msg_graph(ls, st, true);
    hipStream_t s = ls->stream;
    const int64_t n = ls->nrows;
    const int B = std::min(batch, d->npairs);
    const int spe = d->npairs / B;
    const int64_t total_steps = (int64_t)total_epochs * spe;
    const int64_t np = (int64_t)g->theta.n;
    Work w;
    w.alloc(n, B, g->L);
    DevBuf<int32_t> idx(d->npairs);
    DevBuf<double> losses(spe);
    std::vector<int32_t> order(d->npairs);
    std::iota(order.begin(), order.end(), 0);
    std::mt19937_64 rng(seed);
    std::vector<double> lh(spe);
    for (int ep = 0; ep < first_epoch; ++ep) std::shuffle(order.begin(), order.end(), rng);
    int64_t step = (int64_t)first_epoch * spe;
    for (int ep = 0; ep < num_epochs; ++ep) {
      std::shuffle(order.begin(), order.end(), rng);
      NALU_HIP_CHECK(hipMemcpyAsync(idx.p, order.data(), sizeof(int32_t) * d->npairs, hipMemcpyHostToDevice, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      for (int k = 0; k < spe; ++k, ++step) {
        const int32_t* bi = idx.p + (int64_t)k * B;
        hipLaunchKernelGGL(k_gather_rows, dim3(nblocks(n * B)), dim3(kThreads), 0, s, d->R.p, bi, n, B, w.R.p);
        hipLaunchKernelGGL(k_gather_rows, dim3(nblocks(n * B)), dim3(kThreads), 0, s, d->ZS.p, bi, n, B, w.ZS.p);
        hipLaunchKernelGGL(k_gather_s, dim3(1), dim3(64), 0, s, d->S.p, bi, B, w.S.p);
        fwd_bwd(ls, st, g, w, losses.p + k);
        ++g->adam_t;
        const double frac = total_steps > 1 ? (double)step / (double)(total_steps - 1) : 1.0;
        const double lr_t = lr * (0.1 + 0.9 * 0.5 * (1.0 + std::cos(M_PI * frac)));
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.f - std::pow(b1, (float)g->adam_t), bc2 = 1.f - std::pow(b2, (float)g->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nblocks(np)), dim3(kThreads), 0, s, g->theta.p, g->m1.p, g->m2.p, g->grad.p,
                           np, (float)lr_t, b1, b2, bc1, bc2, 1e-8f, g->wbf.p);
      }
      NALU_HIP_CHECK(hipMemcpyAsync(lh.data(), losses.p, sizeof(double) * spe, hipMemcpyDeviceToHost, s));
      NALU_HIP_CHECK(hipStreamSynchronize(s));
      double m = 0.0;
      for (double v : lh) m += v;
      m /= spe;
      NALU_REQUIRE(std::isfinite(m), NALU_ERR_NUMERIC, "gnn_train: non-finite loss");
      if (epoch_loss_host) epoch_loss_host[ep] = m;
    }
    return NALU_OK;
  });
}

extern "C" int nalu_gnn_train(nalu_gnn* g, nalu_linsys* ls, const nalu_gnn_data* d, int epochs, int batch,
                              double lr, uint64_t seed, double* epoch_loss_host) {
  return nalu_gnn_train_epochs(g, ls, d, 0, epochs, epochs, batch, lr, seed, epoch_loss_host);
}
