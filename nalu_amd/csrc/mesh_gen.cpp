// Host-side synthetic Hex8 lattice + random field generator.
//
// Node numbering follows the STK generated mesh ("generated:NxNxN"): node id
// = 1 + i + (n+1)(j + (n+1)k), element nodes in Hex8 order, which is what the
// reference's unit tests rely on (unit_tests/UnitTestTpetra.C:97-98).
// Randomness comes from the reference's machine-independent LCG
// (unit_tests/UnitTestUtils.C:41-53) so that every platform produces the
// same mesh for the same seed.
#include <cmath>
#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace {

// x_{k+1} = (a x_k + c) mod m, mapped to [-1, 1) as 2 x/m - 1.
struct Lcg {
  explicit Lcg(uint32_t seed) : prev(seed) {}
  double operator()() {
    prev = (uint32_t)((1103515245ull * (uint64_t)prev + 12345ull) % 2147483647ull);
    return 2.0 * ((double)prev / 2147483647.0) - 1.0;
  }
  uint32_t prev;
};

// Standard normal by Box-Muller from two LCG draws (one sample per pair).
double lcg_normal(Lcg& g) {
  const double u1 = 1.0 - 0.5 * (g() + 1.0);  // (0, 1]
  const double u2 = 0.5 * (g() + 1.0);        // [0, 1)
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// 1-D node positions on [0,1] with n intervals.  Uniform when stretch <= 1;
// otherwise intervals grow by `stretch` from both ends toward the middle,
// growth capped at max_aspect times the wall spacing.
std::vector<double> axis_positions(int n, double stretch, double max_aspect) {
  std::vector<double> w(n, 1.0);
  if (stretch > 1.0) {
    for (int i = 0; i < n; ++i) {
      const int d = std::min(i, n - 1 - i);  // distance (in intervals) from the nearest wall
      double s = std::pow(stretch, (double)d);
      if (max_aspect > 0.0 && s > max_aspect) s = max_aspect;
      w[i] = s;
    }
  }
  double tot = 0.0;
  for (double v : w) tot += v;
  std::vector<double> x(n + 1, 0.0);
  for (int i = 0; i < n; ++i) x[i + 1] = x[i] + w[i] / tot;
  x[n] = 1.0;
  return x;
}

}  // namespace

extern "C" int nalu_lattice_generate(int n, double jitter, double stretch, double max_aspect,
                                     uint32_t seed, double* coords, int32_t* conn) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(n >= 1 && n <= 1024, NALU_ERR_ARG, "lattice: n out of range");
    NALU_REQUIRE(coords && conn, NALU_ERR_ARG, "lattice: null output");
    const int64_t n1 = n + 1;
    const std::vector<double> ax = axis_positions(n, stretch, max_aspect);
    // local spacing bound per lattice index: min of the two adjacent intervals
    std::vector<double> hl(n1, 0.0);
    for (int i = 1; i < n; ++i) hl[i] = std::min(ax[i] - ax[i - 1], ax[i + 1] - ax[i]);
    Lcg g(seed);
    for (int64_t k = 0; k < n1; ++k)
      for (int64_t j = 0; j < n1; ++j)
        for (int64_t i = 0; i < n1; ++i) {
          const int64_t idx = i + n1 * (j + n1 * k);
          const int64_t ijk[3] = {i, j, k};
          const bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
          for (int d = 0; d < 3; ++d) {
            const double r = g();  // three draws per node, in node-id order
            double x = ax[ijk[d]];
            if (interior) x += jitter * hl[ijk[d]] * r;
            coords[3 * idx + d] = x;
          }
        }
    for (int64_t k = 0; k < n; ++k)
      for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
          const int64_t e = i + (int64_t)n * (j + (int64_t)n * k);
          auto nid = [&](int64_t a, int64_t b, int64_t c) { return (int32_t)(a + n1 * (b + n1 * c)); };
          int32_t* c8 = conn + 8 * e;
          c8[0] = nid(i, j, k);
          c8[1] = nid(i + 1, j, k);
          c8[2] = nid(i + 1, j + 1, k);
          c8[3] = nid(i, j + 1, k);
          c8[4] = nid(i, j, k + 1);
          c8[5] = nid(i + 1, j, k + 1);
          c8[6] = nid(i + 1, j + 1, k + 1);
          c8[7] = nid(i, j + 1, k + 1);
        }
    return NALU_OK;
  });
}

extern "C" int nalu_normal_field(int64_t n, double sigma, uint32_t seed, double* out) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(out && n >= 0, NALU_ERR_ARG, "normal_field: bad args");
    Lcg g(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = sigma * lcg_normal(g);
    return NALU_OK;
  });
}

extern "C" int nalu_lognormal_field(int64_t n, double sigma, uint32_t seed, double* out) {
  return nalu::guarded([&]() -> int {
    NALU_REQUIRE(out && n >= 0, NALU_ERR_ARG, "lognormal_field: bad args");
    Lcg g(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = std::exp(sigma * lcg_normal(g));
    return NALU_OK;
  });
}
