/* Runs every entry of the CPU oracle (oracle/oracle.c) once on a small jittered
 * Hex8 lattice, for the AddressSanitizer / UndefinedBehaviorSanitizer build of
 * tests/sanitize/Makefile.  Each output is checked for finiteness and a few
 * invariants (graph symmetry, SpMV of the solved system), so the run also fails
 * on a NaN that a sanitizer would not report.  Exit status 0 means clean. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/oracle.h"

static int g_fail = 0;
#define CHECK(c, msg)                                        \
  do {                                                       \
    if (!(c)) {                                              \
      fprintf(stderr, "oracle_driver: %s (%s:%d)\n", msg, __FILE__, __LINE__); \
      g_fail = 1;                                            \
    }                                                        \
  } while (0)

static uint32_t g_lcg = 12345u;
static double urand(void) { /* [-1, 1) */
  g_lcg = (uint32_t)((1103515245ull * g_lcg + 12345ull) % 2147483647ull);
  return 2.0 * (double)g_lcg / 2147483647.0 - 1.0;
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "oracle_driver: out of memory\n");
    exit(2);
  }
  return p;
}

static int finite_all(const double *a, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(a[i])) return 0;
  return 1;
}

/* n^3 lattice on [0,1]^3, interior nodes jittered by 0.2 h */
static void lattice(int n, double **X, int32_t **conn, int64_t *nn, int64_t *ne) {
  const int n1 = n + 1;
  *nn = (int64_t)n1 * n1 * n1;
  *ne = (int64_t)n * n * n;
  *X = xcalloc((size_t)(3 * *nn), sizeof(double));
  *conn = xcalloc((size_t)(8 * *ne), sizeof(int32_t));
  const double h = 1.0 / n;
  for (int k = 0; k < n1; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i) {
        const int64_t id = i + (int64_t)n1 * (j + (int64_t)n1 * k);
        const int in = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) (*X)[3 * id + d] = ijk[d] * h + (in ? 0.2 * h * urand() : 0.0);
      }
  int64_t e = 0;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i, ++e) {
        const int32_t b = (int32_t)(i + n1 * (j + n1 * k));
        const int32_t v[8] = {b, b + 1, b + 1 + n1, b + n1, b + n1 * n1, b + 1 + n1 * n1, b + 1 + n1 + n1 * n1,
                              b + n1 + n1 * n1};
        memcpy(*conn + 8 * e, v, sizeof v);
      }
}

typedef struct {
  int nd;
  int64_t nrows, nnz;
  int64_t *rp;
  int32_t *ci, *lid;
} Graph;

static Graph graph(int64_t nn, const int64_t *ids, int64_t ne, const int32_t *conn, int nd) {
  Graph g;
  g.nd = nd;
  g.nrows = nn * nd;
  g.nnz = or_graph_build(nn, ids, ne, 8, conn, nd, NULL, NULL, NULL);
  g.rp = xcalloc((size_t)(g.nrows + 1), sizeof(int64_t));
  g.ci = xcalloc((size_t)g.nnz, sizeof(int32_t));
  g.lid = xcalloc((size_t)nn, sizeof(int32_t));
  CHECK(or_graph_build(nn, ids, ne, 8, conn, nd, g.rp, g.ci, g.lid) == g.nnz, "graph nnz differs between passes");
  CHECK(g.rp[g.nrows] == g.nnz, "row_ptr end");
  /* structural symmetry, sorted columns */
  for (int64_t r = 0; r < g.nrows; ++r) {
    for (int64_t q = g.rp[r]; q < g.rp[r + 1]; ++q) {
      if (q > g.rp[r]) CHECK(g.ci[q] > g.ci[q - 1], "columns not strictly increasing");
      const int32_t c = g.ci[q];
      int found = 0;
      for (int64_t t = g.rp[c]; t < g.rp[c + 1]; ++t) found |= g.ci[t] == (int32_t)r;
      CHECK(found, "graph not structurally symmetric");
    }
  }
  return g;
}

static void graph_free(Graph *g) {
  free(g->rp);
  free(g->ci);
  free(g->lid);
}

int main(void) {
  /* master-element tables and per-element kernels */
  double coords[24], q8[8], k8[8], rho8[8], p8[8], u24[24], g24[24], mu8[8];
  const double unit[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int a = 0; a < 8; ++a)
    for (int d = 0; d < 3; ++d) coords[3 * a + d] = unit[a][d] + 0.05 * urand();
  for (int a = 0; a < 8; ++a) {
    q8[a] = urand();
    k8[a] = 1.0 + 0.5 * urand();
    rho8[a] = 1.0 + 0.1 * urand();
    p8[a] = urand();
    mu8[a] = 0.1 + 0.05 * urand();
    for (int d = 0; d < 3; ++d) {
      u24[3 * a + d] = urand();
      g24[3 * a + d] = urand();
    }
  }
  for (int s = 0; s < 2; ++s) {
    CHECK(or_hexscs_intgloc(s) && or_hexscv_intgloc(s), "tables");
    double shape[12 * 8], deriv[12 * 24], dndx[12 * 24], dder[12 * 24], vols[8];
    or_hex8_shape_fcn(12, or_hexscs_intgloc(s), shape);
    or_hex8_derivative(12, or_hexscs_intgloc(s), deriv);
    or_generic_grad_op(12, deriv, coords, dndx);
    or_hexscs_grad_op(coords, s, dndx, dder);
    or_hexscv_grad_op(coords, s, dndx, dder);
    or_hexscv_volume(coords, vols);
    CHECK(finite_all(shape, 96) && finite_all(dndx, 8 * 24) && finite_all(vols, 8), "master element outputs");
  }
  CHECK(or_hexscs_lrscv() != NULL, "lrscv");
  {
    double coordv[27 * 3], areav[36], qa[4 * 3], qv[12];
    or_subdivide_hex8(coords, coordv);
    or_hex_scs_areav(coords, areav);
    memcpy(qa, coords, sizeof qa);
    or_quad_area_by_triangulation(qa, qv);
    or_quad3d_scs_areav(qa, qv);
    CHECK(finite_all(coordv, 81) && finite_all(areav, 36) && finite_all(qv, 12), "geometry outputs");
    CHECK(or_hex_volume_grandy(coordv) != 0.0, "grandy volume");
  }
  {
    double lhs[24 * 24], rhs[24], md[12];
#define ZERO() (memset(lhs, 0, sizeof lhs), memset(rhs, 0, sizeof rhs))
    for (int s = 0; s < 2; ++s) {
      ZERO();
      or_scalar_diff_elem(coords, q8, k8, s, lhs, rhs);
      CHECK(finite_all(lhs, 64) && finite_all(rhs, 8), "scalar diff");
      ZERO();
      or_continuity_adv_elem(coords, rho8, p8, u24, g24, 0.5, s, s, 1 - s, lhs, rhs);
      CHECK(finite_all(lhs, 64) && finite_all(rhs, 8), "continuity");
      or_mdot_elem(coords, rho8, p8, u24, g24, 0.5, s, s, md);
      or_mdot_rhou_elem(coords, rho8, u24, md);
      CHECK(finite_all(md, 12), "mdot");
      ZERO();
      or_momentum_advdiff_elem(coords, u24, mu8, md, 1.0, 1 - s, s, lhs, rhs);
      CHECK(finite_all(lhs, 576) && finite_all(rhs, 24), "momentum");
    }
    const double gam[3] = {1.5, -2.0, 0.5};
    const double r3[24] = {1, 1, 1, 1, 1, 1, 1, 1, 1.1, 1.1, 1.1, 1.1, 1.1, 1.1, 1.1, 1.1,
                           1.2, 1.2, 1.2, 1.2, 1.2, 1.2, 1.2, 1.2};
    double u3[72];
    for (int i = 0; i < 72; ++i) u3[i] = urand();
    for (int l = 0; l < 2; ++l) {
      ZERO();
      or_momentum_mass_elem(coords, r3, u3, g24, 0.1, gam, l, 1 - l, lhs, rhs);
      CHECK(finite_all(lhs, 576) && finite_all(rhs, 24), "momentum mass");
      ZERO();
      or_continuity_mass_elem(coords, r3, 0.1, gam, l, rhs);
      CHECK(finite_all(rhs, 8), "continuity mass");
    }
  }

  /* whole-mesh assembly, graph with permuted ids, boundary closure, solvers */
  double *X;
  int32_t *conn;
  int64_t nn, ne;
  lattice(4, &X, &conn, &nn, &ne);
  int64_t *ids = xcalloc((size_t)nn, sizeof(int64_t)), *ids0 = xcalloc((size_t)nn, sizeof(int64_t));
  for (int64_t i = 0; i < nn; ++i) {
    ids0[i] = 1 + i;
    ids[i] = 1 + ((i * 37) % nn); /* a permutation: gcd(37, 125) = 1 */
  }
  Graph g1 = graph(nn, ids0, ne, conn, 1);
  Graph g1p = graph(nn, ids, ne, conn, 1);
  Graph g3 = graph(nn, ids0, ne, conn, 3);
  CHECK(g1p.nnz == g1.nnz, "permuted ids change nnz");
  double *rho = xcalloc((size_t)nn, 8), *p = xcalloc((size_t)nn, 8), *mu = xcalloc((size_t)nn, 8);
  double *u = xcalloc((size_t)(3 * nn), 8), *gp = xcalloc((size_t)(3 * nn), 8), *q = xcalloc((size_t)nn, 8);
  for (int64_t i = 0; i < nn; ++i) {
    rho[i] = 1.0 + 0.1 * urand();
    p[i] = urand();
    q[i] = urand();
    mu[i] = 0.1;
    for (int d = 0; d < 3; ++d) {
      u[3 * i + d] = urand();
      gp[3 * i + d] = urand();
    }
  }
  double *v1 = xcalloc((size_t)g1.nnz, 8), *r1 = xcalloc((size_t)g1.nrows, 8);
  double *v3 = xcalloc((size_t)g3.nnz, 8), *r3 = xcalloc((size_t)g3.nrows, 8);
  double *mdot = xcalloc((size_t)(12 * ne), 8);
  or_assemble_scalar_diff(ne, conn, X, q, rho, 1, g1p.rp, g1p.ci, g1p.lid, v1, r1);
  or_compute_mdot(ne, conn, X, rho, p, u, gp, 1.0, 0, 0, mdot);
  CHECK(finite_all(mdot, 12 * ne), "mesh mdot");
  memset(v1, 0, (size_t)g1.nnz * 8);
  memset(r1, 0, (size_t)g1.nrows * 8);
  or_assemble_continuity(ne, conn, X, rho, p, u, gp, 1.0, 0, 0, 0, g1.rp, g1.ci, g1.lid, v1, r1);
  or_assemble_momentum(ne, conn, X, u, mu, mdot, 1.0, 1, 0, g3.rp, g3.ci, g3.lid, v3, r3);
  const double gam[3] = {1.5, -2.0, 0.5};
  or_assemble_momentum_mass(ne, conn, X, rho, rho, rho, u, u, u, gp, 0.1, gam, 0, 1, g3.rp, g3.ci, g3.lid, v3, r3);
  CHECK(finite_all(v1, g1.nnz) && finite_all(v3, g3.nnz) && finite_all(r3, g3.nrows), "mesh assembly");
  double *vol = xcalloc((size_t)nn, 8), *dq = xcalloc((size_t)(3 * nn), 8);
  or_dual_nodal_volume(ne, conn, X, nn, vol);
  or_nodal_grad_elem(ne, conn, X, q, vol, 0, nn, dq);
  const int64_t nf = or_exposed_faces(ne, conn, nn, NULL);
  CHECK(nf == 6 * 16, "exposed faces of a 4^3 block");
  int32_t *faces = xcalloc((size_t)(4 * nf), sizeof(int32_t));
  CHECK(or_exposed_faces(ne, conn, nn, faces) == nf, "exposed face count differs between passes");
  or_nodal_grad_boundary(nf, faces, X, q, vol, 1, dq);
  CHECK(finite_all(vol, nn) && finite_all(dq, 3 * nn), "nodal gradient");
  /* pure diffusion system with a Dirichlet face: SPD, solvable by every method */
  double *vd = xcalloc((size_t)g1.nnz, 8), *rd = xcalloc((size_t)g1.nrows, 8);
  or_assemble_scalar_diff(ne, conn, X, q, mu, 0, g1.rp, g1.ci, g1.lid, vd, rd);
  int32_t bc[25];
  double *sol = xcalloc((size_t)nn, 8), *bcv = xcalloc((size_t)nn, 8);
  for (int i = 0; i < 25; ++i) bc[i] = i;
  or_apply_dirichlet(1, g1.nrows, g1.rp, g1.ci, vd, rd, g1.lid, 25, bc, sol, bcv, 0, 1);
  int32_t *color = xcalloc((size_t)g1.nrows, sizeof(int32_t));
  const int nc = or_greedy_color(g1.nrows, g1.rp, g1.ci, color);
  CHECK(nc >= 2, "coloring");
  int64_t *cp = xcalloc((size_t)(nc + 1), sizeof(int64_t));
  int32_t *cr = xcalloc((size_t)g1.nrows, sizeof(int32_t));
  for (int64_t r = 0; r < g1.nrows; ++r) cp[color[r] + 1]++;
  for (int c = 0; c < nc; ++c) cp[c + 1] += cp[c];
  {
    int64_t *fill = xcalloc((size_t)nc, sizeof(int64_t));
    for (int64_t r = 0; r < g1.nrows; ++r) cr[cp[color[r]] + fill[color[r]]++] = (int32_t)r;
    free(fill);
  }
  double *z = xcalloc((size_t)g1.nrows, 8), *x = xcalloc((size_t)g1.nrows, 8), *y = xcalloc((size_t)g1.nrows, 8);
  or_sgs_apply(g1.nrows, g1.rp, g1.ci, vd, rd, z);
  or_mcsgs_apply(g1.nrows, g1.rp, g1.ci, vd, nc, cp, cr, rd, z);
  or_jacobi_apply(g1.nrows, g1.rp, g1.ci, vd, rd, z);
  CHECK(finite_all(z, g1.nrows), "preconditioners");
  double bn = 0.0;
  for (int64_t r = 0; r < g1.nrows; ++r) bn += rd[r] * rd[r];
  bn = sqrt(bn);
  for (int pc = 0; pc < 4; ++pc) {
    for (int m = 0; m < 2; ++m) {
      double res = 0.0;
      memset(x, 0, (size_t)g1.nrows * 8);
      const int it = m == 0 ? or_gmres(g1.nrows, g1.rp, g1.ci, vd, rd, x, 20, 2000, 1e-10, pc, nc, cp, cr, &res)
                            : or_cg(g1.nrows, g1.rp, g1.ci, vd, rd, x, 2000, 1e-10, pc, nc, cp, cr, &res);
      or_spmv(g1.nrows, g1.rp, g1.ci, vd, x, y);
      double e = 0.0;
      for (int64_t r = 0; r < g1.nrows; ++r) e += (rd[r] - y[r]) * (rd[r] - y[r]);
      CHECK(it > 0 && sqrt(e) <= 1e-8 * bn, "solver did not reach the tolerance");
    }
  }
  (void)or_wtime();

  graph_free(&g1);
  graph_free(&g1p);
  graph_free(&g3);
  free(X), free(conn), free(ids), free(ids0), free(rho), free(p), free(mu), free(u), free(gp), free(q);
  free(v1), free(r1), free(v3), free(r3), free(mdot), free(vol), free(dq), free(faces), free(vd), free(rd);
  free(sol), free(bcv), free(color), free(cp), free(cr), free(z), free(x), free(y);
  if (!g_fail) printf("oracle_driver: ok\n");
  return g_fail;
}
