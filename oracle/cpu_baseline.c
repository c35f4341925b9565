/*
 * Threaded CPU baseline of the low-Mach assembly + pressure solve.
 *
 * TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it beside the
 * GPU path; the product library never links, loads or calls it.
 *
 * It is the CPU oracle's arithmetic (oracle.c: the element kernels of the
 * ElemKernel.C files under src/kernel and ComputeMdotElemAlgorithm.C) compiled for speed
 * (-O3 -march=x86-64-v3 -fopenmp, FP contraction allowed) and driven the way the
 * reference runs on a multicore host:
 *   - graph: the Tpetra-identical node graph, built per row in parallel from
 *     the node -> element incidence (rows sorted, columns sorted), numDof
 *     rows interleaved (src/TpetraLinearSystem.C:1227-1246); set-up, untimed
 *   - assembly: elements in parallel (Kokkos team loop over buckets,
 *     include/AssembleElemSolverAlgorithm.h:84-103), each element's matrix
 *     added by the sumInto row walk with atomic additions (the reference's
 *     non-Serial path, src/TpetraLinearSystem.C:1376-1453)
 *   - solve: CG (src/LinearSolver.C, Belos) with the SGS relaxation applied
 *     per thread block of rows, couplings to other blocks dropped -- what
 *     Ifpack2's per-rank SGS computes when the rows are split over as many
 *     MPI ranks as threads (src/LinearSolverConfig.C:71-85); SpMV, dots and
 *     updates threaded.
 * Timed phases follow the reference's per-equation timers
 * (src/EquationSystem.C:270-292): assemble (mdot + continuity + momentum) and
 * solve (preconditioner set-up + Krylov).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double now(void) { return omp_get_wtime(); }

int cb_max_threads(void) { return omp_get_max_threads(); }
void cb_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

/* Node graph (numDof 1) from the element connectivity: row a holds every node
 * sharing an element with a (itself included), ascending.  Rows are node
 * indices (meshes without explicit ids: lid = index).  Returns nnz. */
int64_t cb_graph(int64_t nn, int64_t ne, const int32_t *conn, int64_t *rp, int32_t *ci) {
  int64_t *ip = (int64_t *)calloc((size_t)nn + 1, sizeof(int64_t));
  int32_t *ie = (int32_t *)malloc(sizeof(int32_t) * (size_t)(8 * ne));
  for (int64_t q = 0; q < 8 * ne; ++q) ip[conn[q] + 1]++;
  for (int64_t a = 0; a < nn; ++a) ip[a + 1] += ip[a];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)nn);
  memcpy(fill, ip, sizeof(int64_t) * (size_t)nn);
  for (int64_t q = 0; q < 8 * ne; ++q) ie[fill[conn[q]]++] = (int32_t)(q / 8);
  free(fill);
  /* row lengths, then rows */
  for (int pass = 0; pass < 2; ++pass) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t a = 0; a < nn; ++a) {
      int32_t cand[512];
      int nc = 0;
      for (int64_t q = ip[a]; q < ip[a + 1] && nc + 8 <= 512; ++q)
        for (int k = 0; k < 8; ++k) cand[nc++] = conn[8 * (int64_t)ie[q] + k];
      for (int i = 1; i < nc; ++i) { /* insertion sort + unique */
        const int32_t v = cand[i];
        int j = i - 1;
        while (j >= 0 && cand[j] > v) {
          cand[j + 1] = cand[j];
          --j;
        }
        cand[j + 1] = v;
      }
      int u = 0;
      for (int i = 0; i < nc; ++i)
        if (u == 0 || cand[u - 1] != cand[i]) cand[u++] = cand[i];
      if (pass == 0) rp[a + 1] = u;
      else memcpy(ci + rp[a], cand, sizeof(int32_t) * (size_t)u);
    }
    if (pass == 0) {
      rp[0] = 0;
      for (int64_t a = 0; a < nn; ++a) rp[a + 1] += rp[a];
    }
  }
  free(ip);
  free(ie);
  return rp[nn];
}

/* numDof expansion of a node graph (fill_in_extra_dof_rows_per_node). */
void cb_expand(int64_t nn, const int64_t *rp1, const int32_t *ci1, int nd, int64_t *rp, int32_t *ci) {
  rp[0] = 0;
  for (int64_t a = 0; a < nn; ++a) {
    const int64_t L = rp1[a + 1] - rp1[a];
    for (int d = 0; d < nd; ++d) rp[a * nd + d + 1] = nd * nd * rp1[a] + (d + 1) * nd * L;
  }
#pragma omp parallel for schedule(static)
  for (int64_t a = 0; a < nn; ++a) {
    const int64_t L = rp1[a + 1] - rp1[a];
    for (int d = 0; d < nd; ++d) {
      int32_t *row = ci + rp[a * nd + d];
      for (int64_t p = 0; p < L; ++p)
        for (int e = 0; e < nd; ++e) row[p * nd + e] = nd * ci1[rp1[a] + p] + e;
    }
  }
}

/* sumInto row walk with atomic additions (sum_into_row, src/TpetraLinearSystem.C:1415-1453). */
static void sum_into_atomic(int nd, const int64_t *rp, const int32_t *ci, double *val, double *rhs,
                            const int32_t *nodes, const double *lhs, const double *r) {
  const int R = 8 * nd;
  int32_t ids[24], perm[24];
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < nd; ++d) {
      ids[i * nd + d] = nd * nodes[i] + d;
      perm[i * nd + d] = i * nd + d;
    }
  for (int i = 1; i < R; ++i) {
    const int32_t k = ids[i], v = perm[i];
    int j = i - 1;
    while (j >= 0 && ids[j] > k) {
      ids[j + 1] = ids[j];
      perm[j + 1] = perm[j];
      --j;
    }
    ids[j + 1] = k;
    perm[j + 1] = v;
  }
  for (int rr = 0; rr < R; ++rr) {
    const int64_t row = ids[rr];
    const double *cur = lhs + (size_t)perm[rr] * R;
    const int64_t b = rp[row], len = rp[row + 1] - b;
    int64_t off = 0;
    for (int j = 0; j < R; ++j) {
      while (off < len && ci[b + off] != ids[j]) ++off;
      if (off < len) {
#pragma omp atomic
        val[b + off] += cur[perm[j]];
      }
    }
#pragma omp atomic
    rhs[row] += r[perm[rr]];
  }
}

static void gather8(const int32_t *nd, const double *f, int nc, double *out) {
  for (int n = 0; n < 8; ++n)
    for (int c = 0; c < nc; ++c) out[nc * n + c] = f[(size_t)nc * nd[n] + c];
}

/* One timed assembly: mdot, continuity (values1 / rhs1), momentum (values3 /
 * rhs3), each zeroed first (zeroSystem).  Elements [e0, e1) only (a bounded
 * sample); returns seconds. */
double cb_assemble(int64_t e0, int64_t e1, const int32_t *conn, const double *x, const double *rho,
                   const double *p, const double *u, const double *gp, const double *mu, double tau,
                   double *mdot, int64_t nrows1, const int64_t *rp1, const int32_t *ci1, double *v1, double *r1,
                   const int64_t *rp3, const int32_t *ci3, double *v3, double *r3) {
  (void)or_hexscs_intgloc(1); /* the oracle's lazily built tables, before the threads start */
  const double t0 = now();
#pragma omp parallel
  {
#pragma omp for schedule(static) nowait
    for (int64_t q = 0; q < rp1[nrows1]; ++q) v1[q] = 0.0;
#pragma omp for schedule(static) nowait
    for (int64_t q = 0; q < rp3[3 * nrows1]; ++q) v3[q] = 0.0;
#pragma omp for schedule(static) nowait
    for (int64_t i = 0; i < nrows1; ++i) r1[i] = 0.0;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < 3 * nrows1; ++i) r3[i] = 0.0;
#pragma omp for schedule(static)
    for (int64_t e = e0; e < e1; ++e) {
      const int32_t *nd = conn + 8 * e;
      double xe[24], re[8], pe[8], ue[24], ge[24];
      gather8(nd, x, 3, xe);
      gather8(nd, rho, 1, re);
      gather8(nd, p, 1, pe);
      gather8(nd, u, 3, ue);
      gather8(nd, gp, 3, ge);
      or_mdot_elem(xe, re, pe, ue, ge, tau, 0, 0, mdot + 12 * e);
    }
#pragma omp for schedule(static)
    for (int64_t e = e0; e < e1; ++e) {
      const int32_t *nd = conn + 8 * e;
      double xe[24], re[8], pe[8], ue[24], ge[24], lhs[64], r[8];
      gather8(nd, x, 3, xe);
      gather8(nd, rho, 1, re);
      gather8(nd, p, 1, pe);
      gather8(nd, u, 3, ue);
      gather8(nd, gp, 3, ge);
      memset(lhs, 0, sizeof(lhs));
      memset(r, 0, sizeof(r));
      or_continuity_adv_elem(xe, re, pe, ue, ge, tau, 0, 0, 0, lhs, r);
      sum_into_atomic(1, rp1, ci1, v1, r1, nd, lhs, r);
    }
#pragma omp for schedule(static)
    for (int64_t e = e0; e < e1; ++e) {
      const int32_t *nd = conn + 8 * e;
      double xe[24], ue[24], me[8], lhs[576], r[24];
      gather8(nd, x, 3, xe);
      gather8(nd, u, 3, ue);
      gather8(nd, mu, 1, me);
      memset(lhs, 0, sizeof(lhs));
      memset(r, 0, sizeof(r));
      or_momentum_advdiff_elem(xe, ue, me, mdot + 12 * e, 0.0, 0, 0, lhs, r);
      sum_into_atomic(3, rp3, ci3, v3, r3, nd, lhs, r);
    }
  }
  return now() - t0;
}

/* ---------------------------------------------------------------- solve */
static double pdot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static void pspmv(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t q = rp[i]; q < rp[i + 1]; ++q) s += v[q] * x[ci[q]];
    y[i] = s;
  }
}

/* SGS per block of rows (one block per thread), couplings outside the block dropped. */
static void block_sgs(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *r, double *z) {
#pragma omp parallel
  {
    const int nt = omp_get_num_threads(), t = omp_get_thread_num();
    const int64_t a = n * t / nt, b = n * (t + 1) / nt;
    for (int64_t i = a; i < b; ++i) z[i] = 0.0;
    for (int sweep = 0; sweep < 2; ++sweep)
      for (int64_t k = 0; k < b - a; ++k) {
        const int64_t i = sweep == 0 ? a + k : b - 1 - k;
        double s = 0.0, dg = 0.0;
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
          const int32_t j = ci[q];
          if (j == i) dg = v[q];
          else if (j >= a && j < b) s += v[q] * z[j];
        }
        z[i] = (r[i] - s) / dg;
      }
  }
}

/* Preconditioned CG to ||r|| <= tol ||r0|| from x = 0; returns iterations, the
 * true relative residual in *res and the seconds in *secs. */
int cb_cg_sgs(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *b, double *x,
              int maxit, double tol, double *res, double *secs) {
  double *r = (double *)malloc(sizeof(double) * (size_t)n), *z = (double *)malloc(sizeof(double) * (size_t)n);
  double *pp = (double *)malloc(sizeof(double) * (size_t)n), *q = (double *)malloc(sizeof(double) * (size_t)n);
  const double t0 = now();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    x[i] = 0.0;
    r[i] = b[i];
  }
  const double r0 = sqrt(pdot(n, r, r));
  block_sgs(n, rp, ci, v, r, z);
  memcpy(pp, z, sizeof(double) * (size_t)n);
  double rz = pdot(n, r, z), rn = r0;
  int it = 0;
  while (it < maxit && r0 > 0.0 && rn > tol * r0) {
    ++it;
    pspmv(n, rp, ci, v, pp, q);
    const double alpha = rz / pdot(n, pp, q);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * pp[i];
      r[i] -= alpha * q[i];
    }
    rn = sqrt(pdot(n, r, r));
    if (rn <= tol * r0) break;
    block_sgs(n, rp, ci, v, r, z);
    const double rz1 = pdot(n, r, z);
    const double beta = rz1 / rz;
    rz = rz1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) pp[i] = z[i] + beta * pp[i];
  }
  *secs = now() - t0;
  pspmv(n, rp, ci, v, x, q);
  double s = 0.0, bb = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    s += (b[i] - q[i]) * (b[i] - q[i]);
    bb += b[i] * b[i];
  }
  *res = sqrt(s) / (bb > 0 ? sqrt(bb) : 1.0);
  free(r);
  free(z);
  free(pp);
  free(q);
  return it;
}
