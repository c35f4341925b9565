/*
 * CPU oracle: FP64 restatement of the reference's Hex8 CVFEM assembly and
 * linear-solve path.  TEST INFRASTRUCTURE ONLY -- see oracle.h for the list
 * of checks that pin each function.  Compile with -ffp-contract=off.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Master-element tables                                                     */
/* ------------------------------------------------------------------------ */

/* HexSCS::intgLoc_ / intgLocShift_ (src/master_element/Hex8CVFEM.C:318-345):
 * ip k sits on the sub-control surface between its two lrscv nodes; the
 * coordinate along the edge direction is 0, the other two are +-1/4 (+-1/2
 * when shifted to the element edges). */
static const double kScsLoc[12][3] = {
    {0.00, -0.25, -0.25}, {0.25, 0.00, -0.25}, {0.00, 0.25, -0.25}, {-0.25, 0.00, -0.25},
    {0.00, -0.25, 0.25},  {0.25, 0.00, 0.25},  {0.00, 0.25, 0.25},  {-0.25, 0.00, 0.25},
    {-0.25, -0.25, 0.00}, {0.25, -0.25, 0.00}, {0.25, 0.25, 0.00},  {-0.25, 0.25, 0.00}};
static double gScsLocShift[12][3];
/* HexSCV::intgLoc_ (Hex8CVFEM.C:88-96): sub-volume centroids at +-1/4. */
static const double kScvLoc[8][3] = {
    {-0.25, -0.25, -0.25}, {0.25, -0.25, -0.25}, {0.25, 0.25, -0.25}, {-0.25, 0.25, -0.25},
    {-0.25, -0.25, 0.25},  {0.25, -0.25, 0.25},  {0.25, 0.25, 0.25},  {-0.25, 0.25, 0.25}};
static double gScvLocShift[8][3];
/* HexSCS::lrscv_ (Hex8CVFEM.C:264-276): left/right node of each scs. */
static const int kLrscv[24] = {0, 1, 1, 2, 2, 3, 0, 3, 4, 5, 5, 6, 6, 7, 4, 7, 0, 4, 1, 5, 2, 6, 3, 7};
/* Hex8 node corner signs in the reference element [-1/2,1/2]^3 (node order
 * of Hex8CVFEM.C nodeLoc_, :420-436). */
static const double kNodeSign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                       {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

static int gTablesReady = 0;
static void init_tables(void) {
  if (gTablesReady) return;
  /* shifted locations double the non-zero coordinates (Hex8CVFEM.C:333-345, :99-107) */
  for (int i = 0; i < 12; ++i)
    for (int d = 0; d < 3; ++d) gScsLocShift[i][d] = 2.0 * kScsLoc[i][d];
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < 3; ++d) gScvLocShift[i][d] = 2.0 * kScvLoc[i][d];
  gTablesReady = 1;
}

const double *or_hexscs_intgloc(int shifted) {
  init_tables();
  return shifted ? &gScsLocShift[0][0] : &kScsLoc[0][0];
}
const double *or_hexscv_intgloc(int shifted) {
  init_tables();
  return shifted ? &gScvLocShift[0][0] : &kScvLoc[0][0];
}
const int *or_hexscs_lrscv(void) { return kLrscv; }

/* HexSCS::hex8_shape_fcn (Hex8CVFEM.C:487-518).  With (x,y,z) the node's
 * corner signs, N = 1/8 + 1/4 (x s1 + y s2 + z s3)
 *                 + 1/2 (yz s2 s3 + zx s3 s1 + xy s1 s2) + xyz s1 s2 s3,
 * evaluated in the reference's left-to-right order (sign flips are exact). */
void or_hex8_shape_fcn(int npts, const double *iso, double *shape) {
  const double half = 0.5, one4th = 0.25, one8th = 0.125;
  for (int j = 0; j < npts; ++j) {
    const double s1 = iso[3 * j], s2 = iso[3 * j + 1], s3 = iso[3 * j + 2];
    for (int n = 0; n < 8; ++n) {
      const double x = kNodeSign[n][0], y = kNodeSign[n][1], z = kNodeSign[n][2];
      const double lin = (x * s1 + y * s2) + z * s3;
      const double bil = ((y * z) * (s2 * s3) + (z * x) * (s3 * s1)) + (x * y) * (s1 * s2);
      const double tri = (x * y * z) * ((s1 * s2) * s3);
      shape[8 * j + n] = ((one8th + one4th * lin) + half * bil) + tri;
    }
  }
}

/* hex8_derivative (Hex8CVFEM.C:24-70).  dN/ds1 = x (1/2 + y s2)(1/2 + z s3)
 * expanded as half*(xz s3 + xy s2) + xyz s2s3 + x/4 (and cyclic), which is
 * term-for-term the reference expression. */
void or_hex8_derivative(int npts, const double *loc, double *deriv) {
  const double half = 0.5, one4th = 0.25;
  for (int ip = 0; ip < npts; ++ip) {
    const double s1 = loc[3 * ip], s2 = loc[3 * ip + 1], s3 = loc[3 * ip + 2];
    const double s1s2 = s1 * s2, s2s3 = s2 * s3, s1s3 = s1 * s3;
    for (int n = 0; n < 8; ++n) {
      const double x = kNodeSign[n][0], y = kNodeSign[n][1], z = kNodeSign[n][2];
      const double xyz = x * y * z;
      double *d = deriv + 24 * ip + 3 * n;
      d[0] = (half * ((x * z) * s3 + (x * y) * s2) + xyz * s2s3) + x * one4th;
      d[1] = (half * ((y * z) * s3 + (y * x) * s1) + xyz * s1s3) + y * one4th;
      d[2] = (half * ((z * y) * s2 + (z * x) * s1) + xyz * s1s2) + z * one4th;
    }
  }
}

/* subdivide_hex_8 (include/master_element/Hex8GeometryFunctions.h:263-321):
 * 8 corners, 12 edge midpoints, 6 face centroids, 1 volume centroid. */
static const int kEdgeMid[12][3] = {/* {point, a, b} */
                                    {8, 0, 1},   {9, 1, 2},   {10, 2, 3},  {11, 3, 0},
                                    {13, 4, 5},  {14, 5, 6},  {15, 6, 7},  {16, 7, 4},
                                    {18, 1, 5},  {19, 0, 4},  {21, 3, 7},  {22, 2, 6}};
static const int kFaceMid[6][5] = {/* {point, a, b, c, d}, summed in this order */
                                   {12, 0, 1, 2, 3}, {17, 4, 5, 6, 7}, {20, 0, 1, 4, 5},
                                   {23, 2, 3, 6, 7}, {24, 1, 2, 5, 6}, {25, 0, 3, 4, 7}};
void or_subdivide_hex8(const double *c, double *cv) {
  for (int n = 0; n < 8; ++n)
    for (int d = 0; d < 3; ++d) cv[3 * n + d] = c[3 * n + d];
  for (int d = 0; d < 3; ++d) {
    for (int e = 0; e < 12; ++e)
      cv[3 * kEdgeMid[e][0] + d] = 0.5 * (c[3 * kEdgeMid[e][1] + d] + c[3 * kEdgeMid[e][2] + d]);
    for (int f = 0; f < 6; ++f) {
      const int *t = kFaceMid[f];
      cv[3 * t[0] + d] =
          0.25 * (((c[3 * t[1] + d] + c[3 * t[2] + d]) + c[3 * t[3] + d]) + c[3 * t[4] + d]);
    }
    double s = 0.0;
    for (int n = 0; n < 8; ++n) s += c[3 * n + d];
    cv[3 * 26 + d] = s * 0.125;
  }
}

/* quad_area_by_triangulation (Hex8GeometryFunctions.h:30-76): four triangles
 * fanned around the quad centroid, half the summed cross products. */
void or_quad_area_by_triangulation(const double *p, double *area) {
  double xm[3];
  for (int d = 0; d < 3; ++d) xm[d] = 0.25 * (((p[d] + p[3 + d]) + p[6 + d]) + p[9 + d]);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  double r1[3] = {p[0] - xm[0], p[1] - xm[1], p[2] - xm[2]};
  for (int t = 0; t < 4; ++t) {
    const int q = (t + 1) % 4;
    const double r2[3] = {p[3 * q] - xm[0], p[3 * q + 1] - xm[1], p[3 * q + 2] - xm[2]};
    a0 += r1[1] * r2[2] - r2[1] * r1[2];
    a1 += r1[2] * r2[0] - r2[2] * r1[0];
    a2 += r1[0] * r2[1] - r2[0] * r1[1];
    r1[0] = r2[0];
    r1[1] = r2[1];
    r1[2] = r2[2];
  }
  area[0] = a0 * 0.5;
  area[1] = a1 * 0.5;
  area[2] = a2 * 0.5;
}

/* HexSCS::determinant (Hex8CVFEM.C:547-581, facet table :551-564). */
static const int kScsFacet[12][4] = {{20, 8, 12, 26},  {24, 9, 12, 26},  {10, 12, 26, 23},
                                     {11, 25, 26, 12}, {13, 20, 26, 17}, {17, 14, 24, 26},
                                     {17, 15, 23, 26}, {16, 17, 26, 25}, {19, 20, 26, 25},
                                     {20, 18, 24, 26}, {22, 23, 26, 24}, {21, 25, 26, 23}};
void or_hex_scs_areav(const double *coords, double *areav) {
  double cv[27 * 3], q[12];
  or_subdivide_hex8(coords, cv);
  for (int ics = 0; ics < 12; ++ics) {
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) q[3 * k + d] = cv[3 * kScsFacet[ics][k] + d];
    or_quad_area_by_triangulation(q, areav + 3 * ics);
  }
}

/* generic_grad_op + cofactorMatrix (include/master_element/MasterElementFunctions.h:31-108). */
void or_generic_grad_op(int npts, const double *deriv, const double *coords, double *dndx) {
  for (int ip = 0; ip < npts; ++ip) {
    const double *rg = deriv + 24 * ip;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int n = 0; n < 8; ++n)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) J[i][j] += rg[3 * n + j] * coords[3 * n + i];
    double A[3][3];
    A[0][0] = J[1][1] * J[2][2] - J[2][1] * J[1][2];
    A[0][1] = J[1][2] * J[2][0] - J[2][2] * J[1][0];
    A[0][2] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
    A[1][0] = J[0][2] * J[2][1] - J[2][2] * J[0][1];
    A[1][1] = J[0][0] * J[2][2] - J[2][0] * J[0][2];
    A[1][2] = J[0][1] * J[2][0] - J[2][1] * J[0][0];
    A[2][0] = J[0][1] * J[1][2] - J[1][1] * J[0][2];
    A[2][1] = J[0][2] * J[1][0] - J[1][2] * J[0][0];
    A[2][2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
    double det = 0.0;
    for (int i = 0; i < 3; ++i) det += J[i][0] * A[i][0];
    const double inv = 1.0 / det;
    for (int n = 0; n < 8; ++n)
      for (int i = 0; i < 3; ++i) {
        double w = 0.0;
        for (int j = 0; j < 3; ++j) w += A[i][j] * rg[3 * n + j];
        dndx[24 * ip + 3 * n + i] = w * inv;
      }
  }
}

void or_hexscs_grad_op(const double *coords, int shifted, double *dndx, double *deriv) {
  or_hex8_derivative(12, or_hexscs_intgloc(shifted), deriv);
  or_generic_grad_op(12, deriv, coords, dndx);
}
void or_hexscv_grad_op(const double *coords, int shifted, double *dndx, double *deriv) {
  or_hex8_derivative(8, or_hexscv_intgloc(shifted), deriv);
  or_generic_grad_op(8, deriv, coords, dndx);
}

/* hex_volume_grandy (Hex8GeometryFunctions.h:78-163). */
static const int kGrandyFace[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                      {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 3, 7}};
static const int kGrandyTri[24][3] = {
    {0, 8, 1},  {8, 2, 1},  {3, 2, 8},  {3, 8, 0},  {6, 9, 5},  {7, 9, 6},  {4, 9, 7},  {4, 5, 9},
    {10, 0, 1}, {5, 10, 1}, {4, 10, 5}, {4, 0, 10}, {7, 6, 11}, {6, 2, 11}, {2, 3, 11}, {3, 7, 11},
    {6, 12, 2}, {5, 12, 6}, {5, 1, 12}, {1, 2, 12}, {0, 4, 13}, {4, 7, 13}, {7, 3, 13}, {3, 0, 13}};
double or_hex_volume_grandy(const double *s) {
  double cv[14][3];
  for (int n = 0; n < 8; ++n)
    for (int d = 0; d < 3; ++d) cv[n][d] = s[3 * n + d];
  for (int k = 0; k < 6; ++k)
    for (int d = 0; d < 3; ++d)
      cv[8 + k][d] = 0.25 * (((cv[kGrandyFace[k][0]][d] + cv[kGrandyFace[k][1]][d]) +
                              cv[kGrandyFace[k][2]][d]) + cv[kGrandyFace[k][3]][d]);
  double vol = 0.0;
  for (int k = 0; k < 24; ++k) {
    const int p = kGrandyTri[k][0], q = kGrandyTri[k][1], r = kGrandyTri[k][2];
    const double m0 = (cv[p][0] + cv[q][0]) + cv[r][0];
    const double m1 = (cv[p][1] + cv[q][1]) + cv[r][1];
    const double m2 = (cv[p][2] + cv[q][2]) + cv[r][2];
    const double d0 = (cv[q][1] - cv[p][1]) * (cv[r][2] - cv[p][2]) -
                      (cv[r][1] - cv[p][1]) * (cv[q][2] - cv[p][2]);
    const double d1 = (cv[r][0] - cv[p][0]) * (cv[q][2] - cv[p][2]) -
                      (cv[q][0] - cv[p][0]) * (cv[r][2] - cv[p][2]);
    const double d2 = (cv[q][0] - cv[p][0]) * (cv[r][1] - cv[p][1]) -
                      (cv[r][0] - cv[p][0]) * (cv[q][1] - cv[p][1]);
    vol += (m0 * d0 + m1 * d1) + m2 * d2;
  }
  return vol / 18.0;
}

/* HexSCV::determinant (Hex8CVFEM.C:149-178). */
static const int kScvSub[8][8] = {{0, 8, 12, 11, 19, 20, 26, 25},  {8, 1, 9, 12, 20, 18, 24, 26},
                                  {12, 9, 2, 10, 26, 24, 22, 23},  {11, 12, 10, 3, 25, 26, 23, 21},
                                  {19, 20, 26, 25, 4, 13, 17, 16}, {20, 18, 24, 26, 13, 5, 14, 17},
                                  {26, 24, 22, 23, 17, 14, 6, 15}, {25, 26, 23, 21, 16, 17, 15, 7}};
void or_hexscv_volume(const double *coords, double *volume) {
  double cv[27 * 3], h[24];
  or_subdivide_hex8(coords, cv);
  for (int ip = 0; ip < 8; ++ip) {
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d) h[3 * n + d] = cv[3 * kScvSub[ip][n] + d];
    volume[ip] = or_hex_volume_grandy(h);
  }
}

/* ------------------------------------------------------------------------ */
/* Element kernels                                                           */
/* ------------------------------------------------------------------------ */

static void scs_shape(int shifted, double *shape) {
  or_hex8_shape_fcn(12, or_hexscs_intgloc(shifted), shape);
}

/* ScalarDiffElemKernel<Hex8>::execute (src/kernel/ScalarDiffElemKernel.C:66-111). */
void or_scalar_diff_elem(const double *coords, const double *q, const double *kc, int shifted,
                         double *lhs, double *rhs) {
  double shape[96], areav[36], dndx[288], deriv[288];
  scs_shape(0, shape);
  or_hex_scs_areav(coords, areav);
  or_hexscs_grad_op(coords, shifted, dndx, deriv);
  for (int ip = 0; ip < 12; ++ip) {
    const int il = kLrscv[2 * ip], ir = kLrscv[2 * ip + 1];
    double kIp = 0.0;
    for (int ic = 0; ic < 8; ++ic) kIp += shape[8 * ip + ic] * kc[ic];
    double qDiff = 0.0;
    for (int ic = 0; ic < 8; ++ic) {
      double f = 0.0;
      for (int j = 0; j < 3; ++j) f += -kIp * dndx[24 * ip + 3 * ic + j] * areav[3 * ip + j];
      qDiff += f * q[ic];
      lhs[8 * il + ic] += f;
      lhs[8 * ir + ic] -= f;
    }
    rhs[il] -= qDiff;
    rhs[ir] += qDiff;
  }
}

/* ContinuityAdvElemKernel<Hex8>::execute (src/kernel/ContinuityAdvElemKernel.C:88-148). */
void or_continuity_adv_elem(const double *coords, const double *rho, const double *p,
                            const double *u, const double *Gpdx, double tau, int shiftMdot,
                            int shiftPoisson, int reducedSens, double *lhs, double *rhs) {
  double shape[96], areav[36], dndx[288], dndxs[288], deriv[288];
  scs_shape(shiftMdot, shape);
  or_hex_scs_areav(coords, areav);
  or_hexscs_grad_op(coords, 0, dndx, deriv);
  or_hexscs_grad_op(coords, 1, dndxs, deriv);
  const double *dn = shiftPoisson ? dndxs : dndx;
  const double *dnl = (shiftPoisson || reducedSens) ? dndxs : dndx;
  for (int ip = 0; ip < 12; ++ip) {
    const int il = kLrscv[2 * ip], ir = kLrscv[2 * ip + 1];
    double rhoU[3] = {0, 0, 0}, gp[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
    for (int ic = 0; ic < 8; ++ic) {
      const double r = shape[8 * ip + ic];
      double lhsfac = 0.0;
      for (int j = 0; j < 3; ++j) {
        gp[j] += r * Gpdx[3 * ic + j];
        rhoU[j] += r * rho[ic] * u[3 * ic + j];
        dp[j] += dn[24 * ip + 3 * ic + j] * p[ic];
        lhsfac += -dnl[24 * ip + 3 * ic + j] * areav[3 * ip + j];
      }
      lhs[8 * il + ic] += lhsfac;
      lhs[8 * ir + ic] -= lhsfac;
    }
    double mdot = 0.0;
    for (int j = 0; j < 3; ++j) mdot += (rhoU[j] - tau * (dp[j] - gp[j])) * areav[3 * ip + j];
    rhs[il] -= mdot / tau;
    rhs[ir] += mdot / tau;
  }
}

/* ComputeMdotElemAlgorithm::execute (src/ComputeMdotElemAlgorithm.C:85-250).
 * The reference evaluates the geometry through the Fortran hex_scs_det /
 * hex_gradient_operator (MasterElementWork.F), which uses the same facet
 * table and formulas; rounding may differ by a few ulp from this C order. */
void or_mdot_elem(const double *coords, const double *rho, const double *p, const double *u,
                  const double *Gpdx, double tau, int shiftMdot, int shiftPoisson, double *mdot) {
  double shape[96], areav[36], dndx[288], deriv[288];
  scs_shape(shiftMdot, shape);
  or_hex_scs_areav(coords, areav);
  or_hexscs_grad_op(coords, shiftPoisson, dndx, deriv);
  for (int ip = 0; ip < 12; ++ip) {
    double rhoU[3] = {0, 0, 0}, gp[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
    for (int ic = 0; ic < 8; ++ic) {
      const double r = shape[8 * ip + ic];
      for (int j = 0; j < 3; ++j) {
        gp[j] += r * Gpdx[3 * ic + j];
        rhoU[j] += r * rho[ic] * u[3 * ic + j];
        dp[j] += dndx[24 * ip + 3 * ic + j] * p[ic];
      }
    }
    double t = 0.0;
    for (int j = 0; j < 3; ++j) t += (rhoU[j] - tau * (dp[j] - gp[j])) * areav[3 * ip + j];
    mdot[ip] = t;
  }
}

/* calc_mass_flow_rate_scs (unit_tests/kernels/UnitTestKernelUtils.C:547-628):
 * the unit-test producer of mdot for the momentum golds (rho u . A only). */
void or_mdot_rhou_elem(const double *coords, const double *rho, const double *u, double *mdot) {
  double shape[96], areav[36];
  scs_shape(0, shape);
  or_hex_scs_areav(coords, areav);
  for (int ip = 0; ip < 12; ++ip) {
    double rhoU[3] = {0, 0, 0};
    for (int ic = 0; ic < 8; ++ic) {
      const double r = shape[8 * ip + ic];
      for (int j = 0; j < 3; ++j) rhoU[j] += r * rho[ic] * u[3 * ic + j];
    }
    double t = 0.0;
    for (int j = 0; j < 3; ++j) t += rhoU[j] * areav[3 * ip + j];
    mdot[ip] = t;
  }
}

/* MomentumAdvDiffElemKernel<Hex8>::execute (src/kernel/MomentumAdvDiffElemKernel.C:72-186). */
void or_momentum_advdiff_elem(const double *coords, const double *u, const double *mu,
                              const double *mdot, double includeDivU, int skewSymmetric,
                              int shiftedGradOp, double *lhs, double *rhs) {
  double shape[96], ashape[96], areav[36], dndx[288], deriv[288];
  scs_shape(0, shape);
  scs_shape(skewSymmetric, ashape);
  or_hex_scs_areav(coords, areav);
  or_hexscs_grad_op(coords, shiftedGradOp, dndx, deriv);
  const int n = 24;
  for (int ip = 0; ip < 12; ++ip) {
    const int il = kLrscv[2 * ip], ir = kLrscv[2 * ip + 1];
    const int ilN = 3 * il, irN = 3 * ir;
    const double tmdot = mdot[ip];
    double muIp = 0.0, divU = 0.0, uIp[3] = {0, 0, 0};
    for (int ic = 0; ic < 8; ++ic) {
      const double r = shape[8 * ip + ic];
      const double rA = ashape[8 * ip + ic];
      muIp += r * mu[ic];
      for (int j = 0; j < 3; ++j) {
        const double uj = u[3 * ic + j];
        uIp[j] += rA * uj;
        divU += uj * dndx[24 * ip + 3 * ic + j];
      }
    }
    for (int i = 0; i < 3; ++i) {
      const double aflux = tmdot * uIp[i];
      const double divUstress = 2.0 / 3.0 * muIp * divU * areav[3 * ip + i] * includeDivU;
      rhs[ilN + i] -= aflux + divUstress;
      rhs[irN + i] += aflux + divUstress;
    }
    for (int ic = 0; ic < 8; ++ic) {
      const int icN = 3 * ic;
      const double lhsfacAdv = ashape[8 * ip + ic] * tmdot;
      for (int i = 0; i < 3; ++i) {
        const int L = ilN + i, R = irN + i;
        lhs[n * L + icN + i] += lhsfacAdv;
        lhs[n * R + icN + i] -= lhsfacAdv;
        double lhs_riC_i = 0.0;
        for (int j = 0; j < 3; ++j) {
          const double axj = areav[3 * ip + j];
          const double uj = u[3 * ic + j];
          const double fi = -muIp * dndx[24 * ip + 3 * ic + j] * axj;
          lhs_riC_i += fi;
          const double fj = -muIp * dndx[24 * ip + 3 * ic + i] * axj;
          lhs[n * L + icN + j] += fj;
          lhs[n * R + icN + j] -= fj;
          rhs[L] -= fj * uj;
          rhs[R] += fj * uj;
        }
        lhs[n * L + icN + i] += lhs_riC_i;
        lhs[n * R + icN + i] -= lhs_riC_i;
        const double ui = u[3 * ic + i];
        rhs[L] -= lhs_riC_i * ui;
        rhs[R] += lhs_riC_i * ui;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Graph build + scatter (src/TpetraLinearSystem.C)                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  int32_t *v;
  int n, cap;
} ivec;

static void ivec_push_unique(ivec *a, int32_t x) {
  for (int i = 0; i < a->n; ++i)
    if (a->v[i] == x) return; /* insert_connection's std::find dedup (:338-366) */
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 8;
    a->v = (int32_t *)realloc(a->v, sizeof(int32_t) * (size_t)a->cap);
  }
  a->v[a->n++] = x;
}

static const int64_t *g_sort_ids;
static int cmp_idx_by_id(const void *a, const void *b) {
  const int64_t x = g_sort_ids[*(const int64_t *)a], y = g_sort_ids[*(const int64_t *)b];
  return (x > y) - (x < y);
}

/* LocalGraphArrays::insertIndices (include/LinearSolver.h:77-89, insert :103-112):
 * sorted insert of numDof consecutive column ids per index, skipping present ones. */
static void lga_insert(int32_t *row, int64_t rowLen, int64_t numInds, const int32_t *inds,
                       int numDof) {
  const int32_t INV = -1;
  int64_t end = 0;
  while (end < rowLen && row[end] != INV) ++end;
  for (int64_t i = 0; i < numInds; ++i) {
    int64_t lo = 0, hi = end;
    while (lo < hi) { /* std::lower_bound */
      const int64_t mid = (lo + hi) / 2;
      if (row[mid] < inds[i]) lo = mid + 1;
      else hi = mid;
    }
    if (lo < end && row[lo] == inds[i]) continue;
    for (int64_t k = end + numDof - 1; k >= lo + numDof; --k) row[k] = row[k - numDof];
    for (int d = 0; d < numDof; ++d) row[lo + d] = inds[i] + d;
    end += numDof;
  }
}

int64_t or_graph_build(int64_t nnodes, const int64_t *node_ids, int64_t nelem, int npe,
                       const int32_t *conn, int numDof, int64_t *row_ptr, int32_t *col_idx,
                       int32_t *node_lid) {
  /* beginLinearSystemConstruction (:189-327): owned nodes sorted by global id,
   * LID = numDof * rank. */
  int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)nnodes);
  int32_t *rank = (int32_t *)malloc(sizeof(int32_t) * (size_t)nnodes);
  for (int64_t i = 0; i < nnodes; ++i) order[i] = i;
  g_sort_ids = node_ids;
  qsort(order, (size_t)nnodes, sizeof(int64_t), cmp_idx_by_id);
  for (int64_t r = 0; r < nnodes; ++r) rank[order[r]] = (int32_t)r;
  if (node_lid)
    for (int64_t i = 0; i < nnodes; ++i) node_lid[i] = numDof * rank[i];

  /* addConnections (:351-367): self + (min id, max id) pairs per element */
  ivec *cx = (ivec *)calloc((size_t)nnodes, sizeof(ivec));
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + (size_t)npe * e;
    for (int a = 0; a < npe; ++a) {
      ivec_push_unique(&cx[rank[nd[a]]], rank[nd[a]]);
      for (int b = a + 1; b < npe; ++b) {
        const int32_t ra = rank[nd[a]], rb = rank[nd[b]];
        const int a_then_b = node_ids[nd[a]] < node_ids[nd[b]];
        if (a_then_b) ivec_push_unique(&cx[ra], rb);
        else ivec_push_unique(&cx[rb], ra);
      }
    }
  }

  /* compute_graph_row_lengths (:802-865) */
  int64_t *len = (int64_t *)calloc((size_t)nnodes, sizeof(int64_t));
  for (int64_t a = 0; a < nnodes; ++a) {
    len[a] += (int64_t)numDof * cx[a].n;
    for (int k = 0; k < cx[a].n; ++k)
      if (cx[a].v[k] != a) len[cx[a].v[k]] += numDof;
  }
  int64_t nnz = 0;
  for (int64_t a = 0; a < nnodes; ++a) nnz += len[a] * numDof;

  if (row_ptr && col_idx) {
    /* LocalGraphArrays ctor (:62-75): per dof row, INVALID-filled */
    const int64_t nrows = nnodes * numDof;
    int64_t off = 0;
    for (int64_t a = 0; a < nnodes; ++a)
      for (int d = 0; d < numDof; ++d) {
        row_ptr[a * numDof + d] = off;
        off += len[a];
      }
    row_ptr[nrows] = off;
    for (int64_t i = 0; i < off; ++i) col_idx[i] = -1;
    /* insert_graph_connections (:877-918): first dof row of each node only */
    int32_t *cols = NULL;
    int cap = 0;
    for (int64_t a = 0; a < nnodes; ++a) {
      if (cx[a].n > cap) {
        cap = cx[a].n;
        cols = (int32_t *)realloc(cols, sizeof(int32_t) * (size_t)cap);
      }
      for (int k = 0; k < cx[a].n; ++k) cols[k] = numDof * cx[a].v[k];
      const int64_t r = a * numDof;
      lga_insert(col_idx + row_ptr[r], row_ptr[r + 1] - row_ptr[r], cx[a].n, cols, numDof);
      const int32_t self = numDof * (int32_t)a;
      for (int k = 0; k < cx[a].n; ++k) {
        const int64_t b = cx[a].v[k];
        if (b == a) continue;
        const int64_t rb = b * numDof;
        lga_insert(col_idx + row_ptr[rb], row_ptr[rb + 1] - row_ptr[rb], 1, &self, numDof);
      }
    }
    free(cols);
    /* fill_in_extra_dof_rows_per_node (:1227-1246) */
    for (int64_t a = 0; a < nnodes; ++a) {
      const int64_t r0 = a * numDof;
      const int64_t L = row_ptr[r0 + 1] - row_ptr[r0];
      for (int d = 1; d < numDof; ++d)
        memcpy(col_idx + row_ptr[r0 + d], col_idx + row_ptr[r0], sizeof(int32_t) * (size_t)L);
    }
    /* remove_invalid_indices (:1193-1225) is a no-op in serial: lengths are exact */
  }
  for (int64_t a = 0; a < nnodes; ++a) free(cx[a].v);
  free(cx);
  free(len);
  free(order);
  free(rank);
  return nnz;
}

/* Tpetra::Details::shellSortKeysAndValues restated: shell sort of keys with
 * a companion permutation.  Keys within one element are distinct, so any
 * correct sort yields the same permutation. */
static void shell_sort_kv(int32_t *keys, int32_t *vals, int n) {
  for (int gap = n / 2; gap > 0; gap /= 2)
    for (int i = gap; i < n; ++i)
      for (int j = i - gap; j >= 0 && keys[j] > keys[j + gap]; j -= gap) {
        int32_t t = keys[j];
        keys[j] = keys[j + gap];
        keys[j + gap] = t;
        t = vals[j];
        vals[j] = vals[j + gap];
        vals[j + gap] = t;
      }
}

/* TpetraLinearSystem::sumInto (:1457-1522) + sum_into_row (:1415-1453). */
void or_sum_into(int numDof, const int64_t *row_ptr, const int32_t *col_idx, double *values,
                 double *rhs, const int32_t *node_lid, int numEntities, const int32_t *entities,
                 const double *lhs_e, const double *rhs_e) {
  const int numRows = numEntities * numDof;
  int32_t ids[96], perm[96];
  for (int i = 0; i < numEntities; ++i)
    for (int d = 0; d < numDof; ++d) ids[i * numDof + d] = node_lid[entities[i]] + d;
  for (int i = 0; i < numRows; ++i) perm[i] = i;
  shell_sort_kv(ids, perm, numRows);
  for (int r = 0; r < numRows; ++r) {
    const int i = perm[r] / numDof;
    const int64_t rowLid = node_lid[entities[i]] + perm[r] % numDof;
    const double *cur = lhs_e + (size_t)perm[r] * numRows;
    const int64_t b = row_ptr[rowLid], length = row_ptr[rowLid + 1] - b;
    int64_t off = 0;
    for (int j = 0; j < numRows; ++j) {
      while (off < length && col_idx[b + off] != ids[j]) ++off;
      if (off < length) values[b + off] += cur[perm[j]];
    }
    rhs[rowLid] += rhs_e[perm[r]];
  }
}

/* TpetraLinearSystem::applyDirichletBCs (:1579-1653), serial (all rows owned). */
void or_apply_dirichlet(int numDof, int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx,
                        double *values, double *rhs, const int32_t *node_lid, int64_t nbc,
                        const int32_t *bc_nodes, const double *solution, const double *bcvalues,
                        int beginPos, int endPos) {
  (void)nrows;
  for (int64_t k = 0; k < nbc; ++k) {
    const int32_t node = bc_nodes[k];
    for (int d = beginPos; d < endPos; ++d) {
      const int64_t lid = node_lid[node] + d;
      for (int64_t q = row_ptr[lid]; q < row_ptr[lid + 1]; ++q)
        values[q] = (col_idx[q] == lid) ? 1.0 : 0.0;
      rhs[lid] = bcvalues[(size_t)node * numDof + d] - solution[(size_t)node * numDof + d];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Assembly drivers: AssembleElemSolverAlgorithm::execute order              */
/* (src/AssembleElemSolverAlgorithm.C:75-100): elements in mesh order, lhs/rhs */
/* zeroed per element, kernel, then apply_coeff -> sumInto.                   */
/* ------------------------------------------------------------------------ */

static void gather(const int32_t *nd, const double *f, int ncomp, double *out) {
  for (int n = 0; n < 8; ++n)
    for (int c = 0; c < ncomp; ++c) out[ncomp * n + c] = f[(size_t)ncomp * nd[n] + c];
}

void or_assemble_scalar_diff(int64_t nelem, const int32_t *conn, const double *coords,
                             const double *q, const double *k, int shifted, const int64_t *row_ptr,
                             const int32_t *col_idx, const int32_t *node_lid, double *values,
                             double *rhs) {
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + 8 * e;
    double x[24], qe[8], ke[8], lhs[64], r[8];
    gather(nd, coords, 3, x);
    gather(nd, q, 1, qe);
    gather(nd, k, 1, ke);
    memset(lhs, 0, sizeof(lhs));
    memset(r, 0, sizeof(r));
    or_scalar_diff_elem(x, qe, ke, shifted, lhs, r);
    or_sum_into(1, row_ptr, col_idx, values, rhs, node_lid, 8, nd, lhs, r);
  }
}

void or_compute_mdot(int64_t nelem, const int32_t *conn, const double *coords, const double *rho,
                     const double *p, const double *u, const double *Gpdx, double tau,
                     int shiftMdot, int shiftPoisson, double *mdot) {
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + 8 * e;
    double x[24], re[8], pe[8], ue[24], ge[24];
    gather(nd, coords, 3, x);
    gather(nd, rho, 1, re);
    gather(nd, p, 1, pe);
    gather(nd, u, 3, ue);
    gather(nd, Gpdx, 3, ge);
    or_mdot_elem(x, re, pe, ue, ge, tau, shiftMdot, shiftPoisson, mdot + 12 * e);
  }
}

void or_assemble_continuity(int64_t nelem, const int32_t *conn, const double *coords,
                            const double *rho, const double *p, const double *u, const double *Gpdx,
                            double tau, int shiftMdot, int shiftPoisson, int reducedSens,
                            const int64_t *row_ptr, const int32_t *col_idx, const int32_t *node_lid,
                            double *values, double *rhs) {
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + 8 * e;
    double x[24], re[8], pe[8], ue[24], ge[24], lhs[64], r[8];
    gather(nd, coords, 3, x);
    gather(nd, rho, 1, re);
    gather(nd, p, 1, pe);
    gather(nd, u, 3, ue);
    gather(nd, Gpdx, 3, ge);
    memset(lhs, 0, sizeof(lhs));
    memset(r, 0, sizeof(r));
    or_continuity_adv_elem(x, re, pe, ue, ge, tau, shiftMdot, shiftPoisson, reducedSens, lhs, r);
    or_sum_into(1, row_ptr, col_idx, values, rhs, node_lid, 8, nd, lhs, r);
  }
}

void or_assemble_momentum(int64_t nelem, const int32_t *conn, const double *coords,
                          const double *u, const double *mu, const double *mdot, double includeDivU,
                          int skewSymmetric, int shiftedGradOp, const int64_t *row_ptr,
                          const int32_t *col_idx, const int32_t *node_lid, double *values,
                          double *rhs) {
  double *lhs = (double *)malloc(sizeof(double) * 576);
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + 8 * e;
    double x[24], ue[24], me[8], r[24];
    gather(nd, coords, 3, x);
    gather(nd, u, 3, ue);
    gather(nd, mu, 1, me);
    memset(lhs, 0, sizeof(double) * 576);
    memset(r, 0, sizeof(r));
    or_momentum_advdiff_elem(x, ue, me, mdot + 12 * e, includeDivU, skewSymmetric, shiftedGradOp,
                             lhs, r);
    or_sum_into(3, row_ptr, col_idx, values, rhs, node_lid, 8, nd, lhs, r);
  }
  free(lhs);
}

/* ------------------------------------------------------------------------ */
/* Kernel-input producers: dual nodal volume and projected nodal gradient   */
/* ------------------------------------------------------------------------ */

/* ComputeGeometryInteriorAlgorithm::execute (src/ComputeGeometryInteriorAlgorithm.C:77-120):
 * every element adds its scv volumes to their nodes (ipNodeMap: ip n -> node n).
 * dual_vol [nnodes] is overwritten. */
void or_dual_nodal_volume(int64_t nelem, const int32_t *conn, const double *coords, int64_t nnodes,
                          double *dual_vol) {
  memset(dual_vol, 0, sizeof(double) * nnodes);
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + 8 * e;
    double x[24], vol[8];
    gather(nd, coords, 3, x);
    or_hexscv_volume(x, vol);
    for (int ip = 0; ip < 8; ++ip) dual_vol[nd[ip]] += vol[ip];
  }
}

/* AssembleNodalGradElemAlgorithm (src/AssembleNodalGradElemAlgorithm.C:85-140):
 * Green-Gauss over the dual control volume, interior (element) part: per scs,
 * q_ip = sum_c N_c(ip) q_c (shifted: N at the shifted ip), then
 * dqdx[left] += q_ip A / V_left, dqdx[right] -= q_ip A / V_right.
 * dqdx [nnodes][3] is overwritten. */
void or_nodal_grad_elem(int64_t nelem, const int32_t *conn, const double *coords, const double *q,
                        const double *dual_vol, int shifted, int64_t nnodes, double *dqdx) {
  double shape[96];
  or_hex8_shape_fcn(12, or_hexscs_intgloc(shifted), shape);
  memset(dqdx, 0, sizeof(double) * 3 * nnodes);
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + 8 * e;
    double x[24], areav[36];
    gather(nd, coords, 3, x);
    or_hex_scs_areav(x, areav);
    for (int ip = 0; ip < 12; ++ip) {
      const int il = kLrscv[2 * ip], ir = kLrscv[2 * ip + 1];
      double qIp = 0.0;
      for (int c = 0; c < 8; ++c) qIp += shape[8 * ip + c] * q[nd[c]];
      const double invL = 1.0 / dual_vol[nd[il]], invR = 1.0 / dual_vol[nd[ir]];
      for (int j = 0; j < 3; ++j) {
        const double fac = qIp * areav[3 * ip + j];
        dqdx[3 * nd[il] + j] += fac * invL;
        dqdx[3 * nd[ir] + j] -= fac * invR;
      }
    }
  }
}

/* HexSCS sideNodeOrdinals_ (src/master_element/Hex8CVFEM.C:415-422): the
 * nodes of each side, counter-clockwise seen from outside. */
static const int kHexSide[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                   {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};

/* Exposed sides of a Hex8 block: element sides that no other element shares,
 * nodes in side order.  Returns their number; faces [count][4] may be NULL to
 * count only.  (The reference receives them as the exposed side parts of the
 * mesh; this builds the same set from the connectivity.) */
int64_t or_exposed_faces(int64_t nelem, const int32_t *conn, int64_t nnodes, int32_t *faces) {
  int64_t *ip = (int64_t *)calloc((size_t)nnodes + 1, sizeof(int64_t));
  int32_t *ie = (int32_t *)malloc(sizeof(int32_t) * (size_t)(8 * nelem));
  for (int64_t q = 0; q < 8 * nelem; ++q) ip[conn[q] + 1]++;
  for (int64_t a = 0; a < nnodes; ++a) ip[a + 1] += ip[a];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)nnodes);
  memcpy(fill, ip, sizeof(int64_t) * (size_t)nnodes);
  for (int64_t q = 0; q < 8 * nelem; ++q) ie[fill[conn[q]]++] = (int32_t)(q / 8);
  free(fill);
  int64_t count = 0;
  for (int64_t e = 0; e < nelem; ++e)
    for (int f = 0; f < 6; ++f) {
      int32_t fn[4];
      for (int k = 0; k < 4; ++k) fn[k] = conn[8 * e + kHexSide[f][k]];
      int shared = 0;
      for (int64_t q = ip[fn[0]]; q < ip[fn[0] + 1] && !shared; ++q) {
        const int64_t o = ie[q];
        if (o == e) continue;
        int hits = 0;
        for (int k = 1; k < 4; ++k)
          for (int n = 0; n < 8; ++n)
            if (conn[8 * o + n] == fn[k]) {
              ++hits;
              break;
            }
        shared = hits == 3;
      }
      if (!shared) {
        if (faces)
          for (int k = 0; k < 4; ++k) faces[4 * count + k] = fn[k];
        ++count;
      }
    }
  free(ip);
  free(ie);
  return count;
}

/* Quad3DSCS::determinant (quad3d_scs_det, src/master_element/MasterElementWork.F:1467-1584):
 * sub-face i of the quad p[4] spans node i, the midpoints of its two edges and
 * the centroid: A_i = 1/2 (c - p_i) x (e_{i-1} - e_i), e_i = (p_i + p_{i+1}) / 2. */
void or_quad3d_scs_areav(const double *p, double *areav) {
  double e[4][3], c[3];
  for (int d = 0; d < 3; ++d) {
    for (int i = 0; i < 4; ++i) e[i][d] = (p[3 * i + d] + p[3 * ((i + 1) % 4) + d]) * 0.5;
    c[d] = (p[d] + p[3 + d] + p[6 + d] + p[9 + d]) * 0.25;
  }
  for (int i = 0; i < 4; ++i) {
    const int im = (i + 3) % 4;
    const double dx13 = c[0] - p[3 * i], dy13 = c[1] - p[3 * i + 1], dz13 = c[2] - p[3 * i + 2];
    const double dx24 = e[im][0] - e[i][0], dy24 = e[im][1] - e[i][1], dz24 = e[im][2] - e[i][2];
    areav[3 * i + 0] = 0.5 * (dz24 * dy13 - dz13 * dy24);
    areav[3 * i + 1] = 0.5 * (dx24 * dz13 - dx13 * dz24);
    areav[3 * i + 2] = 0.5 * (dy24 * dx13 - dy13 * dx24);
  }
}

/* AssembleNodalGradBoundaryAlgorithm::execute (src/AssembleNodalGradBoundaryAlgorithm.C:54-152):
 * for every exposed face and its ip i (nearest node i, Quad3DSCS ipNodeMap_),
 * dqdx[node i] += q_ip A_i / dual_vol[node i], q_ip from the quad shape
 * functions (quad3d_shape_fcn, MasterElementWork.F:3651-3695) at (+-1/4, +-1/4),
 * or at the nodes themselves when shifted (+-1/2).  Adds into dqdx. */
void or_nodal_grad_boundary(int64_t nfaces, const int32_t *faces, const double *coords, const double *q,
                            const double *dual_vol, int shifted, double *dqdx) {
  const double h = shifted ? 0.5 : 0.25;
  const double loc[4][2] = {{-h, -h}, {h, -h}, {h, h}, {-h, h}};
  double shape[4][4];
  for (int j = 0; j < 4; ++j) {
    const double s1 = loc[j][0], s2 = loc[j][1];
    shape[j][0] = 0.25 + 0.5 * (-s1 - s2) + s1 * s2;
    shape[j][1] = 0.25 + 0.5 * (s1 - s2) - s1 * s2;
    shape[j][2] = 0.25 + 0.5 * (s1 + s2) + s1 * s2;
    shape[j][3] = 0.25 + 0.5 * (-s1 + s2) - s1 * s2;
  }
  for (int64_t f = 0; f < nfaces; ++f) {
    const int32_t *fn = faces + 4 * f;
    double x[12], areav[12], qf[4];
    for (int k = 0; k < 4; ++k) {
      for (int d = 0; d < 3; ++d) x[3 * k + d] = coords[3 * (size_t)fn[k] + d];
      qf[k] = q[fn[k]];
    }
    or_quad3d_scs_areav(x, areav);
    for (int ip = 0; ip < 4; ++ip) {
      double qIp = 0.0;
      for (int c = 0; c < 4; ++c) qIp += shape[ip][c] * qf[c];
      const double inv = 1.0 / dual_vol[fn[ip]];
      for (int j = 0; j < 3; ++j) dqdx[3 * (size_t)fn[ip] + j] += qIp * areav[3 * ip + j] * inv;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Time-derivative (mass) kernels                                           */
/* ------------------------------------------------------------------------ */

/* MomentumMassElemKernel<Hex8>::execute (src/kernel/MomentumMassElemKernel.C:95-170):
 * per scv ip (nearest node nn = ip, HexSCV::ipNodeMap_), fields interpolated
 * with the scv shape functions (the node values themselves when lumped, the
 * shifted scv locations being the nodes):
 *   rhs(nn, j) -= (g1 rhoNp1 uNp1_j + g2 rhoN uN_j + g3 rhoNm1 uNm1_j) scV / dt
 *                 + Gjp_j scV densFac,  densFac = balanced ? rhoNp1 : 1,
 *   lhs((nn, j), (c, j)) += N_c g1 rhoNp1 scV / dt.
 * rho3 = [Nm1 | N | Np1] x 8, u3 = [Nm1 | N | Np1] x 8 x 3, gjp 8 x 3,
 * gamma = (g1, g2, g3); lhs 24 x 24 and rhs 24 are accumulated into. */
void or_momentum_mass_elem(const double *coords, const double *rho3, const double *u3,
                           const double *gjp, double dt, const double *gamma, int lumped,
                           int balanced, double *lhs, double *rhs) {
  double shape[64], vol[8];
  or_hex8_shape_fcn(8, or_hexscv_intgloc(lumped), shape);
  or_hexscv_volume(coords, vol);
  for (int ip = 0; ip < 8; ++ip) {
    const int nn = ip;
    double rNm1 = 0.0, rN = 0.0, rNp1 = 0.0, uNm1[3] = {0, 0, 0}, uN[3] = {0, 0, 0},
           uNp1[3] = {0, 0, 0}, g[3] = {0, 0, 0};
    for (int c = 0; c < 8; ++c) {
      const double r = shape[8 * ip + c];
      rNm1 += r * rho3[c];
      rN += r * rho3[8 + c];
      rNp1 += r * rho3[16 + c];
      for (int j = 0; j < 3; ++j) {
        uNm1[j] += r * u3[3 * c + j];
        uN[j] += r * u3[24 + 3 * c + j];
        uNp1[j] += r * u3[48 + 3 * c + j];
        g[j] += r * gjp[3 * c + j];
      }
    }
    const double scV = vol[ip];
    const double densFac = balanced ? rNp1 : 1.0;
    for (int j = 0; j < 3; ++j)
      rhs[3 * nn + j] += -(gamma[0] * rNp1 * uNp1[j] + gamma[1] * rN * uN[j] + gamma[2] * rNm1 * uNm1[j]) *
                             scV / dt -
                         g[j] * scV * densFac;
    for (int c = 0; c < 8; ++c) {
      const double lf = shape[8 * ip + c] * gamma[0] * rNp1 * scV / dt;
      for (int j = 0; j < 3; ++j) lhs[(3 * nn + j) * 24 + 3 * c + j] += lf;
    }
  }
}

/* ContinuityMassElemKernel<Hex8>::execute (src/kernel/ContinuityMassElemKernel.C:84-118):
 * rhs(nn) -= (g1 rhoNp1 + g2 rhoN + g3 rhoNm1) scV / dt / (dt / g1); no lhs. */
void or_continuity_mass_elem(const double *coords, const double *rho3, double dt,
                             const double *gamma, int lumped, double *rhs) {
  double shape[64], vol[8];
  or_hex8_shape_fcn(8, or_hexscv_intgloc(lumped), shape);
  or_hexscv_volume(coords, vol);
  const double projTimeScale = dt / gamma[0];
  for (int ip = 0; ip < 8; ++ip) {
    double rNm1 = 0.0, rN = 0.0, rNp1 = 0.0;
    for (int c = 0; c < 8; ++c) {
      const double r = shape[8 * ip + c];
      rNm1 += r * rho3[c];
      rN += r * rho3[8 + c];
      rNp1 += r * rho3[16 + c];
    }
    rhs[ip] += -(gamma[0] * rNp1 + gamma[1] * rN + gamma[2] * rNm1) * vol[ip] / dt / projTimeScale;
  }
}

/* Whole-mesh assembly of the two mass kernels into a graph's CSR (the field
 * states are nodal arrays: rho [N] each, u [N][3] each, gjp [N][3]). */
void or_assemble_momentum_mass(int64_t nelem, const int32_t *conn, const double *coords,
                               const double *rhoNm1, const double *rhoN, const double *rhoNp1,
                               const double *uNm1, const double *uN, const double *uNp1,
                               const double *gjp, double dt, const double *gamma, int lumped,
                               int balanced, const int64_t *row_ptr, const int32_t *col_idx,
                               const int32_t *node_lid, double *values, double *rhs) {
  double *lhs = (double *)malloc(sizeof(double) * 576);
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + 8 * e;
    double x[24], r3[24], u3[72], ge[24], r[24];
    gather(nd, coords, 3, x);
    gather(nd, rhoNm1, 1, r3);
    gather(nd, rhoN, 1, r3 + 8);
    gather(nd, rhoNp1, 1, r3 + 16);
    gather(nd, uNm1, 3, u3);
    gather(nd, uN, 3, u3 + 24);
    gather(nd, uNp1, 3, u3 + 48);
    gather(nd, gjp, 3, ge);
    memset(lhs, 0, sizeof(double) * 576);
    memset(r, 0, sizeof(r));
    or_momentum_mass_elem(x, r3, u3, ge, dt, gamma, lumped, balanced, lhs, r);
    or_sum_into(3, row_ptr, col_idx, values, rhs, node_lid, 8, nd, lhs, r);
  }
  free(lhs);
}

void or_assemble_continuity_mass(int64_t nelem, const int32_t *conn, const double *coords,
                                 const double *rhoNm1, const double *rhoN, const double *rhoNp1,
                                 double dt, const double *gamma, int lumped, const int64_t *row_ptr,
                                 const int32_t *col_idx, const int32_t *node_lid, double *values,
                                 double *rhs) {
  double lhs[64];
  memset(lhs, 0, sizeof(lhs));
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t *nd = conn + 8 * e;
    double x[24], r3[24], r[8];
    gather(nd, coords, 3, x);
    gather(nd, rhoNm1, 1, r3);
    gather(nd, rhoN, 1, r3 + 8);
    gather(nd, rhoNp1, 1, r3 + 16);
    memset(r, 0, sizeof(r));
    or_continuity_mass_elem(x, r3, dt, gamma, lumped, r);
    or_sum_into(1, row_ptr, col_idx, values, rhs, node_lid, 8, nd, lhs, r);
  }
}

/* ------------------------------------------------------------------------ */
/* Krylov + relaxation: textbook restatements (Saad, Iterative Methods for   */
/* Sparse Linear Systems, 2nd ed.: GMRES Alg. 6.9 with Givens rotations,     */
/* right preconditioning Alg. 9.5; PCG Alg. 9.1; SGS M=(D+L)D^-1(D+U) 4.1).  */
/* Configuration mirrors src/LinearSolverConfig.C:34-108.                    */
/* ------------------------------------------------------------------------ */

void or_spmv(int64_t nrows, const int64_t *rp, const int32_t *ci, const double *v,
             const double *x, double *y) {
  for (int64_t i = 0; i < nrows; ++i) {
    double s = 0.0;
    for (int64_t q = rp[i]; q < rp[i + 1]; ++q) s += v[q] * x[ci[q]];
    y[i] = s;
  }
}

static void gs_row(int64_t i, const int64_t *rp, const int32_t *ci, const double *v,
                   const double *r, double *z) {
  double s = 0.0, dg = 0.0;
  for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
    if (ci[q] == i) dg = v[q];
    else s += v[q] * z[ci[q]];
  }
  z[i] = (r[i] - s) / dg;
}

void or_sgs_apply(int64_t nrows, const int64_t *rp, const int32_t *ci, const double *v,
                  const double *r, double *z) {
  for (int64_t i = 0; i < nrows; ++i) z[i] = 0.0;
  for (int64_t i = 0; i < nrows; ++i) gs_row(i, rp, ci, v, r, z);
  for (int64_t i = nrows - 1; i >= 0; --i) gs_row(i, rp, ci, v, r, z);
}

void or_mcsgs_apply(int64_t nrows, const int64_t *rp, const int32_t *ci, const double *v,
                    int ncolors, const int64_t *cp, const int32_t *cr, const double *r, double *z) {
  for (int64_t i = 0; i < nrows; ++i) z[i] = 0.0;
  for (int c = 0; c < ncolors; ++c)
    for (int64_t k = cp[c]; k < cp[c + 1]; ++k) gs_row(cr[k], rp, ci, v, r, z);
  for (int c = ncolors - 1; c >= 0; --c)
    for (int64_t k = cp[c]; k < cp[c + 1]; ++k) gs_row(cr[k], rp, ci, v, r, z);
}

void or_jacobi_apply(int64_t nrows, const int64_t *rp, const int32_t *ci, const double *v,
                     const double *r, double *z) {
  for (int64_t i = 0; i < nrows; ++i) {
    double dg = 0.0;
    for (int64_t q = rp[i]; q < rp[i + 1]; ++q)
      if (ci[q] == i) dg = v[q];
    z[i] = r[i] / dg;
  }
}

typedef struct {
  int64_t n;
  const int64_t *rp;
  const int32_t *ci;
  const double *v;
  int precond, ncolors;
  const int64_t *cp;
  const int32_t *cr;
} sys_t;

static void prec_apply(const sys_t *s, const double *r, double *z) {
  switch (s->precond) {
    case 1: or_jacobi_apply(s->n, s->rp, s->ci, s->v, r, z); break;
    case 2: or_sgs_apply(s->n, s->rp, s->ci, s->v, r, z); break;
    case 3: or_mcsgs_apply(s->n, s->rp, s->ci, s->v, s->ncolors, s->cp, s->cr, r, z); break;
    default: memcpy(z, r, sizeof(double) * (size_t)s->n);
  }
}

static double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static double true_rel_resid(const sys_t *s, const double *b, const double *x, double *w) {
  or_spmv(s->n, s->rp, s->ci, s->v, x, w);
  for (int64_t i = 0; i < s->n; ++i) w[i] = b[i] - w[i];
  const double nb = sqrt(dot(s->n, b, b));
  return sqrt(dot(s->n, w, w)) / (nb > 0 ? nb : 1.0);
}

int or_gmres(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *b,
             double *x, int m, int maxit, double tol, int precond, int ncolors, const int64_t *cp,
             const int32_t *cr, double *resid_out) {
  sys_t s = {n, rp, ci, v, precond, ncolors, cp, cr};
  double *V = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(m + 1));
  double *Z = (double *)malloc(sizeof(double) * (size_t)n);
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  double *H = (double *)calloc((size_t)(m + 1) * (size_t)m, sizeof(double));
  double *cs = (double *)malloc(sizeof(double) * (size_t)m);
  double *sn = (double *)malloc(sizeof(double) * (size_t)m);
  double *g = (double *)malloc(sizeof(double) * (size_t)(m + 1));
  double *y = (double *)malloc(sizeof(double) * (size_t)m);
  double *hcol = (double *)malloc(sizeof(double) * (size_t)(m + 1));
  int iters = 0;
  /* implicit residual scaling: norm of the initial residual */
  or_spmv(n, rp, ci, v, x, w);
  for (int64_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
  const double r0 = sqrt(dot(n, w, w));
  double rn = r0;
  while (iters < maxit && r0 > 0.0 && rn > tol * r0) {
    /* restart cycle */
    or_spmv(n, rp, ci, v, x, w);
    for (int64_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
    const double beta = sqrt(dot(n, w, w));
    for (int64_t i = 0; i < n; ++i) V[i] = w[i] / beta;
    for (int k = 0; k <= m; ++k) g[k] = 0.0;
    g[0] = beta;
    int k = 0;
    for (; k < m && iters < maxit; ++k) {
      ++iters;
      double *vk = V + (size_t)n * k, *vk1 = V + (size_t)n * (k + 1);
      prec_apply(&s, vk, Z);
      or_spmv(n, rp, ci, v, Z, vk1);
      /* ICGS: two classical Gram-Schmidt passes */
      for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] = 0.0;
      for (int pass = 0; pass < 2; ++pass) {
        for (int j = 0; j <= k; ++j) hcol[j] = dot(n, V + (size_t)n * j, vk1);
        for (int j = 0; j <= k; ++j) {
          const double h = hcol[j];
          const double *vj = V + (size_t)n * j;
          for (int64_t i = 0; i < n; ++i) vk1[i] -= h * vj[i];
          H[(size_t)j * m + k] += h;
        }
      }
      const double hn = sqrt(dot(n, vk1, vk1));
      H[(size_t)(k + 1) * m + k] = hn;
      if (hn > 0.0)
        for (int64_t i = 0; i < n; ++i) vk1[i] /= hn;
      for (int j = 0; j < k; ++j) {
        const double a = H[(size_t)j * m + k], c = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * a + sn[j] * c;
        H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * c;
      }
      const double a = H[(size_t)k * m + k], c = H[(size_t)(k + 1) * m + k];
      const double den = sqrt(a * a + c * c);
      cs[k] = den > 0 ? a / den : 1.0;
      sn[k] = den > 0 ? c / den : 0.0;
      H[(size_t)k * m + k] = den;
      H[(size_t)(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      rn = fabs(g[k + 1]);
      if (rn <= tol * r0) {
        ++k;
        break;
      }
    }
    /* y = H^-1 g; x += M^-1 V y */
    for (int i = k - 1; i >= 0; --i) {
      double t = g[i];
      for (int j = i + 1; j < k; ++j) t -= H[(size_t)i * m + j] * y[j];
      y[i] = t / H[(size_t)i * m + i];
    }
    for (int64_t i = 0; i < n; ++i) w[i] = 0.0;
    for (int j = 0; j < k; ++j) {
      const double *vj = V + (size_t)n * j;
      for (int64_t i = 0; i < n; ++i) w[i] += y[j] * vj[i];
    }
    prec_apply(&s, w, Z);
    for (int64_t i = 0; i < n; ++i) x[i] += Z[i];
  }
  if (resid_out) *resid_out = true_rel_resid(&s, b, x, w);
  free(V); free(Z); free(w); free(H); free(cs); free(sn); free(g); free(y); free(hcol);
  return iters;
}

int or_cg(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *b,
          double *x, int maxit, double tol, int precond, int ncolors, const int64_t *cp,
          const int32_t *cr, double *resid_out) {
  sys_t s = {n, rp, ci, v, precond, ncolors, cp, cr};
  double *r = (double *)malloc(sizeof(double) * (size_t)n);
  double *z = (double *)malloc(sizeof(double) * (size_t)n);
  double *p = (double *)malloc(sizeof(double) * (size_t)n);
  double *q = (double *)malloc(sizeof(double) * (size_t)n);
  or_spmv(n, rp, ci, v, x, q);
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  const double r0 = sqrt(dot(n, r, r));
  prec_apply(&s, r, z);
  memcpy(p, z, sizeof(double) * (size_t)n);
  double rz = dot(n, r, z);
  int it = 0;
  double rn = r0;
  while (it < maxit && r0 > 0.0 && rn > tol * r0) {
    ++it;
    or_spmv(n, rp, ci, v, p, q);
    const double alpha = rz / dot(n, p, q);
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rn = sqrt(dot(n, r, r));
    if (rn <= tol * r0) break;
    prec_apply(&s, r, z);
    const double rz1 = dot(n, r, z);
    const double beta = rz1 / rz;
    rz = rz1;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  if (resid_out) *resid_out = true_rel_resid(&s, b, x, q);
  free(r); free(z); free(p); free(q);
  return it;
}

int or_greedy_color(int64_t nrows, const int64_t *rp, const int32_t *ci, int32_t *color) {
  int ncol = 0;
  unsigned char used[256];
  for (int64_t i = 0; i < nrows; ++i) color[i] = -1;
  for (int64_t i = 0; i < nrows; ++i) {
    memset(used, 0, sizeof(used));
    for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
      const int32_t c = color[ci[q]];
      if (ci[q] != i && c >= 0 && c < 256) used[c] = 1;
    }
    int c = 0;
    while (c < 255 && used[c]) ++c;
    color[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  return ncol;
}

double or_wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
