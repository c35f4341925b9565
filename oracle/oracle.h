/*
 * CPU oracle for the Hex8 CVFEM assembly + linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference's arithmetic, used as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * library (nalu_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Floating-point
 * operation order follows the reference (compiled with -ffp-contract=off so
 * that no FMA contraction changes the rounding).
 *
 * Pinning status (what checks each function; see tests/test_oracle_*.py):
 *   shape fns / derivatives   partition of unity, closed-form trilinear, exact gradients of linear fields
 *   scs areav / subdivide     unit-cube closed form (|A| = 1/4), closure of every SCV, affine covariance
 *   generic grad_op           golds hex8_scv_grad_op[192] / hex8_scv_shifted_grad_op[192] (1e-10)
 *   scv volume (Grandy)       0.125 on unit cube; exact for affine maps; sum = element volume
 *   ScalarDiff kernel         gold 8x8 lhs, rhs = 0 (1e-15)
 *   ContinuityAdv kernel      golds default + reduced sensitivities (1e-15)
 *   MomentumAdvDiff kernel    gold rhs[24], lhs[24][24] (1e-15)
 *   mdot                      same formula as ContinuityAdv (its rhs = -+mdot/tau), closed form on affine fields
 *   graph + sumInto           Tpetra 1x1x2 golds (row lengths, 12x12 values, 1e-9), lattice closed form nnz=(3n+1)^3 nd^2
 *   Dirichlet                 closed form (row = e_i, rhs = bc - current)
 *   SGS / Jacobi / GMRES / CG textbook identities (M = (D+L) D^-1 (D+U) on unit vectors), dense LU solve on tiny meshes,
 *                             FP64 true residual <= tol.  Belos/Ifpack2 iterate histories: parity unpinned (not runnable).
 */
#ifndef NALU_ORACLE_H
#define NALU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- master element tables (Hex8CVFEM.C) ---- */
const double *or_hexscs_intgloc(int shifted);   /* [12][3] */
const double *or_hexscv_intgloc(int shifted);   /* [8][3]  */
const int *or_hexscs_lrscv(void);               /* [12][2] */

void or_hex8_shape_fcn(int npts, const double *isoParCoord, double *shape);
void or_hex8_derivative(int npts, const double *intgLoc, double *deriv);
void or_subdivide_hex8(const double *coords, double *coordv);
void or_quad_area_by_triangulation(const double *areacoords, double *area);
void or_hex_scs_areav(const double *coords, double *areav);
void or_generic_grad_op(int npts, const double *deriv, const double *coords, double *dndx);
void or_hexscs_grad_op(const double *coords, int shifted, double *dndx, double *deriv);
void or_hexscv_grad_op(const double *coords, int shifted, double *dndx, double *deriv);
double or_hex_volume_grandy(const double *scvcoords);
void or_hexscv_volume(const double *coords, double *volume);

/* ---- element kernels; lhs/rhs are ACCUMULATED into (caller zeroes) ---- */
void or_scalar_diff_elem(const double *coords, const double *q, const double *diffFluxCoeff,
                         int shiftedGradOp, double *lhs, double *rhs);
void or_continuity_adv_elem(const double *coords, const double *rho, const double *p,
                            const double *u, const double *Gpdx, double projTimeScale,
                            int shiftMdot, int shiftPoisson, int reducedSens,
                            double *lhs, double *rhs);
void or_mdot_elem(const double *coords, const double *rho, const double *p, const double *u,
                  const double *Gpdx, double projTimeScale, int shiftMdot, int shiftPoisson,
                  double *mdot);
void or_mdot_rhou_elem(const double *coords, const double *rho, const double *u, double *mdot);
void or_momentum_advdiff_elem(const double *coords, const double *u, const double *mu,
                              const double *mdot, double includeDivU, int skewSymmetric,
                              int shiftedGradOp, double *lhs, double *rhs);

/* ---- Tpetra-equivalent serial graph + scatter (TpetraLinearSystem.C) ---- */
/* node_lid[n] (out) = numDof * rank of node_ids[n] in ascending id order.
 * Returns nnz.  Pass row_ptr/col_idx = NULL to query nnz only. */
int64_t or_graph_build(int64_t nnodes, const int64_t *node_ids, int64_t nelem, int npe,
                       const int32_t *conn, int numDof, int64_t *row_ptr, int32_t *col_idx,
                       int32_t *node_lid);
void or_sum_into(int numDof, const int64_t *row_ptr, const int32_t *col_idx, double *values,
                 double *rhs, const int32_t *node_lid, int numEntities, const int32_t *entities,
                 const double *lhs_e, const double *rhs_e);
void or_apply_dirichlet(int numDof, int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx,
                        double *values, double *rhs, const int32_t *node_lid, int64_t nbc,
                        const int32_t *bc_nodes, const double *solution, const double *bcvalues,
                        int beginPos, int endPos);

/* ---- whole-mesh assembly drivers (AssembleElemSolverAlgorithm::execute order) ---- */
void or_assemble_scalar_diff(int64_t nelem, const int32_t *conn, const double *coords,
                             const double *q, const double *k, int shifted, const int64_t *row_ptr,
                             const int32_t *col_idx, const int32_t *node_lid, double *values, double *rhs);
void or_compute_mdot(int64_t nelem, const int32_t *conn, const double *coords, const double *rho,
                     const double *p, const double *u, const double *Gpdx, double projTimeScale,
                     int shiftMdot, int shiftPoisson, double *mdot);
void or_assemble_continuity(int64_t nelem, const int32_t *conn, const double *coords,
                            const double *rho, const double *p, const double *u, const double *Gpdx,
                            double projTimeScale, int shiftMdot, int shiftPoisson, int reducedSens,
                            const int64_t *row_ptr, const int32_t *col_idx, const int32_t *node_lid,
                            double *values, double *rhs);
void or_assemble_momentum(int64_t nelem, const int32_t *conn, const double *coords,
                          const double *u, const double *mu, const double *mdot, double includeDivU,
                          int skewSymmetric, int shiftedGradOp, const int64_t *row_ptr,
                          const int32_t *col_idx, const int32_t *node_lid, double *values, double *rhs);
/* dual nodal volume (ComputeGeometryInteriorAlgorithm.C:77-120) and the interior
 * Green-Gauss nodal gradient (AssembleNodalGradElemAlgorithm.C:85-140); pinned by
 * closed forms (volumes of a lattice, exact gradients of linear fields) */
void or_dual_nodal_volume(int64_t nelem, const int32_t *conn, const double *coords, int64_t nnodes,
                          double *dual_vol);
void or_nodal_grad_elem(int64_t nelem, const int32_t *conn, const double *coords, const double *q,
                        const double *dual_vol, int shifted, int64_t nnodes, double *dqdx);
/* exposed sides of the block (nodes in HexSCS side order; faces may be NULL to
 * count), the Quad3DSCS sub-face area vectors, and the boundary closure of the
 * nodal gradient (AssembleNodalGradBoundaryAlgorithm.C:54-152, adds into dqdx);
 * pinned together with the interior part by exact gradients of linear fields at
 * every node */
int64_t or_exposed_faces(int64_t nelem, const int32_t *conn, int64_t nnodes, int32_t *faces);
void or_quad3d_scs_areav(const double *coords, double *areav);
void or_nodal_grad_boundary(int64_t nfaces, const int32_t *faces, const double *coords, const double *q,
                            const double *dual_vol, int shifted, double *dqdx);
/* mass kernels (MomentumMassElemKernel.C:95-170, ContinuityMassElemKernel.C:84-118);
 * pinned by the reference's unit-test golds (consistent and lumped) */
void or_momentum_mass_elem(const double *coords, const double *rho3, const double *u3,
                           const double *gjp, double dt, const double *gamma, int lumped,
                           int balanced, double *lhs, double *rhs);
void or_continuity_mass_elem(const double *coords, const double *rho3, double dt,
                             const double *gamma, int lumped, double *rhs);
void or_assemble_momentum_mass(int64_t nelem, const int32_t *conn, const double *coords,
                               const double *rhoNm1, const double *rhoN, const double *rhoNp1,
                               const double *uNm1, const double *uN, const double *uNp1,
                               const double *gjp, double dt, const double *gamma, int lumped,
                               int balanced, const int64_t *row_ptr, const int32_t *col_idx,
                               const int32_t *node_lid, double *values, double *rhs);
void or_assemble_continuity_mass(int64_t nelem, const int32_t *conn, const double *coords,
                                 const double *rhoNm1, const double *rhoN, const double *rhoNp1,
                                 double dt, const double *gamma, int lumped, const int64_t *row_ptr,
                                 const int32_t *col_idx, const int32_t *node_lid, double *values,
                                 double *rhs);

/* ---- Krylov + relaxation (Belos / Ifpack2 semantics restated from the textbook) ---- */
void or_spmv(int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx, const double *vals,
             const double *x, double *y);
/* Ifpack2 Relaxation "Symmetric Gauss-Seidel", 1 sweep, zero initial guess */
void or_sgs_apply(int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx, const double *vals,
                  const double *r, double *z);
/* multicolor SGS with a given row coloring (rows of one color are mutually independent) */
void or_mcsgs_apply(int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx,
                    const double *vals, int ncolors, const int64_t *color_ptr,
                    const int32_t *color_rows, const double *r, double *z);
void or_jacobi_apply(int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx,
                     const double *vals, const double *r, double *z);

/* precond: 0 none, 1 jacobi, 2 sgs (sequential), 3 multicolor sgs.  Returns iterations;
 * *resid_out = final true relative residual ||b-Ax||/||b||.  x is the initial guess (in/out). */
int or_gmres(int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx, const double *vals,
             const double *b, double *x, int kspace, int maxit, double tol, int precond,
             int ncolors, const int64_t *color_ptr, const int32_t *color_rows, double *resid_out);
int or_cg(int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx, const double *vals,
          const double *b, double *x, int maxit, double tol, int precond, int ncolors,
          const int64_t *color_ptr, const int32_t *color_rows, double *resid_out);

/* greedy row coloring of a CSR graph in row order (first-fit) */
int or_greedy_color(int64_t nrows, const int64_t *row_ptr, const int32_t *col_idx, int32_t *color);

/* wall-clock helper for the CPU baseline */
double or_wtime(void);

#ifdef __cplusplus
}
#endif
#endif
