/*
 * nalu_hip.h -- C ABI of the MI355X-native Hex8 CVFEM assembly + linear-solve
 * path.  This is the drop-in boundary that replaces Nalu's Tpetra-backed
 * LinearSystem / LinearSolver and the consolidated element-assembly loop.
 *
 * Conventions
 *   - Every function returns an int status unless stated: 0 ok, >0 "not
 *     converged" (solve only), <0 error.  nalu_last_error() returns the
 *     message of the last error on the calling thread.  (The reference throws
 *     on logic errors, src/TpetraLinearSystem.C:1626-1629; the C++ wrapper in
 *     nalu_hip.hpp turns negative codes back into exceptions.)
 *   - Pointers named *_dev are device (HBM) pointers; *_host are host
 *     pointers, borrowed only for the duration of the call.
 *   - Node arrays are node-major: a scalar field is [N], a vector field is
 *     [N][3] (the STK field layout, src/ScratchViews.C:408-483).  Element
 *     fields are element-major: mdot is [E][12] (mass_flow_rate_scs).
 *   - Connectivity is [E][8] int32 node indices in Hex8 order
 *     (src/master_element/Hex8CVFEM.C nodeLoc_ :420-436).
 *   - One handle is not re-entrant; all work of a handle runs in order on the
 *     handle's HIP stream (default: the null stream).
 */
#ifndef NALU_HIP_H
#define NALU_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NALU_OK 0
#define NALU_NOT_CONVERGED 1
#define NALU_ERR_ARG (-1)
#define NALU_ERR_STATE (-2)
#define NALU_ERR_HIP (-3)
#define NALU_ERR_UNSUPPORTED (-4)
#define NALU_ERR_NUMERIC (-5)

/* Solver methods: LinearSolverConfig "method" (src/LinearSolverConfig.C:34-53). */
#define NALU_METHOD_GMRES 0
#define NALU_METHOD_CG 1
#define NALU_METHOD_FGMRES 2
/* Preconditioners: LinearSolverConfig "preconditioner" (src/LinearSolverConfig.C:71-85).
 * SGS is the multicolor ("mt_sgs") symmetric Gauss-Seidel, 1 sweep. */
#define NALU_PREC_NONE 0
#define NALU_PREC_JACOBI 1
#define NALU_PREC_SGS 2
#define NALU_PREC_LEARNED 3 /* learned message-passing net, SGS fallback; no reference counterpart */
#define NALU_PREC_MUELU 4   /* smoothed-aggregation AMG V-cycle ("muelu", src/LinearSolver.C:124-152) */
/* Arithmetic of the Krylov inner loop. */
#define NALU_PRECISION_FP64 0
#define NALU_PRECISION_MIXED 1 /* FP32 inner Krylov/SpMV/SGS inside FP64 iterative refinement */

typedef struct nalu_mesh nalu_mesh;
typedef struct nalu_linsys nalu_linsys;

typedef struct nalu_solver_config {
  int method;             /* NALU_METHOD_* */
  int precond;            /* NALU_PREC_* */
  double tolerance;       /* "tolerance" */
  double final_tolerance; /* "final_tolerance" (used when final_outer_iter != 0) */
  int max_iterations;     /* "max_iterations" */
  int kspace;             /* "kspace" (GMRES restart length) */
  int precision;          /* NALU_PRECISION_* */
  int final_outer_iter;   /* Realm::isFinalOuterIter_ */
} nalu_solver_config;

/* The MueLu parameter-list entries the AMG preconditioner honours (the keys of
 * reg_tests/xml/matches_ml_default.xml and milestone.xml; defaults are that
 * first file's values plus MueLu's / Ifpack2's own defaults for the rest). */
typedef struct nalu_muelu_params {
  int max_coarse_size; /* "coarse: max size" (1000, at most 2048); the coarsest level is solved densely */
  int max_levels;      /* "max levels" (10) */
  int cheb_degree;     /* "chebyshev: degree" (4) */
  double cheb_ratio;   /* "chebyshev: ratio eigenvalue" (20): lambda_min = lambda_max / ratio */
  double cheb_boost;   /* "chebyshev: boost factor" (1.1) applied to the estimate */
  int cheb_eig_iters;  /* "chebyshev: eigenvalue max iterations" (30) */
  double drop_tol;     /* "aggregation: drop tol" (0) */
  double sa_damping;   /* "sa: damping factor" (4/3): omega = damping / lambda_max(D^-1 A) */
} nalu_muelu_params;

/* ---------------------------------------------------------------- errors */
const char *nalu_last_error(void);
/* 1 when a HIP device is present and usable. */
int nalu_device_available(void);
/* Set the HIP stream (hipStream_t) used by a linear system; NULL = null stream. */
int nalu_linsys_set_stream(nalu_linsys *ls, void *stream);

/* ------------------------------------------------------------------ mesh */
/* A device-resident Hex8 block: coordinates [N][3], connectivity [E][8] and
 * optional global ids [N] (NULL: id = index + 1).  Replaces the STK bulk data
 * the reference's algorithms iterate (src/AssembleElemSolverAlgorithm.C:75-100).
 * Arrays are copied from the host. */
int nalu_mesh_create(int64_t num_nodes, int64_t num_elems, const int32_t *conn_host,
                     const double *coords_host, const int64_t *node_ids_host,
                     nalu_mesh **out);
int nalu_mesh_destroy(nalu_mesh *m);
/* Device pointers owned by the mesh (coords [N][3], conn [E][8]). */
int nalu_mesh_device_ptrs(const nalu_mesh *m, const double **coords_dev, const int32_t **conn_dev);

/* Synthetic structured Hex8 lattice on [0,1]^3 with STK generated-mesh
 * numbering: node id = 1 + i + (n+1)(j + (n+1)k); element nodes in Hex8 order
 * (unit_tests/UnitTestTpetra.C:97-98).  Interior nodes are jittered uniformly
 * by +-jitter*h with the reference's machine-independent LCG
 * (unit_tests/UnitTestUtils.C:41-53), three draws per node in node-id order.
 * stretch > 1 grades the spacing geometrically from all six faces toward the
 * centre (ratio = stretch per interval), the growth stopping once the spacing
 * reaches max_aspect times the wall spacing (max_aspect <= 0: no cap).
 * coords_host [(n+1)^3][3], conn_host [n^3][8]. */
int nalu_lattice_generate(int n, double jitter, double stretch, double max_aspect,
                          uint32_t seed, double *coords_host, int32_t *conn_host);
/* n lognormal(0, sigma) samples, Box-Muller on the same LCG. */
int nalu_lognormal_field(int64_t n, double sigma, uint32_t seed, double *out_host);
/* n N(0, sigma^2) samples (Box-Muller on the LCG). */
int nalu_normal_field(int64_t n, double sigma, uint32_t seed, double *out_host);

/* --------------------------------------------------------- linear system */
/* LinearSystem::create (src/LinearSystem.C:85-97) with PetraType PT_HIP. */
int nalu_linsys_create(int num_dof, const nalu_solver_config *cfg, nalu_linsys **out);
/* ~TpetraLinearSystem (src/TpetraLinearSystem.C:97-102). */
int nalu_linsys_destroy(nalu_linsys *ls);
/* buildElemToNodeGraph (src/TpetraLinearSystem.C:430-435). */
int nalu_linsys_build_elem_to_node_graph(nalu_linsys *ls, const nalu_mesh *m);
/* finalizeLinearSystem (src/TpetraLinearSystem.C:1256-1353): CSR built on the
 * GPU, rows ordered by global id, columns sorted, numDof rows interleaved. */
int nalu_linsys_finalize(nalu_linsys *ls);
/* zeroSystem (src/TpetraLinearSystem.C:1355-1372): values, rhs and solution. */
int nalu_linsys_zero(nalu_linsys *ls);
/* sumInto, Kokkos overload (TpetraLinearSystem::sumInto, src/TpetraLinearSystem.C:1457-1522,
 * called per element by SolverAlgorithm::apply_coeff, src/SolverAlgorithm.C:58-68),
 * batched over num_elems elements on the device.  Per element e:
 *   entities_dev [E][n] node indices (n = num_entities, n*numDof <= 64),
 *   rhs_dev [E][R], lhs_dev [E][R][R] row-major (R = n*numDof),
 *   local_ids_dev / sort_permutation_dev [E][R]: caller scratch, overwritten
 *   with the element's sorted column LIDs and their permutation, as the
 *   reference overwrites its team scratch views.
 * Each row walks its sorted CSR row once (sum_into_row, :1376-1453); entries
 * outside the graph are dropped; additions are atomic (the reference's
 * non-Serial path). */
int nalu_linsys_sum_into_elements(nalu_linsys *ls, int64_t num_elems, int num_entities,
                                  const int32_t *entities_dev, const double *rhs_dev,
                                  const double *lhs_dev, int32_t *local_ids_dev,
                                  int32_t *sort_permutation_dev);
/* No counterpart in the reference: the sumInto above for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results without the
 * caller scratch: the localIds / sortPermutation tables and the error flag live
 * in a 32 MiB file-scope device array of the library, one per GPU, kept until
 * the process ends.  At most (32 MiB - 256 B) / (8 R) elements per call
 * (NALU_ERR_ARG beyond); NALU_ERR_STATE while host sumInto calls are queued
 * (applying them would allocate).  Calls that overlap on one GPU share the
 * array and race. */
int nalu_linsys_sum_into_elements_noalloc(nalu_linsys *ls, int64_t num_elems, int num_entities,
                                          const int32_t *entities_dev, const double *rhs_dev,
                                          const double *lhs_dev);
/* The same for the graph mesh's own elements, looked up through the element
 * scatter table built with the graph (no sort, no row walk): conn_dev must be
 * the mesh's connectivity (nalu_mesh_device_ptrs), lhs_dev [E][R][R], rhs_dev
 * [E][R], R = 8*numDof, elements 0 .. num_elems-1.  Atomic scatter. */
int nalu_linsys_sum_into_batch(nalu_linsys *ls, int64_t num_elems, const int32_t *conn_dev,
                               const double *lhs_dev, const double *rhs_dev);
/* sumInto, std::vector overload (src/TpetraLinearSystem.C:1524-1577), one
 * element from host arrays: entities_host [n] node indices, rhs_host [R],
 * lhs_host [R][R].  The arrays are copied into a host queue and the call
 * returns; queued elements are applied on the device, in one batch, before
 * any other operation on the system reads or writes it (zeroSystem discards
 * them). */
int nalu_linsys_sum_into(nalu_linsys *ls, int num_entities, const int32_t *entities_host,
                         const double *rhs_host, const double *lhs_host);
/* applyDirichletBCs (src/TpetraLinearSystem.C:1579-1653): rows of the listed
 * nodes become e_i, rhs = bc - current for dofs [begin_pos, end_pos). */
int nalu_linsys_apply_dirichlet(nalu_linsys *ls, int64_t num_bc_nodes,
                                const int32_t *bc_nodes_dev, const double *solution_dev,
                                const double *bc_values_dev, int begin_pos, int end_pos);
/* resetRows (include/LinearSystem.h:116-120): rows of the listed nodes
 * become diag_value*e_i with rhs_residual in the rhs, for dofs [begin, end). */
int nalu_linsys_reset_rows(nalu_linsys *ls, int64_t num_nodes, const int32_t *nodes_dev,
                           int begin_pos, int end_pos, double diag_value, double rhs_residual);
/* Subtract the mean of the rhs: projects the rhs of a pure-Neumann (all-wall)
 * pressure-Poisson system onto the range of its matrix, whose null space is
 * the constants (the reference instead pins one node, FixPressureAtNodeAlgorithm,
 * src/FixPressureAtNodeAlgorithm.C). */
int nalu_linsys_remove_rhs_mean(nalu_linsys *ls);
/* loadComplete (src/TpetraLinearSystem.C:1757-1778): serial, no export. */
int nalu_linsys_load_complete(nalu_linsys *ls);
/* solve (src/TpetraLinearSystem.C:1780-1847): preconditioner compute + Krylov
 * solve from x0 = 0; writes the delta-solution [rows] to delta_out_dev.
 * Returns 0 converged, NALU_NOT_CONVERGED when max_iterations was reached. */
int nalu_linsys_solve(nalu_linsys *ls, double *delta_out_dev);
int nalu_linsys_iterations(const nalu_linsys *ls);
double nalu_linsys_linear_residual(const nalu_linsys *ls);     /* ||b - A x||_2 */
double nalu_linsys_nonlinear_residual(const nalu_linsys *ls);  /* ||rhs||_2 (l2Scaling = 1) */
double nalu_linsys_scaled_nonlinear_residual(const nalu_linsys *ls);
/* Set Realm::l2Scaling_ (default 1). */
int nalu_linsys_set_l2_scaling(nalu_linsys *ls, double l2_scaling);

/* Sizes and copies of the assembled system (for checks and export). */
int nalu_linsys_sizes(const nalu_linsys *ls, int64_t *num_rows, int64_t *nnz);
int nalu_linsys_copy_csr(nalu_linsys *ls, int64_t *row_ptr_host, int32_t *col_idx_host,
                         double *values_host, double *rhs_host);
/* Device pointers to the CSR arrays (row_ptr/col_idx int32, values/rhs FP64). */
int nalu_linsys_device_ptrs(nalu_linsys *ls, const int32_t **row_ptr_dev,
                            const int32_t **col_idx_dev, double **values_dev, double **rhs_dev);
/* Multicolor ordering used by the SGS preconditioner: colors, and rows per color. */
int nalu_linsys_coloring(nalu_linsys *ls, int *num_colors, int64_t *color_ptr_host,
                         int32_t *color_rows_host);

/* Per-phase accumulated times in seconds (hipEvent), the reference's timer
 * split (src/EquationSystem.C:270-292): [0] graph/init, [1] zero, [2] assemble,
 * [3] dirichlet, [4] load_complete, [5] precond setup, [6] solve. */
int nalu_linsys_phase_times(const nalu_linsys *ls, double *times7);

/* Solver building blocks on the handle's matrix (device vectors of length rows). */
int nalu_linsys_spmv(nalu_linsys *ls, const double *x_dev, double *y_dev);
int nalu_linsys_spmv_f32(nalu_linsys *ls, const float *x_dev, float *y_dev);
int nalu_linsys_precond_setup(nalu_linsys *ls);
int nalu_linsys_precond_apply(nalu_linsys *ls, const double *r_dev, double *z_dev);
/* The solvers work on a copy of the matrix in color-contiguous row order:
 * new row k is original row perm[k] (perm_host [rows]).  These apply M^-1 and
 * A to vectors in that order (after nalu_linsys_precond_setup). */
int nalu_linsys_permutation(nalu_linsys *ls, int32_t *perm_host);
int nalu_linsys_precond_apply_permuted(nalu_linsys *ls, const double *r_dev, double *z_dev);
/* The same in FP32 arithmetic: the preconditioner as the mixed-precision inner
 * solve applies it.  No counterpart in the reference. */
int nalu_linsys_precond_apply_permuted_f32(nalu_linsys *ls, const float *r_dev, float *z_dev);
int nalu_linsys_spmv_permuted(nalu_linsys *ls, const double *x_dev, double *y_dev);
/* Solve A x = b for a caller-provided rhs (x in/out is the initial guess). */
int nalu_linsys_solve_rhs(nalu_linsys *ls, const double *b_dev, double *x_dev, int *iters,
                          double *final_rel_resid);
/* k right-hand sides sharing the matrix and one preconditioner set-up (the
 * three momentum components of the segregated form): B, X [k][nrows] device
 * arrays (X holds the initial guesses), iters / rel: [k] host arrays or NULL.
 * Returns the worst status (0 converged, 1 iteration cap). */
int nalu_linsys_solve_multi(nalu_linsys *ls, int k, const double *B_dev, double *X_dev, int *iters,
                            double *rel);
/* The same right-hand sides solved as one block-diagonal stacked system
 * diag(A, ..., A) [X_0; ...; X_{k-1}] = [B_0; ...; B_{k-1}] in a single Krylov
 * space: every SpMV and SGS sweep reads the matrix once for all k vectors
 * (batched kernels for k <= 4), and convergence is measured on the stacked
 * residual norm -- the way the reference measures its numDof 3 momentum
 * system over all three components.  iters: Krylov iterations of the one
 * solve; rel [k] (optional): each right-hand side's FP64 true relative
 * residual.  No counterpart in the reference (its momentum system is coupled). */
int nalu_linsys_solve_stacked(nalu_linsys *ls, int k, const double *B, double *X, int *iters, double *rel);

/* ------------------------------------------------------ assembly kernels */
/* Each fused kernel performs, for every element: gather, Hex8 SCS geometry,
 * the element kernel, and the scatter into the system (the consolidated
 * AssembleElemSolverAlgorithm::execute, src/AssembleElemSolverAlgorithm.C:75-100).
 * Contributions are ADDED to the system (call nalu_linsys_zero first).
 * mode: 0 = owner-computes row gather (deterministic, no atomics, scalar
 * systems), 1 = element-centric atomic scatter. */
/* ScalarDiffElemKernel (src/kernel/ScalarDiffElemKernel.C:66-111). */
int nalu_assemble_scalar_diff(nalu_linsys *ls, const nalu_mesh *m, const double *q_dev,
                              const double *diff_coeff_dev, int shifted_grad_op, int mode);
/* ComputeMdotElemAlgorithm::execute (src/ComputeMdotElemAlgorithm.C:85-250):
 * mdot_dev [E][12], 16-byte aligned (NALU_ERR_ARG otherwise). */
int nalu_compute_mdot(const nalu_mesh *m, const double *rho_dev, const double *p_dev,
                      const double *u_dev, const double *gpdx_dev, double proj_time_scale,
                      int shift_mdot, int shift_poisson, double *mdot_dev, void *stream);
/* The same with FP32 arithmetic (Jacobians and determinants in FP64, nodal
 * coordinates taken relative to each element's first node); the mixed-precision
 * path.  No counterpart in the reference. */
int nalu_compute_mdot_f32(const nalu_mesh *m, const double *rho_dev, const double *p_dev,
                          const double *u_dev, const double *gpdx_dev, double proj_time_scale,
                          int shift_mdot, int shift_poisson, double *mdot_dev, void *stream);
/* Dual nodal volume: the scv volumes of every element summed to their nodes
 * (ComputeGeometryInteriorAlgorithm, src/ComputeGeometryInteriorAlgorithm.C:77-120);
 * dual_vol_dev [num_nodes] is overwritten.  stream: hipStream_t or NULL. */
int nalu_compute_dual_volume(const nalu_mesh *m, double *dual_vol_dev, void *stream);
/* Projected nodal gradient of a nodal scalar q: dqdx_dev [num_nodes][3] is
 * overwritten with the interior Green-Gauss part (AssembleNodalGradElemAlgorithm,
 * src/AssembleNodalGradElemAlgorithm.C:85-140: sum_scs (+-) q_ip A / V_node)
 * plus the exposed-face closure (AssembleNodalGradBoundaryAlgorithm,
 * src/AssembleNodalGradBoundaryAlgorithm.C:54-152) over every exposed side of
 * the block; shifted uses the shifted shape functions in both. */
int nalu_compute_nodal_gradient(const nalu_mesh *m, const double *q_dev, const double *dual_vol_dev,
                                int shifted, double *dqdx_dev, void *stream);
/* The interior part alone (overwrites dqdx_dev). */
int nalu_compute_nodal_gradient_interior(const nalu_mesh *m, const double *q_dev, const double *dual_vol_dev,
                                         int shifted, double *dqdx_dev, void *stream);
/* The exposed-face closure alone (adds into dqdx_dev). */
int nalu_nodal_gradient_boundary(const nalu_mesh *m, const double *q_dev, const double *dual_vol_dev,
                                 int shifted, double *dqdx_dev, void *stream);
/* Exposed sides of the block (sides no other element shares), nodes in HexSCS
 * side order (src/master_element/Hex8CVFEM.C:415-422, outward right-hand
 * normal): *num_faces, and faces_host [num_faces][4] when not NULL. */
int nalu_mesh_exposed_faces(const nalu_mesh *m, int64_t *num_faces, int32_t *faces_host);
/* ContinuityAdvElemKernel (src/kernel/ContinuityAdvElemKernel.C:88-148).
 * Assembly kernels evaluate with FP32 arithmetic (FP64 Jacobians, FP64
 * accumulation into the matrix) when the system was created with precision
 * NALU_PRECISION_MIXED, in FP64 otherwise. */
int nalu_assemble_continuity(nalu_linsys *ls, const nalu_mesh *m, const double *rho_dev,
                             const double *p_dev, const double *u_dev, const double *gpdx_dev,
                             double proj_time_scale, int shift_mdot, int shift_poisson,
                             int reduced_sens, int mode);
/* ContinuityAdvElemKernel and ComputeMdotElemAlgorithm
 * (src/ComputeMdotElemAlgorithm.C:85-250) from one evaluation of the scs
 * geometry, as the consolidated algorithm's kernels share their element data:
 * the continuity system as nalu_assemble_continuity (row gather) and mdot_dev
 * [E][12] as nalu_compute_mdot (same formula; sums in another order). */
int nalu_assemble_continuity_mdot(nalu_linsys *ls, const nalu_mesh *m, const double *rho_dev,
                                  const double *p_dev, const double *u_dev, const double *gpdx_dev,
                                  double proj_time_scale, int shift_mdot, int shift_poisson,
                                  int reduced_sens, double *mdot_dev);
/* MomentumAdvDiffElemKernel (src/kernel/MomentumAdvDiffElemKernel.C:72-186), numDof 3. */
int nalu_assemble_momentum(nalu_linsys *ls, const nalu_mesh *m, const double *u_dev,
                           const double *mu_dev, const double *mdot_dev, double include_divu,
                           int skew_symmetric, int shifted_grad_op, int mode);
/* Time-derivative (mass) terms added into the system (BDF: gamma = (g1, g2, g3),
 * e.g. (1, -1, 0) for backward Euler; the reference's TimeIntegrator gammas).
 * Fields are nodal device arrays of the three states Nm1, N, Np1.
 * MomentumMassElemKernel (src/kernel/MomentumMassElemKernel.C:27-170), numDof 3:
 * lumped selects the shifted (nodal) scv shape functions; balanced_force scales
 * the Gjp term by rhoNp1 (SolutionOptions balancedForce_). */
int nalu_assemble_momentum_mass(nalu_linsys *ls, const nalu_mesh *m, const double *rho_nm1,
                                const double *rho_n, const double *rho_np1, const double *u_nm1,
                                const double *u_n, const double *u_np1, const double *gjp, double dt,
                                const double *gamma, int lumped, int balanced_force);
/* MomentumAdvDiffElemKernel and MomentumMassElemKernel in one pass over the
 * elements, as AssembleElemSolverAlgorithm::execute runs its kernel list
 * (src/AssembleElemSolverAlgorithm.C:75-100 with both kernels registered):
 * the arguments of nalu_assemble_momentum (node-tile scatter) and of
 * nalu_assemble_momentum_mass; one LDS row segment per tile, one write of the
 * matrix (the mass part is FP64 arithmetic on FP32-assembly systems too, as in
 * nalu_assemble_momentum_mass).  With skew or shifted set, or rows too long for
 * a node tile, the two kernels run as two passes. */
int nalu_assemble_momentum_with_mass(nalu_linsys *ls, const nalu_mesh *m, const double *u, const double *mu,
                                     const double *mdot, double include_divu, int skew, int shifted,
                                     const double *rho_nm1, const double *rho_n, const double *rho_np1,
                                     const double *u_nm1, const double *u_n, const double *u_np1,
                                     const double *gjp, double dt, const double *gamma, int lumped,
                                     int balanced_force);
/* Momentum as three right-hand sides sharing one scalar matrix (the segregated
 * form of MomentumAdvDiffElemKernel, src/kernel/MomentumAdvDiffElemKernel.C:72-186):
 * ls is a numDof 1 system on the mesh; the matrix gets the advection-diffusion
 * operator a_c I of every 3x3 block (the transpose-gradient stress coupling is
 * left out of the lhs), rhs3 [3][nrows] (row order) is added the full residual of
 * each velocity component -- the same rhs as the coupled numDof 3 system.  The
 * caller zeroes rhs3; the matrix follows zeroSystem like the other kernels. */
int nalu_assemble_momentum_segregated(nalu_linsys *ls, const nalu_mesh *m, const double *u_dev,
                                      const double *mu_dev, const double *mdot_dev, double include_divu,
                                      int skew_symmetric, int shifted_grad_op, double *rhs3_dev);
/* MomentumMassElemKernel into the segregated system (scalar lhs, rhs3 [3][nrows]). */
int nalu_assemble_momentum_mass_segregated(nalu_linsys *ls, const nalu_mesh *m, const double *rho_nm1,
                                           const double *rho_n, const double *rho_np1, const double *u_nm1,
                                           const double *u_n, const double *u_np1, const double *gjp,
                                           double dt, const double *gamma, int lumped, int balanced_force,
                                           double *rhs3_dev);
/* The two segregated passes above in one pass over each row's elements (the
 * advection-diffusion and mass kernels of one element algorithm): the
 * arguments of nalu_assemble_momentum_segregated and of
 * nalu_assemble_momentum_mass_segregated (the mass part is FP64 arithmetic on
 * FP32-assembly systems too).  With skew or shifted set, the two kernels run as
 * two passes. */
int nalu_assemble_momentum_with_mass_segregated(nalu_linsys *ls, const nalu_mesh *m, const double *u,
                                                const double *mu, const double *mdot, double include_divu,
                                                int skew, int shifted, const double *rho_nm1, const double *rho_n,
                                                const double *rho_np1, const double *u_nm1, const double *u_n,
                                                const double *u_np1, const double *gjp, double dt,
                                                const double *gamma, int lumped, int balanced_force,
                                                double *rhs3_dev);
/* ContinuityMassElemKernel (src/kernel/ContinuityMassElemKernel.C:27-118), numDof 1:
 * rhs only, scaled by 1 / (dt / g1) like the continuity system. */
int nalu_assemble_continuity_mass(nalu_linsys *ls, const nalu_mesh *m, const double *rho_nm1,
                                  const double *rho_n, const double *rho_np1, double dt,
                                  const double *gamma, int lumped);

/* ----------------------------------------- consolidated element algorithm */
/* AssembleElemSolverAlgorithm with its kernel list
 * (include/AssembleElemSolverAlgorithm.h:52-105, src/AssembleElemSolverAlgorithm.C:75-100;
 * Kernel::setup / execute, include/kernel/Kernel.h:95-125).  Kernel kinds and
 * the device fields each one reads (field[i], nodal arrays; mdot [E][12]): */
#define NALU_KERNEL_SCALAR_DIFF 0      /* ScalarDiffElemKernel: q, diffusion coefficient */
#define NALU_KERNEL_CONTINUITY_ADV 1   /* ContinuityAdvElemKernel: rho, p, u, Gpdx */
#define NALU_KERNEL_MOMENTUM_ADVDIFF 2 /* MomentumAdvDiffElemKernel: u, mu, mdot */
#define NALU_KERNEL_CONTINUITY_MASS 3  /* ContinuityMassElemKernel: rho Nm1, N, Np1 */
#define NALU_KERNEL_MOMENTUM_MASS 4    /* MomentumMassElemKernel: rho Nm1, N, Np1, u Nm1, N, Np1, Gjp */
typedef struct nalu_kernel_desc {
  int kind;             /* NALU_KERNEL_* */
  int shifted_grad_op;  /* SolutionOptions shiftedGradOp (diffusion / Poisson / viscous gradients) */
  int shift_mdot;       /* continuity: shifted mdot shape functions */
  int reduced_sens;     /* continuity: reduced-sensitivity (shifted) lhs */
  int skew_symmetric;   /* momentum: skew-symmetric advection shape functions */
  int lumped;           /* mass kernels: lumped mass */
  int balanced_force;   /* momentum mass: balanced-force pressure-gradient density */
  double include_divu;  /* momentum: divU scaling (SolutionOptions includeDivU, default 0) */
  const double *field[8];
} nalu_kernel_desc;
typedef struct nalu_elem_algorithm nalu_elem_algorithm;
/* An algorithm assembling into ls over the elements of m (ls finalized on m). */
int nalu_elem_algorithm_create(nalu_linsys *ls, const nalu_mesh *m, nalu_elem_algorithm **out);
/* The algorithm owns its kernels (src/Algorithm.C:46-55). */
int nalu_elem_algorithm_destroy(nalu_elem_algorithm *alg);
/* activeKernels_.push_back: kernels run in the order they were added. */
int nalu_elem_algorithm_add_kernel(nalu_elem_algorithm *alg, const nalu_kernel_desc *kernel);
/* The TimeIntegrator every kernel's setup() reads: dt and the BDF gammas
 * (default dt = 1, gamma = 1, -1, 0). */
int nalu_elem_algorithm_set_time_integrator(nalu_elem_algorithm *alg, double dt, const double *gamma3);
/* execute(): setup of every kernel (ContinuityAdvElemKernel: projTimeScale =
 * dt / gamma1), then the kernels' element contributions assembled into the
 * system, each kernel as one fused gather + geometry + physics + scatter pass
 * (element-visit tiles for the advection / diffusion kernels). */
int nalu_elem_algorithm_execute(nalu_elem_algorithm *alg);

/* LinearSolverConfig "write_matrix_files" (src/LinearSolverConfig.C:102): when
 * base_name is not NULL, every solve writes <base>-O-<k>.mm.1 (MatrixMarket
 * coordinate, 1-based Tpetra GIDs), <base>-O-<k>.rhs.1 and, after the solve,
 * <base>-O-<k>.sln.1 (dense arrays), k counting solves
 * (TpetraLinearSystem::writeToFile / writeSolutionToFile,
 * src/TpetraLinearSystem.C:1794-1817,1969-2090).  NULL turns it off. */
int nalu_linsys_set_write_matrix_files(nalu_linsys *ls, const char *base_name);
/* TpetraLinearSystem::writeToFile / writeSolutionToFile (src/TpetraLinearSystem.C:1969-1994,
 * 2070-2090), written now under the same names as the solve-time output:
 * <base_name>-O-<counter>.mm.1 / .rhs.1 and <base_name>-O-<counter>.sln.1. */
int nalu_linsys_write_to_file(nalu_linsys *ls, const char *base_name);
int nalu_linsys_write_solution_to_file(nalu_linsys *ls, const char *base_name);

/* ------------------------------------------------- learned preconditioner
 * No counterpart in the reference: a message-passing network over the matrix
 * graph (BF16 weights, FP32 accumulation, hidden width 32) used inside FGMRES
 * when the config's precond is NALU_PREC_LEARNED, with multicolor SGS as the
 * automatic fallback (no network attached, a stagnating cycle, or a
 * non-finite output).  Scalar (numDof 1) systems only. */
typedef struct nalu_gnn nalu_gnn;
typedef struct nalu_gnn_data nalu_gnn_data;
/* Seeded network whose correction starts at zero (its output is then SGS(r)). */
int nalu_gnn_create(int num_layers, int hidden, uint64_t seed, nalu_gnn **out);
int nalu_gnn_destroy(nalu_gnn *g);
int64_t nalu_gnn_num_params(const nalu_gnn *g);
int nalu_gnn_num_layers(const nalu_gnn *g);
/* FP32 master parameters (the kernels read their BF16 rounding). */
int nalu_gnn_get_params(const nalu_gnn *g, float *params_host, int64_t n);
int nalu_gnn_set_params(nalu_gnn *g, const float *params_host, int64_t n);
int nalu_gnn_get_grad(const nalu_gnn *g, float *grad_host, int64_t n);
/* Attach (g != NULL) or detach the network used by a system's solve. */
int nalu_linsys_set_learned_precond(nalu_linsys *ls, nalu_gnn *g);
/* No counterpart in the reference.  The network's base preconditioner:
 * NALU_PREC_SGS (default) or NALU_PREC_MUELU (the AMG V-cycle, built at the
 * preconditioner set-up with the system's MueLu parameters).  The base gives
 * the network's second input feature and its initial output (the untrained
 * network returns the base's output) and is the fallback; a network trained
 * with one base is meant for that base.  Takes effect at the next set-up; set
 * it before harvesting pairs too (their stored feature is the base's output). */
int nalu_linsys_set_learned_base(nalu_linsys *ls, int precond);
/* 1 when the last solve fell back from the network to its base preconditioner. */
int nalu_linsys_learned_fallback(const nalu_linsys *ls);
/* z = net(r) for one right-hand side, FP32 device arrays in row order; needs a
 * set-up preconditioner (nalu_linsys_precond_setup or a solve). */
int nalu_gnn_apply(nalu_gnn *g, nalu_linsys *ls, const float *r_dev, float *z_dev);
/* Loss L = mean_b [1 - <out,r>^2 / (<out,A out> <e,r>)] and its parameter
 * gradient for `batch` pairs r_dev/e_dev [batch][rows] (FP32, row order);
 * optional outputs: the SGS feature zs_dev and the raw output out_dev. */
int nalu_gnn_loss_grad(nalu_gnn *g, nalu_linsys *ls, int batch, const float *r_dev, const float *e_dev,
                       float *zs_dev, float *out_dev, double *loss_host, float *grad_host);
/* Harvest num_rhs * pairs_per_rhs (residual, error) pairs from plain FP64 CG +
 * SGS solves (to relative residual tol) of the system's assembled matrix with
 * seeded unit-variance Gaussian zero-mean right-hand sides (pure-Neumann systems:
 * each error is stored with its mean removed). */
int nalu_gnn_harvest(nalu_linsys *ls, int num_rhs, int pairs_per_rhs, uint64_t seed, double tol,
                     int max_iterations, nalu_gnn_data **out);
int nalu_gnn_data_destroy(nalu_gnn_data *d);
int nalu_gnn_data_size(const nalu_gnn_data *d);
int nalu_gnn_data_cg_iterations(const nalu_gnn_data *d, int *iters_host, int count);
int nalu_gnn_data_pair(const nalu_gnn_data *d, int idx, float *r_dev, float *e_dev);
/* No counterpart in the reference (training data of the learned preconditioner).
 * Overwrites pair idx with the caller's (r, e), FP64 device vectors in row
 * order with A e = r, stored like a harvested pair (scaled by 1/|r|, e off its
 * mean, the base preconditioner's output of r as the feature); ls is the system
 * the pairs were harvested from. */
int nalu_gnn_data_set_pair(nalu_gnn_data *d, nalu_linsys *ls, int idx, const double *r_dev, const double *e_dev);
/* Adam (cosine learning-rate decay to lr/10) on the pairs harvested from the
 * same system; epoch_loss_host[epochs] receives the mean loss per epoch. */
int nalu_gnn_train(nalu_gnn *g, nalu_linsys *ls, const nalu_gnn_data *d, int epochs, int batch, double lr,
                   uint64_t seed, double *epoch_loss_host);
/* Epochs [first_epoch, first_epoch + num_epochs) of a run of total_epochs (the
 * schedule and pair order of the whole run), so that a caller can validate the
 * network between epochs; epoch_loss_host[num_epochs]. */
int nalu_gnn_train_epochs(nalu_gnn *g, nalu_linsys *ls, const nalu_gnn_data *d, int first_epoch,
                          int num_epochs, int total_epochs, int batch, double lr, uint64_t seed,
                          double *epoch_loss_host);

/* ------------------------------------------- AMG (MueLu-equivalent) */
/* Parameters of the AMG preconditioner (NALU_PREC_MUELU) of this system: the
 * muelu_xml_file_name parameter list of LinearSolverConfig
 * (src/LinearSolverConfig.C:86-100) given as a struct.  Takes effect at the
 * next preconditioner set-up. */
int nalu_linsys_set_muelu_params(nalu_linsys *ls, const nalu_muelu_params *p);
/* Hierarchy summary after set-up: num_levels; per level (up to max_levels
 * entries) rows, nonzeros and the lambda_max(D^-1 A) estimate; set-up seconds;
 * skipped pivots of the dense coarse solve (1 for a pure Neumann operator);
 * |A_c B_c| / (scale |B_c|) of the coarsest operator along its near-null
 * vector (<= 1e-5: treated as an exact null vector).  Any output pointer may
 * be NULL.  No counterpart in the reference (MueLu's verbose hierarchy
 * printout). */
int nalu_amg_info(const nalu_linsys *ls, int *num_levels, int64_t *rows, int64_t *nnz, double *lmax,
                  int max_levels, double *setup_s, int *coarse_free, double *coarse_null_rel);
/* Host copy of level `level`'s operator (which = 0) or prolongator (which = 1)
 * as CSR in the solver's row order (level 0 = the color permutation of
 * nalu_linsys_permutation); rp [n+1], ci / v [nnz]; n, m, nnz are returned
 * (call with NULL arrays first to size them).  Test and inspection aid, no
 * counterpart in the reference. */
int nalu_amg_level_csr(const nalu_linsys *ls, int level, int which, int32_t *rp, int32_t *ci, double *v,
                       int64_t *n, int64_t *m, int64_t *nnz);
/* Aggregate id of every row of level `level` (not the coarsest). */
int nalu_amg_aggregates(const nalu_linsys *ls, int level, int32_t *agg);

#ifdef __cplusplus
}
#endif
#endif
