"""Low-Mach time step driver on the HIP kernels (host orchestration only).

Mirrors LowMachEquationSystem::solve_and_update (src/LowMachEquationSystem.C:779-870)
for one backward-Euler step (gammas 1, -1, 0):

  for each Picard iteration:
    momentum:   mass + advection-diffusion with the current mdot, Dirichlet
                velocity on the wall nodes, solve for du, u += du -- either as
                the reference solves it, one numDof-3 system with the full
                stress coupling (MomentumEquationSystem, :1029), or in the
                three-right-hand-side form (one scalar matrix, 3 rhs)
    continuity: advection (projTimeScale = dt / g1) + density time derivative,
                solve for dp (pure Neumann: zero-mean rhs), p += dp
    mdot:       recomputed from the new u and p with the previous Gjp
    projection: Gjp_new = nodal gradient of p;
                u -= (dt / g1) / rho (Gjp_new - Gjp_old) on the non-wall nodes
                (LowMachEquationSystem::project_nodal_velocity, :884-975)

The projected nodal gradient includes the exposed-face closure
(AssembleNodalGradBoundaryAlgorithm), so Gjp is consistent at wall nodes too;
wall nodes keep their Dirichlet velocities.  Meshes are used without explicit
node ids (row = node).
"""
import numpy as np

from . import (LinearSystem, NaluError, SolverConfig, assemble_continuity, assemble_continuity_mass,
               assemble_momentum_with_mass, assemble_momentum_with_mass_segregated, compute_dual_volume, compute_mdot,
               compute_nodal_gradient)


class LowMachStep:
    """State (torch device tensors): u [N][3], p [N], rho [N] (constant in time),
    mu [N], Gjp [N][3]; wall_nodes int32 device tensor with wall_u [len][3].
    pressure_precond: "sgs", or "muelu" for the AMG preconditioner the reference
    configures for the continuity system (CG + MueLu in its regression inputs,
    e.g. reg_tests/test_files/elemClosedDomain/elemClosedDomain.i:15-25); with
    "muelu" the FP64 pressure solve is CG, the mixed one FGMRES.
    momentum_solve: "coupled" (the reference's form: one numDof-3 system whose
    lhs carries the transpose-gradient stress coupling), "stacked" (the three
    components of the segregated form in one Krylov space, converged on their
    stacked residual norm) or "separate" (three segregated solves, each to the
    tolerance).  The segregated lhs lags the coupling, so its Picard iterates
    differ, but the converged velocities are the same."""

    def __init__(self, mesh, u, p, rho, mu, dt, wall_nodes, wall_u, picard=2, tol=1e-8, precision="fp64",
                 stream=None, pressure_precond="sgs", muelu_params=None, momentum_solve="stacked"):
        import torch
        self.torch = torch
        if getattr(mesh, "has_ids", False):
            # rhs3 / du3 are in row order, which equals node order only without explicit ids
            raise NaluError("LowMachStep: meshes with explicit node ids are not supported (rows would not be nodes)")
        self.mesh, self.dt, self.picard, self.stream = mesh, float(dt), int(picard), stream
        if momentum_solve not in ("coupled", "stacked", "separate"):
            raise NaluError("LowMachStep: momentum_solve is 'coupled', 'stacked' or 'separate'")
        self.momentum_solve = momentum_solve
        self.g = (1.0, -1.0, 0.0)
        self.u, self.p, self.rho, self.mu = u.clone(), p.clone(), rho, mu
        self.wall = wall_nodes.long()
        self.wall_i32 = wall_nodes
        self.wall_u = wall_u
        self.u[self.wall] = wall_u
        n_nodes = u.shape[0]
        dev = u.device
        self.vol = torch.zeros(n_nodes, dtype=torch.float64, device=dev)
        compute_dual_volume(mesh, self.vol, stream)
        self.Gjp = torch.zeros((n_nodes, 3), dtype=torch.float64, device=dev)
        compute_nodal_gradient(mesh, self.p, self.vol, self.Gjp, stream=stream)
        self.mdot = torch.zeros((mesh.num_elems, 12), dtype=torch.float64, device=dev)
        self.interior = torch.ones(n_nodes, dtype=torch.bool, device=dev)
        self.interior[self.wall] = False
        method = "fgmres" if precision == "mixed" else "gmres"
        self.mom = LinearSystem(3 if momentum_solve == "coupled" else 1,
                                SolverConfig(method, "sgs", tolerance=tol, max_iterations=2000, kspace=50,
                                             precision=precision))
        cont_method = "cg" if (pressure_precond == "muelu" and precision == "fp64") else method
        self.cont = LinearSystem(1, SolverConfig(cont_method, pressure_precond, tolerance=tol, max_iterations=5000,
                                                 kspace=50, precision=precision))
        if pressure_precond == "muelu" and muelu_params is not None:
            self.cont.set_muelu_params(muelu_params)
        for ls in (self.mom, self.cont):
            ls.buildElemToNodeGraph(mesh)
            ls.finalizeLinearSystem()
            if stream is not None:
                ls.set_stream(stream)
        if self.mom.sizes()[0] != self.mom.num_dof * n_nodes:
            raise NaluError("LowMachStep: numDof rows per node expected")
        self.rhs3 = torch.zeros((3, n_nodes), dtype=torch.float64, device=dev)
        self.du3 = torch.zeros_like(self.rhs3)
        self.dp = torch.zeros_like(self.p)
        self.log = []

    def _mdot(self):
        compute_mdot(self.mesh, self.rho, self.p, self.u, self.Gjp, self.mdot, tau=self.dt / self.g[0],
                     stream=self.stream)

    def continuity_residual(self):
        """|rhs| of the continuity system (mass imbalance incl. the time term) at the current state."""
        self._mdot()
        self.cont.zeroSystem()
        assemble_continuity(self.cont, self.mesh, self.rho, self.p, self.u, self.Gjp, tau=self.dt / self.g[0],
                            mode="gather")
        assemble_continuity_mass(self.cont, self.mesh, (self.rho, self.rho, self.rho), self.dt, self.g)
        self.cont.loadComplete()
        return float(np.linalg.norm(self.cont.csr()[3]))

    def _momentum_segregated(self, uN):
        """Three right-hand sides sharing one scalar matrix."""
        self.mom.zeroSystem()
        self.rhs3.zero_()
        # advection-diffusion and mass in one pass over each row's elements
        assemble_momentum_with_mass_segregated(self.mom, self.mesh, self.u, self.mu, self.mdot,
                                               (self.rho, self.rho, self.rho), (uN, uN, self.u), self.Gjp, self.dt,
                                               self.rhs3, gamma=self.g)
        # Dirichlet rows (the field arrays are nodal: component 0 of the velocity
        # and its boundary value; rhs3 then takes all three components)
        u0 = self.u[:, 0].contiguous()
        bc0 = u0.clone()
        bc0[self.wall] = self.wall_u[:, 0]
        self.mom.applyDirichletBCs(self.wall_i32, u0, bc0)
        self.rhs3[:, self.wall] = (self.wall_u - self.u[self.wall]).T
        self.mom.loadComplete()
        self.du3.zero_()
        if self.momentum_solve == "stacked":
            st_m, its_m, rel_m = self.mom.solve_stacked(self.rhs3, self.du3)
        else:
            st_m, its_m, rel_m = self.mom.solve_multi(self.rhs3, self.du3)
        self.u += self.du3.T
        self._du_max = float(self.du3.abs().max())
        return st_m, its_m, rel_m

    def _momentum_coupled(self, uN):
        """The numDof-3 system (rows interleaved node-major), Dirichlet on all
        three components of the wall nodes, delta = du."""
        self.mom.zeroSystem()
        # advection-diffusion and mass in one element pass (the momentum kernel list)
        assemble_momentum_with_mass(self.mom, self.mesh, self.u, self.mu, self.mdot, (self.rho, self.rho, self.rho),
                                    (uN, uN, self.u), self.Gjp, self.dt, gamma=self.g)
        u_flat = self.u.reshape(-1).contiguous()
        bc = self.u.clone()
        bc[self.wall] = self.wall_u
        self.mom.applyDirichletBCs(self.wall_i32, u_flat, bc.reshape(-1).contiguous())
        self.mom.loadComplete()
        du = self.torch.zeros_like(u_flat)
        st_m = self.mom.solve(du)
        self.u += du.reshape(-1, 3)
        self._du_max = float(du.abs().max())
        rel = self.mom.linearResidual() / max(self.mom.nonLinearResidual(), 1e-300)
        return st_m, [self.mom.linearSolveIterations()], [rel]

    def step(self):
        """One time step; returns the per-iteration record."""
        uN = self.u.clone()
        rec = []
        self._mdot()
        for k in range(self.picard):
            if self.momentum_solve == "coupled":
                st_m, its_m, rel_m = self._momentum_coupled(uN)
            else:
                st_m, its_m, rel_m = self._momentum_segregated(uN)
            # continuity
            self.cont.zeroSystem()
            assemble_continuity(self.cont, self.mesh, self.rho, self.p, self.u, self.Gjp, tau=self.dt / self.g[0],
                                mode="gather")
            assemble_continuity_mass(self.cont, self.mesh, (self.rho, self.rho, self.rho), self.dt, self.g)
            self.cont.loadComplete()
            self.cont.removeRhsMean()
            self.dp.zero_()
            st_c = self.cont.solve(self.dp)
            its_c = self.cont.linearSolveIterations()
            self.p += self.dp
            self.p -= self.p.mean()
            # mdot with the new velocity and pressure, then the nodal-velocity projection
            self._mdot()
            g_old = self.Gjp.clone()
            compute_nodal_gradient(self.mesh, self.p, self.vol, self.Gjp, stream=self.stream)
            corr = (self.dt / self.g[0]) / self.rho[:, None] * (self.Gjp - g_old)
            self.u[self.interior] -= corr[self.interior]
            rec.append({"momentum_status": st_m, "momentum_iters": its_m, "momentum_rel": rel_m,
                        "continuity_status": st_c, "continuity_iters": its_c,
                        "du_max": self._du_max, "dp_max": float(self.dp.abs().max())})
        self.log.append(rec)
        return rec
