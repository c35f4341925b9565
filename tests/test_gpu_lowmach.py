"""GPU integration test of the low-Mach time-step driver (momentum as the coupled
numDof-3 system or in its three-right-hand-side form, continuity, mdot,
nodal-velocity projection) built
on the parity-tested kernels: every linear solve reaches its tolerance and the
projection lowers the continuity residual (the discrete mass imbalance)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nalu_amd as na  # noqa: E402
from nalu_amd.lowmach import LowMachStep  # noqa: E402
from nalu_amd.problems import lattice_ijk, lowmach_problem  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


@pytest.mark.parametrize("precision,pprec,msolve", [("fp64", "sgs", "separate"), ("mixed", "sgs", "stacked"),
                                                   ("fp64", "muelu", "stacked"), ("mixed", "muelu", "separate"),
                                                   ("fp64", "sgs", "coupled"), ("mixed", "sgs", "coupled")])
def test_lowmach_step(precision, pprec, msolve):
    n = 8
    print("case: low-Mach step n=%d precision=%s pressure %s momentum %s" % (n, precision, pprec, msolve), flush=True)
    P = lowmach_problem(n, seed=11, mu=0.01, noise=0.05)
    mesh = na.Mesh(P["coords"], P["conn"])
    i, j, k = lattice_ijk(n)
    wall = np.nonzero((i == 0) | (i == n) | (j == 0) | (j == n) | (k == 0) | (k == n))[0].astype(np.int32)
    u0 = dev(P["u"])
    wall_t = torch.as_tensor(wall, device=DEV)
    step = LowMachStep(mesh, u0, dev(P["p"]), dev(P["rho"]), dev(P["mu"]), dt=0.25 / n, wall_nodes=wall_t,
                       wall_u=u0[wall_t.long()].clone(), picard=3, precision=precision, pressure_precond=pprec,
                       muelu_params=na.MueLuParams(max_coarse_size=50, cheb_degree=2), momentum_solve=msolve)
    r0 = step.continuity_residual()
    rec = step.step()
    r1 = step.continuity_residual()
    for it in rec:
        print("  ", {k: (v if not isinstance(v, list) else v) for k, v in it.items()}, flush=True)
        assert it["momentum_status"] == 0 and it["continuity_status"] == 0
        # separate / coupled: to 1e-8 of the rhs; stacked: the stacked residual to
        # 1e-8, a component to that bar scaled by its share of the rhs
        assert max(it["momentum_rel"]) <= (1e-6 if msolve == "stacked" else 1e-8)
    print("  continuity residual %.3e -> %.3e" % (r0, r1), flush=True)
    assert np.isfinite(r1) and r1 < 0.5 * r0
    # Picard corrections shrink
    assert rec[-1]["du_max"] < rec[0]["du_max"]
    # wall velocities stay at their Dirichlet values
    np.testing.assert_allclose(step.u[wall_t.long()].cpu().numpy(), P["u"][wall], rtol=0, atol=1e-12)


def test_coupled_and_segregated_momentum_converge_to_the_same_state():
    """The reference's coupled numDof-3 momentum system and the three-rhs
    segregated form share their right-hand side (the full residual); only the
    segregated lhs lags the transpose-gradient coupling.  With mdot, p and Gjp
    frozen the momentum residual is linear in u, so the coupled Newton step
    solves it at once, while the segregated iteration converges to the same
    velocity linearly."""
    n = 6
    print("case: coupled vs stacked momentum iterations, n=%d" % n, flush=True)
    P = lowmach_problem(n, seed=12, mu=0.05, noise=0.05)
    mesh = na.Mesh(P["coords"], P["conn"])
    i, j, k = lattice_ijk(n)
    wall = np.nonzero((i == 0) | (i == n) | (j == 0) | (j == n) | (k == 0) | (k == n))[0].astype(np.int32)
    wall_t = torch.as_tensor(wall, device=DEV)
    out = {}
    for msolve, iters in (("coupled", 3), ("stacked", 30)):
        u0 = dev(P["u"])
        step = LowMachStep(mesh, u0, dev(P["p"]), dev(P["rho"]), dev(P["mu"]), dt=0.25 / n, wall_nodes=wall_t,
                           wall_u=u0[wall_t.long()].clone(), tol=1e-13, momentum_solve=msolve)
        step._mdot()
        uN = step.u.clone()
        hist = []
        for _ in range(iters):  # momentum iterations alone (the first half of a Picard iteration)
            st, _, _ = step._momentum_coupled(uN) if msolve == "coupled" else step._momentum_segregated(uN)
            assert st == 0
            hist.append(step._du_max)
        print("  %s: du_max %s" % (msolve, ["%.1e" % h for h in hist]), flush=True)
        assert hist[-1] <= 1e-11 * hist[0]
        out[msolve] = step.u.cpu().numpy()
    d = np.abs(out["coupled"] - out["stacked"]).max() / np.abs(out["stacked"]).max()
    print("  max rel velocity difference %.2e" % d, flush=True)
    assert d <= 1e-10
