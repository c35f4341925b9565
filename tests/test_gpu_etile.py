"""GPU parity of the element-visit tile assembly (mode "etile") against the oracle.

Every option branch of the three element kernels is compared with the whole-mesh
oracle: continuity with all eight (shifted mdot, shifted Poisson, reduced
sensitivities) combinations, scalar diffusion shifted / not, momentum with skew
advection, shifted gradients and the divU term.  Meshes run without and with
permuted global ids: ids scatter the rows, so a tile then owns arbitrary subsets
of an element's nodes and every visit class (node, edge, face, any) and every
rotation of the reference cube is exercised.  Lattice sizes whose lines are
shorter and longer than a tile cover tiles that hold whole lines and tiles that
end inside a line.

Bars: FP64 matrix and rhs within 1e-12 of the oracle relative to each row's
largest entry (rhs: to the largest rhs entry); FP32 arithmetic within 3e-5
(the mixed-precision bar of tests/test_gpu_mixed.py); the graph bit-exact.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nalu_amd as na  # noqa: E402
from nalu_amd.problems import heat_problem, lowmach_problem  # noqa: E402
from tests.fixtures import rel_err_rowwise  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = 1e-12
TOL32 = 3e-5


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def system(P, nd, ids, precision="fp64"):
    mesh = na.Mesh(P["coords"], P["conn"], ids)
    ls = na.LinearSystem(nd, na.SolverConfig("gmres", "sgs", precision=precision))
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    return mesh, ls


def perm_ids(nn, seed):
    return np.random.default_rng(seed).permutation(nn).astype(np.int64) + 1


def check(ls, g, ov, orr, tol):
    rp, ci, v, r = ls.csr()
    np.testing.assert_array_equal(rp, g.row_ptr)
    np.testing.assert_array_equal(ci, g.col_idx)
    em = rel_err_rowwise(v, ov, g.row_ptr)
    er = float(np.abs(r - orr).max() / max(np.abs(orr).max(), 1e-300))
    print("  matrix %.2e rhs %.2e" % (em, er), flush=True)
    assert em <= tol and er <= tol
    return em


@pytest.mark.parametrize("key", range(8))
@pytest.mark.parametrize("n,ids", [(7, False), (5, True), (70, False)])
def test_continuity_options(oracle, key, n, ids):
    sm, sp, rs = bool(key & 4), bool(key & 2), bool(key & 1)
    print("case: etile continuity n=%d ids=%s shift_mdot=%s shift_poisson=%s reduced=%s" % (n, ids, sm, sp, rs),
          flush=True)
    if n > 20 and key not in (0, 7):
        pytest.skip("large mesh: default and all-options cases only")
    P = lowmach_problem(n, seed=40 + key, rho_sigma=0.3)
    nn = P["coords"].shape[0]
    I = perm_ids(nn, key) if ids else None
    mesh, ls = system(P, 1, I)
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=P["tau"],
                           shift_mdot=sm, shift_poisson=sp, reduced_sens=rs, mode="etile")
    g = oracle.graph_build(nn, P["conn"], 1, node_ids=I)
    ov, orr = oracle.assemble_continuity(g, P["conn"], P["coords"], P["rho"], P["p"], P["u"], P["gpdx"], P["tau"],
                                         shift_mdot=sm, shift_poisson=sp, reduced_sens=rs)
    check(ls, g, ov, orr, TOL)


@pytest.mark.parametrize("shifted", [False, True])
@pytest.mark.parametrize("n,ids", [(6, False), (4, True), (65, False)])
def test_scalar_diff(oracle, shifted, n, ids):
    print("case: etile scalar diffusion n=%d ids=%s shifted=%s" % (n, ids, shifted), flush=True)
    P = heat_problem(n, seed=5)
    nn = P["coords"].shape[0]
    I = perm_ids(nn, 3) if ids else None
    q = np.sin(3 * P["coords"][:, 0]) + P["coords"][:, 1] * P["coords"][:, 2]
    mesh, ls = system(P, 1, I)
    ls.zeroSystem()
    na.assemble_scalar_diff(ls, mesh, dev(q), dev(P["k"]), shifted=shifted, mode="etile")
    g = oracle.graph_build(nn, P["conn"], 1, node_ids=I)
    ov, orr = oracle.assemble_scalar_diff(g, P["conn"], P["coords"], q, P["k"], shifted=shifted)
    check(ls, g, ov, orr, TOL)


@pytest.mark.parametrize("skew,shifted,divu", [(False, False, 0.0), (True, False, 0.0), (False, True, 0.0),
                                               (True, True, 1.0), (False, False, 1.0)])
@pytest.mark.parametrize("n,ids", [(5, False), (4, True), (66, False)])
def test_momentum_options(oracle, skew, shifted, divu, n, ids):
    print("case: etile momentum n=%d ids=%s skew=%s shifted=%s divU=%g" % (n, ids, skew, shifted, divu), flush=True)
    if n > 20 and (skew or shifted or divu):
        pytest.skip("large mesh: default options only")
    P = lowmach_problem(n, seed=31, mu=0.05)
    nn = P["coords"].shape[0]
    I = perm_ids(nn, 7) if ids else None
    mesh, ls = system(P, 3, I)
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    mdot = torch.zeros((P["conn"].shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
    ls.zeroSystem()
    na.assemble_momentum(ls, mesh, u, dev(P["mu"]), mdot, include_divu=divu, skew=skew, shifted=shifted,
                         mode="etile")
    g = oracle.graph_build(nn, P["conn"], 3, node_ids=I)
    ov, orr = oracle.assemble_momentum(g, P["conn"], P["coords"], P["u"], P["mu"], mdot.cpu().numpy(),
                                       include_divu=divu, skew=skew, shifted=shifted)
    check(ls, g, ov, orr, TOL)


@pytest.mark.parametrize("what", ["continuity", "momentum", "scalar"])
def test_fp32_etile(oracle, what):
    n = 6
    print("case: etile FP32 %s n=%d, stretched" % (what, n), flush=True)
    P = lowmach_problem(n, seed=21, rho_sigma=0.3, mu=0.05, stretch=1.25, max_aspect=100.0)
    nn = P["coords"].shape[0]
    nd = 3 if what == "momentum" else 1
    mesh, ls = system(P, nd, None, precision="mixed")
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    g = oracle.graph_build(nn, P["conn"], nd)
    ls.zeroSystem()
    if what == "continuity":
        na.assemble_continuity(ls, mesh, rho, p, u, gp, tau=P["tau"], mode="etile")
        ov, orr = oracle.assemble_continuity(g, P["conn"], P["coords"], P["rho"], P["p"], P["u"], P["gpdx"],
                                             P["tau"])
    elif what == "momentum":
        mdot = torch.zeros((P["conn"].shape[0], 12), dtype=torch.float64, device=DEV)
        na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
        na.assemble_momentum(ls, mesh, u, dev(P["mu"]), mdot, mode="etile")
        ov, orr = oracle.assemble_momentum(g, P["conn"], P["coords"], P["u"], P["mu"], mdot.cpu().numpy())
    else:
        na.assemble_scalar_diff(ls, mesh, p, rho, mode="etile")
        ov, orr = oracle.assemble_scalar_diff(g, P["conn"], P["coords"], P["p"], P["rho"])
    em = check(ls, g, ov, orr, TOL32)
    assert em > 0.0  # FP32 arithmetic is in use


def test_etile_accumulates_after_other_algorithms(oracle):
    """A second assembly adds to the first (zeroSystem folds into the first one only)."""
    print("case: etile accumulation after the node-tile kernel", flush=True)
    P = lowmach_problem(6, seed=8, mu=0.05)
    nn = P["coords"].shape[0]
    mesh, ls = system(P, 3, None)
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    mdot = torch.zeros((P["conn"].shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
    ls.zeroSystem()
    na.assemble_momentum(ls, mesh, u, dev(P["mu"]), mdot, mode="tile")
    na.assemble_momentum(ls, mesh, u, dev(P["mu"]), mdot, mode="etile")
    g = oracle.graph_build(nn, P["conn"], 3)
    ov, orr = oracle.assemble_momentum(g, P["conn"], P["coords"], P["u"], P["mu"], mdot.cpu().numpy())
    check(ls, g, 2 * ov, 2 * orr, TOL)
    ls.zeroSystem()
    na.assemble_momentum(ls, mesh, u, dev(P["mu"]), mdot, mode="etile")
    check(ls, g, ov, orr, TOL)
