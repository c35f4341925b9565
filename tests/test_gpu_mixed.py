"""GPU tests of the mixed-precision path: FP32-arithmetic assembly against the
FP64 oracle, on uniform and on highly stretched meshes, and mixed-precision
solves against FP64 solves.

Bars: FP32 assembly (FP64 Jacobians and determinants, coordinates relative to
each element's first node, FP64 accumulation) agrees with the FP64 oracle within
3e-5 of each row's largest entry (matrix) and of the largest entry (rhs, mdot) --
a few hundred FP32 operations per entry at unit roundoff 6e-8 -- including
elements with aspect ratio up to 100; the graph is unchanged (bit-exact);
mixed-precision solves reach the FP64 true residual 1e-8.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nalu_amd as na  # noqa: E402
from nalu_amd.problems import heat_problem, lowmach_problem, poisson_rhs  # noqa: E402
from tests.fixtures import rel_err_rowwise  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL32 = 3e-5


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def mixed_system(coords, conn, nd=1):
    mesh = na.Mesh(coords, conn)
    ls = na.LinearSystem(nd, na.SolverConfig("gmres", "sgs", precision="mixed"))
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    return mesh, ls


@pytest.mark.parametrize("stretch", [1.0, 1.25])
@pytest.mark.parametrize("mode", ["gather", "tile", "etile"])
def test_fp32_continuity_and_mdot(oracle, mode, stretch):
    n = 8
    print("case: fp32 continuity + mdot n=%d mode=%s stretch=%g" % (n, mode, stretch), flush=True)
    P = lowmach_problem(n, seed=21, rho_sigma=0.3, stretch=stretch, max_aspect=100.0 if stretch > 1 else 0.0)
    mesh, ls = mixed_system(P["coords"], P["conn"])
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, rho, p, u, gp, tau=P["tau"], mode=mode)
    rp, ci, v, r = ls.csr()
    g = oracle.graph_build(P["coords"].shape[0], P["conn"])
    ov, orr = oracle.assemble_continuity(g, P["conn"], P["coords"], P["rho"], P["p"], P["u"], P["gpdx"], P["tau"])
    np.testing.assert_array_equal(ci, g.col_idx)
    e_mat = rel_err_rowwise(v, ov, g.row_ptr)
    e_rhs = np.abs(r - orr).max() / np.abs(orr).max()
    mdot = torch.zeros((P["conn"].shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"], precision="fp32")
    om = oracle.compute_mdot(P["conn"], P["coords"], P["rho"], P["p"], P["u"], P["gpdx"], P["tau"])
    e_md = np.abs(mdot.cpu().numpy() - om).max() / np.abs(om).max()
    print("  rel err: matrix %.2e rhs %.2e mdot %.2e" % (e_mat, e_rhs, e_md), flush=True)
    assert e_mat <= TOL32 and e_rhs <= TOL32 and e_md <= TOL32
    # FP32 arithmetic really is in use: the entries differ from the FP64 path
    assert e_mat > 0.0


@pytest.mark.parametrize("mode", ["atomic", "tile", "etile"])
def test_fp32_momentum(oracle, mode):
    n = 6
    print("case: fp32 momentum n=%d %s" % (n, mode), flush=True)
    P = lowmach_problem(n, seed=31, mu=0.05)
    mesh, ls = mixed_system(P["coords"], P["conn"], nd=3)
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    mdot = torch.zeros((P["conn"].shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
    ls.zeroSystem()
    na.assemble_momentum(ls, mesh, u, dev(P["mu"]), mdot, mode=mode)
    rp, ci, v, r = ls.csr()
    g = oracle.graph_build(P["coords"].shape[0], P["conn"], 3)
    np.testing.assert_array_equal(ci, g.col_idx)
    ov, orr = oracle.assemble_momentum(g, P["conn"], P["coords"], P["u"], P["mu"], mdot.cpu().numpy())
    e_mat = rel_err_rowwise(v, ov, g.row_ptr)
    e_rhs = np.abs(r - orr).max() / np.abs(orr).max()
    print("  rel err: matrix %.2e rhs %.2e" % (e_mat, e_rhs), flush=True)
    assert e_mat <= TOL32 and e_rhs <= TOL32


def test_fp32_heat_assembly(oracle):
    print("case: fp32 heat n=8", flush=True)
    P = heat_problem(8, seed=3)
    mesh, ls = mixed_system(P["coords"], P["conn"])
    ls.zeroSystem()
    q = np.sin(3 * P["coords"][:, 0]) + P["coords"][:, 1]
    na.assemble_scalar_diff(ls, mesh, dev(q), dev(P["k"]), mode="tile")
    rp, ci, v, r = ls.csr()
    g = oracle.graph_build(P["coords"].shape[0], P["conn"])
    ov, orr = oracle.assemble_scalar_diff(g, P["conn"], P["coords"], q, P["k"])
    e_mat = rel_err_rowwise(v, ov, g.row_ptr)
    e_rhs = np.abs(r - orr).max() / np.abs(orr).max()
    print("  rel err: matrix %.2e rhs %.2e" % (e_mat, e_rhs), flush=True)
    assert e_mat <= TOL32 and e_rhs <= TOL32


def test_mixed_pressure_solve_matches_fp64(oracle):
    """FP32 assembly + mixed FGMRES against FP64 assembly + FP64 GMRES: both reach
    1e-8 on their own systems; the pressures agree to the FP32 assembly error."""
    n = 12
    print("case: mixed vs fp64 pressure n=%d" % n, flush=True)
    P = lowmach_problem(n, seed=5)
    sols = {}
    for prec in ("fp64", "mixed"):
        mesh = na.Mesh(P["coords"], P["conn"])
        ls = na.LinearSystem(1, na.SolverConfig("fgmres" if prec == "mixed" else "gmres", "sgs", tolerance=1e-8,
                                                max_iterations=2000, precision=prec))
        ls.buildElemToNodeGraph(mesh)
        ls.finalizeLinearSystem()
        ls.zeroSystem()
        na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=1.0,
                               mode="gather")
        ls.loadComplete()
        ls.removeRhsMean()
        x = torch.zeros(ls.sizes()[0], dtype=torch.float64, device=DEV)
        st = ls.solve(x)
        rel = ls.linearResidual() / ls.nonLinearResidual()
        assert st == 0 and rel <= 1e-8
        xs = x.cpu().numpy()
        sols[prec] = xs - xs.mean()
    d = np.linalg.norm(sols["mixed"] - sols["fp64"]) / np.linalg.norm(sols["fp64"])
    print("  pressure rel diff %.2e" % d, flush=True)
    assert d <= 2e-3


def test_fp32_segregated_momentum(oracle):
    """FP32-arithmetic segregated momentum (3 rhs, one scalar matrix) against the
    oracle's FP64 derivation, to the FP32 assembly bar."""
    n = 6
    print("case: fp32 segregated momentum n=%d" % n, flush=True)
    P = lowmach_problem(n, seed=31, mu=0.05)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    mesh = na.Mesh(X, conn)
    mdot = torch.zeros((conn.shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), mdot, tau=P["tau"])
    ls = na.LinearSystem(1, na.SolverConfig("gmres", "sgs", precision="mixed"))
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    rhs3 = torch.zeros((3, N), dtype=torch.float64, device=DEV)
    na.assemble_momentum_segregated(ls, mesh, dev(P["u"]), dev(P["mu"]), mdot, rhs3)
    rp, ci, v, _ = ls.csr()
    g1 = oracle.graph_build(N, conn, 1)
    g3 = oracle.graph_build(N, conn, 3)
    ov, orr = oracle.momentum_segregated(g1, g3, conn, X, P["u"], P["mu"], mdot.cpu().numpy())
    e_mat = rel_err_rowwise(v, ov, g1.row_ptr)
    e_rhs = np.abs(rhs3.cpu().numpy() - orr).max() / np.abs(orr).max()
    print("  rel err: matrix %.2e rhs %.2e" % (e_mat, e_rhs), flush=True)
    assert e_mat <= TOL32 and e_rhs <= TOL32 and e_mat > 0.0
