"""GPU parity of the time-derivative (mass) kernels against the reference's
unit-test golds and the CPU oracle.

Bars: golds within 1e-14 absolute (the reference tests use 1e-15 on the same
values, our sub-volume arithmetic differs in summation order); whole-mesh matrix
and rhs within 1e-12 of the oracle relative to each row's / the vector's largest
entry; the graph untouched (bit-exact).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nalu_amd as na  # noqa: E402
from nalu_amd.problems import lowmach_problem  # noqa: E402
from tests.fixtures import UNIT_HEX, rel_err_rowwise, trig_dpdx, trig_velocity  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = 1e-12


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def dense_from_csr(rp, ci, v, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    return A


def system(coords, conn, nd):
    mesh = na.Mesh(coords, conn)
    ls = na.LinearSystem(nd, na.SolverConfig())
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    return mesh, ls


@pytest.mark.parametrize("lumped", [False, True])
def test_momentum_mass_gold(golds, lumped):
    print("case: momentum mass gold lumped=%s" % lumped, flush=True)
    g = golds["momentum_mass_lumped" if lumped else "momentum_mass"]
    X = UNIT_HEX
    mesh, ls = system(X, np.arange(8, dtype=np.int32)[None, :], 3)
    z1, z3 = np.zeros(8), np.zeros((8, 3))
    na.assemble_momentum_mass(ls, mesh, (dev(z1), dev(z1), dev(np.ones(8))), (dev(z3), dev(z3), dev(trig_velocity(X))),
                              dev(trig_dpdx(X)), g["dt"], g["gamma"], lumped=lumped)
    rp, ci, v, r = ls.csr()
    want = np.eye(24) * g["lhs_diag"] if lumped else np.array(g["lhs"]).reshape(24, 24)
    np.testing.assert_allclose(dense_from_csr(rp, ci, v, 24), want, rtol=0, atol=1e-14)
    np.testing.assert_allclose(r, g["rhs"], rtol=0, atol=1e-14)


@pytest.mark.parametrize("lumped", [False, True])
def test_continuity_mass_gold(golds, lumped):
    print("case: continuity mass gold lumped=%s" % lumped, flush=True)
    g = golds["continuity_mass"]
    mesh, ls = system(UNIT_HEX, np.arange(8, dtype=np.int32)[None, :], 1)
    z1 = np.zeros(8)
    na.assemble_continuity_mass(ls, mesh, (dev(z1), dev(z1), dev(np.ones(8))), g["dt"], g["gamma"], lumped=lumped)
    rp, ci, v, r = ls.csr()
    np.testing.assert_allclose(v, 0.0, atol=0)
    np.testing.assert_allclose(r, g["rhs_all"], rtol=0, atol=1e-14)


@pytest.mark.parametrize("lumped,balanced", [(False, False), (True, False), (False, True)])
def test_mass_whole_mesh_parity(oracle, lumped, balanced):
    n = 7
    print("case: mass whole mesh n=%d lumped=%s balanced=%s" % (n, lumped, balanced), flush=True)
    P = lowmach_problem(n, seed=13, rho_sigma=0.3)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    rng = np.random.default_rng(17)
    rho = [P["rho"] * (1 + 0.1 * rng.random(N)) for _ in range(3)]
    u = [P["u"] + 0.1 * rng.standard_normal((N, 3)) for _ in range(3)]
    gjp = P["gpdx"] + 0.1 * rng.standard_normal((N, 3))
    dt, gam = 0.013, (1.5, -2.0, 0.5)
    # momentum: mass added on top of the advection-diffusion system
    mesh, ls = system(X, conn, 3)
    mdot = torch.zeros((conn.shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, dev(rho[2]), dev(P["p"]), dev(u[2]), dev(gjp), mdot, tau=P["tau"])
    na.assemble_momentum(ls, mesh, dev(u[2]), dev(P["mu"]), mdot, mode="tile")
    na.assemble_momentum_mass(ls, mesh, [dev(a) for a in rho], [dev(a) for a in u], dev(gjp), dt, gam,
                              lumped=lumped, balanced_force=balanced)
    rp, ci, v, r = ls.csr()
    g3 = oracle.graph_build(N, conn, 3)
    np.testing.assert_array_equal(ci, g3.col_idx)
    ov, orr = oracle.assemble_momentum(g3, conn, X, u[2], P["mu"], mdot.cpu().numpy())
    mv, mr = oracle.assemble_momentum_mass(g3, conn, X, rho, u, gjp, dt, gam, lumped=lumped, balanced=balanced)
    e_mat = rel_err_rowwise(v, ov + mv, g3.row_ptr)
    e_rhs = np.abs(r - (orr + mr)).max() / np.abs(orr + mr).max()
    # the mass part alone is not lost in the sum
    assert np.abs(mv).max() > 1e-3 * np.abs(ov).max()
    # continuity: mass rhs only
    mesh1, ls1 = system(X, conn, 1)
    na.assemble_continuity_mass(ls1, mesh1, [dev(a) for a in rho], dt, gam, lumped=lumped)
    _, _, v1, r1 = ls1.csr()
    g1 = oracle.graph_build(N, conn, 1)
    _, cr = oracle.assemble_continuity_mass(g1, conn, X, rho, dt, gam, lumped=lumped)
    e_c = np.abs(r1 - cr).max() / np.abs(cr).max()
    print("  rel err: momentum matrix %.2e rhs %.2e continuity rhs %.2e" % (e_mat, e_rhs, e_c), flush=True)
    assert e_mat <= TOL and e_rhs <= TOL and e_c <= TOL
    assert not np.any(v1)


@pytest.mark.parametrize("lumped,balanced,skew,ids", [(False, False, False, False), (True, False, False, False),
                                                      (False, True, True, False), (False, False, False, True)])
def test_momentum_with_mass_one_pass(oracle, lumped, balanced, skew, ids):
    """Advection-diffusion and mass in one node-tile pass against the oracle's
    two kernels summed (first assembly after zeroSystem, and on top of an
    earlier assembly)."""
    n = 6
    print("case: momentum + mass one pass n=%d lumped=%s balanced=%s skew=%s ids=%s" % (n, lumped, balanced, skew, ids),
          flush=True)
    P = lowmach_problem(n, seed=31, rho_sigma=0.3)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    rng = np.random.default_rng(37)
    rho = [P["rho"] * (1 + 0.1 * rng.random(N)) for _ in range(3)]
    u = [P["u"] + 0.1 * rng.standard_normal((N, 3)) for _ in range(3)]
    gjp = P["gpdx"] + 0.1 * rng.standard_normal((N, 3))
    dt, gam = 0.017, (1.5, -2.0, 0.5)
    node_ids = np.random.default_rng(5).permutation(N).astype(np.int64) + 1 if ids else None
    mesh = na.Mesh(X, conn, node_ids=node_ids)
    ls = na.LinearSystem(3, na.SolverConfig())
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    mdot = torch.zeros((conn.shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, dev(rho[2]), dev(P["p"]), dev(u[2]), dev(gjp), mdot, tau=P["tau"])
    g3 = oracle.graph_build(N, conn, 3, node_ids=node_ids)
    ov, orr = oracle.assemble_momentum(g3, conn, X, u[2], P["mu"], mdot.cpu().numpy(), skew=skew)
    mv, mr = oracle.assemble_momentum_mass(g3, conn, X, rho, u, gjp, dt, gam, lumped=lumped, balanced=balanced)
    for repeat in (1, 2):  # 1: stores into a pending zero; 2: adds onto the first pass
        if repeat == 1:
            ls.zeroSystem()
        na.assemble_momentum_with_mass(ls, mesh, dev(u[2]), dev(P["mu"]), mdot, [dev(a) for a in rho],
                                       [dev(a) for a in u], dev(gjp), dt, gam, lumped=lumped, balanced_force=balanced,
                                       skew=skew)
        rp, ci, v, r = ls.csr()
        np.testing.assert_array_equal(ci, g3.col_idx)
        e_mat = rel_err_rowwise(v, repeat * (ov + mv), g3.row_ptr)
        e_rhs = np.abs(r - repeat * (orr + mr)).max() / np.abs(repeat * (orr + mr)).max()
        print("  pass %d rel err: matrix %.2e rhs %.2e" % (repeat, e_mat, e_rhs), flush=True)
        assert e_mat <= TOL and e_rhs <= TOL


@pytest.mark.parametrize("lumped", [False, True])
def test_momentum_mass_tile_alone_with_global_ids(oracle, lumped):
    """The numDof 3 mass tile as the first assembly after zeroSystem (its
    store path) on a mesh with permuted global ids (rows are not nodes)."""
    n = 5
    print("case: momentum mass alone, permuted ids, n=%d lumped=%s" % (n, lumped), flush=True)
    P = lowmach_problem(n, seed=21, rho_sigma=0.3)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    ids = np.random.default_rng(8).permutation(N).astype(np.int64) + 1
    rng = np.random.default_rng(23)
    rho = [P["rho"] * (1 + 0.1 * rng.random(N)) for _ in range(3)]
    u = [P["u"] + 0.1 * rng.standard_normal((N, 3)) for _ in range(3)]
    gjp = P["gpdx"] + 0.1 * rng.standard_normal((N, 3))
    dt, gam = 0.02, (1.5, -2.0, 0.5)
    mesh = na.Mesh(X, conn, node_ids=ids)
    ls = na.LinearSystem(3, na.SolverConfig())
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    na.assemble_momentum_mass(ls, mesh, [dev(a) for a in rho], [dev(a) for a in u], dev(gjp), dt, gam,
                              lumped=lumped)
    rp, ci, v, r = ls.csr()
    g3 = oracle.graph_build(N, conn, 3, node_ids=ids)
    np.testing.assert_array_equal(ci, g3.col_idx)
    mv, mr = oracle.assemble_momentum_mass(g3, conn, X, rho, u, gjp, dt, gam, lumped=lumped)
    e_mat = rel_err_rowwise(v, mv, g3.row_ptr)
    e_rhs = np.abs(r - mr).max() / np.abs(mr).max()
    print("  rel err: matrix %.2e rhs %.2e" % (e_mat, e_rhs), flush=True)
    assert e_mat <= TOL and e_rhs <= TOL


def test_mass_rejects_bad_arguments():
    mesh, ls = system(UNIT_HEX, np.arange(8, dtype=np.int32)[None, :], 1)
    z = dev(np.zeros(8))
    with pytest.raises(na.NaluError):
        na.assemble_continuity_mass(ls, mesh, (z, z, z), 0.0, (1.0, -1.0, 0.0))
    with pytest.raises(na.NaluError):
        na.assemble_continuity_mass(ls, mesh, (z, z, z), 0.1, (0.0, -1.0, 0.0))
    with pytest.raises(na.NaluError):  # numDof 1 system for the momentum kernel
        na.assemble_momentum_mass(ls, mesh, (z, z, z), (z, z, z), z, 0.1)


@pytest.mark.parametrize("skew,shifted", [(False, False), (True, True)])
def test_momentum_segregated_parity(oracle, skew, shifted):
    """Segregated momentum (3 rhs, one scalar matrix) against the oracle's
    derivation from the coupled system and the scalar diffusion operator; its
    rhs equals the coupled system's rhs component by component."""
    n = 6
    print("case: segregated momentum n=%d skew=%s shifted=%s" % (n, skew, shifted), flush=True)
    P = lowmach_problem(n, seed=31, mu=0.05)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    mesh = na.Mesh(X, conn)
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    mdot = torch.zeros((conn.shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
    ls = na.LinearSystem(1, na.SolverConfig())
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    rhs3 = torch.zeros((3, N), dtype=torch.float64, device=DEV)
    na.assemble_momentum_segregated(ls, mesh, u, dev(P["mu"]), mdot, rhs3, skew=skew, shifted=shifted)
    rp, ci, v, _ = ls.csr()
    g1 = oracle.graph_build(N, conn, 1)
    g3 = oracle.graph_build(N, conn, 3)
    np.testing.assert_array_equal(ci, g1.col_idx)
    ov, orr = oracle.momentum_segregated(g1, g3, conn, X, P["u"], P["mu"], mdot.cpu().numpy(), skew=skew,
                                         shifted=shifted)
    e_mat = rel_err_rowwise(v, ov, g1.row_ptr)
    e_rhs = np.abs(rhs3.cpu().numpy() - orr).max() / np.abs(orr).max()
    # against the GPU's own coupled assembly: identical residual
    ls3 = na.LinearSystem(3, na.SolverConfig())
    ls3.buildElemToNodeGraph(mesh)
    ls3.finalizeLinearSystem()
    ls3.zeroSystem()
    na.assemble_momentum(ls3, mesh, u, dev(P["mu"]), mdot, skew=skew, shifted=shifted, mode="tile")
    r3 = ls3.csr()[3].reshape(N, 3).T
    e_self = np.abs(rhs3.cpu().numpy() - r3).max() / np.abs(r3).max()
    print("  rel err: matrix %.2e rhs %.2e rhs vs coupled %.2e" % (e_mat, e_rhs, e_self), flush=True)
    assert e_mat <= TOL and e_rhs <= TOL and e_self <= TOL


@pytest.mark.parametrize("lumped,balanced,precision", [(False, False, "fp64"), (True, True, "fp64"),
                                                       (False, False, "mixed")])
def test_segregated_momentum_with_mass_one_pass(lumped, balanced, precision):
    """The segregated advection-diffusion and mass kernels in one pass give the
    sums of the two oracle-checked passes (to rounding order; with FP32
    assembly too, whose contributions are summed in FP64 either way)."""
    n = 6
    print("case: segregated momentum + mass one pass n=%d lumped=%s balanced=%s %s" % (n, lumped, balanced, precision),
          flush=True)
    P = lowmach_problem(n, seed=41, mu=0.05, rho_sigma=0.3)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    rng = np.random.default_rng(43)
    rho = [dev(P["rho"] * (1 + 0.1 * rng.random(N))) for _ in range(3)]
    us = [dev(P["u"] + 0.1 * rng.standard_normal((N, 3))) for _ in range(3)]
    gjp = dev(P["gpdx"] + 0.1 * rng.standard_normal((N, 3)))
    dt, gam = 0.02, (1.5, -2.0, 0.5)
    mesh = na.Mesh(X, conn)
    mdot = torch.zeros((conn.shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho[2], dev(P["p"]), us[2], gjp, mdot, tau=P["tau"])
    out = []
    for fused in (True, False):
        ls = na.LinearSystem(1, na.SolverConfig(precision=precision))
        ls.buildElemToNodeGraph(mesh)
        ls.finalizeLinearSystem()
        ls.zeroSystem()
        rhs3 = torch.zeros((3, N), dtype=torch.float64, device=DEV)
        if fused:
            na.assemble_momentum_with_mass_segregated(ls, mesh, us[2], dev(P["mu"]), mdot, rho, us, gjp, dt, rhs3,
                                                      gamma=gam, lumped=lumped, balanced_force=balanced)
        else:
            na.assemble_momentum_segregated(ls, mesh, us[2], dev(P["mu"]), mdot, rhs3)
            na.assemble_momentum_mass_segregated(ls, mesh, rho, us, gjp, dt, rhs3, gamma=gam, lumped=lumped,
                                                 balanced_force=balanced)
        rp, ci, v, _ = ls.csr()
        out.append((rp, v, rhs3.cpu().numpy()))
    (rp, vf, bf), (_, v2, b2) = out
    e_mat = rel_err_rowwise(vf, v2, rp)
    e_rhs = np.abs(bf - b2).max() / np.abs(b2).max()
    print("  one pass vs two: matrix %.2e rhs %.2e" % (e_mat, e_rhs), flush=True)
    assert e_mat <= 1e-13 and e_rhs <= 1e-13


def test_segregated_transient_momentum_solve(oracle):
    """Backward-Euler momentum step in the segregated form: mass + advection-
    diffusion, three right-hand sides solved with one matrix and one SGS set-up
    to 1e-8; each solution satisfies its own system (checked with the oracle's
    SpMV) and the same step solved as three separate solve_rhs calls agrees."""
    n = 8
    print("case: segregated transient momentum n=%d" % n, flush=True)
    P = lowmach_problem(n, seed=5, mu=0.01)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    mesh = na.Mesh(X, conn)
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    mdot = torch.zeros((conn.shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
    ls = na.LinearSystem(1, na.SolverConfig("gmres", "sgs", tolerance=1e-8, max_iterations=500, kspace=50))
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    rhs3 = torch.zeros((3, N), dtype=torch.float64, device=DEV)
    dt = 0.2 / n
    na.assemble_momentum_segregated(ls, mesh, u, dev(P["mu"]), mdot, rhs3)
    uN = dev(P["u"] * 0.9)
    na.assemble_momentum_mass_segregated(ls, mesh, (rho, rho, rho), (uN, uN, u), gp, dt, rhs3)
    ls.loadComplete()
    X3 = torch.zeros_like(rhs3)
    st, its, rel = ls.solve_multi(rhs3, X3)
    print("  iterations", its, "rel", ["%.1e" % r for r in rel], flush=True)
    assert st == 0 and max(rel) <= 1e-8 and max(its) < 200
    rp, ci, v, _ = ls.csr()
    rp64 = rp.astype(np.int64)
    B, Xh = rhs3.cpu().numpy(), X3.cpu().numpy()
    for i in range(3):
        res = oracle.spmv(rp64, ci, v, Xh[i]) - B[i]
        assert np.linalg.norm(res) <= 1.5e-8 * np.linalg.norm(B[i])
        x1 = torch.zeros(N, dtype=torch.float64, device=DEV)
        ls.solve_rhs(rhs3[i].contiguous(), x1)
        np.testing.assert_allclose(x1.cpu().numpy(), Xh[i], rtol=0, atol=1e-12 * np.abs(Xh[i]).max())


@pytest.mark.parametrize("precond,precision,k", [("sgs", "fp64", 3), ("sgs", "mixed", 3), ("sgs", "fp64", 2),
                                                  ("sgs", "fp64", 4), ("jacobi", "fp64", 3), ("sgs", "fp64", 5)])
def test_stacked_solve(oracle, precond, precision, k):
    """The segregated momentum step's right-hand sides solved as one stacked
    block-diagonal system (batched SpMV / SGS kernels for k <= 4, one slice at a
    time otherwise): the stacked residual meets 1e-8 of the stacked rhs, each
    component satisfies its own system to that bar (oracle SpMV), and the
    solutions agree with separate solves."""
    n = 8
    print("case: stacked solve n=%d %s %s k=%d" % (n, precond, precision, k), flush=True)
    P = lowmach_problem(n, seed=5, mu=0.01)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    mesh = na.Mesh(X, conn)
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    mdot = torch.zeros((conn.shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
    method = "fgmres" if precision == "mixed" else "gmres"
    ls = na.LinearSystem(1, na.SolverConfig(method, precond, tolerance=1e-8, max_iterations=800, kspace=50,
                                            precision=precision))
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    rhs3 = torch.zeros((3, N), dtype=torch.float64, device=DEV)
    dt = 0.2 / n
    na.assemble_momentum_segregated(ls, mesh, u, dev(P["mu"]), mdot, rhs3)
    uN = dev(P["u"] * 0.9)
    na.assemble_momentum_mass_segregated(ls, mesh, (rho, rho, rho), (uN, uN, u), gp, dt, rhs3)
    ls.loadComplete()
    extra = torch.as_tensor(np.random.default_rng(3).standard_normal((max(0, k - 3), N)) * 0.01, device=DEV)
    B = torch.cat([rhs3, extra])[:k].contiguous()
    Xs = torch.zeros_like(B)
    st, its, rel = ls.solve_stacked(B, Xs)
    print("  stacked iterations", its, "rel", ["%.1e" % r for r in rel], flush=True)
    assert st == 0
    rp, ci, v, _ = ls.csr()
    rp64 = rp.astype(np.int64)
    Bh, Xh = B.cpu().numpy(), Xs.cpu().numpy()
    bstack = np.linalg.norm(Bh)
    res = [oracle.spmv(rp64, ci, v, Xh[i]) - Bh[i] for i in range(k)]
    assert np.sqrt(sum(np.linalg.norm(r) ** 2 for r in res)) <= 1.5e-8 * bstack
    for i in range(k):
        assert abs(np.linalg.norm(res[i]) / np.linalg.norm(Bh[i]) - rel[i]) <= 1e-3 * rel[i] + 1e-14
    Xm = torch.zeros_like(B)
    st_m, its_m, _ = ls.solve_multi(B, Xm)
    assert st_m == 0
    np.testing.assert_allclose(Xh, Xm.cpu().numpy(), rtol=0, atol=1e-6 * np.abs(Xh).max())


@pytest.mark.parametrize("shifted", [False, True])
def test_dual_volume_and_nodal_gradient_parity(oracle, shifted):
    """Dual nodal volume and the projected nodal gradient (interior Green-Gauss
    part, then with the exposed-face closure) against the oracle on a jittered
    lattice (1e-12 of the largest entry), the exposed sides identical to the
    oracle's; the gradient of the pressure then feeds continuity assembly as Gpdx."""
    n = 8
    print("case: dual volume + nodal gradient n=%d shifted=%s" % (n, shifted), flush=True)
    P = lowmach_problem(n, seed=21, rho_sigma=0.3)
    X, conn = P["coords"], P["conn"]
    N = X.shape[0]
    mesh = na.Mesh(X, conn)
    V = torch.zeros(N, dtype=torch.float64, device=DEV)
    na.compute_dual_volume(mesh, V)
    oV = oracle.dual_nodal_volume(conn, X)
    np.testing.assert_allclose(V.cpu().numpy(), oV, rtol=0, atol=1e-12 * oV.max())
    G = torch.zeros((N, 3), dtype=torch.float64, device=DEV)
    # interior part alone, then the full projected gradient (with the exposed-face closure)
    na.compute_nodal_gradient(mesh, dev(P["p"]), V, G, shifted=shifted, boundary=False)
    oGi = oracle.nodal_grad_elem(conn, X, P["p"], oV, shifted=shifted)
    err_i = np.abs(G.cpu().numpy() - oGi).max() / np.abs(oGi).max()
    na.compute_nodal_gradient(mesh, dev(P["p"]), V, G, shifted=shifted)
    oG = oracle.nodal_grad(conn, X, P["p"], oV, shifted=shifted)
    err = np.abs(G.cpu().numpy() - oG).max() / np.abs(oG).max()
    np.testing.assert_array_equal(mesh.exposed_faces(), oracle.exposed_faces(conn, N))
    print("  rel err: dual volume %.2e gradient interior %.2e with closure %.2e"
          % (np.abs(V.cpu().numpy() - oV).max() / oV.max(), err_i, err), flush=True)
    assert err_i <= TOL and err <= TOL
    # the produced gradient drives the continuity assembly like the analytic one
    ls = na.LinearSystem(1, na.SolverConfig())
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), G, tau=P["tau"], mode="gather")
    _, _, v, r = ls.csr()
    g = oracle.graph_build(N, conn)
    ov, orr = oracle.assemble_continuity(g, conn, X, P["rho"], P["p"], P["u"], oG, P["tau"])
    assert rel_err_rowwise(v, ov, g.row_ptr) <= TOL
    assert np.abs(r - orr).max() <= TOL * np.abs(orr).max()


def test_nodal_gradient_exact_on_affine_mesh_all_nodes():
    """The projected gradient with the closure reproduces the gradient of a linear
    field at every node, boundary nodes included, on a sheared (affine) lattice."""
    print("case: nodal gradient of a linear field, sheared lattice n=6", flush=True)
    n = 6
    P = lowmach_problem(n, seed=3, jitter=0.0)
    X = P["coords"] @ np.array([[1.0, 0.2, 0.0], [0.1, 0.9, 0.3], [0.0, -0.1, 1.2]]).T
    mesh = na.Mesh(X, P["conn"])
    N = X.shape[0]
    a = np.array([0.7, -1.3, 2.1])
    V = torch.zeros(N, dtype=torch.float64, device=DEV)
    na.compute_dual_volume(mesh, V)
    G = torch.zeros((N, 3), dtype=torch.float64, device=DEV)
    na.compute_nodal_gradient(mesh, dev(X @ a + 0.4), V, G)
    np.testing.assert_allclose(G.cpu().numpy(), np.broadcast_to(a, (N, 3)), rtol=0, atol=1e-11)
