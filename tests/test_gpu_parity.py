"""GPU parity: the HIP path against the CPU oracle and the reference golds.

Bars: graph (row_ptr / col_idx) bit-exact; matrix and rhs entries within
1e-12 of the oracle relative to the largest entry of their row (the kernels
evaluate the same formulas in a different floating-point order); reference
element golds within 1e-14 absolute (the reference tests use 1e-15 on the
same values); solves reach the FP64 true relative residual they are asked for.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nalu_amd as na  # noqa: E402
from nalu_amd.problems import heat_problem, lowmach_problem, poisson_rhs  # noqa: E402
from tests.fixtures import (UNIT_HEX, lattice_1x1xN, rel_err_rowwise, trig_dpdx, trig_pressure,  # noqa: E402
                            trig_temperature, trig_velocity)

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = 1e-12


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def dense_from_csr(rp, ci, v, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    return A


def make_system(coords, conn, nd=1, cfg=None, ids=None):
    mesh = na.Mesh(coords, conn, ids)
    ls = na.LinearSystem(nd, cfg or na.SolverConfig())
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    return mesh, ls


# ------------------------------------------------------------- reference golds on one element

@pytest.mark.parametrize("mode", ["gather", "atomic", "tile", "etile"])
def test_scalar_diff_gold_gpu(golds, mode):
    print("case: scalar_diff gold, mode", mode, flush=True)
    mesh, ls = make_system(UNIT_HEX, np.arange(8, dtype=np.int32)[None, :])
    ls.zeroSystem()
    na.assemble_scalar_diff(ls, mesh, dev(trig_temperature(UNIT_HEX)), dev(np.ones(8)), mode=mode)
    rp, ci, v, r = ls.csr()
    A = dense_from_csr(rp, ci, v, 8)
    np.testing.assert_allclose(A, np.array(golds["scalar_diff"]["lhs"]).reshape(8, 8), rtol=0, atol=1e-14)
    np.testing.assert_allclose(r, 0.0, atol=1e-14)


@pytest.mark.parametrize("mode", ["gather", "atomic", "tile", "etile"])
@pytest.mark.parametrize("name,shift,reduced", [("continuity_default", False, False),
                                                ("continuity_reduced_sens", False, True),
                                                ("continuity_reduced_sens", True, True)])
def test_continuity_gold_gpu(golds, mode, name, shift, reduced):
    print("case: continuity gold", name, shift, reduced, mode, flush=True)
    X = UNIT_HEX
    mesh, ls = make_system(X, np.arange(8, dtype=np.int32)[None, :])
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, dev(np.ones(8)), dev(trig_pressure(X)), dev(trig_velocity(X)),
                           dev(trig_dpdx(X)), tau=1.0, shift_poisson=shift, reduced_sens=reduced, mode=mode)
    rp, ci, v, r = ls.csr()
    np.testing.assert_allclose(dense_from_csr(rp, ci, v, 8), np.array(golds[name]["lhs"]).reshape(8, 8),
                               rtol=0, atol=1e-14)
    np.testing.assert_allclose(r, golds[name]["rhs"], rtol=0, atol=1e-14)


@pytest.mark.parametrize("mode", ["atomic", "tile", "etile"])
def test_momentum_gold_gpu(golds, mode):
    print("case: momentum gold", mode, flush=True)
    X = UNIT_HEX
    g = golds["momentum_advdiff"]
    mesh, ls = make_system(X, np.arange(8, dtype=np.int32)[None, :], nd=3)
    u = dev(trig_velocity(X))
    mdot = torch.zeros(12, dtype=torch.float64, device=DEV)
    # mdot of the unit-test producer is rho u . A: no pressure terms
    na.compute_mdot(mesh, dev(np.ones(8)), dev(np.zeros(8)), u, dev(np.zeros((8, 3))), mdot, tau=1.0)
    ls.zeroSystem()
    na.assemble_momentum(ls, mesh, u, dev(np.full(8, g["mu"])), mdot, mode=mode)
    rp, ci, v, r = ls.csr()
    np.testing.assert_allclose(dense_from_csr(rp, ci, v, 24), np.array(g["lhs"]).reshape(24, 24), rtol=0,
                               atol=1e-14)
    np.testing.assert_allclose(r, g["rhs"], rtol=0, atol=1e-14)


def test_tpetra_gold_sum_into_batch(golds):
    print("case: tpetra 1x1x2 sum_into_batch", flush=True)
    gold = golds["tpetra_1x1x2"]
    coords, conn = lattice_1x1xN(2)
    mesh, ls = make_system(coords, conn)
    rp, ci, _, _ = ls.csr()
    np.testing.assert_array_equal(np.diff(rp), gold["row_lengths"])
    ev = np.array(gold["elem_vals"]).reshape(8, 8)
    ls.zeroSystem()
    ls.sumIntoBatch(dev(np.stack([ev, ev])), dev(np.zeros((2, 8))))
    rp, ci, v, r = ls.csr()
    np.testing.assert_allclose(dense_from_csr(rp, ci, v, 12), np.array(gold["lhs_vals"]).reshape(12, 12),
                               atol=gold["tol"])
    # host overload, one element at a time
    ls.zeroSystem()
    for e in range(2):
        ls.sumInto(conn[e], np.zeros(8), ev)
    rp, ci, v2, _ = ls.csr()
    np.testing.assert_array_equal(v2, v)


# ------------------------------------------------------------- whole-mesh parity

@pytest.mark.parametrize("mode", ["gather", "atomic", "tile", "etile"])
def test_heat_assembly_parity(oracle, mode):
    n = 8
    print("case: heat assembly n=%d mode=%s" % (n, mode), flush=True)
    P = heat_problem(n, seed=11)
    mesh, ls = make_system(P["coords"], P["conn"])
    T, k = dev(P["T"] + 0.3 * np.sin(P["coords"][:, 1])), dev(P["k"])
    ls.zeroSystem()
    na.assemble_scalar_diff(ls, mesh, T, k, mode=mode)
    ls.applyDirichletBCs(dev(P["bc_nodes"], torch.int32), T, dev(P["bc_values"]))
    ls.loadComplete()
    rp, ci, v, r = ls.csr()
    g = oracle.graph_build(P["coords"].shape[0], P["conn"])
    ov, orr = oracle.assemble_scalar_diff(g, P["conn"], P["coords"], T.cpu().numpy(), P["k"])
    oracle.apply_dirichlet(g, ov, orr, P["bc_nodes"], T.cpu().numpy(), P["bc_values"])
    np.testing.assert_array_equal(rp, g.row_ptr)
    np.testing.assert_array_equal(ci, g.col_idx)
    assert rel_err_rowwise(v, ov, g.row_ptr) <= TOL
    np.testing.assert_allclose(r, orr, rtol=0, atol=TOL * np.abs(orr).max())


def test_gather_assembly_is_deterministic():
    P = heat_problem(8, seed=3)
    mesh, ls = make_system(P["coords"], P["conn"])
    T, k = dev(np.random.default_rng(0).normal(size=len(P["T"]))), dev(P["k"])
    out = []
    for _ in range(2):
        ls.zeroSystem()
        na.assemble_scalar_diff(ls, mesh, T, k, mode="gather")
        out.append(ls.csr())
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_gather_accumulates_after_another_algorithm(oracle):
    # two algorithms contributing to one system: the second adds to the first
    P = heat_problem(6, seed=5)
    mesh, ls = make_system(P["coords"], P["conn"])
    T, k = dev(np.cos(P["coords"][:, 0])), dev(P["k"])
    ls.zeroSystem()
    na.assemble_scalar_diff(ls, mesh, T, k, mode="atomic")
    na.assemble_scalar_diff(ls, mesh, T, k, mode="gather")
    _, _, v2, r2 = ls.csr()
    ls.zeroSystem()
    na.assemble_scalar_diff(ls, mesh, T, k, mode="gather")
    rp, _, v1, r1 = ls.csr()
    assert rel_err_rowwise(v2, 2 * v1, rp) <= TOL
    np.testing.assert_allclose(r2, 2 * r1, rtol=0, atol=TOL * np.abs(r1).max())


@pytest.mark.parametrize("mode", ["gather", "atomic", "tile", "etile"])
def test_continuity_and_mdot_parity(oracle, mode):
    n = 8
    print("case: continuity + mdot n=%d mode=%s" % (n, mode), flush=True)
    P = lowmach_problem(n, seed=21, rho_sigma=0.3)
    mesh, ls = make_system(P["coords"], P["conn"])
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, rho, p, u, gp, tau=P["tau"], mode=mode)
    rp, ci, v, r = ls.csr()
    g = oracle.graph_build(P["coords"].shape[0], P["conn"])
    ov, orr = oracle.assemble_continuity(g, P["conn"], P["coords"], P["rho"], P["p"], P["u"], P["gpdx"], P["tau"])
    np.testing.assert_array_equal(ci, g.col_idx)
    assert rel_err_rowwise(v, ov, g.row_ptr) <= TOL
    np.testing.assert_allclose(r, orr, rtol=0, atol=TOL * np.abs(orr).max())
    mdot = torch.zeros((P["conn"].shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
    om = oracle.compute_mdot(P["conn"], P["coords"], P["rho"], P["p"], P["u"], P["gpdx"], P["tau"])
    np.testing.assert_allclose(mdot.cpu().numpy(), om, rtol=0, atol=TOL * np.abs(om).max())


@pytest.mark.parametrize("shift_mdot,shift_poisson,reduced", [(False, False, False), (True, False, False),
                                                               (False, True, True), (True, True, False)])
def test_continuity_with_fused_mdot(oracle, shift_mdot, shift_poisson, reduced):
    """The row gather that also writes mdot: the continuity system bit-identical to
    the plain row gather, mdot against the oracle and within round-off of the
    separate mdot pass, every (element, scs) written."""
    n = 7
    print("case: continuity + fused mdot n=%d shift_mdot=%d shift_poisson=%d reduced=%d"
          % (n, shift_mdot, shift_poisson, reduced), flush=True)
    P = lowmach_problem(n, seed=23, rho_sigma=0.3)
    mesh, ls = make_system(P["coords"], P["conn"])
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    E = P["conn"].shape[0]
    kw = dict(tau=P["tau"], shift_mdot=shift_mdot, shift_poisson=shift_poisson, reduced_sens=reduced)
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, rho, p, u, gp, mode="gather", **kw)
    _, _, v0, r0 = ls.csr()
    md = torch.full((E, 12), float("nan"), dtype=torch.float64, device=DEV)
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, rho, p, u, gp, mode="gather", mdot_out=md, **kw)
    _, _, v1, r1 = ls.csr()
    np.testing.assert_array_equal(v1, v0)
    np.testing.assert_array_equal(r1, r0)
    ref = torch.zeros_like(md)
    na.compute_mdot(mesh, rho, p, u, gp, ref, tau=P["tau"], shift_mdot=shift_mdot, shift_poisson=shift_poisson)
    m1, m0 = md.cpu().numpy(), ref.cpu().numpy()
    assert np.isfinite(m1).all()
    np.testing.assert_allclose(m1, m0, rtol=0, atol=1e-13 * np.abs(m0).max())
    om = oracle.compute_mdot(P["conn"], P["coords"], P["rho"], P["p"], P["u"], P["gpdx"], P["tau"],
                             shift_mdot=shift_mdot, shift_poisson=shift_poisson)
    np.testing.assert_allclose(m1, om, rtol=0, atol=TOL * np.abs(om).max())
    with pytest.raises(na.NaluError, match="gather"):
        na.assemble_continuity(ls, mesh, rho, p, u, gp, mode="tile", mdot_out=md, **kw)


@pytest.mark.parametrize("mode", ["atomic", "tile", "etile"])
def test_momentum_parity(oracle, mode):
    n = 6
    print("case: momentum n=%d %s" % (n, mode), flush=True)
    P = lowmach_problem(n, seed=31, mu=0.05)
    mesh, ls = make_system(P["coords"], P["conn"], nd=3)
    rho, p, u, gp = dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"])
    mdot = torch.zeros((P["conn"].shape[0], 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"])
    ls.zeroSystem()
    na.assemble_momentum(ls, mesh, u, dev(P["mu"]), mdot, mode=mode)
    rp, ci, v, r = ls.csr()
    g = oracle.graph_build(P["coords"].shape[0], P["conn"], 3)
    np.testing.assert_array_equal(rp, g.row_ptr)
    np.testing.assert_array_equal(ci, g.col_idx)
    ov, orr = oracle.assemble_momentum(g, P["conn"], P["coords"], P["u"], P["mu"], mdot.cpu().numpy())
    assert rel_err_rowwise(v, ov, g.row_ptr) <= TOL
    np.testing.assert_allclose(r, orr, rtol=0, atol=TOL * np.abs(orr).max())


def test_explicit_global_ids(oracle):
    print("case: permuted global ids", flush=True)
    P = heat_problem(4, seed=9)
    nn = P["coords"].shape[0]
    ids = np.random.default_rng(4).permutation(nn).astype(np.int64) + 1
    mesh, ls = make_system(P["coords"], P["conn"], ids=ids)
    ls.zeroSystem()
    na.assemble_scalar_diff(ls, mesh, dev(np.sin(P["coords"][:, 2])), dev(P["k"]), mode="tile")
    rp, ci, v, r = ls.csr()
    g = oracle.graph_build(nn, P["conn"], 1, node_ids=ids)
    np.testing.assert_array_equal(rp, g.row_ptr)
    np.testing.assert_array_equal(ci, g.col_idx)
    ov, orr = oracle.assemble_scalar_diff(g, P["conn"], P["coords"], np.sin(P["coords"][:, 2]), P["k"])
    assert rel_err_rowwise(v, ov, g.row_ptr) <= TOL


# ------------------------------------------------------------- sparse kernels and solves

def _heat_ls(n, cfg, seed=13):
    P = heat_problem(n, seed=seed)
    mesh, ls = make_system(P["coords"], P["conn"], cfg=cfg)
    T, k = dev(P["T"]), dev(P["k"])
    ls.zeroSystem()
    na.assemble_scalar_diff(ls, mesh, T, k)
    ls.applyDirichletBCs(dev(P["bc_nodes"], torch.int32), T, dev(P["bc_values"]))
    ls.loadComplete()
    return P, mesh, ls


def test_spmv_and_preconditioners(oracle):
    print("case: spmv / sgs / jacobi n=10", flush=True)
    for pc in ("sgs", "jacobi"):
        P, mesh, ls = _heat_ls(10, na.SolverConfig("gmres", pc))
        rp, ci, v, _ = ls.csr()
        x = np.random.default_rng(1).normal(size=len(rp) - 1)
        y = torch.zeros(len(x), dtype=torch.float64, device=DEV)
        ls.spmv(dev(x), y)
        oy = oracle.spmv(rp, ci, v, x)
        np.testing.assert_allclose(y.cpu().numpy(), oy, rtol=0, atol=1e-13 * np.abs(oy).max())
        ls.precond_setup()
        z = torch.zeros_like(y)
        ls.precond_apply(dev(x), z)
        if pc == "sgs":
            cp, cr = ls.coloring()
            oz = oracle.mcsgs_apply(rp, ci, v, cp, cr, x)
            # rows of one color must be independent
            color = np.zeros(len(x), dtype=int)
            for c in range(len(cp) - 1):
                color[cr[cp[c]:cp[c + 1]]] = c
            for i in range(len(x)):
                nb = ci[rp[i]:rp[i + 1]]
                assert np.all((color[nb] != color[i]) | (nb == i))
        else:
            oz = oracle.jacobi_apply(rp, ci, v, x)
        np.testing.assert_allclose(z.cpu().numpy(), oz, rtol=0, atol=1e-12 * np.abs(oz).max())
        # the solver-order kernels must not read what their output buffer held before
        xp = dev(x[ls.permutation()])
        z0, zn = torch.zeros_like(y), torch.full_like(y, float("nan"))
        ls.precond_apply_permuted(xp, z0)
        ls.precond_apply_permuted(xp, zn)
        assert torch.equal(z0, zn)
        y0, yn = torch.zeros_like(y), torch.full_like(y, float("nan"))
        ls.spmv_permuted(xp, y0)
        ls.spmv_permuted(xp, yn)
        assert torch.equal(y0, yn)


@pytest.mark.parametrize("method,prec", [("gmres", "fp64"), ("gmres", "mixed"), ("fgmres", "fp64")])
def test_heat_solve(oracle, method, prec):
    n = 16
    print("case: heat solve n=%d %s %s" % (n, method, prec), flush=True)
    P, mesh, ls = _heat_ls(n, na.SolverConfig(method, "sgs", tolerance=1e-8, kspace=50, max_iterations=2000,
                                              precision=prec))
    rp, ci, v, r = ls.csr()
    dT = torch.zeros(len(r), dtype=torch.float64, device=DEV)
    st = ls.solve(dT)
    x = dT.cpu().numpy()
    res = np.linalg.norm(r - oracle.spmv(rp, ci, v, x)) / np.linalg.norm(r)
    print("  status %d iters %d true resid %.3e" % (st, ls.linearSolveIterations(), res), flush=True)
    assert st == 0 and res <= 1e-8
    xo, it, reso = oracle.gmres(rp, ci, v, r, kspace=50, maxit=2000, tol=1e-10, precond="sgs")
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-6 * np.abs(xo).max())
    assert abs(ls.nonLinearResidual() - np.linalg.norm(r)) <= 1e-12 * np.linalg.norm(r)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_pressure_poisson_cg(oracle, prec):
    n = 16
    print("case: pressure poisson CG+SGS n=%d %s" % (n, prec), flush=True)
    P = lowmach_problem(n, seed=41)
    method = "cg" if prec == "fp64" else "gmres"
    mesh, ls = make_system(P["coords"], P["conn"], cfg=na.SolverConfig(method, "sgs", tolerance=1e-8,
                                                                       max_iterations=3000, precision=prec))
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=1.0)
    ls.loadComplete()
    rp, ci, v, _ = ls.csr()
    b = poisson_rhs(len(rp) - 1, 77)
    x = torch.zeros(len(b), dtype=torch.float64, device=DEV)
    st, it, rel = ls.solve_rhs(dev(b), x)
    res = np.linalg.norm(b - oracle.spmv(rp, ci, v, x.cpu().numpy())) / np.linalg.norm(b)
    print("  status %d iters %d true resid %.3e" % (st, it, res), flush=True)
    assert st == 0 and res <= 1e-8 and abs(res - rel) <= 1e-12


def test_fused_cg_matches_unfused(oracle, monkeypatch):
    """The fused CG (dot products inside the SpMV, update and sweep kernels, the
    convergence test read every 8 iterations) against the kernel-per-step CG on the
    same system: the same iteration count within one (sums in another order), the
    same true residual bound, and an iteration cap that is not a multiple of 8."""
    n = 14
    print("case: fused vs unfused CG+SGS, continuity n=%d" % n, flush=True)
    P = lowmach_problem(n, seed=43)
    mesh, ls = make_system(P["coords"], P["conn"], cfg=na.SolverConfig("cg", "sgs", tolerance=1e-8,
                                                                       max_iterations=3000))
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=1.0)
    ls.loadComplete()
    rp, ci, v, _ = ls.csr()
    b = poisson_rhs(len(rp) - 1, 78)
    out = {}
    for fused in (True, False):
        if fused:
            monkeypatch.delenv("NALU_CG_UNFUSED", raising=False)
        else:
            monkeypatch.setenv("NALU_CG_UNFUSED", "1")
        x = torch.zeros(len(b), dtype=torch.float64, device=DEV)
        st, it, rel = ls.solve_rhs(dev(b), x)
        xs = x.cpu().numpy()
        res = np.linalg.norm(b - oracle.spmv(rp, ci, v, xs)) / np.linalg.norm(b)
        print("  fused=%d status %d iters %d true resid %.3e" % (fused, st, it, res), flush=True)
        assert st == 0 and res <= 1e-8 and abs(res - rel) <= 1e-12
        out[fused] = (it, xs)
    assert abs(out[True][0] - out[False][0]) <= 1
    np.testing.assert_allclose(out[True][1], out[False][1], rtol=0, atol=1e-6 * np.abs(out[False][1]).max())
    monkeypatch.delenv("NALU_CG_UNFUSED", raising=False)
    mesh, ls = make_system(P["coords"], P["conn"], cfg=na.SolverConfig("cg", "sgs", tolerance=1e-14,
                                                                       max_iterations=13))
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=1.0)
    ls.loadComplete()
    x = torch.zeros(len(b), dtype=torch.float64, device=DEV)
    st, it, _ = ls.solve_rhs(dev(b), x)
    assert st == 1 and it == 13


def test_nan_rhs_is_reported():
    print("case: NaN rhs", flush=True)
    P, mesh, ls = _heat_ls(4, na.SolverConfig("gmres", "sgs", max_iterations=50))
    _, _, _, rhs_ptr = ls.device_ptrs()
    b = torch.full((ls.sizes()[0],), float("nan"), dtype=torch.float64, device=DEV)
    x = torch.zeros_like(b)
    with pytest.raises(na.NaluError, match="non-finite"):
        ls.solve_rhs(b, x)


def test_iteration_cap_reports_not_converged():
    P, mesh, ls = _heat_ls(12, na.SolverConfig("gmres", "jacobi", tolerance=1e-12, max_iterations=5, kspace=5))
    dT = torch.zeros(ls.sizes()[0], dtype=torch.float64, device=DEV)
    assert ls.solve(dT) == 1
    assert ls.linearSolveIterations() == 5


def test_parity_coloring_on_lattice_and_fallback():
    print("case: coloring", flush=True)
    P = heat_problem(6, seed=2)
    for nd, expect in ((1, 8), (3, 24)):
        mesh, ls = make_system(P["coords"], P["conn"], nd=nd)
        cp, cr = ls.coloring()
        assert len(cp) - 1 == expect
        np.testing.assert_array_equal(np.sort(cr), np.arange(ls.sizes()[0]))
    # an element with rotated local axes admits no parity labelling: fallback coloring stays valid
    coords, conn = lattice_1x1xN(2)
    conn = conn.copy()
    conn[1] = conn[1][[1, 2, 3, 0, 5, 6, 7, 4]]
    mesh, ls = make_system(coords, conn)
    rp, ci, _, _ = ls.csr()
    cp, cr = ls.coloring()
    color = np.zeros(len(rp) - 1, dtype=int)
    for c in range(len(cp) - 1):
        color[cr[cp[c]:cp[c + 1]]] = c
    for i in range(len(rp) - 1):
        nb = ci[rp[i]:rp[i + 1]]
        assert np.all((color[nb] != color[i]) | (nb == i))


def test_wrong_numdof_is_rejected():
    P = heat_problem(3)
    mesh, ls = make_system(P["coords"], P["conn"], nd=3)
    with pytest.raises(na.NaluError, match="numDof"):
        na.assemble_scalar_diff(ls, mesh, dev(P["T"]), dev(P["k"]))


def test_matrix_market_files(tmp_path, oracle):
    """write_matrix_files: the MatrixMarket matrix uses 1-based Tpetra GIDs
    (numDof (id - 1) + dof + 1) and round-trips the CSR values exactly."""
    import scipy.io
    print("case: matrix market dump", flush=True)
    P = heat_problem(3, seed=9)
    nn = P["coords"].shape[0]
    ids = np.random.default_rng(4).permutation(nn).astype(np.int64) + 1
    mesh, ls = make_system(P["coords"], P["conn"], nd=3, ids=ids, cfg=na.SolverConfig("gmres", "sgs"))
    ls.zeroSystem()
    rng = np.random.default_rng(2)
    E = P["conn"].shape[0]
    lhs = rng.standard_normal((E, 24, 24))
    lhs += 48 * np.eye(24)[None]
    rhs = rng.standard_normal((E, 24))
    ls.sumIntoBatch(dev(lhs), dev(rhs))
    ls.loadComplete()
    base = str(tmp_path / "momentum")
    ls.set_write_matrix_files(base)
    x = torch.zeros(ls.sizes()[0], dtype=torch.float64, device=DEV)
    ls.solve(x)
    rp, ci, v, r = ls.csr()
    A = scipy.io.mmread(base + "-O-0.mm.1").tocsr()
    b = scipy.io.mmread(base + "-O-0.rhs.1").ravel()
    s = scipy.io.mmread(base + "-O-0.sln.1").ravel()
    # GID of each row: rows are sorted by node id, so row l belongs to the l-th smallest id
    sid = np.sort(ids)
    gid = (3 * (np.repeat(sid, 3) - 1) + np.tile(np.arange(3), nn)).astype(np.int64)  # 0-based
    for row in range(len(r)):
        cols = ci[rp[row]:rp[row + 1]]
        np.testing.assert_array_equal(A[gid[row], gid[cols]].toarray().ravel(), v[rp[row]:rp[row + 1]])
    assert A.nnz == len(v)
    np.testing.assert_array_equal(b, r)
    np.testing.assert_array_equal(s, x.cpu().numpy())
