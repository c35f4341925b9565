"""GPU tests of the drop-in boundary: both LinearSystem::sumInto overloads, the
consolidated element algorithm with a kernel list, and the C++ driver built on
include/nalu_hip.hpp.

Bars: entries within 1e-12 of the oracle relative to each row's largest entry
(1e-9 absolute against the reference's Tpetra gold, its own tolerance); the
scratch arrays of the Kokkos overload hold each element's sorted column LIDs and
the permutation that sorts them; the host-queue overload gives the same system
as the device overload (atomic additions in a different order: 1e-13).
"""
import json
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nalu_amd as na  # noqa: E402
from nalu_amd.problems import heat_problem, lowmach_problem  # noqa: E402
from tests.fixtures import lattice_1x1xN, rel_err_rowwise  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = 1e-12
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def system(coords, conn, nd, ids=None, cfg=None):
    mesh = na.Mesh(coords, conn, ids)
    ls = na.LinearSystem(nd, cfg or na.SolverConfig())
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    return mesh, ls


def dense_from_csr(rp, ci, v, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    return A


def test_sum_into_elements_tpetra_gold(golds):
    print("case: Kokkos-overload sumInto, tpetra 1x1x2 gold", flush=True)
    gold = golds["tpetra_1x1x2"]
    coords, conn = lattice_1x1xN(2)
    mesh, ls = system(coords, conn, 1)
    ev = np.array(gold["elem_vals"]).reshape(8, 8)
    ls.zeroSystem()
    ids = torch.zeros((2, 8), dtype=torch.int32, device=DEV)
    perm = torch.zeros_like(ids)
    ls.sumIntoElements(dev(conn, torch.int32), dev(np.zeros((2, 8))), dev(np.stack([ev, ev])), ids, perm)
    rp, ci, v, _ = ls.csr()
    np.testing.assert_allclose(dense_from_csr(rp, ci, v, 12), np.array(gold["lhs_vals"]).reshape(12, 12),
                               atol=gold["tol"])


@pytest.mark.parametrize("nd", [1, 3])
def test_sum_into_overloads_match_oracle(oracle, nd):
    n = 4 if nd == 1 else 3
    print("case: sumInto overloads, numDof %d, n=%d, permuted ids" % (nd, n), flush=True)
    P = lowmach_problem(n, seed=17, mu=0.05)
    X, conn = P["coords"], P["conn"]
    nn, E = X.shape[0], conn.shape[0]
    ids = np.random.default_rng(5).permutation(nn).astype(np.int64) + 1
    R = 8 * nd
    lhs = np.zeros((E, R, R))
    rhs = np.zeros((E, R))
    md = oracle.compute_mdot(conn, X, P["rho"], P["p"], P["u"], P["gpdx"], P["tau"])
    for e in range(E):
        nd_e = conn[e]
        if nd == 1:
            lhs[e], rhs[e] = oracle.scalar_diff_elem(X[nd_e], P["p"][nd_e], P["rho"][nd_e])
        else:
            lhs[e], rhs[e] = oracle.momentum_advdiff_elem(X[nd_e], P["u"][nd_e], P["mu"][nd_e], md[e])
    mesh, ls = system(X, conn, nd, ids)
    g = oracle.graph_build(nn, conn, nd, node_ids=ids)
    ov = np.zeros(g.nnz)
    orr = np.zeros(g.nrows)
    for e in range(E):
        oracle.sum_into(g, ov, orr, conn[e], lhs[e], rhs[e])
    # device Kokkos overload with caller scratch
    ls.zeroSystem()
    sc_ids = torch.full((E, R), -7, dtype=torch.int32, device=DEV)
    sc_perm = torch.full((E, R), -7, dtype=torch.int32, device=DEV)
    ls.sumIntoElements(dev(conn, torch.int32), dev(rhs), dev(lhs), sc_ids, sc_perm)
    rp, ci, v, r = ls.csr()
    np.testing.assert_array_equal(rp, g.row_ptr)
    np.testing.assert_array_equal(ci, g.col_idx)
    err = rel_err_rowwise(v, ov, g.row_ptr)
    print("  device overload: matrix %.2e" % err, flush=True)
    assert err <= TOL
    np.testing.assert_allclose(r, orr, rtol=0, atol=TOL * np.abs(orr).max())
    # scratch: sorted LIDs and the permutation that sorts them
    lid = g.node_lid
    sid, sperm = sc_ids.cpu().numpy(), sc_perm.cpu().numpy()
    for e in range(E):
        keys = np.array([lid[conn[e][i // nd]] + i % nd for i in range(R)])
        assert np.all(np.diff(sid[e]) > 0)
        np.testing.assert_array_equal(sid[e], keys[sperm[e]])
    # host (std::vector) overload: queued per element, applied before csr()
    ls.zeroSystem()
    for e in range(E):
        ls.sumInto(conn[e], rhs[e], lhs[e])
    _, _, v2, r2 = ls.csr()
    assert rel_err_rowwise(v2, v, g.row_ptr) <= 1e-13
    np.testing.assert_allclose(r2, r, rtol=0, atol=1e-13 * np.abs(r).max())
    # zeroSystem discards queued calls
    ls.zeroSystem()
    ls.sumInto(conn[0], rhs[0], lhs[0])
    ls.zeroSystem()
    _, _, v3, r3 = ls.csr()
    assert not v3.any() and not r3.any()


def test_sum_into_rejects_bad_entities(oracle):
    print("case: sumInto argument checks", flush=True)
    coords, conn = lattice_1x1xN(2)
    mesh, ls = system(coords, conn, 1)
    bad = conn.copy()
    bad[1, 3] = 99
    ids = torch.zeros((2, 8), dtype=torch.int32, device=DEV)
    with pytest.raises(na.NaluError):
        ls.sumIntoElements(dev(bad, torch.int32), dev(np.zeros((2, 8))), dev(np.zeros((2, 8, 8))), ids, ids.clone())
    with pytest.raises(na.NaluError):
        ls.sumInto(np.array([0, 1, 99]), np.zeros(3), np.zeros(9))
    with pytest.raises(na.NaluError):
        ls.sumInto(conn[0], np.zeros(7), np.zeros(64))


def test_elem_algorithm_continuity(oracle):
    n, dt, gam = 6, 0.05, (1.5, -2.0, 0.5)
    print("case: AssembleElemSolverAlgorithm continuity adv + mass, n=%d, BDF2" % n, flush=True)
    P = lowmach_problem(n, seed=23, rho_sigma=0.3)
    X, conn = P["coords"], P["conn"]
    nn = X.shape[0]
    mesh, ls = system(X, conn, 1)
    rho3 = [P["rho"] * s for s in (0.97, 1.0, 1.02)]
    rho_d = [dev(a) for a in rho3]
    fields = (dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]))
    alg = na.AssembleElemSolverAlgorithm(ls, mesh)
    alg.add_kernel(na.ContinuityAdvElemKernel(*fields, shifted_grad_op=True))
    alg.add_kernel(na.ContinuityMassElemKernel(*rho_d))
    alg.set_time_integrator(dt, gam)
    ls.zeroSystem()
    alg.execute()
    ls.loadComplete()
    g = oracle.graph_build(nn, conn, 1)
    ov, orr = oracle.assemble_continuity(g, conn, X, P["rho"], P["p"], P["u"], P["gpdx"], dt / gam[0],
                                         shift_poisson=True)
    mv, mr = oracle.assemble_continuity_mass(g, conn, X, rho3, dt, np.array(gam))
    rp, ci, v, r = ls.csr()
    err = rel_err_rowwise(v, ov + mv, g.row_ptr)
    print("  matrix %.2e" % err, flush=True)
    assert err <= TOL
    np.testing.assert_allclose(r, orr + mr, rtol=0, atol=TOL * np.abs(orr + mr).max())


def test_elem_algorithm_momentum(oracle):
    n, dt, gam = 5, 0.1, (1.0, -1.0, 0.0)
    print("case: AssembleElemSolverAlgorithm momentum adv-diff + mass, n=%d" % n, flush=True)
    P = lowmach_problem(n, seed=29, mu=0.05)
    X, conn = P["coords"], P["conn"]
    nn, E = X.shape[0], conn.shape[0]
    mesh, ls = system(X, conn, 3)
    rho, u = dev(P["rho"]), dev(P["u"])
    md = torch.zeros((E, 12), dtype=torch.float64, device=DEV)
    na.compute_mdot(mesh, rho, dev(P["p"]), u, dev(P["gpdx"]), md, tau=P["tau"])
    u_old = P["u"] * 0.9
    alg = na.AssembleElemSolverAlgorithm(ls, mesh)
    alg.add_kernel(na.MomentumAdvDiffElemKernel(u, dev(P["mu"]), md, skew_symmetric=True))
    alg.add_kernel(na.MomentumMassElemKernel(rho, rho, rho, dev(u_old), dev(u_old), u, dev(P["gpdx"])))
    alg.set_time_integrator(dt, gam)
    ls.zeroSystem()
    alg.execute()
    ls.loadComplete()
    g = oracle.graph_build(nn, conn, 3)
    ov, orr = oracle.assemble_momentum(g, conn, X, P["u"], P["mu"], md.cpu().numpy(), skew=True)
    mv, mr = oracle.assemble_momentum_mass(g, conn, X, (P["rho"],) * 3, (u_old, u_old, P["u"]), P["gpdx"], dt,
                                           np.array(gam))
    rp, ci, v, r = ls.csr()
    err = rel_err_rowwise(v, ov + mv, g.row_ptr)
    print("  matrix %.2e" % err, flush=True)
    assert err <= TOL
    np.testing.assert_allclose(r, orr + mr, rtol=0, atol=TOL * np.abs(orr + mr).max())


def test_elem_algorithm_rejects_mismatches():
    print("case: AssembleElemSolverAlgorithm argument checks", flush=True)
    P = heat_problem(3, seed=2)
    mesh, ls = system(P["coords"], P["conn"], 1)
    alg = na.AssembleElemSolverAlgorithm(ls, mesh)
    u = dev(np.zeros((P["coords"].shape[0], 3)))
    with pytest.raises(na.NaluError):  # numDof 3 kernel on a scalar system
        alg.add_kernel(na.MomentumAdvDiffElemKernel(u, dev(P["k"]), dev(np.zeros((27, 12)))))
    with pytest.raises(na.NaluError):  # wrong field size
        alg.add_kernel(na.ScalarDiffElemKernel(u, dev(P["k"])))
    with pytest.raises(na.NaluError):
        alg.set_time_integrator(0.0)


def test_heat_cond_driver():
    """The C++ driver (include/nalu_hip.hpp, the reference's call order) on the 16^3 heat case."""
    exe = os.path.join(ROOT, "nalu_amd", "heat_cond_driver")
    if not os.path.exists(exe):
        pytest.fail("heat_cond_driver is not built (run __graft_entry__.build())")
    print("case: C++ heat_cond_driver n=16", flush=True)
    out = subprocess.run([exe, "16", "0.1", "1234"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    print("  " + json.dumps(rec), flush=True)
    assert rec["status"] == 0 and rec["iterations"] > 0
    assert rec["linear_residual"] <= 1e-8 * rec["nonlinear_residual"] * 1.0001
    # maximum principle of the steady heat solution between the Dirichlet values 0 and 1
    assert -1e-6 <= rec["T_min"] and rec["T_max"] <= 1.0 + 1e-6
