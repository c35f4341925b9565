"""The sampled-row oracle used by the full-size GPU tests must reproduce the
oracle's whole-mesh assembly row for row (CPU only)."""
import numpy as np

from nalu_amd.problems import lowmach_problem
from tests.sampling import compare_rows, node_incidence, oracle_rows, sample_nodes


def test_sampled_rows_equal_whole_mesh_assembly(oracle):
    n = 5
    P = lowmach_problem(n, seed=17, rho_sigma=0.3, mu=0.05)
    conn, X = P["conn"], P["coords"]
    nn = X.shape[0]
    md = oracle.compute_mdot(conn, X, P["rho"], P["p"], P["u"], P["gpdx"])
    inc = node_incidence(conn, nn)
    nodes = sample_nodes(n, 60, seed=3)
    assert nodes.size > 40 and nodes.min() == 0 and nodes.max() == nn - 1
    # momentum, numDof 3
    g3 = oracle.graph_build(nn, conn, 3)
    v3, r3 = oracle.assemble_momentum(g3, conn, X, P["u"], P["mu"], md)
    exp3 = oracle_rows(nodes, inc, conn, 3, lambda e: oracle.momentum_advdiff_elem(
        X[conn[e]], P["u"][conn[e]], P["mu"][conn[e]], md[e]))
    err, err_r, k = compare_rows(exp3, g3.row_ptr, g3.col_idx, v3, r3)
    assert k == 3 * nodes.size and err <= 1e-13 and err_r <= 1e-13 * np.abs(r3).max()
    # continuity, numDof 1
    g1 = oracle.graph_build(nn, conn, 1)
    v1, r1 = oracle.assemble_continuity(g1, conn, X, P["rho"], P["p"], P["u"], P["gpdx"])
    exp1 = oracle_rows(nodes, inc, conn, 1, lambda e: oracle.continuity_adv_elem(
        X[conn[e]], P["rho"][conn[e]], P["p"][conn[e]], P["u"][conn[e]], P["gpdx"][conn[e]]))
    err, err_r, k = compare_rows(exp1, g1.row_ptr, g1.col_idx, v1, r1)
    assert k == nodes.size and err <= 1e-13 and err_r <= 1e-13 * np.abs(r1).max()
    # a planted mistake (one wrong element entry) must be caught
    bad = v3.copy()
    row = 3 * int(nodes[5])
    bad[g3.row_ptr[row]] += 1e-6 * np.abs(v3).max()
    assert compare_rows(exp3, g3.row_ptr, g3.col_idx, bad, r3)[0] > 1e-9
