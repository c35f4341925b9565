"""Pins the CPU oracle of the learned preconditioner (oracle/gnn.py).

The network has no reference counterpart; its oracle is pinned by
mathematics: central finite differences of the loss pin the analytic
gradient, the zero-correction network must return its SGS input exactly,
a one-layer network with hand-set weights must return one Jacobi-weighted
neighbour average, the operator A = D (I - W) must reproduce the CSR
matrix-vector product, and the loss must vanish when the output is the exact
error and lie in [0, 1] otherwise.
"""
import numpy as np
import pytest

from oracle import gnn as G
from oracle import pyoracle as O

from nalu_amd import lattice
from nalu_amd.problems import lowmach_problem


@pytest.fixture(scope="module")
def system():
    P = lowmach_problem(4, rho_sigma=0.3)
    g = O.graph_build(P["coords"].shape[0], P["conn"])
    vals, rhs = O.assemble_continuity(g, P["conn"], P["coords"], P["rho"], P["p"], P["u"], P["gpdx"])
    return g, vals, G.Graph(g.row_ptr, g.col_idx, vals)


def _inputs(gr, B=2, seed=3):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((B, gr.n)).astype(np.float32).astype(np.float64)
    zs = rng.standard_normal((B, gr.n)).astype(np.float32).astype(np.float64)
    e = rng.standard_normal((B, gr.n)).astype(np.float32).astype(np.float64)
    # a consistent pair needs <e, r> > 0: take r = A e (+ the constant-free part)
    r = gr.aop(e)
    return r, zs, e


def test_aop_matches_csr(system):
    g, vals, gr = system
    x = np.random.default_rng(1).standard_normal(gr.n)
    y = O.spmv(g.row_ptr, g.col_idx, vals, x)
    assert np.linalg.norm(gr.aop(x[None])[0] - y) <= 1e-6 * np.linalg.norm(y)
    # (A + A^T) x against the explicit transpose
    import scipy.sparse as sp
    A = sp.csr_matrix((vals, g.col_idx, g.row_ptr), shape=(gr.n, gr.n))
    ys = (A + A.T) @ x
    assert np.linalg.norm(gr.aop_sym(x[None])[0] - ys) <= 1e-6 * np.linalg.norm(ys)


def test_layout_sizes():
    o = G.layout(4)
    assert o["total"] == 32 * 3 + 32 + 4 * (32 * 64 + 32) + 32 + 2 == 8482


@pytest.mark.parametrize("rnd", [False, True])
def test_zero_correction_is_sgs(system, rnd):
    _, _, gr = system
    th = G.init_params(4, seed=5)
    r, zs, _ = _inputs(gr)
    z = G.apply(th, gr, r, zs, rnd=rnd)
    # the pressure system is pure Neumann: the output is projected off the constants
    assert gr.neumann
    want = zs - zs.mean(axis=-1, keepdims=True)
    assert np.allclose(z, want, rtol=1e-12, atol=1e-12 * np.abs(zs).max())


def test_no_projection_without_a_null_space(system):
    """Only a pure-Neumann system (every row summing to zero) has its output
    projected: with the diagonal raised the zero-correction network is the
    base feature exactly."""
    g, vals, _ = system
    v2 = vals.copy()
    rp, ci = g.row_ptr, g.col_idx
    for i in range(len(rp) - 1):
        for k in range(rp[i], rp[i + 1]):
            if ci[k] == i:
                v2[k] *= 1.5
    gr2 = G.Graph(rp, ci, v2)
    assert not gr2.neumann
    th = G.init_params(4, seed=5)
    r, zs, _ = _inputs(gr2)
    z = G.apply(th, gr2, r, zs, rnd=False)
    assert np.allclose(z, zs, rtol=1e-12, atol=1e-12 * np.abs(zs).max())


def test_one_layer_neighbour_average(system):
    """h0[0] = f0, layer: u[1] = m[0], out = h1[1] => z = W D^-1 r where positive."""
    g, vals, gr = system
    o = G.layout(1)
    th = np.zeros(o["total"])
    th[o["w_in"] + 0 * 3 + 0] = 1.0                    # h0[0] = f0
    th[o["w0"] + 1 * 64 + 32 + 0] = 1.0                # u[1] = m[0]
    th[o["w_out"] + 1] = 1.0                           # out = h1[1]
    n = gr.n
    r = np.abs(np.random.default_rng(2).standard_normal((1, n))) + 0.1   # positive residual -> relu inactive
    z = G.apply(th, gr, r, np.zeros_like(r), num_layers=1, rnd=False)[0]
    # closed form from the CSR: z_i = sum_{j != i} (-a_ij / a_ii) r_j / a_jj
    rp, ci = g.row_ptr, g.col_idx
    d = np.array([vals[rp[i]:rp[i + 1]][ci[rp[i]:rp[i + 1]] == i][0] for i in range(n)])
    want = np.zeros(n)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            j = ci[k]
            if j != i:
                want[i] += (-vals[k] / d[i]) * r[0, j] / d[j]
    want -= want.mean()  # pure-Neumann system: projected off the constants
    assert np.linalg.norm(z - want) <= 1e-6 * np.linalg.norm(want)


def test_loss_bounds(system):
    _, _, gr = system
    r, zs, e = _inputs(gr)
    # exact error -> zero loss (the optimally scaled direction is e itself)
    L, _ = G.loss(e, gr, r, e)
    assert abs(L) < 1e-12
    L2, _ = G.loss(3.7 * e, gr, r, e)
    assert abs(L2) < 1e-12
    th = G.init_params(4, seed=1)
    out, _ = G.forward(th, gr, r, zs, rnd=False)
    L3, _ = G.loss(out, gr, r, e)
    assert 0.0 <= L3 <= 1.0


def test_gradient_matches_finite_differences(system):
    _, _, gr = system
    L = 2
    o = G.layout(L)
    rng = np.random.default_rng(7)
    th = G.init_params(L, seed=11)
    th[o["w_out"]:o["c"]] = rng.normal(0, 0.3, 32)     # make every gradient non-trivial
    th[o["c"]] = 0.2
    for l in range(L):
        th[o[f"b{l}"]:o[f"b{l}"] + 32] = rng.normal(0, 0.1, 32)
    r, zs, e = _inputs(gr, B=2, seed=4)
    f, g = G.loss_and_grad(th, gr, r, e, zs, num_layers=L, rnd=False)
    idx = np.concatenate([rng.choice(o["total"], 40, replace=False),
                          [o["c"], o["c"] + 1, o["w_out"] + 3, o["b_in"] + 5, o["w_in"] + 7]])
    for k in idx:
        h = 1e-6 * max(1.0, abs(th[k]))
        tp = th.copy()
        tp[k] += h
        tm = th.copy()
        tm[k] -= h
        fp = G.loss(G.forward(tp, gr, r, zs, L, rnd=False)[0], gr, r, e)[0]
        fm = G.loss(G.forward(tm, gr, r, zs, L, rnd=False)[0], gr, r, e)[0]
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, np.abs(g).max()), (k, fd, g[k])
