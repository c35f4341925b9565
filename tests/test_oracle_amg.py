"""Pins of the AMG oracle (CPU): each check reaches the answer by a route other
than the oracle's own formula -- a sequential greedy MIS for the synchronous
MIS-2 rounds, graph distances for the aggregate structure, orthonormality and
null-space identities for the prolongators, the closed-form Chebyshev error
polynomial on the eigenvalues, numpy's inverse for the coarse solve, and the
symmetry and convergence a symmetric V-cycle must show."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import amg as M


def lattice_laplacian(k, seed=0, shift=0.0):
    """Graph Laplacian of the 27-point lattice k^3 with random positive weights
    (zero row sums), plus shift * I."""
    rng = np.random.default_rng(seed)
    idx = np.arange(k ** 3).reshape(k, k, k)
    rows, cols = [], []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if (dx, dy, dz) == (0, 0, 0):
                    continue
                a = idx[max(0, -dx):k - max(0, dx), max(0, -dy):k - max(0, dy), max(0, -dz):k - max(0, dz)]
                b = idx[max(0, dx):k - max(0, -dx) or None, max(0, dy):k - max(0, -dy) or None,
                        max(0, dz):k - max(0, -dz) or None]
                rows.append(a.ravel())
                cols.append(b.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    keep = r < c
    r, c = r[keep], c[keep]
    w = rng.uniform(0.5, 1.5, len(r))
    n = k ** 3
    W = sp.csr_matrix((np.concatenate([w, w]), (np.concatenate([r, c]), np.concatenate([c, r]))), shape=(n, n))
    L = sp.diags(np.asarray(W.sum(axis=1)).ravel()) - W
    return (L + shift * sp.identity(n)).tocsr()


def test_hashes_are_32bit_and_in_range():
    i = np.arange(1000, dtype=np.uint64)
    p = M.prio_hash(i)
    assert p.max() < 2 ** 31 and len(np.unique(p)) > 990
    v = M.hash_vector(1000)
    assert v.min() > 0.5 - 1e-12 and v.max() < 1.5


def test_mis2_rounds_equal_greedy_mis_of_squared_graph():
    for tol, seed in ((0.0, 1), (0.08, 2)):
        A = lattice_laplacian(9, seed=seed)
        st, _ = M.strength(A, tol)
        state = M.mis2(A, st)
        assert np.array_equal(state == 1, M.mis2_greedy(A, st))


def test_aggregate_structure():
    A = lattice_laplacian(10, seed=3)
    st, _ = M.strength(A, 0.0)
    agg, nc, state = M.aggregate(A, 0.0)
    n = A.shape[0]
    assert agg.min() == 0 and agg.max() == nc - 1 and len(np.unique(agg)) == nc
    roots = np.nonzero(state == 1)[0]
    assert len(roots) == nc and np.array_equal(agg[roots], np.arange(nc))
    S = M._strong_graph(A, st) + sp.identity(n)
    S2 = (S @ S).tocsr()
    # roots at least three steps apart
    R = S2[roots][:, roots]
    assert R.nnz == len(roots)
    # every row within two steps of its aggregate's root
    for i in range(n):
        assert roots[agg[i]] in S2.indices[S2.indptr[i]:S2.indptr[i + 1]]
    # a 27-point lattice coarsens by roughly 3 per direction
    assert 8 <= n / nc <= 40


def test_tentative_prolongator_orthonormal_and_reproduces_B():
    A = lattice_laplacian(8, seed=4)
    agg, nc, _ = M.aggregate(A, 0.0)
    B = np.random.default_rng(5).uniform(0.5, 2.0, A.shape[0])
    Pt, Bc = M.tentative(agg, nc, B)
    assert np.allclose((Pt.T @ Pt).toarray(), np.eye(nc), atol=1e-14)
    assert np.allclose(Pt @ Bc, B, rtol=1e-14)


def test_smoothed_prolongator_keeps_constants_and_coarse_null_space():
    A = lattice_laplacian(9, seed=6)
    for tol in (0.0, 0.03, 0.08):
        agg, nc, _ = M.aggregate(A, tol)
        Pt, Bc = M.tentative(agg, nc, np.ones(A.shape[0]))
        lmax = M.power_lmax(A, 30)
        P = M.smoothed_prolongator(A, Pt, tol, (4.0 / 3.0) / lmax)
        assert np.allclose(P @ Bc, 1.0, atol=1e-12)
        Ac = P.T @ A @ P
        assert np.abs(Ac @ Bc).max() <= 1e-12 * abs(Ac).max()
        # Galerkin product against a dense product
        Ad = P.toarray().T @ A.toarray() @ P.toarray()
        assert np.abs(Ac.toarray() - Ad).max() <= 1e-12 * np.abs(Ad).max()


def test_power_iteration_brackets_lambda_max():
    A = lattice_laplacian(6, seed=7, shift=0.5)
    d = A.diagonal()
    Ds = sp.diags(1.0 / np.sqrt(d))
    lam = np.linalg.eigvalsh((Ds @ A @ Ds).toarray()).max()
    est = M.power_lmax(A, 30)
    assert 0.9 * lam <= est <= lam * (1 + 1e-12)


def test_chebyshev_matches_closed_form_error_polynomial():
    A = lattice_laplacian(5, seed=8, shift=0.3)
    n = A.shape[0]
    d = A.diagonal()
    Ds = np.diag(1.0 / np.sqrt(d))
    lam, V = np.linalg.eigh(Ds @ A.toarray() @ Ds)
    prm = M.Params(cheb_degree=4, cheb_ratio=20.0, cheb_boost=1.1)
    lev = {"A": A, "invd": 1.0 / d, "lmax": lam.max()}
    b = np.random.default_rng(9).standard_normal(n)
    x = M.chebyshev(lev, b, None, prm, True)
    lmax = lam.max() * 1.1
    lmin = lmax / 20.0
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    k = prm.cheb_degree
    T = np.polynomial.chebyshev.Chebyshev.basis(k)
    p = T((theta - lam) / delta) / T(theta / delta)
    e0 = np.linalg.solve(A.toarray(), b)
    # p(D^-1 A) = D^-1/2 V p(L) V^T D^1/2
    ek = np.diag(1.0 / np.sqrt(d)) @ V @ np.diag(p) @ V.T @ np.diag(np.sqrt(d)) @ e0
    assert np.linalg.norm((e0 - x) - ek) <= 1e-10 * np.linalg.norm(e0)
    # the same polynomial from a nonzero start
    x0 = np.random.default_rng(10).standard_normal(n)
    x1 = M.chebyshev(lev, b, x0, prm, False)
    ek = np.diag(1.0 / np.sqrt(d)) @ V @ np.diag(p) @ V.T @ np.diag(np.sqrt(d)) @ (e0 - x0)
    assert np.linalg.norm((e0 - x1) - ek) <= 1e-10 * np.linalg.norm(e0)


def test_coarse_pseudo_inverse():
    rng = np.random.default_rng(11)
    G = rng.standard_normal((40, 40))
    Aspd = G @ G.T + 40 * np.eye(40)
    X, free = M.coarse_pinv(Aspd)
    assert free == 0 and np.allclose(X, np.linalg.inv(Aspd), rtol=1e-10, atol=1e-13)
    L = lattice_laplacian(4, seed=12).toarray()
    X, free = M.coarse_pinv(L)
    assert free == 1
    b = rng.standard_normal(L.shape[0])
    b -= b.mean()
    assert np.linalg.norm(L @ (X @ b) - b) <= 1e-9 * np.linalg.norm(b)


def test_vcycle_symmetric_and_converges():
    A = lattice_laplacian(12, seed=13, shift=1e-3)
    prm = M.Params(max_coarse_size=20, cheb_degree=2)
    levels = M.hierarchy(A, prm)
    assert len(levels) >= 3
    rng = np.random.default_rng(14)
    u, v = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    Mu, Mv = M.vcycle(levels, u, prm), M.vcycle(levels, v, prm)
    assert abs(Mu @ v - u @ Mv) <= 1e-12 * abs(Mu @ v)
    b = rng.standard_normal(A.shape[0])
    its = []
    Mop = spla.LinearOperator(A.shape, matvec=lambda r: M.vcycle(levels, r, prm))
    x, info = spla.cg(A, b, M=Mop, rtol=1e-8, maxiter=100, callback=lambda xk: its.append(1))
    assert info == 0 and len(its) <= 20
    assert np.linalg.norm(b - A @ x) <= 1e-8 * np.linalg.norm(b)


def test_single_aggregate_coarse_level_is_treated_as_null():
    # a pure-Neumann operator coarsened to one aggregate: the 1x1 Galerkin
    # operator is round-off, which the scale from the level above must expose
    A = lattice_laplacian(6, seed=1)
    prm = M.Params(max_coarse_size=1, cheb_degree=2)
    levels = M.hierarchy(A, prm)
    assert levels[-1]["A"].shape[0] == 1 and levels[-1]["free"] == 1
    b = np.random.default_rng(2).standard_normal(A.shape[0])
    b -= b.mean()
    z = M.vcycle(levels, b, prm)
    assert np.all(np.isfinite(z)) and np.linalg.norm(z) < 1e3 * np.linalg.norm(b) / A.diagonal().min()


def test_near_singular_coarse_operator_is_deflated():
    # row sums that vanish only to FP32 round-off (an FP32-assembled Neumann
    # operator): without removing the residual rank-one part the coarse solve
    # would add a near-null component of size 1 / round-off
    A = lattice_laplacian(10, seed=5)
    rng = np.random.default_rng(6)
    A = A.copy()
    A.data = A.data * (1.0 + 6e-8 * rng.standard_normal(A.nnz))
    prm = M.Params(max_coarse_size=60, cheb_degree=2)
    levels = M.hierarchy(A, prm)
    top = levels[-1]
    assert 1e-12 < top["null_rel"] <= 1e-5 and top["free"] == 1
    b = rng.standard_normal(A.shape[0])
    b -= b.mean()
    z = M.vcycle(levels, b, prm)
    assert np.linalg.norm(z) < 1e3 * np.linalg.norm(b) / A.diagonal().min()
