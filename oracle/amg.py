"""CPU oracle of the smoothed-aggregation AMG preconditioner (NALU_PREC_MUELU).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product).

The reference delegates this preconditioner to MueLu (TpetraLinearSolver::setMueLu,
src/LinearSolver.C:124-152) with a parameter list such as
reg_tests/xml/matches_ml_default.xml; MueLu is a third-party library absent from
the reference tree (Trilinos, no pinned version: Nalu master tracks Trilinos
develop).  Its published algorithm -- smoothed aggregation (Vanek, Mandel, Brezina
1996) with uncoupled aggregation, a tentative prolongator from the near null
space, prolongator smoothing with omega = 4/3 / lambda_max(D^-1 A), Galerkin
coarse operators, Chebyshev smoothing with lambda_min = lambda_max / ratio
(Ifpack2), a direct coarse solve -- is restated here in float64 numpy/scipy, in
the form the HIP kernels implement it:

  strength:     j strong for row i when a_ij != 0 and a_ij^2 > tol^2 |a_ii a_jj|
                (every nonzero when tol = 0); weak entries are lumped into the
                diagonal of the filtered operator A_f used for smoothing P
  aggregation:  roots = the maximal distance-2 independent set of the strength
                graph selected by fixed hash priorities (synchronous rounds; the
                same set as the greedy pass in decreasing priority order);
                non-roots join the root they are most strongly connected to,
                then rows two steps away join their strongest aggregated
                neighbour (ties: larger priority key)
  P_tent:       P_tent(i, agg(i)) = B_i / |B restricted to agg(i)|, B_c = that norm
  P:            (I - omega D_f^-1 A_f) P_tent; rows without strong entries (or
                with |d_f| <= 1e-12 |a_ii|) keep their P_tent row
  A_c:          P^T A P
  coarse:       Gauss-Jordan (pseudo-)inverse, partial pivoting, pivots below
                1e-10 max(max|a_ii|, max|a_ii above| max_I |P(:, I)|^2) skipped
                (their unknown set to 0); when |A_c B_c| <= 1e-5 scale |B_c| the
                rank-one residual (A_c B_c) B_c^T / |B_c|^2 is removed first
  V-cycle:      Chebyshev(degree k, zero start) ; r = b - A x ; x_c = V(P^T r) ;
                x += P x_c ; Chebyshev(degree k, from x)
  Chebyshev:    Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1 on
                D^-1 A, lambda_max = boost x power-iteration estimate

Parity pins (tests/test_oracle_amg.py): the synchronous MIS-2 rounds equal a
sequential greedy MIS of the squared strength graph; roots are pairwise >= 3
steps apart and every row lies within 2 steps of its aggregate's root; P_tent
has orthonormal columns and P_tent B_c = B; P B_c = B and A_c B_c = 0 for a
zero-row-sum operator; the Chebyshev error polynomial matches the closed form
T_k((theta - lambda)/delta) / T_k(theta/delta) on the eigenvalues of D^-1 A;
the coarse pseudo-inverse equals inv() for a nonsingular matrix and solves
consistent singular systems; the V-cycle is symmetric for a symmetric operator.
Against MueLu itself the hierarchy is parity unpinned (the library cannot be
built or run here).
"""
import numpy as np
import scipy.sparse as sp

M32 = np.uint64(0xFFFFFFFF)


def _u32(x):
    return np.asarray(x, dtype=np.uint64) & M32


def prio_hash(i):
    """Priority hash of row i (31 bits); the same integer arithmetic as the kernels."""
    h = _u32(_u32(i) * np.uint64(0x9E3779B1))
    h = h ^ (h >> np.uint64(15))
    h = _u32(h * np.uint64(0x2C1B3C6D))
    h = h ^ (h >> np.uint64(12))
    return h & np.uint64(0x7FFFFFFF)


def mis_key(state, idx):
    """Key of a row: class (root 3, undecided 2, covered 1) | priority | index."""
    idx = np.asarray(idx, dtype=np.uint64)
    cls = np.where(state == 1, 3, np.where(state == 0, 2, 1)).astype(np.uint64)
    return (cls << np.uint64(62)) | (prio_hash(idx) << np.uint64(31)) | idx


def hash_vector(n):
    """Fixed start vector of the power iteration, entries in (0.5, 1.5)."""
    i = np.arange(n, dtype=np.uint64)
    h = _u32(i * np.uint64(2654435761) + np.uint64(0x9E3779B9))
    h = h ^ (h >> np.uint64(16))
    h = _u32(h * np.uint64(0x85EBCA6B))
    h = h ^ (h >> np.uint64(13))
    return 0.5 + (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(0x1000000)


def csr(rp, ci, v, n, m=None):
    return sp.csr_matrix((np.asarray(v, dtype=np.float64), np.asarray(ci), np.asarray(rp)),
                         shape=(n, n if m is None else m))


def diagonal(A):
    return np.asarray(A.diagonal(), dtype=np.float64)


def strength(A, tol):
    """(strong mask over A's entries in CSR order, lumped weak sums per row)."""
    A = A.tocsr()
    n = A.shape[0]
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    d = diagonal(A)
    a = A.data
    off = A.indices != rows
    st = off & (a != 0.0)
    if tol > 0.0:
        st &= a * a > tol * tol * np.abs(d[rows] * d[A.indices])
    weak = off & ~st
    lump = np.bincount(rows[weak], weights=a[weak], minlength=n)
    return st, lump


def _strong_graph(A, st):
    n = A.shape[0]
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    return sp.csr_matrix((np.ones(int(st.sum())), (rows[st], A.indices[st])), shape=(n, n))


def _row_max(S, vals):
    """max(vals[i], max over S's row-i columns j of vals[j]) for uint64 vals."""
    out = vals.copy()
    nz = np.diff(S.indptr) > 0
    if S.nnz:
        g = vals[S.indices]
        m = np.maximum.reduceat(g, S.indptr[:-1][nz])
        out[nz] = np.maximum(out[nz], m)
    return out


def mis2(A, st):
    """Synchronous distance-2 MIS rounds; returns the state (1 root, 2 covered)."""
    n = A.shape[0]
    S = _strong_graph(A, st)
    state = np.zeros(n, dtype=np.int8)
    idx = np.arange(n, dtype=np.uint64)
    while True:
        key = mis_key(state, idx)
        m2 = _row_max(S, _row_max(S, key))
        und = state == 0
        root = und & (m2 == key)
        cov = und & ~root & ((m2 >> np.uint64(62)) == 3)
        state[root] = 1
        state[cov] = 2
        if not np.any(state == 0):
            return state


def mis2_greedy(A, st):
    """Sequential greedy MIS of the squared strength graph in decreasing key order
    (a second route to the same root set, used to pin mis2)."""
    n = A.shape[0]
    S = _strong_graph(A, st)
    S2 = ((S + sp.identity(n)) @ (S + sp.identity(n))).tocsr()
    key = mis_key(np.zeros(n, dtype=np.int8), np.arange(n, dtype=np.uint64))
    order = np.argsort(key)[::-1]
    blocked = np.zeros(n, dtype=bool)
    root = np.zeros(n, dtype=bool)
    for i in order:
        if blocked[i]:
            continue
        root[i] = True
        blocked[S2.indices[S2.indptr[i]:S2.indptr[i + 1]]] = True
        blocked[i] = True
    return root


def aggregate(A, tol):
    """(aggregate id per row, number of aggregates, state) as the kernels form them."""
    A = A.tocsr()
    n = A.shape[0]
    st, _ = strength(A, tol)
    state = mis2(A, st)
    rid = np.cumsum(state == 1) - 1
    nc = int(np.sum(state == 1))
    key1 = mis_key(np.ones(n, dtype=np.int8), np.arange(n, dtype=np.uint64))
    key2 = mis_key(np.full(n, 2, dtype=np.int8), np.arange(n, dtype=np.uint64))
    agg1 = np.full(n, -1, dtype=np.int64)
    agg1[state == 1] = rid[state == 1]
    for i in np.nonzero(state != 1)[0]:
        lo, hi = A.indptr[i], A.indptr[i + 1]
        best = None
        for q in range(lo, hi):
            j = A.indices[q]
            if not st[q] or state[j] != 1:
                continue
            cand = (abs(A.data[q]), int(key1[j]))
            if best is None or cand > best[0]:
                best = (cand, rid[j])
        if best is not None:
            agg1[i] = best[1]
    agg = agg1.copy()
    for i in np.nonzero(agg1 < 0)[0]:
        lo, hi = A.indptr[i], A.indptr[i + 1]
        best = None
        for q in range(lo, hi):
            j = A.indices[q]
            if not st[q] or agg1[j] < 0:
                continue
            cand = (abs(A.data[q]), int(key2[j]))
            if best is None or cand > best[0]:
                best = (cand, agg1[j])
        if best is None:
            raise ValueError("row %d left without an aggregate" % i)
        agg[i] = best[1]
    return agg, nc, state


def tentative(agg, nc, B):
    n = len(agg)
    Bc = np.sqrt(np.bincount(agg, weights=B * B, minlength=nc))
    v = np.where(Bc[agg] > 0, B / np.where(Bc[agg] > 0, Bc[agg], 1.0), 0.0)
    return sp.csr_matrix((v, (np.arange(n), agg)), shape=(n, nc)), Bc


def smoothed_prolongator(A, Pt, tol, omega):
    A = A.tocsr()
    n = A.shape[0]
    st, lump = strength(A, tol)
    d = diagonal(A)
    df = d + lump
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    has = np.bincount(rows[st], minlength=n) > 0
    use = has & (np.abs(df) > 1e-12 * np.abs(d))
    w = np.where(use, omega / np.where(use, df, 1.0), 0.0)
    sv = np.where(A.indices == rows, 1.0 - w[rows] * df[rows], np.where(st, -w[rows] * A.data, 0.0))
    S = sp.csr_matrix((sv, A.indices, A.indptr), shape=A.shape)
    return (S @ Pt).tocsr()


def power_lmax(A, iters):
    """lambda_max(D^-1 A): x <- x/|x|, y = D^-1 A x, lambda = x . y."""
    invd = 1.0 / diagonal(A)
    x = hash_vector(A.shape[0])
    lam = 0.0
    for _ in range(max(1, iters)):
        nx = np.sqrt(x @ x)
        y = invd * (A @ x) / nx
        lam = (x @ y) / nx
        x = y
    return lam


def coarse_pinv(Ac, rel_tiny=1e-10, scale=None):
    """Gauss-Jordan (pseudo-)inverse with partial pivoting; (X, number of skipped
    pivots).  Pivots at or below rel_tiny * scale are skipped (scale: max|a_ii|
    unless given)."""
    Ad = np.asarray(Ac.todense() if sp.issparse(Ac) else Ac, dtype=np.float64)
    n = Ad.shape[0]
    M = np.hstack([Ad, np.eye(n)])
    tiny = rel_tiny * (np.max(np.abs(np.diag(Ad))) if scale is None else scale)
    ok = np.zeros(n, dtype=bool)
    for k in range(n):
        col = np.abs(M[k:, k])
        p = k + int(np.argmax(col))
        if col[p - k] > tiny:
            if p != k:
                M[[k, p]] = M[[p, k]]
            M[k] /= M[k, k]
            f = M[:, k].copy()
            f[k] = 0.0
            M -= np.outer(f, M[k])
            ok[k] = True
    X = M[:, n:].copy()
    X[~ok] = 0.0
    return X, int(np.sum(~ok))


class Params:
    def __init__(self, max_coarse_size=1000, max_levels=10, cheb_degree=4, cheb_ratio=20.0, cheb_boost=1.1,
                 cheb_eig_iters=30, drop_tol=0.0, sa_damping=4.0 / 3.0):
        self.max_coarse_size = max_coarse_size
        self.max_levels = max_levels
        self.cheb_degree = cheb_degree
        self.cheb_ratio = cheb_ratio
        self.cheb_boost = cheb_boost
        self.cheb_eig_iters = cheb_eig_iters
        self.drop_tol = drop_tol
        self.sa_damping = sa_damping


def hierarchy(A, prm):
    """List of level dicts {A, invd, lmax, agg, P}; the last holds the dense pseudo-inverse X."""
    levels = []
    A = A.tocsr()
    B = np.ones(A.shape[0])
    stalled = False
    while True:
        lev = {"A": A, "invd": 1.0 / diagonal(A), "lmax": power_lmax(A, prm.cheb_eig_iters)}
        levels.append(lev)
        n = A.shape[0]
        if stalled or n <= prm.max_coarse_size or len(levels) >= prm.max_levels:
            scale = np.max(np.abs(diagonal(A)))
            if len(levels) >= 2:
                up = levels[-2]
                colsq = np.asarray(up["P"].multiply(up["P"]).sum(axis=0)).ravel()
                scale = max(scale, np.max(np.abs(diagonal(up["A"]))) * colsq.max())
            y = A @ B
            lev["null_rel"] = np.linalg.norm(y) / (scale * np.linalg.norm(B))
            Ad = A.toarray()
            if lev["null_rel"] <= 1e-5:
                Ad = Ad - np.outer(y, B) / (B @ B)
            lev["X"], lev["free"] = coarse_pinv(Ad, scale=scale)
            return levels
        agg, nc, _ = aggregate(A, prm.drop_tol)
        Pt, Bc = tentative(agg, nc, B)
        P = smoothed_prolongator(A, Pt, prm.drop_tol, prm.sa_damping / lev["lmax"])
        lev["agg"], lev["P"] = agg, P
        A = (P.T @ (A @ P)).tocsr()
        B = Bc
        stalled = nc >= n


def chebyshev(lev, b, x, prm, zero_start):
    A, invd = lev["A"], lev["invd"]
    lmax = lev["lmax"] * prm.cheb_boost
    lmin = lmax / prm.cheb_ratio
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho = 1.0 / sigma
    x = np.zeros_like(b) if zero_start else x.copy()
    r = b if zero_start else b - A @ x
    d = invd * r / theta
    x = x + d
    for _ in range(1, prm.cheb_degree):
        rn = 1.0 / (2.0 * sigma - rho)
        r = b - A @ x
        d = rn * rho * d + (2.0 * rn / delta) * invd * r
        x = x + d
        rho = rn
    return x


def vcycle(levels, b, prm, l=0):
    lev = levels[l]
    if l + 1 == len(levels):
        return lev["X"] @ b
    x = chebyshev(lev, b, None, prm, True)
    r = b - lev["A"] @ x
    xc = vcycle(levels, lev["P"].T @ r, prm, l + 1)
    x = x + lev["P"] @ xc
    return chebyshev(lev, b, x, prm, False)
