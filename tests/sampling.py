"""Row sampling for parity checks at sizes the whole-mesh oracle cannot afford.

A sampled row is assembled from scratch by the oracle: every element incident
to the row's node is evaluated with the oracle's element kernel and the element
row is summed into a dictionary keyed by global column, exactly what
sum_into_row does for that row (src/TpetraLinearSystem.C:1415-1453).  The
result is compared with the same row of the device CSR.  Meshes here carry no
explicit ids, so node a owns rows numDof*a .. numDof*a+numDof-1 and columns are
numDof*node + component.
"""
import numpy as np


def node_incidence(conn, num_nodes):
    """CSR of node -> (element, local index) pairs."""
    flat = conn.ravel().astype(np.int64)
    order = np.argsort(flat, kind="stable")
    ptr = np.zeros(num_nodes + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(np.bincount(flat, minlength=num_nodes))
    return ptr, (order // 8).astype(np.int64), (order % 8).astype(np.int64)


def sample_nodes(n, count, seed, boundary_share=0.25):
    """Node indices of an (n+1)^3 lattice: the 8 corners, then a seeded mix of
    boundary (face / edge) and interior nodes, sorted and unique."""
    n1 = n + 1
    rng = np.random.default_rng(seed)
    corners = [i + n1 * (j + n1 * k) for i in (0, n) for j in (0, n) for k in (0, n)]
    nb = int(count * boundary_share)
    ijk = rng.integers(0, n1, size=(nb, 3))
    face = rng.integers(0, 3, size=nb)
    ijk[np.arange(nb), face] = rng.choice([0, n], size=nb)
    bnd = ijk[:, 0] + n1 * (ijk[:, 1] + n1 * ijk[:, 2])
    inner = rng.integers(1, n, size=(count - nb - 8, 3))
    inn = inner[:, 0] + n1 * (inner[:, 1] + n1 * inner[:, 2])
    return np.unique(np.concatenate([corners, bnd, inn]).astype(np.int64))


def oracle_rows(nodes, inc, conn, num_dof, elem_fn):
    """Expected rows of the sampled nodes.  elem_fn(e) -> (lhs [R][R], rhs [R])
    with R = 8 num_dof, the oracle's element matrix of element e (cached here).
    Returns {row: (dict column -> value, rhs value)}."""
    ptr, elems, local = inc
    cache = {}
    out = {}
    for a in nodes:
        acc = [dict() for _ in range(num_dof)]
        rr = np.zeros(num_dof)
        for q in range(ptr[a], ptr[a + 1]):
            e, la = int(elems[q]), int(local[q])
            if e not in cache:
                cache[e] = elem_fn(e)
            lhs, rhs = cache[e]
            for d in range(num_dof):
                row = lhs[num_dof * la + d]
                rr[d] += rhs[num_dof * la + d]
                for c in range(8):
                    node = int(conn[e, c])
                    for j in range(num_dof):
                        col = num_dof * node + j
                        acc[d][col] = acc[d].get(col, 0.0) + row[num_dof * c + j]
        for d in range(num_dof):
            out[num_dof * int(a) + d] = (acc[d], rr[d])
    return out


def compare_rows(expect, row_ptr, col_idx, values, rhs):
    """(max row-relative matrix error, max |rhs error|, rows compared); raises
    AssertionError when a row's column set differs from the oracle's."""
    worst, worst_r = 0.0, 0.0
    for row, (cols, r) in expect.items():
        s, t = int(row_ptr[row]), int(row_ptr[row + 1])
        ci = col_idx[s:t]
        ck = np.array(sorted(cols))
        assert np.array_equal(ci, ck), "row %d: columns differ from the oracle" % row
        ev = np.array([cols[c] for c in ck])
        scale = max(np.abs(ev).max(), 1e-300)
        worst = max(worst, float(np.abs(values[s:t] - ev).max() / scale))
        worst_r = max(worst_r, abs(float(rhs[row]) - r))
    return worst, worst_r, len(expect)
