"""Shared inputs of the reference's unit-test fixtures, restated as data.

generated:1x1x1 unit cube (unit_tests/UnitTestUtils.C:30-35) and the
TrigFieldFunction fields of unit_tests/kernels/UnitTestKernelUtils.C:23-275
(constants :238-271).  Components the reference never sets are zero (the
fields are declared with a null initial value).
"""
import numpy as np

PI = np.arccos(-1.0)
A_TRIG = 0.3

UNIT_HEX = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)


def trig_temperature(X):
    # (k/4)(cos 2 aT pi x + cos 2 aT pi y + cos 2 aT pi z), k = aT = 1
    return 0.25 * (np.cos(2 * PI * X[:, 0]) + np.cos(2 * PI * X[:, 1]) + np.cos(2 * PI * X[:, 2]))


def trig_velocity(X):
    u = np.zeros((X.shape[0], 3))
    u[:, 0] = -np.cos(A_TRIG * PI * X[:, 0]) * np.sin(A_TRIG * PI * X[:, 1])
    u[:, 1] = np.sin(A_TRIG * PI * X[:, 0]) * np.cos(A_TRIG * PI * X[:, 1])
    return u


def trig_pressure(X):
    return -0.25 * (np.cos(2 * A_TRIG * PI * X[:, 0]) + np.cos(2 * A_TRIG * PI * X[:, 1]))


def trig_dpdx(X):
    g = np.zeros((X.shape[0], 3))
    g[:, 0] = 0.5 * A_TRIG * PI * np.sin(2 * A_TRIG * PI * X[:, 0])
    g[:, 1] = 0.5 * A_TRIG * PI * np.sin(2 * A_TRIG * PI * X[:, 1])
    return g


def lattice_1x1xN(nz):
    """generated:1x1xN mesh: node id = 1 + i + 2(j + 2k), elements in Hex8 order."""
    coords = np.array([[i, j, k] for k in range(nz + 1) for j in range(2) for i in range(2)], dtype=np.float64)
    nid = lambda i, j, k: i + 2 * (j + 2 * k)
    conn = np.array([[nid(0, 0, k), nid(1, 0, k), nid(1, 1, k), nid(0, 1, k),
                      nid(0, 0, k + 1), nid(1, 0, k + 1), nid(1, 1, k + 1), nid(0, 1, k + 1)]
                     for k in range(nz)], dtype=np.int32)
    return coords, conn


def py_lcg(seed, count):
    """The reference's machine-independent LCG (unit_tests/UnitTestUtils.C:41-53)."""
    out = np.empty(count)
    prev = seed
    for i in range(count):
        prev = (1103515245 * prev + 12345) % 2147483647
        out[i] = 2.0 * (prev / 2147483647.0) - 1.0
    return out


def rel_err_rowwise(a, b, row_ptr):
    """max_i max_j |a_ij - b_ij| / max_j |b_ij|: entry error relative to its row scale."""
    worst = 0.0
    d = np.abs(a - b)
    mag = np.abs(b)
    row_of = np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr))
    scale = np.zeros(len(row_ptr) - 1)
    np.maximum.at(scale, row_of, mag)
    scale[scale == 0] = 1.0
    worst = np.max(d / scale[row_of]) if d.size else 0.0
    return float(worst)
