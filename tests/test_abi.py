"""C-ABI library checks that need no GPU: load, exported symbols, host-side
generators and argument validation."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import nalu_amd as na
from tests.fixtures import py_lcg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nalu_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nalu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = na.lib()
    names = declared_symbols()
    assert len(names) >= 39
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_is_gfx950():
    blob = open(na.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_lattice_numbering_and_jitter_follow_lcg():
    n, jit, seed = 3, 0.1, 1234
    coords, conn = na.lattice(n, jitter=jit, seed=seed)
    n1 = n + 1
    draws = py_lcg(seed, 3 * n1 ** 3).reshape(-1, 3)
    h = 1.0 / n
    for node in range(n1 ** 3):
        ijk = np.array([node % n1, (node // n1) % n1, node // (n1 * n1)])
        base = ijk * h
        interior = np.all((ijk > 0) & (ijk < n))
        expect = base + (jit * h * draws[node] if interior else 0.0)
        np.testing.assert_allclose(coords[node], expect, atol=1e-15)
    # element 0 = nodes 1 2 4 3 5 6 8 7 in 1-based ids (unit_tests/UnitTestTpetra.C:97-98) for n1 = 2
    c2, k2 = na.lattice(1)
    np.testing.assert_array_equal(k2[0] + 1, [1, 2, 4, 3, 5, 6, 8, 7])
    # every element has positive volume orientation: x1-x0, x3-x0, x4-x0 right-handed
    e = coords[conn[:, [1, 3, 4]]] - coords[conn[:, [0]]]
    assert np.all(np.linalg.det(e) > 0)


def test_stretched_lattice_aspect():
    coords, _ = na.lattice(64, jitter=0.0, stretch=1.2, max_aspect=100.0)
    x = np.unique(np.round(coords[:, 0], 14))
    dx = np.diff(x)
    assert abs(dx.max() / dx.min() - 100.0) < 1e-6
    assert abs(x[0]) < 1e-15 and abs(x[-1] - 1.0) < 1e-15
    np.testing.assert_allclose(dx, dx[::-1], rtol=0, atol=1e-13)  # symmetric grading


def test_normal_and_lognormal_fields():
    z = na.normal_field(200000, 1.0, 99)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    k = na.lognormal_field(1000, 0.5, 5)
    np.testing.assert_allclose(k, np.exp(na.normal_field(1000, 0.5, 5)), rtol=1e-15)


def test_config_validation_without_gpu():
    with pytest.raises(na.NaluError):
        na.LinearSystem(0)
    with pytest.raises(KeyError):
        na.SolverConfig(method="bicgstab")._struct()
    ls = na.LinearSystem(3, na.SolverConfig("cg", "jacobi"))
    with pytest.raises(na.NaluError, match="not finalized"):
        ls.zeroSystem()
    with pytest.raises(na.NaluError, match="no graph"):
        ls.finalizeLinearSystem()


def test_mesh_validation_without_gpu():
    # every check runs on the host before any device call
    coords, conn = na.lattice(2)
    with pytest.raises(na.NaluError, match="bad arguments"):
        na.Mesh(coords, conn[:0])  # no elements
    bad = conn.copy()
    bad[3, 5] = coords.shape[0]
    with pytest.raises(na.NaluError, match="out of range"):
        na.Mesh(coords, bad)
    bad[3, 5] = -1
    with pytest.raises(na.NaluError, match="out of range"):
        na.Mesh(coords, bad)
    ids = np.arange(1, coords.shape[0] + 1, dtype=np.int64)
    with pytest.raises(na.NaluError, match="one id per node"):
        na.Mesh(coords, conn, node_ids=ids[:-1])
    ids[0] = 0
    with pytest.raises(na.NaluError, match="ids must be in"):
        na.Mesh(coords, conn, node_ids=ids)
    with pytest.raises(na.NaluError, match=r"\[N,3\]"):
        na.Mesh(coords[:, :2], conn)


def test_new_entry_points_reject_null_handles_without_gpu():
    # the argument checks run before any device call
    L = na.lib()
    g = np.array([1.0, -1.0, 0.0])
    gp = g.ctypes.data_as(na._dp)
    assert L.nalu_assemble_momentum_with_mass(None, None, None, None, None, 0.0, 0, 0, None, None, None, None, None,
                                              None, None, 1.0, gp, 0, 0) < 0
    assert L.nalu_assemble_momentum_with_mass_segregated(None, None, None, None, None, 0.0, 0, 0, None, None, None,
                                                         None, None, None, None, 1.0, gp, 0, 0, None) < 0
    assert L.nalu_gnn_data_set_pair(None, None, 0, None, None) < 0


def test_product_does_not_import_oracle():
    # the shipped package never reaches the checker
    for dirpath, _, files in os.walk(os.path.join(ROOT, "nalu_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                txt = open(os.path.join(dirpath, f), errors="replace").read()
                assert "pyoracle" not in txt and "liboracle" not in txt and "or_graph_build" not in txt, f


def test_learned_preconditioner_requires_fgmres():
    # a nonlinear preconditioner breaks CG and non-flexible GMRES recurrences
    for method in ("gmres", "cg"):
        with pytest.raises(na.NaluError, match="fgmres"):
            na.LinearSystem(1, na.SolverConfig(method, "learned"))
    na.LinearSystem(1, na.SolverConfig("fgmres", "learned"))


def test_device_array_validation_without_gpu():
    torch = pytest.importorskip("torch")
    ok = torch.zeros(6, dtype=torch.float64)
    with pytest.raises(na.NaluError, match="dtype"):
        na._dev(torch.zeros(6, dtype=torch.float32), "x", "f64", numel=6)
    with pytest.raises(na.NaluError, match="entries"):
        na._dev(ok, "x", "f64", numel=7)
    with pytest.raises(na.NaluError, match="shape"):
        na._dev(ok, "x", "f64", shape=(2, 4))
    with pytest.raises(na.NaluError, match="contiguous"):
        na._dev(torch.zeros((4, 6), dtype=torch.float64).t(), "x", "f64", numel=24)
    with pytest.raises(na.NaluError, match="int32"):
        na._dev(torch.zeros(3, dtype=torch.int64), "nodes", "i32")
    with pytest.raises(na.NaluError, match="missing"):
        na._dev(None, "x", "f64")
    assert na._dev(None, "x", "f64", optional=True) is None
    with pytest.raises(na.NaluError, match="raw address"):
        na._dev(12345, "x", "f64")
    # a well-formed host tensor passes every check but the device one
    with pytest.raises(na.NaluError, match="device"):
        na._dev(ok.reshape(2, 3), "x", "f64", shape=(2, 3))
