"""The allocation-free Kokkos sumInto entry (nalu_linsys_sum_into_elements_noalloc)
against the allocating one (nalu_linsys_sum_into_elements), through the C ABI:
two calls in a row, the second with fewer elements, each into a freshly zeroed
system; entries within 1e-13 of each row's largest entry (atomic additions in a
different order), and the size limit rejected with an error."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nalu_amd as na  # noqa: E402
from nalu_amd.problems import lowmach_problem  # noqa: E402
from tests.fixtures import rel_err_rowwise  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _noalloc():
    L = na.lib()
    f = L.nalu_linsys_sum_into_elements_noalloc
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return f


def test_noalloc_sum_into_matches_allocating_entry(oracle):
    print("case: no-alloc sumInto vs sumInto, momentum n=5, two calls", flush=True)
    P = lowmach_problem(5, seed=19, mu=0.05)
    X, conn = P["coords"], P["conn"]
    E = conn.shape[0]
    md = oracle.compute_mdot(conn, X, P["rho"], P["p"], P["u"], P["gpdx"], P["tau"])
    lhs = np.zeros((E, 24, 24))
    rhs = np.zeros((E, 24))
    for e in range(E):
        lhs[e], rhs[e] = oracle.momentum_advdiff_elem(X[conn[e]], P["u"][conn[e]], P["mu"][conn[e]], md[e])
    mesh = na.Mesh(X, conn)
    ls = na.LinearSystem(3, na.SolverConfig())
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    rp = ls.csr()[0]
    f = _noalloc()
    ents = torch.as_tensor(conn, dtype=torch.int32, device=DEV).contiguous()
    L_d = torch.as_tensor(lhs, device=DEV).contiguous()
    R_d = torch.as_tensor(rhs, device=DEV).contiguous()
    for ne in (E, E // 3):
        ls.zeroSystem()
        ids = torch.zeros((ne, 24), dtype=torch.int32, device=DEV)
        ls.sumIntoElements(ents[:ne], R_d[:ne], L_d[:ne], ids, ids.clone())
        _, _, v_ref, r_ref = ls.csr()
        ls.zeroSystem()
        rc = f(ls.handle, ne, 8, ents.data_ptr(), R_d.data_ptr(), L_d.data_ptr())
        assert rc == 0, na.lib().nalu_last_error()
        _, _, v, r = ls.csr()
        err = rel_err_rowwise(v, v_ref, rp)
        print("  %d elements: matrix %.2e" % (ne, err), flush=True)
        assert err <= 1e-13
        np.testing.assert_allclose(r, r_ref, rtol=0, atol=1e-13 * np.abs(r_ref).max())
    # beyond the scratch limit: rejected before any launch
    big = (32 << 20) // (2 * 4 * 24) + 1
    assert f(ls.handle, big, 8, ents.data_ptr(), R_d.data_ptr(), L_d.data_ptr()) == -1  # NALU_ERR_ARG
    assert b"scratch limit" in na.lib().nalu_last_error()
