"""ctypes driver of the threaded CPU baseline (libcpubaseline.so).

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg (and the test that pins
this baseline against the oracle) use it; the product (nalu_amd) never does.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libcpubaseline.so")
        if not os.path.exists(path):
            raise RuntimeError("CPU baseline not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        I, I64, D = C.c_int, C.c_int64, C.c_double
        L.cb_max_threads.restype = I
        L.cb_set_threads.argtypes = [I]
        L.cb_graph.restype = I64
        L.cb_graph.argtypes = [I64, I64, _i32, _i64, _i32]
        L.cb_expand.argtypes = [I64, _i64, _i32, I, _i64, _i32]
        L.cb_assemble.restype = D
        L.cb_assemble.argtypes = [I64, I64, _i32, _d, _d, _d, _d, _d, _d, D, _d, I64, _i64, _i32, _d, _d, _i64,
                                  _i32, _d, _d]
        L.cb_cg_sgs.restype = I
        L.cb_cg_sgs.argtypes = [I64, _i64, _i32, _d, _d, _d, I, D, _d, _d]
        _LIB = L
    return _LIB


def _p(a):
    assert a.flags.c_contiguous
    return a.ctypes.data_as({np.float64: _d, np.int32: _i32, np.int64: _i64}[a.dtype.type])


class LowMachCase:
    """Graphs and buffers of the low-Mach problem P (a nalu_amd.problems.lowmach_problem)."""

    def __init__(self, P):
        self.P = P
        L = lib()
        X, conn = np.ascontiguousarray(P["coords"]), np.ascontiguousarray(P["conn"], dtype=np.int32)
        self.X, self.conn = X, conn
        nn, E = X.shape[0], conn.shape[0]
        self.nn, self.E = nn, E
        self.rp1 = np.zeros(nn + 1, dtype=np.int64)
        nnz1 = 27 * nn + 64  # upper bound of a Hex8 node graph row count is 27 on a lattice
        ci1 = np.zeros(max(nnz1, 64 * nn), dtype=np.int32)
        nnz = L.cb_graph(nn, E, _p(conn), _p(self.rp1), _p(ci1))
        self.ci1 = np.ascontiguousarray(ci1[:nnz])
        self.rp3 = np.zeros(3 * nn + 1, dtype=np.int64)
        self.ci3 = np.zeros(9 * nnz, dtype=np.int32)
        L.cb_expand(nn, _p(self.rp1), _p(self.ci1), 3, _p(self.rp3), _p(self.ci3))
        self.v1, self.r1 = np.zeros(nnz), np.zeros(nn)
        self.v3, self.r3 = np.zeros(9 * nnz), np.zeros(3 * nn)
        self.mdot = np.zeros((E, 12))
        f = lambda k: np.ascontiguousarray(P[k], dtype=np.float64)
        self.f = {k: f(k) for k in ("rho", "p", "u", "gpdx", "mu")}

    def assemble(self, e0=0, e1=None):
        e1 = self.E if e1 is None else e1
        f = self.f
        return lib().cb_assemble(e0, e1, _p(self.conn), _p(self.X), _p(f["rho"]), _p(f["p"]), _p(f["u"]),
                                 _p(f["gpdx"]), _p(f["mu"]), float(self.P["tau"]), _p(self.mdot), self.nn,
                                 _p(self.rp1), _p(self.ci1), _p(self.v1), _p(self.r1), _p(self.rp3), _p(self.ci3),
                                 _p(self.v3), _p(self.r3))

    def pressure_solve(self, tol=1e-8, maxit=20000):
        """CG + per-thread-block SGS on the zero-mean continuity rhs: (iterations, true rel resid, seconds)."""
        b = self.r1 - self.r1.mean()
        x = np.zeros_like(b)
        res, secs = C.c_double(0), C.c_double(0)
        it = lib().cb_cg_sgs(self.nn, _p(self.rp1), _p(self.ci1), _p(self.v1), _p(b), _p(x), int(maxit), float(tol),
                             C.byref(res), C.byref(secs))
        return it, res.value, secs.value


def set_threads(n):
    lib().cb_set_threads(int(n))


def max_threads():
    return lib().cb_max_threads()
