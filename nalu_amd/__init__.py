"""MI355X-native Hex8 CVFEM assembly + linear solve (Nalu hot path).

Python host binding over the C ABI of ``libnalu_hip.so`` (``include/nalu_hip.h``).
The classes mirror the reference's interfaces:

* :class:`LinearSystem` -- ``sierra::nalu::LinearSystem`` (include/LinearSystem.h:39-167)
  as implemented by TpetraLinearSystem (src/TpetraLinearSystem.C); method names
  follow the reference (``buildElemToNodeGraph``, ``finalizeLinearSystem``,
  ``zeroSystem``, ``sumInto``, ``applyDirichletBCs``, ``loadComplete``, ``solve``).
* :class:`SolverConfig` -- the ``linear_solvers`` YAML block
  (src/LinearSolverConfig.C:34-108).
* :func:`assemble_scalar_diff` / :func:`compute_mdot` / :func:`assemble_continuity` /
  :func:`assemble_momentum` -- the consolidated element algorithms with
  ScalarDiffElemKernel, ComputeMdotElemAlgorithm, ContinuityAdvElemKernel and
  MomentumAdvDiffElemKernel.

Device arrays are torch CUDA tensors (HIP on ROCm); torch only provides memory
and streams.  All compute runs in the HIP kernels of libnalu_hip.so; there is no
CPU fallback: without the library every entry point raises.
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "lib", "available", "NaluError", "SolverConfig", "Mesh", "LinearSystem", "lattice", "normal_field",
    "lognormal_field", "assemble_scalar_diff", "compute_mdot", "assemble_continuity", "assemble_momentum",
    "METHOD", "PRECOND", "PRECISION", "LearnedPreconditioner", "HarvestedPairs", "MueLuParams",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnalu_hip.so")

METHOD = {"gmres": 0, "cg": 1, "fgmres": 2}
PRECOND = {"none": 0, "jacobi": 1, "sgs": 2, "mt_sgs": 2, "learned": 3, "muelu": 4}
PRECISION = {"fp64": 0, "mixed": 1}

_i64 = C.c_int64
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class NaluError(RuntimeError):
    pass


class _MueLuParams(C.Structure):
    _fields_ = [("max_coarse_size", C.c_int), ("max_levels", C.c_int), ("cheb_degree", C.c_int),
                ("cheb_ratio", C.c_double), ("cheb_boost", C.c_double), ("cheb_eig_iters", C.c_int),
                ("drop_tol", C.c_double), ("sa_damping", C.c_double)]


class _KernelDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("shifted_grad_op", C.c_int), ("shift_mdot", C.c_int), ("reduced_sens", C.c_int),
                ("skew_symmetric", C.c_int), ("lumped", C.c_int), ("balanced_force", C.c_int),
                ("include_divu", C.c_double), ("field", C.c_void_p * 8)]


class _Config(C.Structure):
    _fields_ = [("method", C.c_int), ("precond", C.c_int), ("tolerance", C.c_double),
                ("final_tolerance", C.c_double), ("max_iterations", C.c_int), ("kspace", C.c_int),
                ("precision", C.c_int), ("final_outer_iter", C.c_int)]


_L = None


def lib():
    """Load libnalu_hip.so (built in-tree by __graft_entry__.build())."""
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise NaluError("libnalu_hip.so is not built (run __graft_entry__.build()): %s" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.nalu_last_error.restype = C.c_char_p
        for name in ("nalu_linsys_linear_residual", "nalu_linsys_nonlinear_residual",
                     "nalu_linsys_scaled_nonlinear_residual"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [_vp]
        L.nalu_linsys_iterations.argtypes = [_vp]
        L.nalu_lattice_generate.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint32, _dp, _i32p]
        L.nalu_normal_field.argtypes = [_i64, C.c_double, C.c_uint32, _dp]
        L.nalu_lognormal_field.argtypes = [_i64, C.c_double, C.c_uint32, _dp]
        L.nalu_mesh_create.argtypes = [_i64, _i64, _i32p, _dp, _i64p, C.POINTER(_vp)]
        L.nalu_mesh_destroy.argtypes = [_vp]
        L.nalu_mesh_device_ptrs.argtypes = [_vp, C.POINTER(_vp), C.POINTER(_vp)]
        L.nalu_linsys_create.argtypes = [C.c_int, C.POINTER(_Config), C.POINTER(_vp)]
        L.nalu_linsys_destroy.argtypes = [_vp]
        L.nalu_linsys_set_stream.argtypes = [_vp, _vp]
        L.nalu_linsys_build_elem_to_node_graph.argtypes = [_vp, _vp]
        L.nalu_linsys_finalize.argtypes = [_vp]
        L.nalu_linsys_zero.argtypes = [_vp]
        L.nalu_linsys_sum_into_batch.argtypes = [_vp, _i64, _vp, _vp, _vp]
        L.nalu_linsys_sum_into.argtypes = [_vp, C.c_int, _i32p, _dp, _dp]
        L.nalu_linsys_sum_into_elements.argtypes = [_vp, _i64, C.c_int, _vp, _vp, _vp, _vp, _vp]
        L.nalu_linsys_apply_dirichlet.argtypes = [_vp, _i64, _vp, _vp, _vp, C.c_int, C.c_int]
        L.nalu_linsys_reset_rows.argtypes = [_vp, _i64, _vp, C.c_int, C.c_int, C.c_double, C.c_double]
        L.nalu_linsys_load_complete.argtypes = [_vp]
        L.nalu_linsys_remove_rhs_mean.argtypes = [_vp]
        L.nalu_linsys_solve.argtypes = [_vp, _vp]
        L.nalu_linsys_set_l2_scaling.argtypes = [_vp, C.c_double]
        L.nalu_linsys_sizes.argtypes = [_vp, _i64p, _i64p]
        L.nalu_linsys_copy_csr.argtypes = [_vp, _i64p, _i32p, _dp, _dp]
        L.nalu_linsys_device_ptrs.argtypes = [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]
        L.nalu_linsys_coloring.argtypes = [_vp, C.POINTER(C.c_int), _i64p, _i32p]
        L.nalu_linsys_phase_times.argtypes = [_vp, _dp]
        L.nalu_linsys_spmv.argtypes = [_vp, _vp, _vp]
        L.nalu_linsys_spmv_f32.argtypes = [_vp, _vp, _vp]
        L.nalu_linsys_precond_setup.argtypes = [_vp]
        L.nalu_linsys_precond_apply.argtypes = [_vp, _vp, _vp]
        L.nalu_linsys_precond_apply_permuted.argtypes = [_vp, _vp, _vp]
        L.nalu_linsys_precond_apply_permuted_f32.argtypes = [_vp, _vp, _vp]
        L.nalu_linsys_spmv_permuted.argtypes = [_vp, _vp, _vp]
        L.nalu_linsys_permutation.argtypes = [_vp, _i32p]
        L.nalu_linsys_set_muelu_params.argtypes = [_vp, C.POINTER(_MueLuParams)]
        L.nalu_amg_info.argtypes = [_vp, C.POINTER(C.c_int), _i64p, _i64p, _dp, C.c_int, _dp, C.POINTER(C.c_int),
                                    _dp]
        L.nalu_amg_level_csr.argtypes = [_vp, C.c_int, C.c_int, _i32p, _i32p, _dp, _i64p, _i64p, _i64p]
        L.nalu_amg_aggregates.argtypes = [_vp, C.c_int, _i32p]
        L.nalu_linsys_solve_rhs.argtypes = [_vp, _vp, _vp, C.POINTER(C.c_int), _dp]
        L.nalu_assemble_scalar_diff.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int]
        L.nalu_compute_mdot.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, _vp, _vp]
        L.nalu_compute_mdot_f32.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, _vp, _vp]
        L.nalu_assemble_continuity_mdot.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int,
                                                    C.c_int, _vp]
        L.nalu_assemble_continuity.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int,
                                               C.c_int, C.c_int]
        L.nalu_assemble_momentum.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, C.c_int]
        L.nalu_assemble_momentum_mass.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _dp,
                                                  C.c_int, C.c_int]
        L.nalu_assemble_continuity_mass.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double, _dp, C.c_int]
        L.nalu_assemble_momentum_with_mass.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, _vp,
                                                       _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _dp, C.c_int,
                                                       C.c_int]
        L.nalu_assemble_momentum_with_mass_segregated.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int,
                                                                  C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                                  C.c_double, _dp, C.c_int, C.c_int, _vp]
        L.nalu_assemble_momentum_segregated.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, _vp]
        L.nalu_assemble_momentum_mass_segregated.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double,
                                                             _dp, C.c_int, C.c_int, _vp]
        L.nalu_linsys_solve_multi.argtypes = [_vp, C.c_int, _vp, _vp, _i32p, _dp]
        L.nalu_linsys_solve_stacked.argtypes = [_vp, C.c_int, _vp, _vp, C.POINTER(C.c_int), _dp]
        L.nalu_gnn_train_epochs.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                            C.c_uint64, _dp]
        L.nalu_compute_dual_volume.argtypes = [_vp, _vp, _vp]
        L.nalu_compute_nodal_gradient.argtypes = [_vp, _vp, _vp, C.c_int, _vp, _vp]
        L.nalu_compute_nodal_gradient_interior.argtypes = [_vp, _vp, _vp, C.c_int, _vp, _vp]
        L.nalu_nodal_gradient_boundary.argtypes = [_vp, _vp, _vp, C.c_int, _vp, _vp]
        L.nalu_mesh_exposed_faces.argtypes = [_vp, _i64p, _i32p]
        # learned preconditioner
        _fp = C.POINTER(C.c_float)
        L.nalu_gnn_create.argtypes = [C.c_int, C.c_int, C.c_uint64, C.POINTER(_vp)]
        L.nalu_gnn_destroy.argtypes = [_vp]
        L.nalu_gnn_num_params.argtypes = [_vp]
        L.nalu_gnn_num_params.restype = C.c_int64
        L.nalu_gnn_num_layers.argtypes = [_vp]
        L.nalu_gnn_get_params.argtypes = [_vp, _fp, _i64]
        L.nalu_gnn_set_params.argtypes = [_vp, _fp, _i64]
        L.nalu_gnn_get_grad.argtypes = [_vp, _fp, _i64]
        L.nalu_gnn_apply.argtypes = [_vp, _vp, _vp, _vp]
        L.nalu_gnn_loss_grad.argtypes = [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp, _dp, _fp]
        L.nalu_gnn_harvest.argtypes = [_vp, C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_int, C.POINTER(_vp)]
        L.nalu_gnn_data_destroy.argtypes = [_vp]
        L.nalu_gnn_data_size.argtypes = [_vp]
        L.nalu_gnn_data_cg_iterations.argtypes = [_vp, C.POINTER(C.c_int), C.c_int]
        L.nalu_gnn_data_pair.argtypes = [_vp, C.c_int, _vp, _vp]
        L.nalu_gnn_data_set_pair.argtypes = [_vp, _vp, C.c_int, _vp, _vp]
        L.nalu_gnn_train.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_double, C.c_uint64, _dp]
        L.nalu_linsys_set_learned_precond.argtypes = [_vp, _vp]
        L.nalu_linsys_set_learned_base.argtypes = [_vp, C.c_int]
        L.nalu_linsys_set_write_matrix_files.argtypes = [_vp, C.c_char_p]
        L.nalu_linsys_learned_fallback.argtypes = [_vp]
        L.nalu_elem_algorithm_create.argtypes = [_vp, _vp, C.POINTER(_vp)]
        L.nalu_elem_algorithm_destroy.argtypes = [_vp]
        L.nalu_elem_algorithm_add_kernel.argtypes = [_vp, C.POINTER(_KernelDesc)]
        L.nalu_elem_algorithm_set_time_integrator.argtypes = [_vp, C.c_double, _dp]
        L.nalu_elem_algorithm_execute.argtypes = [_vp]
        _L = L
    return _L


def available():
    """True when the library is built and a HIP device is usable."""
    try:
        return bool(lib().nalu_device_available())
    except (NaluError, OSError):
        return False


def _check(rc, what=""):
    if rc < 0:
        msg = lib().nalu_last_error().decode(errors="replace")
        raise NaluError("%s failed (%d): %s" % (what, rc, msg))
    return rc


_DTYPE_NAMES = {"f64": "float64", "f32": "float32", "i32": "int32"}


def _dev(t, what, dt, numel=None, shape=None, optional=False):
    """Device pointer of a torch tensor after the checks the kernels rely on:
    element type dt ("f64", "f32" or "i32"), the number of entries (numel) or
    the exact shape, contiguous layout, device memory.  Every wrapper validates
    its arrays here before any launch, so that a caller mistake raises
    NaluError instead of letting a kernel index past a buffer."""
    if t is None:
        if optional:
            return None
        raise NaluError("%s: array missing" % what)
    if isinstance(t, int):
        raise NaluError("%s: a torch tensor is expected, not a raw address" % what)
    name = getattr(getattr(t, "dtype", None), "__repr__", lambda: "")()
    if not name.endswith(_DTYPE_NAMES[dt]):
        raise NaluError("%s: dtype %s, expected torch.%s" % (what, getattr(t, "dtype", type(t)), _DTYPE_NAMES[dt]))
    if shape is not None and tuple(t.shape) != tuple(int(s) for s in shape):
        raise NaluError("%s: shape %s, expected %s" % (what, tuple(t.shape), tuple(int(s) for s in shape)))
    if numel is not None and t.numel() != int(numel):
        raise NaluError("%s: %d entries, expected %d" % (what, t.numel(), int(numel)))
    if not t.is_contiguous():
        raise NaluError("%s: expected a contiguous tensor" % what)
    if not t.is_cuda:
        raise NaluError("%s: expected a device (cuda) tensor" % what)
    return t.data_ptr()


def _np(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------------ generators
def lattice(n, jitter=0.0, stretch=1.0, max_aspect=0.0, seed=1):
    """Structured Hex8 lattice on the unit cube: (coords [(n+1)^3, 3], conn [n^3, 8])."""
    nn = (n + 1) ** 3
    coords = np.zeros((nn, 3), dtype=np.float64)
    conn = np.zeros((n ** 3, 8), dtype=np.int32)
    _check(lib().nalu_lattice_generate(int(n), float(jitter), float(stretch), float(max_aspect), int(seed),
                                       coords.ctypes.data_as(_dp), conn.ctypes.data_as(_i32p)), "lattice")
    return coords, conn


def normal_field(n, sigma=1.0, seed=1):
    out = np.zeros(int(n), dtype=np.float64)
    _check(lib().nalu_normal_field(int(n), float(sigma), int(seed), out.ctypes.data_as(_dp)), "normal_field")
    return out


def lognormal_field(n, sigma=1.0, seed=1):
    out = np.zeros(int(n), dtype=np.float64)
    _check(lib().nalu_lognormal_field(int(n), float(sigma), int(seed), out.ctypes.data_as(_dp)), "lognormal_field")
    return out


# ------------------------------------------------------------------ mesh
class Mesh:
    """Device-resident Hex8 block (coords [N,3], conn [E,8], optional global ids)."""

    def __init__(self, coords, conn, node_ids=None):
        coords = _np(coords, np.float64)
        conn = _np(conn, np.int32)
        if coords.ndim != 2 or coords.shape[1] != 3 or conn.ndim != 2 or conn.shape[1] != 8:
            raise NaluError("Mesh: coords must be [N,3] and conn [E,8]")
        self.num_nodes = coords.shape[0]
        self.num_elems = conn.shape[0]
        ids = None if node_ids is None else _np(node_ids, np.int64)
        if ids is not None and ids.shape != (self.num_nodes,):
            raise NaluError("Mesh: node_ids must hold one id per node")
        self.has_ids = ids is not None
        h = _vp()
        _check(lib().nalu_mesh_create(self.num_nodes, self.num_elems, conn.ctypes.data_as(_i32p),
                                      coords.ctypes.data_as(_dp),
                                      None if ids is None else ids.ctypes.data_as(_i64p), C.byref(h)), "mesh_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def device_ptrs(self):
        cp, kp = _vp(), _vp()
        _check(lib().nalu_mesh_device_ptrs(self._h, C.byref(cp), C.byref(kp)), "mesh_device_ptrs")
        return cp.value, kp.value

    def exposed_faces(self):
        """[nfaces, 4] node indices of the block's exposed sides (HexSCS side order)."""
        n = C.c_int64(0)
        _check(lib().nalu_mesh_exposed_faces(self._h, C.byref(n), None), "mesh_exposed_faces")
        faces = np.zeros((n.value, 4), dtype=np.int32)
        if n.value:
            _check(lib().nalu_mesh_exposed_faces(self._h, C.byref(n), faces.ctypes.data_as(_i32p)),
                   "mesh_exposed_faces")
        return faces

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _L is not None:
            _L.nalu_mesh_destroy(h)
            self._h = None


# ------------------------------------------------------------------ solver config
class SolverConfig:
    """The ``linear_solvers`` block: method, preconditioner, tolerance, final_tolerance,
    max_iterations, kspace (src/LinearSolverConfig.C:34-108), plus the precision of
    the Krylov inner loop."""

    def __init__(self, method="gmres", preconditioner="sgs", tolerance=1e-8, final_tolerance=None,
                 max_iterations=1000, kspace=50, precision="fp64", final_outer_iter=False):
        self.method = method
        self.preconditioner = preconditioner
        self.tolerance = tolerance
        self.final_tolerance = tolerance if final_tolerance is None else final_tolerance
        self.max_iterations = max_iterations
        self.kspace = kspace
        self.precision = precision
        self.final_outer_iter = final_outer_iter

    def _struct(self):
        return _Config(METHOD[self.method], PRECOND[self.preconditioner], float(self.tolerance),
                       float(self.final_tolerance), int(self.max_iterations), int(self.kspace),
                       PRECISION[self.precision], int(bool(self.final_outer_iter)))


class MueLuParams:
    """The MueLu parameter-list keys the AMG preconditioner honours; defaults are
    reg_tests/xml/matches_ml_default.xml (milestone.xml: cheb_degree=2,
    cheb_eig_iters=15, drop_tol=0.02)."""

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

    def _struct(self):
        return _MueLuParams(int(self.max_coarse_size), int(self.max_levels), int(self.cheb_degree),
                            float(self.cheb_ratio), float(self.cheb_boost), int(self.cheb_eig_iters),
                            float(self.drop_tol), float(self.sa_damping))


# ------------------------------------------------------------------ linear system
class LinearSystem:
    """HIP linear system with the reference's LinearSystem call sequence."""

    def __init__(self, num_dof=1, config=None):
        self.num_dof = num_dof
        self.config = config or SolverConfig()
        self._cfg = self.config._struct()
        h = _vp()
        _check(lib().nalu_linsys_create(int(num_dof), C.byref(self._cfg), C.byref(h)), "linsys_create")
        self._h = h
        self.mesh = None

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _L is not None:
            _L.nalu_linsys_destroy(h)
            self._h = None

    def _rows(self):
        """Rows of the finalized system (cached)."""
        n = getattr(self, "_nrows", None)
        if n is None:
            n = self._nrows = self.sizes()[0]
        return n

    def _nodes(self):
        if self.mesh is None:
            raise NaluError("no graph was built")
        return self.mesh.num_nodes

    # --- construction
    def buildElemToNodeGraph(self, mesh):
        self.mesh = mesh
        _check(lib().nalu_linsys_build_elem_to_node_graph(self._h, mesh.handle), "buildElemToNodeGraph")

    def finalizeLinearSystem(self):
        _check(lib().nalu_linsys_finalize(self._h), "finalizeLinearSystem")

    # --- assembly
    def zeroSystem(self):
        _check(lib().nalu_linsys_zero(self._h), "zeroSystem")

    def sumInto(self, entities, rhs, lhs):
        """std::vector overload: one element, host arrays (entities = node indices,
        rhs [R], lhs [R][R] row-major, R = len(entities) * numDof).  Queued; the
        queue is applied on the device before the system is next used."""
        ents = _np(entities, np.int32).ravel()
        R = ents.shape[0] * self.num_dof
        rhs = _np(rhs, np.float64).ravel()
        lhs = _np(lhs, np.float64).ravel()
        if ents.shape[0] < 1 or rhs.shape[0] != R or lhs.shape[0] != R * R:
            raise NaluError("sumInto: %d entities need rhs [%d] and lhs [%d][%d]" % (ents.shape[0], R, R, R))
        _check(lib().nalu_linsys_sum_into(self._h, ents.shape[0], ents.ctypes.data_as(_i32p),
                                          rhs.ctypes.data_as(_dp), lhs.ctypes.data_as(_dp)), "sumInto")

    def sumIntoElements(self, entities, rhs, lhs, local_ids, sort_permutation):
        """Kokkos overload on the device, batched: entities [E, n] int32 node
        indices, rhs [E, R], lhs [E, R, R] (R = n numDof), local_ids and
        sort_permutation [E, R] int32 caller scratch (overwritten with each
        element's sorted column LIDs and their permutation)."""
        if entities.dim() != 2:
            raise NaluError("sumIntoElements: entities must be [E, n]")
        E, n = int(entities.shape[0]), int(entities.shape[1])
        R = n * self.num_dof
        _check(lib().nalu_linsys_sum_into_elements(
            self._h, E, n, _dev(entities, "sumIntoElements entities", "i32", shape=(E, n)),
            _dev(rhs, "sumIntoElements rhs", "f64", shape=(E, R)),
            _dev(lhs, "sumIntoElements lhs", "f64", shape=(E, R, R)),
            _dev(local_ids, "sumIntoElements localIds", "i32", shape=(E, R)),
            _dev(sort_permutation, "sumIntoElements sortPermutation", "i32", shape=(E, R))), "sumIntoElements")

    def sumIntoBatch(self, lhs, rhs, num_elems=None):
        """Kokkos overload batched over the graph mesh's elements: lhs [E,R,R], rhs [E,R] device tensors."""
        _, conn = self.mesh.device_ptrs()
        ne = self.mesh.num_elems if num_elems is None else int(num_elems)
        if not 0 < ne <= self.mesh.num_elems:
            raise NaluError("sumIntoBatch: num_elems must be in [1, %d]" % self.mesh.num_elems)
        R = 8 * self.num_dof
        _check(lib().nalu_linsys_sum_into_batch(self._h, ne, conn, _dev(lhs, "sumIntoBatch lhs", "f64", shape=(ne, R, R)),
                                                _dev(rhs, "sumIntoBatch rhs", "f64", shape=(ne, R))), "sumIntoBatch")

    def applyDirichletBCs(self, bc_nodes, solution, bc_values, begin_pos=0, end_pos=None):
        end_pos = self.num_dof if end_pos is None else end_pos
        nf = self._nodes() * self.num_dof
        _check(lib().nalu_linsys_apply_dirichlet(self._h, int(bc_nodes.numel()),
                                                 _dev(bc_nodes, "applyDirichletBCs nodes", "i32"),
                                                 _dev(solution, "applyDirichletBCs solution", "f64", numel=nf),
                                                 _dev(bc_values, "applyDirichletBCs bc values", "f64", numel=nf),
                                                 begin_pos, end_pos), "applyDirichletBCs")

    def resetRows(self, nodes, begin_pos=0, end_pos=None, diag_value=1.0, rhs_residual=0.0):
        end_pos = self.num_dof if end_pos is None else end_pos
        _check(lib().nalu_linsys_reset_rows(self._h, int(nodes.numel()), _dev(nodes, "resetRows nodes", "i32"),
                                            begin_pos, end_pos, float(diag_value), float(rhs_residual)), "resetRows")

    def loadComplete(self):
        _check(lib().nalu_linsys_load_complete(self._h), "loadComplete")

    def removeRhsMean(self):
        """Project the rhs of a pure-Neumann pressure system onto the matrix range."""
        _check(lib().nalu_linsys_remove_rhs_mean(self._h), "remove_rhs_mean")

    def set_stream(self, stream_handle):
        """Run this system's work on a HIP stream (int handle, e.g. torch's cuda_stream)."""
        _check(lib().nalu_linsys_set_stream(self._h, stream_handle), "set_stream")

    # --- solve
    def solve(self, delta_out=None):
        """Returns 0 when converged, 1 when max_iterations was reached."""
        d = _dev(delta_out, "solve delta", "f64", numel=self._rows(), optional=True)
        return _check(lib().nalu_linsys_solve(self._h, d), "solve")

    def solve_multi(self, B, X):
        """k right-hand sides (rows of B [k][n]) sharing the matrix; X [k][n] holds the
        initial guesses.  Returns (worst status, iterations per rhs, rel. residual per rhs)."""
        k = int(B.shape[0]) if B.dim() == 2 else 0
        if k < 1:
            raise NaluError("solve_multi: B must be [k][rows]")
        shp = (k, self._rows())
        its = (C.c_int32 * k)()
        rel = (C.c_double * k)()
        st = _check(lib().nalu_linsys_solve_multi(self._h, k, _dev(B, "solve_multi B", "f64", shape=shp),
                                                  _dev(X, "solve_multi X", "f64", shape=shp), its, rel), "solve_multi")
        return st, list(its), list(rel)

    def solve_stacked(self, B, X):
        """The rows of B [k][n] solved as one block-diagonal stacked system in one
        Krylov space (one matrix read per SpMV / SGS sweep for all k).  Returns
        (status, Krylov iterations, FP64 true relative residual per rhs)."""
        k = int(B.shape[0]) if B.dim() == 2 else 0
        if k < 1:
            raise NaluError("solve_stacked: B must be [k][rows]")
        shp = (k, self._rows())
        it = C.c_int(0)
        rel = (C.c_double * k)()
        st = _check(lib().nalu_linsys_solve_stacked(self._h, k, _dev(B, "solve_stacked B", "f64", shape=shp),
                                                    _dev(X, "solve_stacked X", "f64", shape=shp), C.byref(it), rel),
                    "solve_stacked")
        return st, it.value, list(rel)

    def solve_rhs(self, b, x):
        it = C.c_int(0)
        rel = C.c_double(0)
        n = self._rows()
        st = _check(lib().nalu_linsys_solve_rhs(self._h, _dev(b, "solve_rhs b", "f64", numel=n),
                                                _dev(x, "solve_rhs x", "f64", numel=n), C.byref(it), C.byref(rel)),
                    "solve_rhs")
        return st, it.value, rel.value

    def linearSolveIterations(self):
        return lib().nalu_linsys_iterations(self._h)

    def linearResidual(self):
        return lib().nalu_linsys_linear_residual(self._h)

    def nonLinearResidual(self):
        return lib().nalu_linsys_nonlinear_residual(self._h)

    def scaledNonLinearResidual(self):
        return lib().nalu_linsys_scaled_nonlinear_residual(self._h)

    # --- introspection
    def sizes(self):
        nr, nz = C.c_int64(0), C.c_int64(0)
        _check(lib().nalu_linsys_sizes(self._h, C.byref(nr), C.byref(nz)), "sizes")
        return nr.value, nz.value

    def csr(self):
        """Host copies (row_ptr int64, col_idx int32, values, rhs)."""
        nr, nz = self.sizes()
        rp = np.zeros(nr + 1, dtype=np.int64)
        ci = np.zeros(nz, dtype=np.int32)
        v = np.zeros(nz, dtype=np.float64)
        r = np.zeros(nr, dtype=np.float64)
        _check(lib().nalu_linsys_copy_csr(self._h, rp.ctypes.data_as(_i64p), ci.ctypes.data_as(_i32p),
                                          v.ctypes.data_as(_dp), r.ctypes.data_as(_dp)), "copy_csr")
        return rp, ci, v, r

    def device_ptrs(self):
        a, b, c, d = _vp(), _vp(), _vp(), _vp()
        _check(lib().nalu_linsys_device_ptrs(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)), "device_ptrs")
        return a.value, b.value, c.value, d.value

    def coloring(self):
        nr, _ = self.sizes()
        nc = C.c_int(0)
        _check(lib().nalu_linsys_coloring(self._h, C.byref(nc), None, None), "coloring")
        cp = np.zeros(nc.value + 1, dtype=np.int64)
        cr = np.zeros(nr, dtype=np.int32)
        _check(lib().nalu_linsys_coloring(self._h, C.byref(nc), cp.ctypes.data_as(_i64p), cr.ctypes.data_as(_i32p)),
               "coloring")
        return cp, cr

    def phase_times(self):
        t = np.zeros(7, dtype=np.float64)
        _check(lib().nalu_linsys_phase_times(self._h, t.ctypes.data_as(_dp)), "phase_times")
        return dict(zip(["init", "zero", "assemble", "dirichlet", "load_complete", "precond", "solve"], t))

    def spmv(self, x, y):
        n = self._rows()
        _check(lib().nalu_linsys_spmv(self._h, _dev(x, "spmv x", "f64", numel=n), _dev(y, "spmv y", "f64", numel=n)),
               "spmv")

    def spmv_f32(self, x, y):
        n = self._rows()
        _check(lib().nalu_linsys_spmv_f32(self._h, _dev(x, "spmv_f32 x", "f32", numel=n),
                                          _dev(y, "spmv_f32 y", "f32", numel=n)), "spmv_f32")

    def precond_setup(self):
        _check(lib().nalu_linsys_precond_setup(self._h), "precond_setup")

    def precond_apply(self, r, z):
        n = self._rows()
        _check(lib().nalu_linsys_precond_apply(self._h, _dev(r, "precond_apply r", "f64", numel=n),
                                               _dev(z, "precond_apply z", "f64", numel=n)), "precond_apply")

    def permutation(self):
        """Solver row order: new row k is original row perm[k]."""
        nr, _ = self.sizes()
        perm = np.zeros(nr, dtype=np.int32)
        _check(lib().nalu_linsys_permutation(self._h, perm.ctypes.data_as(_i32p)), "permutation")
        return perm

    def set_muelu_params(self, params):
        """AMG parameters (a MueLuParams) for the next preconditioner set-up."""
        _check(lib().nalu_linsys_set_muelu_params(self._h, C.byref(params._struct())), "set_muelu_params")

    def amg_info(self):
        """{'rows', 'nnz', 'lmax' per level, 'setup_s', 'coarse_free', 'coarse_null_rel'} of the AMG hierarchy."""
        nl = C.c_int(0)
        cap = 32
        rows = np.zeros(cap, dtype=np.int64)
        nnz = np.zeros(cap, dtype=np.int64)
        lmax = np.zeros(cap, dtype=np.float64)
        st = C.c_double(0.0)
        nf = C.c_int(0)
        nr = C.c_double(0.0)
        _check(lib().nalu_amg_info(self._h, C.byref(nl), rows.ctypes.data_as(_i64p), nnz.ctypes.data_as(_i64p),
                                   lmax.ctypes.data_as(_dp), cap, C.byref(st), C.byref(nf), C.byref(nr)), "amg_info")
        k = nl.value
        return {"rows": rows[:k].tolist(), "nnz": nnz[:k].tolist(), "lmax": lmax[:k].tolist(),
                "setup_s": st.value, "coarse_free": nf.value, "coarse_null_rel": nr.value}

    def amg_level_csr(self, level, which="A"):
        """(row_ptr, col_idx, values, n, m) of a level's operator ("A") or prolongator ("P"), solver order."""
        w = {"A": 0, "P": 1}[which]
        n, m, nz = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        _check(lib().nalu_amg_level_csr(self._h, int(level), w, None, None, None, C.byref(n), C.byref(m),
                                        C.byref(nz)), "amg_level_csr")
        rp = np.zeros(n.value + 1, dtype=np.int32)
        ci = np.zeros(max(nz.value, 1), dtype=np.int32)
        v = np.zeros(max(nz.value, 1), dtype=np.float64)
        _check(lib().nalu_amg_level_csr(self._h, int(level), w, rp.ctypes.data_as(_i32p), ci.ctypes.data_as(_i32p),
                                        v.ctypes.data_as(_dp), C.byref(n), C.byref(m), C.byref(nz)), "amg_level_csr")
        return rp, ci[:nz.value], v[:nz.value], n.value, m.value

    def amg_aggregates(self, level):
        n = self.amg_level_csr(level, "A")[3]
        agg = np.zeros(n, dtype=np.int32)
        _check(lib().nalu_amg_aggregates(self._h, int(level), agg.ctypes.data_as(_i32p)), "amg_aggregates")
        return agg

    def precond_apply_permuted(self, r, z):
        n = self._rows()
        _check(lib().nalu_linsys_precond_apply_permuted(self._h, _dev(r, "precond_apply_permuted r", "f64", numel=n),
                                                        _dev(z, "precond_apply_permuted z", "f64", numel=n)),
               "precond_apply_permuted")

    def precond_apply_permuted_f32(self, r, z):
        """M^-1 r in FP32 (float32 device tensors, solver order)."""
        n = self._rows()
        _check(lib().nalu_linsys_precond_apply_permuted_f32(self._h,
                                                            _dev(r, "precond_apply_permuted_f32 r", "f32", numel=n),
                                                            _dev(z, "precond_apply_permuted_f32 z", "f32", numel=n)),
               "precond_apply_permuted_f32")

    def spmv_permuted(self, x, y):
        n = self._rows()
        _check(lib().nalu_linsys_spmv_permuted(self._h, _dev(x, "spmv_permuted x", "f64", numel=n),
                                               _dev(y, "spmv_permuted y", "f64", numel=n)), "spmv_permuted")

    def set_l2_scaling(self, s):
        _check(lib().nalu_linsys_set_l2_scaling(self._h, float(s)), "set_l2_scaling")

    def set_write_matrix_files(self, base_name):
        """write_matrix_files: every solve writes <base>-O-<k>.mm.1 / .rhs.1 / .sln.1 (None turns it off)."""
        _check(lib().nalu_linsys_set_write_matrix_files(self._h, None if base_name is None else base_name.encode()),
               "set_write_matrix_files")

    def set_learned_preconditioner(self, net):
        """Attach a LearnedPreconditioner (used when the config's preconditioner is
        "learned", inside FGMRES, with SGS as the fallback); None detaches."""
        self._net = net
        _check(lib().nalu_linsys_set_learned_precond(self._h, None if net is None else net.handle),
               "set_learned_precond")

    def set_learned_base(self, base):
        """The learned preconditioner's base: "sgs" (default) or "muelu" (the AMG
        V-cycle with this system's MueLu parameters): the network's input feature,
        its initial output and its fallback.  Set it before the preconditioner
        set-up (and before harvesting training pairs)."""
        _check(lib().nalu_linsys_set_learned_base(self._h, PRECOND[base]), "set_learned_base")

    def learned_fallback(self):
        """True when the last solve fell back from the network to its base."""
        return bool(_check(lib().nalu_linsys_learned_fallback(self._h), "learned_fallback"))


# ------------------------------------------------------------------ learned preconditioner
class HarvestedPairs:
    """(residual, error) pairs of plain FP64 CG + SGS solves, held on the device."""

    def __init__(self, ls, num_rhs, pairs_per_rhs, seed=1, tol=1e-10, max_iterations=5000):
        h = _vp()
        _check(lib().nalu_gnn_harvest(ls.handle, int(num_rhs), int(pairs_per_rhs), int(seed), float(tol),
                                      int(max_iterations), C.byref(h)), "gnn_harvest")
        self._h = h
        self.ls = ls
        self.num_rhs = num_rhs

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return _check(lib().nalu_gnn_data_size(self._h), "gnn_data_size")

    def cg_iterations(self):
        it = (C.c_int * self.num_rhs)()
        _check(lib().nalu_gnn_data_cg_iterations(self._h, it, self.num_rhs), "gnn_data_cg_iterations")
        return list(it)

    def pair(self, idx, r, e):
        """Copy pair idx (original row order) into float32 device tensors r, e."""
        n = self.ls._rows()
        if not 0 <= int(idx) < len(self):
            raise NaluError("pair: index out of range")
        _check(lib().nalu_gnn_data_pair(self._h, int(idx), _dev(r, "pair r", "f32", numel=n),
                                        _dev(e, "pair e", "f32", numel=n)), "gnn_data_pair")

    def set_pair(self, idx, r, e):
        """Overwrite pair idx with (r, e), float64 device tensors in the original
        row order with A e = r (stored scaled by 1/|r|, e off its mean)."""
        n = self.ls._rows()
        if not 0 <= int(idx) < len(self):
            raise NaluError("set_pair: index out of range")
        _check(lib().nalu_gnn_data_set_pair(self._h, self.ls.handle, int(idx), _dev(r, "set_pair r", "f64", numel=n),
                                            _dev(e, "set_pair e", "f64", numel=n)), "gnn_data_set_pair")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _L is not None:
            _L.nalu_gnn_data_destroy(h)
            self._h = None


class LearnedPreconditioner:
    """Message-passing network over the matrix graph (BF16 weights, FP32
    accumulation) trained and applied by the HIP kernels of libnalu_hip.so."""

    def __init__(self, num_layers=4, hidden=32, seed=0):
        h = _vp()
        _check(lib().nalu_gnn_create(int(num_layers), int(hidden), int(seed), C.byref(h)), "gnn_create")
        self._h = h
        self.num_layers = num_layers

    @property
    def handle(self):
        return self._h

    def num_params(self):
        return int(lib().nalu_gnn_num_params(self._h))

    def get_params(self):
        out = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().nalu_gnn_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.shape[0]),
               "gnn_get_params")
        return out

    def set_params(self, th):
        th = _np(th, np.float32)
        _check(lib().nalu_gnn_set_params(self._h, th.ctypes.data_as(C.POINTER(C.c_float)), th.shape[0]),
               "gnn_set_params")

    def apply(self, ls, r, z):
        """z = network(r) for float32 device tensors in the original row order."""
        n = ls._rows()
        _check(lib().nalu_gnn_apply(self._h, ls.handle, _dev(r, "gnn apply r", "f32", numel=n),
                                    _dev(z, "gnn apply z", "f32", numel=n)), "gnn_apply")

    def loss_grad(self, ls, r, e, zs=None, out=None):
        """Loss and parameter gradient of one batch (r, e: float32 device tensors [B, n])."""
        batch = int(r.shape[0]) if r.dim() == 2 else 1
        nb = batch * ls._rows()
        loss = C.c_double(0)
        g = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().nalu_gnn_loss_grad(self._h, ls.handle, batch, _dev(r, "loss_grad r", "f32", numel=nb),
                                        _dev(e, "loss_grad e", "f32", numel=nb),
                                        _dev(zs, "loss_grad zs", "f32", numel=nb, optional=True),
                                        _dev(out, "loss_grad out", "f32", numel=nb, optional=True),
                                        C.byref(loss), g.ctypes.data_as(C.POINTER(C.c_float))), "gnn_loss_grad")
        return loss.value, g

    def train(self, ls, pairs, epochs=4, batch=16, lr=3e-3, seed=0, validate=None):
        """Adam on the harvested pairs; returns the mean loss of each epoch.

        validate: optional callable(net) -> score (lower is better), evaluated
        after every epoch; the parameters of the best-scoring epoch are kept and
        the scores are available as self.validation."""
        hist = np.zeros(int(epochs), dtype=np.float64)
        if validate is None:
            _check(lib().nalu_gnn_train(self._h, ls.handle, pairs.handle, int(epochs), int(batch), float(lr),
                                        int(seed), hist.ctypes.data_as(_dp)), "gnn_train")
            return hist
        self.validation = []
        best, best_th = None, None
        for ep in range(int(epochs)):
            h = np.zeros(1, dtype=np.float64)
            _check(lib().nalu_gnn_train_epochs(self._h, ls.handle, pairs.handle, ep, 1, int(epochs), int(batch),
                                               float(lr), int(seed), h.ctypes.data_as(_dp)), "gnn_train")
            hist[ep] = h[0]
            score = float(validate(self))
            self.validation.append(score)
            if best is None or score < best:
                best, best_th = score, self.get_params()
        self.set_params(best_th)
        self.best_epoch = int(np.argmin(self.validation))
        return hist

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _L is not None:
            _L.nalu_gnn_destroy(h)
            self._h = None


# ------------------------------------------------------------------ assembly algorithms
# scatter strategies: "gather" = one thread per row walks its elements (deterministic),
# "atomic" = element-centric FP64 atomics, "tile" = node-tile LDS accumulation, "etile" =
# element-visit tiles (each incident element evaluated once per tile of rows)
MODE = {"gather": 0, "atomic": 1, "tile": 2, "etile": 3}


def assemble_scalar_diff(ls, mesh, q, k, shifted=False, mode="tile"):
    N = mesh.num_nodes
    _check(lib().nalu_assemble_scalar_diff(ls.handle, mesh.handle, _dev(q, "scalar_diff q", "f64", numel=N),
                                           _dev(k, "scalar_diff k", "f64", numel=N), int(shifted), MODE[mode]),
           "assemble_scalar_diff")


def compute_mdot(mesh, rho, p, u, gpdx, mdot, tau=1.0, shift_mdot=False, shift_poisson=False, stream=None,
                 precision="fp64"):
    """mdot at the 12 scs of every element; precision "fp32" evaluates with FP32
    arithmetic (FP64 Jacobians), as the mixed-precision path does."""
    fn = lib().nalu_compute_mdot_f32 if precision == "fp32" else lib().nalu_compute_mdot
    N, E = mesh.num_nodes, mesh.num_elems
    _check(fn(mesh.handle, _dev(rho, "mdot rho", "f64", numel=N), _dev(p, "mdot p", "f64", numel=N),
              _dev(u, "mdot u", "f64", numel=3 * N), _dev(gpdx, "mdot Gpdx", "f64", numel=3 * N), float(tau),
              int(shift_mdot), int(shift_poisson), _dev(mdot, "mdot", "f64", numel=12 * E), stream), "compute_mdot")


def compute_dual_volume(mesh, dual_vol, stream=None):
    """Dual nodal volume [N] (sum of the scv volumes around each node)."""
    _check(lib().nalu_compute_dual_volume(mesh.handle, _dev(dual_vol, "dual volume", "f64", numel=mesh.num_nodes),
                                          stream), "compute_dual_volume")


def compute_nodal_gradient(mesh, q, dual_vol, dqdx, shifted=False, stream=None, boundary=True):
    """Projected nodal gradient [N][3] of the nodal field q: the interior Green-Gauss
    part plus (boundary=True) the exposed-face closure."""
    N = mesh.num_nodes
    fn = lib().nalu_compute_nodal_gradient if boundary else lib().nalu_compute_nodal_gradient_interior
    _check(fn(mesh.handle, _dev(q, "nodal gradient q", "f64", numel=N),
              _dev(dual_vol, "nodal gradient dual volume", "f64", numel=N), int(shifted),
              _dev(dqdx, "nodal gradient dqdx", "f64", numel=3 * N), stream), "compute_nodal_gradient")


def assemble_continuity(ls, mesh, rho, p, u, gpdx, tau=1.0, shift_mdot=False, shift_poisson=False,
                        reduced_sens=False, mode="tile", mdot_out=None):
    """ContinuityAdvElemKernel.  mdot_out ([E][12] device tensor, row-gather mode
    only) also receives mdot from the same scs evaluations
    (ComputeMdotElemAlgorithm), so no separate compute_mdot pass is needed."""
    N = mesh.num_nodes
    if mdot_out is not None:
        if mode != "gather":
            raise NaluError("assemble_continuity: mdot_out needs mode 'gather'")
        _check(lib().nalu_assemble_continuity_mdot(
            ls.handle, mesh.handle, _dev(rho, "continuity rho", "f64", numel=N), _dev(p, "continuity p", "f64", numel=N),
            _dev(u, "continuity u", "f64", numel=3 * N), _dev(gpdx, "continuity Gpdx", "f64", numel=3 * N),
            float(tau), int(shift_mdot), int(shift_poisson), int(reduced_sens),
            _dev(mdot_out, "continuity mdot_out", "f64", numel=12 * mesh.num_elems)), "assemble_continuity")
        return
    _check(lib().nalu_assemble_continuity(ls.handle, mesh.handle, _dev(rho, "continuity rho", "f64", numel=N),
                                          _dev(p, "continuity p", "f64", numel=N),
                                          _dev(u, "continuity u", "f64", numel=3 * N),
                                          _dev(gpdx, "continuity Gpdx", "f64", numel=3 * N),
                                          float(tau), int(shift_mdot), int(shift_poisson), int(reduced_sens),
                                          MODE[mode]), "assemble_continuity")


def assemble_momentum(ls, mesh, u, mu, mdot, include_divu=0.0, skew=False, shifted=False, mode="tile"):
    N, E = mesh.num_nodes, mesh.num_elems
    _check(lib().nalu_assemble_momentum(ls.handle, mesh.handle, _dev(u, "momentum u", "f64", numel=3 * N),
                                        _dev(mu, "momentum mu", "f64", numel=N),
                                        _dev(mdot, "momentum mdot", "f64", numel=12 * E), float(include_divu),
                                        int(skew), int(shifted), MODE[mode]), "assemble_momentum")


def _states(fields, name, numel):
    if len(fields) != 3:
        raise NaluError("%s: three states (Nm1, N, Np1) expected" % name)
    return [_dev(f, "%s state %d" % (name, i), "f64", numel=numel) for i, f in enumerate(fields)]


def _gammas(gamma):
    g = np.ascontiguousarray(np.asarray(gamma, dtype=np.float64).ravel())
    if g.shape != (3,):
        raise NaluError("gamma: three BDF coefficients expected")
    return g


def assemble_momentum_with_mass(ls, mesh, u, mu, mdot, rho, u_states, gjp, dt, gamma=(1.0, -1.0, 0.0), lumped=False,
                                balanced_force=False, include_divu=0.0, skew=False, shifted=False):
    """MomentumAdvDiffElemKernel + MomentumMassElemKernel in one element pass
    (the kernel list of one element algorithm): u, mu, mdot as for
    assemble_momentum; rho = (rhoNm1, rhoN, rhoNp1), u_states = (uNm1, uN, uNp1),
    gjp, dt, gamma as for assemble_momentum_mass."""
    N, E = mesh.num_nodes, mesh.num_elems
    r, v, g = _states(rho, "momentum_with_mass rho", N), _states(u_states, "momentum_with_mass u", 3 * N), \
        _gammas(gamma)
    _check(lib().nalu_assemble_momentum_with_mass(
        ls.handle, mesh.handle, _dev(u, "momentum_with_mass u", "f64", numel=3 * N),
        _dev(mu, "momentum_with_mass mu", "f64", numel=N), _dev(mdot, "momentum_with_mass mdot", "f64", numel=12 * E),
        float(include_divu), int(skew), int(shifted), r[0], r[1], r[2], v[0], v[1], v[2],
        _dev(gjp, "momentum_with_mass Gjp", "f64", numel=3 * N), float(dt), g.ctypes.data_as(_dp), int(lumped),
        int(balanced_force)), "assemble_momentum_with_mass")


def assemble_momentum_with_mass_segregated(ls, mesh, u, mu, mdot, rho, u_states, gjp, dt, rhs3,
                                           gamma=(1.0, -1.0, 0.0), lumped=False, balanced_force=False,
                                           include_divu=0.0, skew=False, shifted=False):
    """assemble_momentum_segregated + assemble_momentum_mass_segregated in one
    pass over each row's elements; rhs3 [3][N] as for those two."""
    N, E = mesh.num_nodes, mesh.num_elems
    r, v, g = _states(rho, "momentum_with_mass_segregated rho", N), \
        _states(u_states, "momentum_with_mass_segregated u", 3 * N), _gammas(gamma)
    _check(lib().nalu_assemble_momentum_with_mass_segregated(
        ls.handle, mesh.handle, _dev(u, "momentum_with_mass_segregated u", "f64", numel=3 * N),
        _dev(mu, "momentum_with_mass_segregated mu", "f64", numel=N),
        _dev(mdot, "momentum_with_mass_segregated mdot", "f64", numel=12 * E), float(include_divu), int(skew),
        int(shifted), r[0], r[1], r[2], v[0], v[1], v[2],
        _dev(gjp, "momentum_with_mass_segregated Gjp", "f64", numel=3 * N), float(dt), g.ctypes.data_as(_dp),
        int(lumped), int(balanced_force), _dev(rhs3, "momentum_with_mass_segregated rhs3", "f64", shape=(3, ls._rows()))),
        "assemble_momentum_with_mass_segregated")


def assemble_momentum_mass(ls, mesh, rho, u, gjp, dt, gamma=(1.0, -1.0, 0.0), lumped=False, balanced_force=False):
    """MomentumMassElemKernel: rho = (rhoNm1, rhoN, rhoNp1), u = (uNm1, uN, uNp1), nodal device arrays."""
    N = mesh.num_nodes
    r, v, g = _states(rho, "momentum_mass rho", N), _states(u, "momentum_mass u", 3 * N), _gammas(gamma)
    _check(lib().nalu_assemble_momentum_mass(ls.handle, mesh.handle, r[0], r[1], r[2], v[0], v[1], v[2],
                                             _dev(gjp, "momentum_mass Gjp", "f64", numel=3 * N),
                                             float(dt), g.ctypes.data_as(_dp), int(lumped), int(balanced_force)),
           "assemble_momentum_mass")


def assemble_momentum_segregated(ls, mesh, u, mu, mdot, rhs3, include_divu=0.0, skew=False, shifted=False):
    """Momentum as 3 right-hand sides sharing one scalar matrix (numDof 1 system);
    rhs3 [3][N] (device, row order) is added the residual of each component."""
    N, E = mesh.num_nodes, mesh.num_elems
    _check(lib().nalu_assemble_momentum_segregated(ls.handle, mesh.handle, _dev(u, "momentum u", "f64", numel=3 * N),
                                                   _dev(mu, "momentum mu", "f64", numel=N),
                                                   _dev(mdot, "momentum mdot", "f64", numel=12 * E),
                                                   float(include_divu), int(skew), int(shifted),
                                                   _dev(rhs3, "momentum rhs3", "f64", shape=(3, ls._rows()))),
           "assemble_momentum_segregated")


def assemble_momentum_mass_segregated(ls, mesh, rho, u, gjp, dt, rhs3, gamma=(1.0, -1.0, 0.0), lumped=False,
                                      balanced_force=False):
    N = mesh.num_nodes
    r, v, g = _states(rho, "momentum_mass rho", N), _states(u, "momentum_mass u", 3 * N), _gammas(gamma)
    _check(lib().nalu_assemble_momentum_mass_segregated(ls.handle, mesh.handle, r[0], r[1], r[2], v[0], v[1], v[2],
                                                        _dev(gjp, "momentum_mass Gjp", "f64", numel=3 * N), float(dt),
                                                        g.ctypes.data_as(_dp), int(lumped), int(balanced_force),
                                                        _dev(rhs3, "momentum rhs3", "f64", shape=(3, ls._rows()))),
           "assemble_momentum_mass_segregated")


def assemble_continuity_mass(ls, mesh, rho, dt, gamma=(1.0, -1.0, 0.0), lumped=False):
    """ContinuityMassElemKernel: rho = (rhoNm1, rhoN, rhoNp1), nodal device arrays."""
    r, g = _states(rho, "continuity_mass rho", mesh.num_nodes), _gammas(gamma)
    _check(lib().nalu_assemble_continuity_mass(ls.handle, mesh.handle, r[0], r[1], r[2], float(dt),
                                               g.ctypes.data_as(_dp), int(lumped)), "assemble_continuity_mass")


# ------------------------------------------------------------------ consolidated element algorithm
class Kernel:
    """Kernel of the consolidated element algorithm (include/kernel/Kernel.h:95-125):
    a kind, its options and the device fields it reads.  Fields stay referenced
    by the kernel object while the algorithm holds it."""
    KIND = None
    FIELDS = ()

    def __init__(self, *fields, shifted_grad_op=False, shift_mdot=False, reduced_sens=False, skew_symmetric=False,
                 lumped=False, balanced_force=False, include_divu=0.0):
        if len(fields) != len(self.FIELDS):
            raise NaluError("%s: fields %s expected" % (type(self).__name__, ", ".join(self.FIELDS)))
        self.fields = fields
        self.opts = dict(shifted_grad_op=int(shifted_grad_op), shift_mdot=int(shift_mdot),
                         reduced_sens=int(reduced_sens), skew_symmetric=int(skew_symmetric), lumped=int(lumped),
                         balanced_force=int(balanced_force), include_divu=float(include_divu))

    def desc(self, mesh):
        N, E = mesh.num_nodes, mesh.num_elems
        d = _KernelDesc(kind=self.KIND, **self.opts)
        for i, (f, name) in enumerate(zip(self.fields, self.FIELDS)):
            numel = 12 * E if name == "mdot" else (3 * N if name in ("u", "Gpdx", "Gjp", "uNm1", "uN", "uNp1") else N)
            d.field[i] = _dev(f, "%s %s" % (type(self).__name__, name), "f64", numel=numel)
        return d


class ScalarDiffElemKernel(Kernel):
    KIND, FIELDS = 0, ("q", "diffusion coefficient")


class ContinuityAdvElemKernel(Kernel):
    KIND, FIELDS = 1, ("rho", "p", "u", "Gpdx")


class MomentumAdvDiffElemKernel(Kernel):
    KIND, FIELDS = 2, ("u", "mu", "mdot")


class ContinuityMassElemKernel(Kernel):
    KIND, FIELDS = 3, ("rhoNm1", "rhoN", "rhoNp1")


class MomentumMassElemKernel(Kernel):
    KIND, FIELDS = 4, ("rhoNm1", "rhoN", "rhoNp1", "uNm1", "uN", "uNp1", "Gjp")


class AssembleElemSolverAlgorithm:
    """AssembleElemSolverAlgorithm (include/AssembleElemSolverAlgorithm.h:52-105):
    activeKernels_ run in order by execute(), after each kernel's setup with the
    time integrator (dt, gamma1..3)."""

    def __init__(self, linsys, mesh):
        self.linsys, self.mesh = linsys, mesh
        self.activeKernels_ = []
        h = _vp()
        _check(lib().nalu_elem_algorithm_create(linsys.handle, mesh.handle, C.byref(h)), "AssembleElemSolverAlgorithm")
        self._h = h

    def add_kernel(self, kernel):
        _check(lib().nalu_elem_algorithm_add_kernel(self._h, C.byref(kernel.desc(self.mesh))), "add_kernel")
        self.activeKernels_.append(kernel)

    def set_time_integrator(self, dt, gamma=(1.0, -1.0, 0.0)):
        g = _gammas(gamma)
        _check(lib().nalu_elem_algorithm_set_time_integrator(self._h, float(dt), g.ctypes.data_as(_dp)),
               "set_time_integrator")

    def execute(self):
        _check(lib().nalu_elem_algorithm_execute(self._h), "AssembleElemSolverAlgorithm::execute")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _L is not None:
            _L.nalu_elem_algorithm_destroy(h)
            self._h = None
