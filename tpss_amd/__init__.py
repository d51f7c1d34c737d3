"""tpss_amd -- MI355X vertex-patch Schwarz smoother (Python host mirror of the C ABI).

The classes mirror the reference interface of the hot path:

* ``Level``                  -- one multigrid level with its SIPG operator
                               (Laplace::DG::MF::Operator, apps/laplace_integrator.h:578-743)
* ``SchwarzPreconditioner``  -- ``vmult(dst, src)`` (schwarz_preconditioner.templates.h:209-239)
* ``SchwarzSmoother``        -- ``vmult`` / ``step`` (smoother_richardson.h:73-156)
* ``GMGSolver``              -- GMG-preconditioned CG (apps/poisson_problem.h:642-755)

Vectors are device buffers: anything with ``data_ptr()`` (a torch CUDA tensor)
or a raw integer device address.  All work runs in libtpss_amd.so; there is no
CPU fallback -- importing raises if the library has not been built.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtpss_amd.so")

OK, E_INVALID_ARG, E_NONFINITE, E_SINGULAR, E_HIP, E_NOT_CONVERGED, E_UNSUPPORTED = range(7)
ADDITIVE, MULTIPLICATIVE = 0, 1
COLOR_NONE, COLOR_TILED, COLOR_REDBLACK = 0, 1, 2
F64, F32, BF16 = 0, 1, 2
POISSON, ELASTICITY = 0, 1
LOCAL_SURROGATE, LOCAL_FIT = 0, 1
RAS_NONE, RAS_WEIGHTS, RAS_BOOLEAN = 0, 1, 2

#: TPSS::lookup_damping_factor (TPSS.h:120-163) for vertex patches
DEFAULT_DAMPING = {(ADDITIVE, 2): 0.25, (ADDITIVE, 3): 0.15,
                   (MULTIPLICATIVE, 2): 1.0, (MULTIPLICATIVE, 3): 1.0}


class TPSSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tpss status {code}: {msg}")
        self.code = code


class LevelDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("degree", ctypes.c_int), ("n_cells_1d", ctypes.c_int),
                ("ip_factor", ctypes.c_double), ("vertices", ctypes.POINTER(ctypes.c_double)),
                ("vector_precision", ctypes.c_int), ("equation", ctypes.c_int),
                ("mu", ctypes.c_double), ("lambda_", ctypes.c_double)]


class SmootherDesc(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("coloring", ctypes.c_int),
                ("relaxation", ctypes.c_double), ("reverse", ctypes.c_int),
                ("symmetrized", ctypes.c_int), ("damping_factor", ctypes.c_double),
                ("n_steps", ctypes.c_int), ("local_precision", ctypes.c_int),
                ("local_solver", ctypes.c_int), ("fit_sweeps", ctypes.c_int),
                ("ras", ctypes.c_int)]


class TimeInfo(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("description", ctypes.c_char_p),
                ("n_calls", ctypes.c_long)]


class GMGDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("degree", ctypes.c_int), ("n_cells_coarse", ctypes.c_int),
                ("n_levels", ctypes.c_int), ("ip_factor", ctypes.c_double),
                ("vertices", ctypes.POINTER(ctypes.c_double)), ("mg_precision", ctypes.c_int),
                ("smoother", SmootherDesc), ("coarse_tol", ctypes.c_double),
                ("equation", ctypes.c_int), ("mu", ctypes.c_double), ("lambda_", ctypes.c_double)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("initial_residual", ctypes.c_double),
                ("final_residual", ctypes.c_double), ("reduction_rate", ctypes.c_double),
                ("n_frac", ctypes.c_double), ("seconds", ctypes.c_double)]


#: every entry point declared in include/tpss.h (checked by the CPU test-suite)
EXPORTS = [
    "tpss_last_error", "tpss_version", "tpss_level_create", "tpss_level_destroy",
    "tpss_level_n_dofs", "tpss_operator_vmult", "tpss_operator_residual",
    "tpss_smoother_create", "tpss_smoother_clone_shallow", "tpss_smoother_destroy",
    "tpss_precond_vmult", "tpss_precond_vmult_add", "tpss_precond_Tvmult",
    "tpss_precond_Tvmult_add", "tpss_noalloc_init", "tpss_precond_vmult_noalloc",
    "tpss_smoother_step", "tpss_smoother_vmult",
    "tpss_smoother_set_patch_damping", "tpss_smoother_info", "tpss_smoother_color_ptr",
    "tpss_smoother_patch_vertices", "tpss_smoother_patch_dof_indices", "tpss_smoother_patch_h",
    "tpss_smoother_restricted_dof_flags",
    "tpss_smoother_fit_errors",
    "tpss_smoother_get_timings", "tpss_smoother_set_timing", "tpss_smoother_set_check_finite",
    "tpss_smoother_finite_status",
    "tpss_smoother_set_fused_residual",
    "tpss_gmg_create", "tpss_gmg_destroy", "tpss_gmg_level", "tpss_gmg_smoother",
    "tpss_gmg_vcycle", "tpss_gmg_solve", "tpss_transfer_prolongate_add",
    "tpss_transfer_restrict", "tpss_vector_fill_random", "tpss_vector_convert",
    "tpss_device_count", "tpss_device_malloc", "tpss_device_free", "tpss_memcpy_h2d",
    "tpss_memcpy_d2h", "tpss_device_synchronize",
    "tpss_mlp_create", "tpss_mlp_destroy", "tpss_mlp_n_params", "tpss_mlp_set_params",
    "tpss_mlp_get_params", "tpss_mlp_forward", "tpss_mlp_loss_grad", "tpss_mlp_train",
    "tpss_smoother_patch_features", "tpss_smoother_optimal_damping", "tpss_smoother_set_damping_mlp",
    "tpss_smoother_local_lambda_max",
    "tpss_ksvd_batched", "tpss_ksvd_sum_batched", "tpss_kron_vmult", "tpss_kron_table",
    "tpss_kron_pinv", "tpss_kcp_batched",
]

_VP = ctypes.c_void_p
_PVP = ctypes.POINTER(ctypes.c_void_p)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python tpss_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    i64 = ctypes.c_int64
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "tpss_last_error": (ctypes.c_char_p, []),
        "tpss_version": (ctypes.c_char_p, []),
        "tpss_level_create": (ctypes.c_int, [ctypes.POINTER(LevelDesc), _PVP]),
        "tpss_level_destroy": (ctypes.c_int, [_VP]),
        "tpss_level_n_dofs": (ctypes.c_int, [_VP, ctypes.POINTER(i64)]),
        "tpss_operator_vmult": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_operator_residual": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
        "tpss_smoother_create": (ctypes.c_int, [_VP, ctypes.POINTER(SmootherDesc), _PVP]),
        "tpss_smoother_clone_shallow": (ctypes.c_int, [_VP, ctypes.POINTER(SmootherDesc), _PVP]),
        "tpss_smoother_destroy": (ctypes.c_int, [_VP]),
        "tpss_precond_vmult": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_precond_vmult_add": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_precond_Tvmult": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_precond_Tvmult_add": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_noalloc_init": (ctypes.c_int, []),
        "tpss_precond_vmult_noalloc": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_smoother_step": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_smoother_vmult": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_smoother_set_patch_damping": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double)]),
        "tpss_smoother_info": (ctypes.c_int, [_VP, ip, ip, ip]),
        "tpss_smoother_color_ptr": (ctypes.c_int, [_VP, ip]),
        "tpss_smoother_patch_vertices": (ctypes.c_int, [_VP, ip]),
        "tpss_smoother_patch_dof_indices": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint32)]),
        "tpss_smoother_patch_h": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double)]),
        "tpss_smoother_restricted_dof_flags": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint8)]),
        "tpss_smoother_get_timings": (ctypes.c_int, [_VP, ctypes.POINTER(TimeInfo), ip]),
        "tpss_smoother_set_timing": (ctypes.c_int, [_VP, ctypes.c_int]),
        "tpss_smoother_set_check_finite": (ctypes.c_int, [_VP, ctypes.c_int]),
        "tpss_smoother_finite_status": (ctypes.c_int, [_VP, _VP]),
        "tpss_smoother_set_fused_residual": (ctypes.c_int, [_VP, ctypes.c_int]),
        "tpss_gmg_create": (ctypes.c_int, [ctypes.POINTER(GMGDesc), _PVP]),
        "tpss_gmg_destroy": (ctypes.c_int, [_VP]),
        "tpss_gmg_level": (ctypes.c_int, [_VP, ctypes.c_int, _PVP]),
        "tpss_gmg_smoother": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _PVP]),
        "tpss_gmg_vcycle": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tpss_gmg_solve": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_double, ctypes.c_int, _VP,
                                          ctypes.POINTER(SolveInfo),
                                          ctypes.POINTER(ctypes.c_double)]),
        "tpss_transfer_prolongate_add": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
        "tpss_transfer_restrict": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
        "tpss_vector_fill_random": (ctypes.c_int, [_VP, i64, ctypes.c_int, ctypes.c_uint64, _VP]),
        "tpss_vector_convert": (ctypes.c_int, [_VP, ctypes.c_int, _VP, ctypes.c_int, i64, _VP]),
        "tpss_device_count": (ctypes.c_int, [ip]),
        "tpss_device_malloc": (ctypes.c_int, [_PVP, ctypes.c_size_t]),
        "tpss_device_free": (ctypes.c_int, [_VP]),
        "tpss_memcpy_h2d": (ctypes.c_int, [_VP, _VP, ctypes.c_size_t]),
        "tpss_memcpy_d2h": (ctypes.c_int, [_VP, _VP, ctypes.c_size_t]),
        "tpss_device_synchronize": (ctypes.c_int, []),
        "tpss_mlp_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_uint64, _PVP]),
        "tpss_mlp_destroy": (ctypes.c_int, [_VP]),
        "tpss_mlp_n_params": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_size_t)]),
        "tpss_mlp_set_params": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t]),
        "tpss_mlp_get_params": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t]),
        "tpss_mlp_forward": (ctypes.c_int, [_VP, _VP, _VP, i64, _VP]),
        "tpss_mlp_loss_grad": (ctypes.c_int, [_VP, _VP, _VP, i64, _VP,
                                              ctypes.POINTER(ctypes.c_double), _VP]),
        "tpss_mlp_train": (ctypes.c_int, [_VP, _VP, _VP, i64, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_double)]),
        "tpss_smoother_patch_features": (ctypes.c_int, [_VP, _VP, _VP]),
        "tpss_smoother_optimal_damping": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_uint64,
                                                         ctypes.c_double, ctypes.c_double, _VP]),
        "tpss_smoother_set_damping_mlp": (ctypes.c_int, [_VP, _VP]),
        "tpss_smoother_local_lambda_max": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_uint64, _VP]),
        "tpss_ksvd_batched": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _VP,
                                             _VP, _VP]),
        "tpss_ksvd_sum_batched": (ctypes.c_int, [_VP, _VP] + [ctypes.c_int] * 7 + [_VP] * 4),
        "tpss_kron_vmult": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ip, ip, ctypes.c_int,
                                           _VP, _VP, _VP, ctypes.c_int, ctypes.c_int, _VP]),
        "tpss_kron_table": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ip, ip, ctypes.c_int,
                                           _VP, ctypes.c_int, _VP, _VP]),
        "tpss_kron_pinv": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, _VP, _VP]),
        "tpss_kcp_batched": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ip, ip, ctypes.c_int,
                                            ctypes.c_int, _VP, _VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _check(status):
    if status != OK:
        raise TPSSError(status, lib.tpss_last_error().decode())


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    return ctypes.c_void_p(x.data_ptr())


_DTYPES = {F64: "float64", F32: "float32"}


class DeviceVector:
    """A level vector in device memory allocated by libtpss_amd itself."""

    def __init__(self, n, precision=F64):
        import numpy as np
        self.n, self.precision = int(n), precision
        self.dtype = np.dtype(_DTYPES[precision])
        p = ctypes.c_void_p()
        _check(lib.tpss_device_malloc(ctypes.byref(p), self.n * self.dtype.itemsize))
        self._p = p

    @classmethod
    def from_numpy(cls, a, precision=None):
        import numpy as np
        if precision is None:
            precision = F32 if np.asarray(a).dtype == np.float32 else F64
        v = cls(np.asarray(a).size, precision)
        v.upload(a)
        return v

    @classmethod
    def zeros(cls, n, precision=F64):
        import numpy as np
        v = cls(n, precision)
        v.upload(np.zeros(n, dtype=_DTYPES[precision]))
        return v

    def upload(self, a):
        import numpy as np
        h = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        assert h.size == self.n
        _check(lib.tpss_memcpy_h2d(self._p, h.ctypes.data_as(ctypes.c_void_p), h.nbytes))

    def numpy(self):
        import numpy as np
        h = np.empty(self.n, dtype=self.dtype)
        _check(lib.tpss_memcpy_d2h(h.ctypes.data_as(ctypes.c_void_p), self._p, h.nbytes))
        return h

    def data_ptr(self):
        return self._p.value

    def numel(self):
        return self.n

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            lib.tpss_device_free(p)
            self._p = None


def device_count():
    c = ctypes.c_int()
    _check(lib.tpss_device_count(ctypes.byref(c)))
    return c.value


def synchronize():
    _check(lib.tpss_device_synchronize())


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _vertices_ptr(vertices):
    if vertices is None:
        return None, None
    import numpy as np
    arr = np.ascontiguousarray(vertices, dtype=np.float64)
    return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _smoother_desc(variant, coloring, relaxation, reverse, symmetrized, damping_factor, n_steps,
                   local_precision, local_solver=LOCAL_SURROGATE, fit_sweeps=0, ras=RAS_NONE):
    return SmootherDesc(variant, coloring, float(relaxation), int(reverse), int(symmetrized),
                        float(damping_factor), int(n_steps), int(local_precision),
                        int(local_solver), int(fit_sweeps), int(ras))


class Level:
    """Mesh + DG Q_p space + SIPG operator of one level (device).

    elasticity=(mu, lambda) selects the linear elasticity operator with dim
    component blocks instead of the Poisson operator."""

    def __init__(self, dim, degree, n_cells_1d, ip_factor=1.0, vertices=None, precision=F64,
                 elasticity=None):
        self._varr, vp = _vertices_ptr(vertices)
        eq, mu, lam = (POISSON, 0.0, 0.0) if elasticity is None else (ELASTICITY,) + tuple(
            float(v) for v in elasticity)
        desc = LevelDesc(dim, degree, n_cells_1d, float(ip_factor), vp, precision, eq, mu, lam)
        h = ctypes.c_void_p()
        _check(lib.tpss_level_create(ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h
        self.dim, self.degree, self.n, self.precision = dim, degree, n_cells_1d, precision
        n_dofs = ctypes.c_int64()
        _check(lib.tpss_level_n_dofs(h, ctypes.byref(n_dofs)))
        self.n_dofs = n_dofs.value
        self._owned = True

    @classmethod
    def _borrow(cls, handle, dim, degree, n, precision, owner):
        """wraps a level owned by ``owner`` (kept alive as long as the wrapper)"""
        self = cls.__new__(cls)
        self._owner = owner
        self.handle = handle
        self.dim, self.degree, self.n, self.precision = dim, degree, n, precision
        n_dofs = ctypes.c_int64()
        _check(lib.tpss_level_n_dofs(handle, ctypes.byref(n_dofs)))
        self.n_dofs = n_dofs.value
        self._owned = False
        return self

    def vmult(self, dst, src, stream=None):
        _check(lib.tpss_operator_vmult(self.handle, _ptr(dst), _ptr(src), _stream(stream)))

    def residual(self, dst, b, x, stream=None):
        _check(lib.tpss_operator_residual(self.handle, _ptr(dst), _ptr(b), _ptr(x), _stream(stream)))

    def __del__(self):
        if getattr(self, "_owned", False) and self.handle:
            lib.tpss_level_destroy(self.handle)
            self.handle = None


class SchwarzPreconditioner:
    """Vertex-patch Schwarz preconditioner of one level (device)."""

    def __init__(self, level, variant=ADDITIVE, coloring=None, relaxation=None, reverse=False,
                 symmetrized=False, damping_factor=1.0, n_steps=1, local_precision=None,
                 local_solver=LOCAL_SURROGATE, fit_sweeps=0, ras=RAS_NONE, _handle=None,
                 _owner=None):
        if coloring is None:
            coloring = COLOR_NONE if variant == ADDITIVE else COLOR_REDBLACK
        if relaxation is None:
            relaxation = DEFAULT_DAMPING[(variant, level.dim)]
        if local_precision is None:
            local_precision = level.precision
        self.level = level
        self.desc = _smoother_desc(variant, coloring, relaxation, reverse, symmetrized,
                                   damping_factor, n_steps, local_precision, local_solver,
                                   fit_sweeps, ras)
        if _handle is not None:
            # borrowed from _owner (a GMGSolver), which must outlive this wrapper
            self._owner = _owner
            self.handle, self._owned = _handle, False
        else:
            h = ctypes.c_void_p()
            _check(lib.tpss_smoother_create(level.handle, ctypes.byref(self.desc), ctypes.byref(h)))
            self.handle, self._owned = h, True
        p, c, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib.tpss_smoother_info(self.handle, ctypes.byref(p), ctypes.byref(c), ctypes.byref(k)))
        self.n_patches, self.n_colors, self.n_dofs_per_patch = p.value, c.value, k.value

    def shallow_copy(self, reverse=None, symmetrized=None, relaxation=None):
        d = SmootherDesc.from_buffer_copy(self.desc)
        if reverse is not None:
            d.reverse = int(reverse)
        if symmetrized is not None:
            d.symmetrized = int(symmetrized)
        if relaxation is not None:
            d.relaxation = float(relaxation)
        h = ctypes.c_void_p()
        _check(lib.tpss_smoother_clone_shallow(self.handle, ctypes.byref(d), ctypes.byref(h)))
        other = type(self).__new__(type(self))
        other.level, other.desc, other.handle, other._owned = self.level, d, h, True
        other.n_patches, other.n_colors, other.n_dofs_per_patch = (
            self.n_patches, self.n_colors, self.n_dofs_per_patch)
        other._parent = self
        return other

    def vmult(self, dst, src, stream=None):
        _check(lib.tpss_precond_vmult(self.handle, _ptr(dst), _ptr(src), _stream(stream)))

    def vmult_add(self, dst, src, stream=None):
        _check(lib.tpss_precond_vmult_add(self.handle, _ptr(dst), _ptr(src), _stream(stream)))

    def Tvmult(self, dst, src, stream=None):
        _check(lib.tpss_precond_Tvmult(self.handle, _ptr(dst), _ptr(src), _stream(stream)))

    def Tvmult_add(self, dst, src, stream=None):
        _check(lib.tpss_precond_Tvmult_add(self.handle, _ptr(dst), _ptr(src), _stream(stream)))

    def set_patch_damping(self, omega):
        if omega is None:
            _check(lib.tpss_smoother_set_patch_damping(self.handle, None))
            return
        import numpy as np
        w = np.ascontiguousarray(omega, dtype=np.float64)
        assert w.shape == (self.n_patches,)
        _check(lib.tpss_smoother_set_patch_damping(self.handle,
                                                   w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def patch_features(self):
        """Singular values of the cell Jacobians of every patch (scaled by n)."""
        nf = self.level.dim * 2 ** self.level.dim
        v = DeviceVector(self.n_patches * nf, F32)
        _check(lib.tpss_smoother_patch_features(self.handle, _ptr(v), None))
        return v.numpy().reshape(self.n_patches, nf)

    def optimal_damping(self, n_samples=4, seed=5, omega_min=0.2, omega_max=1.2):
        """Per-patch A-norm line-search damping computed with the smoother's own
        local solvers and operator (FP64 levels only)."""
        v = DeviceVector(self.n_patches, F32)
        _check(lib.tpss_smoother_optimal_damping(self.handle, int(n_samples), int(seed),
                                                 float(omega_min), float(omega_max), _ptr(v)))
        return v.numpy()

    def local_lambda_max(self, n_iter=30, seed=7):
        """Largest eigenvalue of M_j^{-1} A_j per patch (local solver M_j against the
        exact local matrix A_j; 1 for exact local solvers) by power iteration
        (FP64 levels only)."""
        v = DeviceVector(self.n_patches, F64)
        _check(lib.tpss_smoother_local_lambda_max(self.handle, int(n_iter), int(seed), _ptr(v)))
        return v.numpy()

    def set_damping_mlp(self, mlp):
        _check(lib.tpss_smoother_set_damping_mlp(self.handle, mlp.handle if mlp else None))

    def set_timing(self, enabled=True):
        _check(lib.tpss_smoother_set_timing(self.handle, int(enabled)))

    def set_fused_residual(self, enabled=True):
        _check(lib.tpss_smoother_set_fused_residual(self.handle, int(enabled)))

    def set_check_finite(self, mode=1):
        """0 off; 1 on, reported at the next call (default); 2 on, synchronous"""
        _check(lib.tpss_smoother_set_check_finite(self.handle, int(mode)))

    def finite_status(self, stream=None):
        """raises TPSSError(E_NONFINITE) if an output since the last report was non-finite"""
        _check(lib.tpss_smoother_finite_status(self.handle, _stream(stream)))

    def get_time_data(self):
        n = ctypes.c_int(8)
        arr = (TimeInfo * 8)()
        _check(lib.tpss_smoother_get_timings(self.handle, arr, ctypes.byref(n)))
        return [(arr[i].description.decode(), arr[i].time, arr[i].n_calls) for i in range(n.value)]

    def color_ptr(self):
        import numpy as np
        out = np.zeros(self.n_colors + 1, dtype=np.int32)
        _check(lib.tpss_smoother_color_ptr(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def patch_vertices(self):
        import numpy as np
        out = np.zeros((self.n_patches, self.level.dim), dtype=np.int32)
        _check(lib.tpss_smoother_patch_vertices(self.handle,
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def patch_dof_indices(self):
        import numpy as np
        out = np.zeros((self.n_patches, self.n_dofs_per_patch), dtype=np.uint32)
        _check(lib.tpss_smoother_patch_dof_indices(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return out

    def restricted_dof_flags(self):
        """(n_patches, n_dofs_per_patch) uint8: 1 where the restricted prolongation skips"""
        import numpy as np
        out = np.zeros((self.n_patches, self.n_dofs_per_patch), dtype=np.uint8)
        _check(lib.tpss_smoother_restricted_dof_flags(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def patch_h(self):
        import numpy as np
        out = np.zeros((self.n_patches, self.level.dim, 2))
        _check(lib.tpss_smoother_patch_h(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def fit_errors(self):
        """Relative Frobenius error of each patch's separable fit (LOCAL_FIT)."""
        import numpy as np
        out = np.zeros(self.n_patches)
        _check(lib.tpss_smoother_fit_errors(self.handle,
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def __del__(self):
        if getattr(self, "_owned", False) and self.handle:
            lib.tpss_smoother_destroy(self.handle)
            self.handle = None


class SchwarzSmoother(SchwarzPreconditioner):
    """SchwarzSmoother = SmootherRichardson around the preconditioner:
    ``step(dst, src)``: dst += omega_R * P (src - A dst), ``vmult``: dst = 0; step."""

    def step(self, dst, src, stream=None):
        _check(lib.tpss_smoother_step(self.handle, _ptr(dst), _ptr(src), _stream(stream)))

    def vmult(self, dst, src, stream=None):  # noqa: D401 -- SmootherRichardson::vmult
        _check(lib.tpss_smoother_vmult(self.handle, _ptr(dst), _ptr(src), _stream(stream)))

    def precond_vmult(self, dst, src, stream=None):
        _check(lib.tpss_precond_vmult(self.handle, _ptr(dst), _ptr(src), _stream(stream)))


class GMGSolver:
    """GMG-preconditioned CG in FP64 with Schwarz smoothers on every level."""

    def __init__(self, dim, degree, n_cells_coarse, n_levels, ip_factor=1.0, vertices=None,
                 mg_precision=F64, variant=MULTIPLICATIVE, coloring=None, relaxation=None,
                 damping_factor=1.0, n_steps=1, local_precision=None, coarse_tol=1e-12,
                 elasticity=None, local_solver=LOCAL_SURROGATE, fit_sweeps=0, ras=RAS_NONE):
        if coloring is None:
            coloring = COLOR_NONE if variant == ADDITIVE else COLOR_REDBLACK
        if relaxation is None:
            relaxation = DEFAULT_DAMPING[(variant, dim)]
        if local_precision is None:
            local_precision = mg_precision
        self._varr, vp = _vertices_ptr(vertices)
        sd = _smoother_desc(variant, coloring, relaxation, False, False, damping_factor, n_steps,
                            local_precision, local_solver, fit_sweeps, ras)
        eq, mu, lam = (POISSON, 0.0, 0.0) if elasticity is None else (ELASTICITY,) + tuple(
            float(v) for v in elasticity)
        self.desc = GMGDesc(dim, degree, n_cells_coarse, n_levels, float(ip_factor), vp,
                            mg_precision, sd, float(coarse_tol), eq, mu, lam)
        h = ctypes.c_void_p()
        _check(lib.tpss_gmg_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self.handle = h
        self.dim, self.degree, self.n_levels = dim, degree, n_levels
        self.n_fine = n_cells_coarse << (n_levels - 1)
        self.n_components = 1 if elasticity is None else dim
        self.n_dofs = self.n_components * self.n_fine ** dim * (degree + 1) ** dim
        self.mg_precision = mg_precision

    def level(self, l):
        h = ctypes.c_void_p()
        _check(lib.tpss_gmg_level(self.handle, l, ctypes.byref(h)))
        return Level._borrow(h, self.dim, self.degree, self.desc.n_cells_coarse << l,
                             self.mg_precision, owner=self)

    def smoother(self, l, post=False):
        h = ctypes.c_void_p()
        _check(lib.tpss_gmg_smoother(self.handle, l, int(post), ctypes.byref(h)))
        return SchwarzSmoother(self.level(l), _handle=h, variant=self.desc.smoother.variant,
                               coloring=self.desc.smoother.coloring,
                               relaxation=self.desc.smoother.relaxation,
                               local_precision=self.desc.smoother.local_precision,
                               _owner=self)

    def vcycle(self, dst, src, stream=None):
        _check(lib.tpss_gmg_vcycle(self.handle, _ptr(dst), _ptr(src), _stream(stream)))

    def solve(self, x, b, rel_tol=1e-8, max_iter=200, stream=None, raise_on_fail=True):
        info = SolveInfo()
        hist = (ctypes.c_double * (max_iter + 2))()
        st = lib.tpss_gmg_solve(self.handle, _ptr(x), _ptr(b), float(rel_tol), int(max_iter),
                                _stream(stream), ctypes.byref(info), hist)
        if st != OK and (raise_on_fail or st != E_NOT_CONVERGED):
            _check(st)
        n = info.iterations if info.iterations >= 0 else max_iter
        return info, [hist[i] for i in range(n + 1)]

    def __del__(self):
        if getattr(self, "handle", None):
            lib.tpss_gmg_destroy(self.handle)
            self.handle = None


def fill_random(x, seed, precision=None, stream=None):
    if precision is None:
        precision = getattr(x, "precision", F64)
    _check(lib.tpss_vector_fill_random(_ptr(x), x.numel(), precision, seed, _stream(stream)))


class DampingMLP:
    """Per-patch damping network (features -> H tanh -> H tanh -> sigmoid-scaled omega)."""

    def __init__(self, n_in=24, n_hidden=64, omega_min=0.2, omega_max=1.2, seed=4):
        h = ctypes.c_void_p()
        _check(lib.tpss_mlp_create(n_in, n_hidden, float(omega_min), float(omega_max), int(seed),
                                   ctypes.byref(h)))
        self.handle = h
        self.n_in, self.n_hidden, self.omega_min, self.omega_max = n_in, n_hidden, omega_min, omega_max
        n = ctypes.c_size_t()
        _check(lib.tpss_mlp_n_params(h, ctypes.byref(n)))
        self.n_params = n.value

    def get_params(self):
        import numpy as np
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(lib.tpss_mlp_get_params(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       out.size))
        return out

    def set_params(self, params):
        import numpy as np
        w = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib.tpss_mlp_set_params(self.handle, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       w.size))

    def forward(self, X):
        """X: host (B, n_in) array -> host omega (B,)."""
        import numpy as np
        X = np.ascontiguousarray(X, dtype=np.float32)
        dX = DeviceVector.from_numpy(X.ravel(), F32)
        out = DeviceVector(X.shape[0], F32)
        _check(lib.tpss_mlp_forward(self.handle, _ptr(dX), _ptr(out), X.shape[0], None))
        return out.numpy()

    def loss_grad(self, X, target):
        import numpy as np
        X = np.ascontiguousarray(X, dtype=np.float32)
        dX = DeviceVector.from_numpy(X.ravel(), F32)
        dt = DeviceVector.from_numpy(np.ascontiguousarray(target, dtype=np.float32), F32)
        g = DeviceVector(self.n_params, F32)
        loss = ctypes.c_double()
        _check(lib.tpss_mlp_loss_grad(self.handle, _ptr(dX), _ptr(dt), X.shape[0], _ptr(g),
                                      ctypes.byref(loss), None))
        return loss.value, g.numpy()

    def train(self, X, target, epochs=30, batch=4096, lr=2e-3, seed=1):
        import numpy as np
        X = np.ascontiguousarray(X, dtype=np.float32)
        dX = DeviceVector.from_numpy(X.ravel(), F32)
        dt = DeviceVector.from_numpy(np.ascontiguousarray(target, dtype=np.float32), F32)
        loss = ctypes.c_double()
        _check(lib.tpss_mlp_train(self.handle, _ptr(dX), _ptr(dt), X.shape[0], int(epochs),
                                  int(batch), float(lr), int(seed), ctypes.byref(loss)))
        return loss.value

    def save(self, path):
        import numpy as np
        np.save(path, self.get_params(), allow_pickle=False)

    def load(self, path):
        import numpy as np
        self.set_params(np.load(path, allow_pickle=False))
        return self

    def __del__(self):
        if getattr(self, "handle", None):
            lib.tpss_mlp_destroy(self.handle)
            self.handle = None


def ksvd(A, m1, n1, m0, n0, rank):
    """Batched Kronecker SVD on the device: A (batch, m1*m0, n1*n0) or a single
    matrix -> (B1 (batch, rank, m1, n1), B0 (batch, rank, m0, n0), sigma)."""
    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.float64)
    single = A.ndim == 2
    if single:
        A = A[None]
    batch = A.shape[0]
    dA = DeviceVector.from_numpy(A.ravel(), F64)
    dB1 = DeviceVector(batch * rank * m1 * n1)
    dB0 = DeviceVector(batch * rank * m0 * n0)
    ds = DeviceVector(batch * rank)
    _check(lib.tpss_ksvd_batched(_ptr(dA), batch, m1, n1, m0, n0, rank, _ptr(dB1), _ptr(dB0),
                                 _ptr(ds), None))
    B1 = dB1.numpy().reshape(batch, rank, m1, n1)
    B0 = dB0.numpy().reshape(batch, rank, m0, n0)
    s = ds.numpy().reshape(batch, rank)
    return (B1[0], B0[0], s[0]) if single else (B1, B0, s)


def _kron_pack(batch_terms):
    """(order, rank, m, n, packed factors) of a batch of Kronecker sums given as
    lists over rank of [A_0, ..., A_{order-1}] (A_0 rightmost)"""
    import numpy as np
    first = batch_terms[0]
    order, rank = len(first[0]), len(first)
    m = (ctypes.c_int * 3)(*[np.shape(A)[0] for A in first[0]])
    n = (ctypes.c_int * 3)(*[np.shape(A)[1] for A in first[0]])
    packed = np.concatenate([np.asarray(A, dtype=np.float64).ravel()
                             for terms in batch_terms for t in terms for A in t])
    return order, rank, m, n, packed


def kron_vmult_batched(batch_terms, X, transpose=False, add_to=None):
    """Y[b] = A_b X[b] (or A_b^T X[b]; or add_to[b] + ...) for the Kronecker sums A_b
    (TensorProductMatrix basic state vmult / Tvmult / vmult_add) on the device."""
    import numpy as np
    order, rank, m, n, packed = _kron_pack(batch_terms)
    batch = len(batch_terms)
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(batch, -1)
    nout = int(np.prod([n[d] if transpose else m[d] for d in range(order)]))
    dF, dX = DeviceVector.from_numpy(packed), DeviceVector.from_numpy(X.ravel())
    if add_to is None:
        dY, add = DeviceVector(batch * nout), 0
    else:
        dY, add = DeviceVector.from_numpy(np.asarray(add_to, dtype=np.float64).ravel()), 1
    _check(lib.tpss_kron_vmult(order, rank, m, n, batch, _ptr(dF), _ptr(dX), _ptr(dY),
                               int(transpose), add, None))
    return dY.numpy().reshape(batch, nout)


def kron_vmult(terms, x, transpose=False, add_to=None):
    return kron_vmult_batched([terms], [x], transpose,
                              None if add_to is None else [add_to])[0]


def kron_table(terms_or_batch, transpose=False):
    """dense matrices (as_table / as_transpose_table) of a batch of Kronecker sums;
    a single sum (list of rank-1 terms) gives a batch of one"""
    import numpy as np
    batch_terms = terms_or_batch if np.ndim(terms_or_batch[0][0][0]) == 2 else [terms_or_batch]
    order, rank, m, n, packed = _kron_pack(batch_terms)
    M = int(np.prod([m[d] for d in range(order)]))
    N = int(np.prod([n[d] for d in range(order)]))
    dF, dT = DeviceVector.from_numpy(packed), DeviceVector(len(batch_terms) * M * N)
    _check(lib.tpss_kron_table(order, rank, m, n, len(batch_terms), _ptr(dF), int(transpose),
                               _ptr(dT), None))
    shape = (N, M) if transpose else (M, N)
    return dT.numpy().reshape((len(batch_terms),) + shape)


def pinv_batched(A, threshold=0.0, kernel_size=-1):
    """SVD pseudo-inverses of a batch of square matrices (InverseTable)"""
    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.float64)
    batch, n = A.shape[0], A.shape[1]
    dA, dP = DeviceVector.from_numpy(A.ravel()), DeviceVector(A.size)
    _check(lib.tpss_kron_pinv(_ptr(dA), n, batch, float(threshold), int(kernel_size), _ptr(dP),
                              None))
    return dP.numpy().reshape(A.shape)


def ksvd_sum(big, small, rank):
    """Kronecker SVD of sum_i big_i (x) small_i (batch, in_rank, m, n) arrays, or of
    one sum given as lists -> (B1, B0, sigma) like ksvd()"""
    import numpy as np
    big = np.ascontiguousarray(big, dtype=np.float64)
    small = np.ascontiguousarray(small, dtype=np.float64)
    single = big.ndim == 3
    if single:
        big, small = big[None], small[None]
    batch, in_rank, m1, n1 = big.shape
    m0, n0 = small.shape[2:]
    dA, dB = DeviceVector.from_numpy(big.ravel()), DeviceVector.from_numpy(small.ravel())
    d1, d0, ds = (DeviceVector(batch * rank * m1 * n1), DeviceVector(batch * rank * m0 * n0),
                  DeviceVector(batch * rank))
    _check(lib.tpss_ksvd_sum_batched(_ptr(dA), _ptr(dB), batch, in_rank, m1, n1, m0, n0, rank,
                                     _ptr(d1), _ptr(d0), _ptr(ds), None))
    B1 = d1.numpy().reshape(batch, rank, m1, n1)
    B0 = d0.numpy().reshape(batch, rank, m0, n0)
    s = ds.numpy().reshape(batch, rank)
    return (B1[0], B0[0], s[0]) if single else (B1, B0, s)


def kcp(batch_terms, rank, iterations=0):
    """CP-ALS (compute_kcp) of a batch of sums of 3-fold Kronecker products, each a
    list over input terms of [C_i, B_i, A_i]; returns rank-R lists in that order"""
    import numpy as np
    shapes = [np.shape(A) for A in batch_terms[0][0]]
    in_rank = len(batch_terms[0])
    packed = np.concatenate([np.asarray(A, dtype=np.float64).ravel()
                             for terms in batch_terms for t in terms for A in t])
    m = (ctypes.c_int * 3)(*[s_[0] for s_ in shapes])
    n = (ctypes.c_int * 3)(*[s_[1] for s_ in shapes])
    term = sum(a * b for a, b in shapes)
    dIn, dOut = DeviceVector.from_numpy(packed), DeviceVector(len(batch_terms) * rank * term)
    _check(lib.tpss_kcp_batched(_ptr(dIn), len(batch_terms), in_rank, m, n, rank,
                                int(iterations), _ptr(dOut), None))
    out = dOut.numpy().reshape(len(batch_terms), rank, term)
    res = []
    for b in range(len(batch_terms)):
        terms = []
        for r in range(rank):
            off, facs = 0, []
            for s_ in shapes:
                facs.append(out[b, r, off:off + s_[0] * s_[1]].reshape(s_))
                off += s_[0] * s_[1]
            terms.append(facs)
        res.append(terms)
    return res


def version():
    return lib.tpss_version().decode()
