"""ctypes view of the CPU oracle (oracle/libtpss_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP path.  Only the tests, the
smoke check and the CPU-baseline leg of the benchmark import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libtpss_oracle.so")

ADDITIVE, MULTIPLICATIVE = 0, 1
COLOR_NONE, COLOR_TILED, COLOR_REDBLACK = 0, 1, 2
LOCAL_SURROGATE, LOCAL_FIT = 0, 1
RAS_NONE, RAS_WEIGHTS, RAS_BOOLEAN = 0, 1, 2

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U = ctypes.POINTER(ctypes.c_uint32)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    sig = {
        "or_gauss_legendre": (None, [ctypes.c_int, _D, _D]),
        "or_gauss_lobatto": (None, [ctypes.c_int, _D]),
        "or_lagrange": (None, [ctypes.c_int, _D, ctypes.c_double, _D, _D]),
        "or_penalty": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double]),
        "or_patch_matrices_1d": (None, [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, _D, _D]),
        "or_patch_matrices_1d_coef": (None, [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             _D, _D]),
        "or_gen_eig": (ctypes.c_int, [ctypes.c_int, _D, _D, _D, _D]),
        "or_fd_apply": (None, [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]),
        "or_create": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int]),
        "or_create_elasticity": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_double, ctypes.c_int]),
        "or_n_components": (ctypes.c_int, [ctypes.c_void_p]),
        "or_set_local_solver": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
        "or_fit_errors": (None, [ctypes.c_void_p, _D]),
        "or_set_ras": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "or_ras_weights": (None, [ctypes.c_void_p, _D]),
        "or_restricted_flags": (None, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ubyte)]),
        "or_fit_fallbacks": (ctypes.c_int, [ctypes.c_void_p]),
        "or_patch_local_matrix": (None, [ctypes.c_void_p, ctypes.c_int, _D]),
        "or_separable_fit": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, _D, _D, _D, ctypes.c_int]),
        "or_destroy": (None, [ctypes.c_void_p]),
        "or_set_vertices": (ctypes.c_int, [ctypes.c_void_p, _D]),
        "or_setup": (ctypes.c_int, [ctypes.c_void_p]),
        "or_set_patch_damping": (None, [ctypes.c_void_p, _D]),
        "or_set_threads": (None, [ctypes.c_void_p, ctypes.c_int]),
        "or_n_dofs": (ctypes.c_long, [ctypes.c_void_p]),
        "or_n_patches": (ctypes.c_int, [ctypes.c_void_p]),
        "or_n_colors": (ctypes.c_int, [ctypes.c_void_p]),
        "or_color_ptr": (None, [ctypes.c_void_p, _I]),
        "or_patch_vertices": (None, [ctypes.c_void_p, _I]),
        "or_patch_dof_indices": (None, [ctypes.c_void_p, _U]),
        "or_patch_h": (None, [ctypes.c_void_p, _D]),
        "or_patch_factors": (None, [ctypes.c_void_p, _D, _D]),
        "or_color_sequence": (None, [ctypes.c_void_p, ctypes.c_int, _I, _I]),
        "or_operator_vmult": (None, [ctypes.c_void_p, _D, _D]),
        "or_local_apply": (None, [ctypes.c_void_p, ctypes.c_int, _D, _D]),
        "or_precond_vmult": (None, [ctypes.c_void_p, _D, _D]),
        "or_smoother_step": (None, [ctypes.c_void_p, _D, _D]),
        "or_smoother_vmult": (None, [ctypes.c_void_p, _D, _D]),
        "or_prolongation_1d": (None, [ctypes.c_int, _D]),
        "or_prolongate_add": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D]),
        "or_restrict": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D]),
        "or_gmg_create": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_double]),
        "or_gmg_create_v": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_int, ctypes.c_double, _D]),
        "or_gmg_create_elasticity": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_int, ctypes.c_double,
                                                       ctypes.c_double, ctypes.c_double,
                                                       ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                       ctypes.c_int, ctypes.c_double]),
        "or_gmg_destroy": (None, [ctypes.c_void_p]),
        "or_gmg_level": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_int]),
        "or_gmg_smoother": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
        "or_gmg_set_threads": (None, [ctypes.c_void_p, ctypes.c_int]),
        "or_gmg_vcycle": (None, [ctypes.c_void_p, _D, _D]),
        "or_gmg_cg": (ctypes.c_int, [ctypes.c_void_p, _D, _D, ctypes.c_double, ctypes.c_int, _D]),
        "or_ksvd": (ctypes.c_int, [ctypes.c_int] * 4 + [_D, ctypes.c_int, _D, _D, _D]),
        "or_ksvd_diag": (None, [ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _D, _D, _D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def ip_(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)


def gauss_legendre(nq):
    x, w = np.zeros(nq), np.zeros(nq)
    lib.or_gauss_legendre(nq, dp(x), dp(w))
    return x, w


def gauss_lobatto(n):
    x = np.zeros(n)
    lib.or_gauss_lobatto(n, dp(x))
    return x


def lagrange(nodes, x):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    v, d = np.zeros(len(nodes)), np.zeros(len(nodes))
    lib.or_lagrange(len(nodes), dp(nodes), float(x), dp(v), dp(d))
    return v, d


def patch_matrices_1d(p, h, bdry_left, bdry_right, ip):
    k = 2 * (p + 1)
    h = np.ascontiguousarray(h, dtype=np.float64)
    M, L = np.zeros((k, k)), np.zeros((k, k))
    lib.or_patch_matrices_1d(p, dp(h), int(bdry_left), int(bdry_right), float(ip), dp(M), dp(L))
    return M, L


def patch_matrices_1d_coef(p, h, bdry_left, bdry_right, ip, a_cons, a_pen):
    k = 2 * (p + 1)
    h = np.ascontiguousarray(h, dtype=np.float64)
    M, L = np.zeros((k, k)), np.zeros((k, k))
    lib.or_patch_matrices_1d_coef(p, dp(h), int(bdry_left), int(bdry_right), float(ip),
                                  float(a_cons), float(a_pen), dp(M), dp(L))
    return M, L


def separable_fit(dim, A, M, L, sweeps):
    """ALS separable fit of A (k^dim x k^dim); M, L: (dim, k, k) initial guess."""
    k = M.shape[-1]
    A = np.ascontiguousarray(A, dtype=np.float64)
    M = np.ascontiguousarray(M, dtype=np.float64).copy()
    L = np.ascontiguousarray(L, dtype=np.float64).copy()
    err = lib.or_separable_fit(dim, k, dp(A), dp(M), dp(L), sweeps)
    return M, L, err


def gen_eig(A, B):
    k = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    S, lam = np.zeros((k, k)), np.zeros(k)
    st = lib.or_gen_eig(k, dp(A), dp(B), dp(S), dp(lam))
    if st:
        raise np.linalg.LinAlgError("mass matrix not positive definite")
    return S, lam


def fd_apply(dim, S, lam, r):
    k = lam.shape[-1]
    S = np.ascontiguousarray(S, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    u = np.zeros_like(r)
    lib.or_fd_apply(dim, k, dp(S), dp(lam), dp(r), dp(u))
    return u


def prolongation_1d(p):
    P = np.zeros((2, p + 1, p + 1))
    lib.or_prolongation_1d(p, dp(P))
    return P


def prolongate_add(dim, p, n_coarse, xc, xf):
    lib.or_prolongate_add(dim, p, n_coarse, dp(xc), dp(xf))


def restrict(dim, p, n_coarse, xf):
    xc = np.zeros(n_coarse ** dim * (p + 1) ** dim)
    lib.or_restrict(dim, p, n_coarse, dp(xf), dp(xc))
    return xc


def ksvd(A, m1, n1, m0, n0, rank):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B1, B0, s = np.zeros((rank, m1, n1)), np.zeros((rank, m0, n0)), np.zeros(rank)
    lib.or_ksvd(m1, n1, m0, n0, dp(A), rank, dp(B1), dp(B0), dp(s))
    return B1, B0, s


def ksvd_diag(d, m1, m0, rank):
    d = np.ascontiguousarray(d, dtype=np.float64)
    B1, B0, s = np.zeros((rank, m1)), np.zeros((rank, m0)), np.zeros(rank)
    lib.or_ksvd_diag(m1, m0, dp(d), rank, dp(B1), dp(B0), dp(s))
    return B1, B0, s


class Smoother:
    """CPU FP64 Schwarz smoother on one level (oracle)."""

    def __init__(self, dim, p, n, ip=1.0, variant=ADDITIVE, coloring=COLOR_NONE,
                 relaxation=1.0, reverse=False, symmetrized=False, damping=1.0,
                 n_steps=1, vertices=None, threads=1, elasticity=None, local_solver=LOCAL_SURROGATE,
                 fit_sweeps=4, ras=0):
        """elasticity=(mu, lambda) selects the linear elasticity operator with
        dim components and block-diagonal local solvers."""
        self.dim, self.p, self.n = dim, p, n
        if elasticity is None:
            self.h = lib.or_create(dim, p, n, float(ip), variant, coloring, float(relaxation),
                                   int(reverse), int(symmetrized), float(damping), n_steps)
        else:
            mu, lam = elasticity
            self.h = lib.or_create_elasticity(dim, p, n, float(ip), float(mu), float(lam), variant,
                                              coloring, float(relaxation), int(reverse),
                                              int(symmetrized), float(damping), n_steps)
        if not self.h:
            raise ValueError("invalid smoother parameters")
        if vertices is not None:
            self._v = np.ascontiguousarray(vertices, dtype=np.float64)
            lib.or_set_vertices(self.h, dp(self._v))
        if threads > 1:
            lib.or_set_threads(self.h, threads)
        if local_solver != LOCAL_SURROGATE:
            lib.or_set_local_solver(self.h, local_solver, fit_sweeps)
        st = lib.or_setup(self.h)
        if st:
            raise np.linalg.LinAlgError("local solver setup failed")
        self.N = lib.or_n_dofs(self.h)
        self.P = lib.or_n_patches(self.h)
        self.n_colors = lib.or_n_colors(self.h)
        self.k = 2 * (p + 1)
        self.ncomp = lib.or_n_components(self.h)
        self.ndp = self.ncomp * self.k ** dim  # patch-DoFs over all components
        if ras and lib.or_set_ras(self.h, int(ras)):
            raise ValueError("restricted additive Schwarz: additive, scalar problems only")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.or_destroy(h)
            self.h = None

    def color_ptr(self):
        out = np.zeros(self.n_colors + 1, dtype=np.int32)
        lib.or_color_ptr(self.h, ip_(out))
        return out

    def patch_vertices(self):
        out = np.zeros((self.P, self.dim), dtype=np.int32)
        lib.or_patch_vertices(self.h, ip_(out))
        return out

    def patch_dof_indices(self):
        out = np.zeros((self.P, self.ndp), dtype=np.uint32)
        lib.or_patch_dof_indices(self.h, out.ctypes.data_as(_U))
        return out

    def patch_h(self):
        out = np.zeros((self.P, self.dim, 2))
        lib.or_patch_h(self.h, dp(out))
        return out

    def patch_factors(self):
        S = np.zeros((self.P, self.ncomp, self.dim, self.k, self.k))
        lam = np.zeros((self.P, self.ncomp, self.dim, self.k))
        lib.or_patch_factors(self.h, dp(S), dp(lam))
        if self.ncomp == 1:
            return S[:, 0], lam[:, 0]
        return S, lam

    def fit_errors(self):
        out = np.zeros(self.P)
        lib.or_fit_errors(self.h, dp(out))
        return out

    def ras_weights(self):
        out = np.zeros(self.N)
        lib.or_ras_weights(self.h, dp(out))
        return out

    def restricted_flags(self):
        """(P, ndp) uint8: 1 where rscatter_add skips the patch DoF"""
        out = np.zeros((self.P, self.ndp), dtype=np.uint8)
        lib.or_restricted_flags(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return out

    def fit_fallbacks(self):
        return lib.or_fit_fallbacks(self.h)

    def patch_local_matrix(self, patch):
        A = np.zeros((self.ndp, self.ndp))
        lib.or_patch_local_matrix(self.h, int(patch), dp(A))
        return A

    def color_sequence(self, transpose=False):
        out = np.zeros(2 * self.n_colors, dtype=np.int32)
        n = ctypes.c_int(0)
        lib.or_color_sequence(self.h, int(transpose), ip_(out), ctypes.byref(n))
        return out[: n.value].copy()

    def set_patch_damping(self, omega):
        if omega is None:
            lib.or_set_patch_damping(self.h, None)
        else:
            self._omega = np.ascontiguousarray(omega, dtype=np.float64)
            lib.or_set_patch_damping(self.h, dp(self._omega))

    def operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.N)
        lib.or_operator_vmult(self.h, dp(y), dp(x))
        return y

    def local_apply(self, patch, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        u = np.zeros(self.ndp)
        lib.or_local_apply(self.h, int(patch), dp(r), dp(u))
        return u

    def precond_vmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros(self.N)
        lib.or_precond_vmult(self.h, dp(dst), dp(src))
        return dst

    def step(self, dst, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        assert dst.dtype == np.float64 and dst.flags["C_CONTIGUOUS"]
        lib.or_smoother_step(self.h, dp(dst), dp(src))
        return dst

    def smoother_vmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros(self.N)
        lib.or_smoother_vmult(self.h, dp(dst), dp(src))
        return dst


class GMG:
    """CPU FP64 geometric multigrid + CG (oracle)."""

    def __init__(self, dim, p, n_coarse, n_levels, ip=1.0, variant=MULTIPLICATIVE,
                 coloring=COLOR_REDBLACK, relaxation=1.0, n_steps=1, coarse_tol=1e-12,
                 vertices=None, elasticity=None, threads=1):
        vp = None
        if vertices is not None:
            self._v = np.ascontiguousarray(vertices, dtype=np.float64)
            vp = dp(self._v)
        if elasticity is None:
            self.h = lib.or_gmg_create_v(dim, p, n_coarse, n_levels, float(ip), variant, coloring,
                                         float(relaxation), n_steps, float(coarse_tol), vp)
            ncomp = 1
        else:
            assert vertices is None, "elasticity: Cartesian meshes only"
            mu, lam = elasticity
            self.h = lib.or_gmg_create_elasticity(dim, p, n_coarse, n_levels, float(ip), float(mu),
                                                  float(lam), variant, coloring, float(relaxation),
                                                  n_steps, float(coarse_tol))
            ncomp = dim
        if threads > 1:
            lib.or_gmg_set_threads(self.h, int(threads))
        self.N = ncomp * (n_coarse << (n_levels - 1)) ** dim * (p + 1) ** dim

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.or_gmg_destroy(h)
            self.h = None

    def set_patch_damping(self, level, omega):
        """per-patch damping of the pre- and post-smoother of one level (oracle patch order)"""
        self._om = getattr(self, "_om", {})
        self._om[level] = np.ascontiguousarray(omega, dtype=np.float64)
        for post in (0, 1):
            lib.or_set_patch_damping(lib.or_gmg_smoother(self.h, level, post), dp(self._om[level]))

    def vcycle(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.N)
        lib.or_gmg_vcycle(self.h, dp(x), dp(b))
        return x

    def cg(self, b, rel_tol=1e-8, max_it=200):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.N)
        hist = np.zeros(max_it + 1)
        it = lib.or_gmg_cg(self.h, dp(x), dp(b), float(rel_tol), max_it, dp(hist))
        return x, it, hist[: (it if it >= 0 else max_it) + 1]


def uniform_vertices(dim, n):
    g = np.linspace(0.0, 1.0, n + 1)
    mesh = np.meshgrid(*([g] * dim), indexing="ij")
    # vertex index x fastest: reverse axes of an ij grid
    coords = np.stack([m.transpose(tuple(range(dim))[::-1]).ravel() for m in mesh], axis=-1)
    return np.ascontiguousarray(coords)


def splitmix_uniform(seed, n):
    """U[-1,1] keyed by (seed, index): SplitMix64 of seed*2^32 + i (shared with the HIP library)."""
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * np.uint64(0x100000000) + i
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def distorted_vertices(dim, n, amplitude, seed):
    """Interior vertices moved by h*amplitude*U[-1,1]^dim (SplitMix64 keyed by
    (seed, vertex*dim + d)); boundary vertices stay (kept boundary)."""
    v = uniform_vertices(dim, n)
    nv = v.shape[0]
    noise = splitmix_uniform(seed, nv * dim).reshape(nv, dim)
    idx = np.arange(nv)
    interior = np.ones(nv, dtype=bool)
    t = idx.copy()
    for _ in range(dim):
        c = t % (n + 1)
        interior &= (c > 0) & (c < n)
        t //= (n + 1)
    v[interior] += (amplitude / n) * noise[interior]
    return v
