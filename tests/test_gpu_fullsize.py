"""Full-size properties of the hot path at the benchmark's sizes (64^3 cells), where
the oracle would take minutes: results that two different device paths must share,
and identities of the operators and smoothers that hold at any size.

Bars: FP64 within 1e-12 (relative, max norm or relative to the norms involved)."""
import numpy as np
import pytest

from test_gpu_parity import _tpss, dev, host, rel_err

pytestmark = pytest.mark.gpu

N_CELLS = 64


def _rand(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


@pytest.fixture(scope="module")
def q3():
    t = _tpss()
    lv = t.Level(3, 3, N_CELLS)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE, relaxation=1.0)
    return t, lv, sm


def test_residual_free_pass_matches_residual_update_path(q3):
    """The benchmark's FP64 step (16 residual-free color passes) against the
    reference's order (localized residual, then the patch corrections) on the
    whole 64^3 Q3 level, forward and transposed color sequences."""
    t, lv, sm = q3
    b = _rand(lv.n_dofs, 1)
    x0 = _rand(lv.n_dofs, 2)
    out = {}
    for fused in (True, False):
        sm.set_fused_residual(fused)
        x = dev(x0)
        sm.step(x, dev(b))
        d = dev(np.zeros(lv.n_dofs))
        sm.Tvmult(d, dev(b))
        out[fused] = (host(x), host(d))
    sm.set_fused_residual(True)
    assert rel_err(out[True][0], out[False][0]) < 1e-12
    assert rel_err(out[True][1], out[False][1]) < 1e-12


def test_preconditioner_linearity_and_adjoint(q3):
    """P(a u + c v) = a P u + c P v, and <P^T u, v> = <u, P v> (Tvmult runs the
    colors in reverse; the local solvers are symmetric)."""
    t, lv, sm = q3
    u, v = _rand(lv.n_dofs, 3), _rand(lv.n_dofs, 4)
    a, c = 0.75, -1.25

    def P(w):
        d = dev(np.zeros(lv.n_dofs))
        sm.precond_vmult(d, dev(w))
        return host(d)

    def PT(w):
        d = dev(np.zeros(lv.n_dofs))
        sm.Tvmult(d, dev(w))
        return host(d)

    Pu, Pv = P(u), P(v)
    assert rel_err(P(a * u + c * v), a * Pu + c * Pv) < 1e-12
    lhs, rhs = PT(u) @ v, u @ Pv
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Pv) * 10


def test_multiplicative_step_contracts_the_error_in_the_energy_norm(q3):
    """With exact local solvers the multiplicative sweep is a block Gauss-Seidel
    iteration on an s.p.d. matrix: with b = 0 a step maps the error e to e' with
    |e'|_A < |e|_A."""
    t, lv, sm = q3
    e = _rand(lv.n_dofs, 5)

    def energy(w):
        y = dev(np.zeros(lv.n_dofs))
        lv.vmult(y, dev(w))
        return float(w @ host(y))

    x = dev(e)
    sm.step(x, dev(np.zeros(lv.n_dofs)))
    e1 = host(x)
    E0, E1 = energy(e), energy(e1)
    assert 0.0 < E1 < 0.5 * E0


def test_moved_vertex_operator_symmetry_and_linear_consistency():
    """The moved-vertex SIPG operator at 64^3 Q3 (vertices moved by 0.2 h): symmetric,
    and zero on the rows of cells without boundary faces for a linear function
    (the mapped Q1 space is contained in the cells' spaces and QGauss(p+1)
    integrates the consistency terms exactly)."""
    t = _tpss()
    import oracle as o
    n, p = N_CELLS, 3
    V = o.distorted_vertices(3, n, 0.2, seed=3)
    lv = t.Level(3, p, n, ip_factor=4.0, vertices=V)
    u, v = _rand(lv.n_dofs, 6), _rand(lv.n_dofs, 7)
    Au, Av = dev(np.zeros(lv.n_dofs)), dev(np.zeros(lv.n_dofs))
    lv.vmult(Au, dev(u))
    lv.vmult(Av, dev(v))
    Au, Av = host(Au), host(Av)
    assert abs(Au @ v - u @ Av) <= 1e-12 * np.abs(Au).max() * np.abs(v).sum()
    # u = 0.3 x - 1.1 y + 0.7 z + 0.2 at the physical DoF positions (trilinear map
    # of each cell's corners at the Gauss-Lobatto nodes, cell-major, local
    # lexicographic with x fastest)
    xi = np.asarray(o.gauss_lobatto(p + 1))
    W = np.stack([xi, 1.0 - xi])  # W[b][i]: weight of the corner with bit b at node i
    Vg = V.reshape(n + 1, n + 1, n + 1, 3)  # [z][y][x]
    X = np.zeros((n, n, n, p + 1, p + 1, p + 1, 3))  # [cz][cy][cx][iz][iy][ix]
    for bz in (0, 1):
        for by in (0, 1):
            for bx in (0, 1):
                C = Vg[bz:bz + n, by:by + n, bx:bx + n]  # corner (bx, by, bz) of every cell
                w = np.einsum("k,j,i->kji", W[1 - bz], W[1 - by], W[1 - bx])
                X += C[:, :, :, None, None, None, :] * w[None, None, None, :, :, :, None]
    X = X.reshape(-1, 3)
    lin = 0.3 * X[:, 0] - 1.1 * X[:, 1] + 0.7 * X[:, 2] + 0.2
    y = dev(np.zeros(lv.n_dofs))
    lv.vmult(y, dev(lin))
    y = host(y)
    scale = np.abs(y).max()
    assert scale > 1e-3  # the rows of the boundary cells do not vanish
    inner = y.reshape(n, n, n, -1)[1:-1, 1:-1, 1:-1]
    assert np.abs(inner).max() < 1e-11 * scale


def test_elasticity_operator_symmetry():
    """The SIPG elasticity operator at 64^3 Q2 (mu 1, lambda 10, ip 10) is symmetric."""
    t = _tpss()
    lv = t.Level(3, 2, N_CELLS, ip_factor=10.0, elasticity=(1.0, 10.0))
    u, v = _rand(lv.n_dofs, 8), _rand(lv.n_dofs, 9)
    Au, Av = dev(np.zeros(lv.n_dofs)), dev(np.zeros(lv.n_dofs))
    lv.vmult(Au, dev(u))
    lv.vmult(Av, dev(v))
    Au, Av = host(Au), host(Av)
    assert abs(Au @ v - u @ Av) <= 1e-12 * np.abs(Au).max() * np.abs(v).sum()
