"""Independent dense assembly of the SIPG Poisson matrix on small meshes (numpy).

Written from the textbook SIPG bilinear form with full dim-dimensional
quadrature and numpy's own Legendre machinery, deliberately unlike the
oracle's 1D-factorised and per-quadrature-point code, so that agreement pins
both.  Form (apps/laplace_integrator.h:613-743):
  cell      int grad u . grad v
  interior  sigma [[u]][[v]] - {d_n u}[[v]] - [[u]]{d_n v},
            sigma = ip * gamma * (1/h_in + 1/h_out) / 2 ... (ip * 0.5 * 2 gamma (0.5/h + 0.5/h))
  boundary  sigma_b u v - d_n u v - u d_n v,  sigma_b = ip * 2 gamma / h
with gamma = p (p + 1).
"""
import itertools

import numpy as np
from numpy.polynomial import legendre as L


def gll_nodes(p):
    """Gauss-Lobatto nodes on [0,1] from numpy's Legendre roots."""
    if p == 0:
        return np.array([0.5])
    inner = L.Legendre.basis(p).deriv().roots()
    z = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    return 0.5 * (z + 1.0)


def lagrange_basis(nodes):
    """Monomial coefficients of the Lagrange basis (columns)."""
    V = np.vander(nodes, increasing=True)
    return np.linalg.inv(V)  # phi_j(x) = sum_m C[m, j] x^m


def eval_basis(C, x):
    x = np.atleast_1d(x)
    m = C.shape[0]
    V = np.vander(x, m, increasing=True)
    dV = np.zeros_like(V)
    for e in range(1, m):
        dV[:, e] = e * x ** (e - 1)
    return V @ C, dV @ C  # (npts, nbasis)


def assemble(dim, p, n, ip):
    """Dense SIPG matrix, DoFs cell-major (cell lexicographic x fastest,
    local lexicographic x fastest)."""
    h = 1.0 / n
    nodes = gll_nodes(p)
    C = lagrange_basis(nodes)
    gx, gw = L.leggauss(p + 1)
    qx, qw = 0.5 * (gx + 1.0), 0.5 * gw
    nb = p + 1
    ndc = nb ** dim
    ncell = n ** dim
    N = ncell * ndc
    A = np.zeros((N, N))
    gamma = p * (p + 1) if p > 0 else 1

    def local_multi(i):
        return [(i // nb ** e) % nb for e in range(dim)]

    loc = [local_multi(i) for i in range(ndc)]

    def shape(xi):
        """values (ndc,), physical gradients (ndc, dim) at reference point xi."""
        v1, d1 = zip(*[eval_basis(C, xi[e]) for e in range(dim)])
        val = np.array([np.prod([v1[e][0, li[e]] for e in range(dim)]) for li in loc])
        grad = np.zeros((ndc, dim))
        for a in range(dim):
            grad[:, a] = [np.prod([(d1[e][0, li[e]] / h) if e == a else v1[e][0, li[e]]
                                   for e in range(dim)]) for li in loc]
        return val, grad

    def cell_index(cc):
        return sum(cc[e] * n ** e for e in range(dim))

    # cell integrals
    Ke = np.zeros((ndc, ndc))
    for q in itertools.product(range(p + 1), repeat=dim):
        xi = [qx[q[e]] for e in range(dim)]
        w = np.prod([qw[q[e]] for e in range(dim)]) * h ** dim
        _, g = shape(xi)
        Ke += w * g @ g.T
    for c in range(ncell):
        s = slice(c * ndc, (c + 1) * ndc)
        A[s, s] += Ke

    sig_i = ip * 0.5 * 2 * gamma * (0.5 / h + 0.5 / h)
    sig_b = ip * 2 * gamma / h
    for cc in itertools.product(range(n), repeat=dim):
        cc = list(cc)[::-1]  # any order; indices computed explicitly
        c = cell_index(cc)
        for d in range(dim):
            for side in (0, 1):
                nrm = np.zeros(dim)
                nrm[d] = 1.0 if side == 1 else -1.0
                neighbour = cc[d] + (1 if side == 1 else -1)
                is_bdry = neighbour < 0 or neighbour >= n
                if not is_bdry and side == 0:
                    continue  # interior faces once, from the lower cell
                for q in itertools.product(range(p + 1), repeat=dim - 1):
                    xi_m, xi_p = [0.0] * dim, [0.0] * dim
                    w = h ** (dim - 1)
                    t = 0
                    for e in range(dim):
                        if e == d:
                            xi_m[e] = float(side)
                            xi_p[e] = 0.0
                        else:
                            xi_m[e] = xi_p[e] = qx[q[t]]
                            w *= qw[q[t]]
                            t += 1
                    vm, gm = shape(xi_m)
                    dnm = gm @ nrm
                    sm = slice(c * ndc, (c + 1) * ndc)
                    if is_bdry:
                        A[sm, sm] += w * (sig_b * np.outer(vm, vm) - np.outer(vm, dnm)
                                          - np.outer(dnm, vm))
                        continue
                    cn = list(cc)
                    cn[d] = neighbour
                    cp = cell_index(cn)
                    vp, gp = shape(xi_p)
                    dnp = gp @ nrm
                    sp = slice(cp * ndc, (cp + 1) * ndc)
                    # jump test functions: [[v]] = (vm, -vp); average of normal derivative (dnm, dnp)/2
                    jv = {0: vm, 1: -vp}
                    av = {0: 0.5 * dnm, 1: 0.5 * dnp}
                    sl = {0: sm, 1: sp}
                    for a in (0, 1):      # test side
                        for b in (0, 1):  # trial side
                            A[sl[a], sl[b]] += w * (sig_i * np.outer(jv[a], jv[b])
                                                    - np.outer(jv[a], av[b])
                                                    - np.outer(av[a], jv[b]))
    return A
