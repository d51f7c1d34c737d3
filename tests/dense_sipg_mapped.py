"""Independent dense assembly of the SIPG Poisson matrix on a multilinear mesh
given by its vertices (numpy), for checking the oracle's moved-vertex operator.

Written from the geometry itself, unlike the oracle's inverse-Jacobian code:
* the cell map is the multilinear interpolation of the 2^dim corner vertices,
  its Jacobian J[a, b] = dx_a / dxhat_b from the corner differences;
* physical gradients solve J^T g = ghat (no explicit inverse);
* a face's normal and surface element come from its tangent vectors (the
  rotated tangent in 2D, the cross product of the two tangents in 3D), the
  orientation from the direction of increasing xhat_d, not from J^{-T} e_d;
* the penalty length of a cell at a face is 1 / |n . grad xhat_d| at the
  face's first quadrature point (all tangential coordinates at the smallest
  Gauss point), grad xhat_d from solve(J^T, e_d)
  (apps/laplace_integrator.h:651-656, 692-695; on a parallelepiped cell this is
  the distance between the face and the opposite one).
SIPG form (apps/laplace_integrator.h:613-743), QGauss(p+1) in every direction:
  cell      int grad u . grad v
  interior  sigma [[u]][[v]] - {d_n u}[[v]] - [[u]]{d_n v},
            sigma = ip * 0.5 * 2 gamma (0.5 / h_in + 0.5 / h_out)
  boundary  sigma_b u v - d_n u v - u d_n v,  sigma_b = ip * 2 gamma / h
with gamma = p (p + 1).  DoFs cell-major, cells and local DoFs lexicographic
(x fastest), vertices ((n+1)^dim, dim) lexicographic (x fastest).
"""
import itertools

import numpy as np
from numpy.polynomial import legendre as L

from dense_sipg import eval_basis, gll_nodes, lagrange_basis


def _corners(V, dim, n, cc):
    """corner vertices of cell cc, corner index = sum_e bit_e 2^e"""
    out = np.zeros((2 ** dim, dim))
    for corner in range(2 ** dim):
        iv = [cc[e] + ((corner >> e) & 1) for e in range(dim)]
        out[corner] = V[sum(iv[e] * (n + 1) ** e for e in range(dim))]
    return out


def _map(X, dim, xi):
    """x(xi) and J[a, b] = dx_a/dxi_b of the multilinear map with corners X"""
    x = np.zeros(dim)
    J = np.zeros((dim, dim))
    for corner in range(2 ** dim):
        bits = [(corner >> e) & 1 for e in range(dim)]
        fac = [xi[e] if bits[e] else 1.0 - xi[e] for e in range(dim)]
        x += np.prod(fac) * X[corner]
        for b in range(dim):
            dfac = np.prod([(1.0 if bits[e] else -1.0) if e == b else fac[e] for e in range(dim)])
            J[:, b] += dfac * X[corner]
    return x, J


def assemble_mapped(dim, p, n, ip, V):
    nodes = gll_nodes(p)
    C = lagrange_basis(nodes)
    gx, gw = L.leggauss(p + 1)
    qx, qw = 0.5 * (gx + 1.0), 0.5 * gw
    nb = p + 1
    ndc = nb ** dim
    N = n ** dim * ndc
    A = np.zeros((N, N))
    gamma = p * (p + 1) if p > 0 else 1
    loc = [[(i // nb ** e) % nb for e in range(dim)] for i in range(ndc)]

    def ref_shape(xi):
        v1, d1 = zip(*[eval_basis(C, xi[e]) for e in range(dim)])
        val = np.array([np.prod([v1[e][0, li[e]] for e in range(dim)]) for li in loc])
        grad = np.zeros((ndc, dim))
        for a in range(dim):
            grad[:, a] = [np.prod([d1[e][0, li[e]] if e == a else v1[e][0, li[e]]
                                   for e in range(dim)]) for li in loc]
        return val, grad

    def phys(X, xi):
        x, J = _map(X, dim, xi)
        val, g = ref_shape(xi)
        G = np.linalg.solve(J.T, g.T).T  # rows: J^{-T} ghat_i
        return val, G, J

    def face_geometry(J, d, side):
        """unit outward normal and surface element of face (d, side)"""
        tang = [J[:, e] for e in range(dim) if e != d]
        if dim == 2:
            t = tang[0]
            nv = np.array([t[1], -t[0]])
        else:
            nv = np.cross(tang[0], tang[1])
        ds = np.linalg.norm(nv)
        nv = nv / ds
        outward = J[:, d] * (1.0 if side == 1 else -1.0)
        if nv @ outward < 0:
            nv = -nv
        return nv, ds

    def pen_length(J, d, nv):
        ed = np.zeros(dim)
        ed[d] = 1.0
        return 1.0 / abs(nv @ np.linalg.solve(J.T, ed))

    def cell_index(cc):
        return sum(cc[e] * n ** e for e in range(dim))

    for cc in itertools.product(range(n), repeat=dim):
        cc = list(cc)
        X = _corners(V, dim, n, cc)
        c = cell_index(cc)
        sm = slice(c * ndc, (c + 1) * ndc)
        for q in itertools.product(range(nb), repeat=dim):
            xi = [qx[q[e]] for e in range(dim)]
            _, G, J = phys(X, xi)
            A[sm, sm] += np.prod([qw[t] for t in q]) * abs(np.linalg.det(J)) * G @ G.T
        for d in range(dim):
            for side in (0, 1):
                nbr = cc[d] + (1 if side == 1 else -1)
                is_bdry = nbr < 0 or nbr >= n
                if not is_bdry and side == 0:
                    continue  # each interior face once, from its lower cell
                if not is_bdry:
                    cn = list(cc)
                    cn[d] = nbr
                    Xp = _corners(V, dim, n, cn)
                    cp = cell_index(cn)
                    sp = slice(cp * ndc, (cp + 1) * ndc)
                # penalty lengths at the first face quadrature point
                xi0 = [float(side) if e == d else qx[0] for e in range(dim)]
                _, J0 = _map(X, dim, xi0)
                nv0, _ = face_geometry(J0, d, side)
                h_in = pen_length(J0, d, nv0)
                if is_bdry:
                    sigma = ip * 2.0 * gamma / h_in
                else:
                    xp0 = [0.0 if e == d else qx[0] for e in range(dim)]
                    _, Jp0 = _map(Xp, dim, xp0)
                    h_out = pen_length(Jp0, d, nv0)
                    sigma = ip * 0.5 * 2.0 * gamma * (0.5 / h_in + 0.5 / h_out)
                for qt in itertools.product(range(nb), repeat=dim - 1):
                    xi_m, xi_p, w, t = [0.0] * dim, [0.0] * dim, 1.0, 0
                    for e in range(dim):
                        if e == d:
                            xi_m[e], xi_p[e] = float(side), 0.0
                        else:
                            xi_m[e] = xi_p[e] = qx[qt[t]]
                            w *= qw[qt[t]]
                            t += 1
                    vm, Gm, Jm = phys(X, xi_m)
                    nv, ds = face_geometry(Jm, d, side)
                    w *= ds
                    dnm = Gm @ nv
                    if is_bdry:
                        A[sm, sm] += w * (sigma * np.outer(vm, vm) - np.outer(vm, dnm)
                                          - np.outer(dnm, vm))
                        continue
                    vp, Gp, _ = phys(Xp, xi_p)
                    dnp = Gp @ nv
                    jv = {0: vm, 1: -vp}
                    av = {0: 0.5 * dnm, 1: 0.5 * dnp}
                    sl = {0: sm, 1: sp}
                    for a in (0, 1):
                        for b in (0, 1):
                            A[sl[a], sl[b]] += w * (sigma * np.outer(jv[a], jv[b])
                                                    - np.outer(jv[a], av[b])
                                                    - np.outer(av[a], jv[b]))
    return A


def affine_vertices(dim, n, M, shift=None):
    """vertices of the uniform n^dim mesh of [0,1]^dim mapped by x -> M x + shift"""
    g = np.linspace(0.0, 1.0, n + 1)
    pts = np.array([[g[i] for i in reversed(t)] for t in itertools.product(range(n + 1), repeat=dim)])
    # itertools.product varies the last index fastest; reversed() makes x fastest
    out = pts @ np.asarray(M, dtype=float).T
    if shift is not None:
        out += np.asarray(shift, dtype=float)
    return out
