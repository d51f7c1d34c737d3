"""Independent dense assembly of the SIPG linear elasticity matrix (numpy).

Written from the textbook vector-valued SIPG form with explicit vector basis
functions phi_{c,A} = phi_A e_c, full symmetric-gradient tensors and explicit
jumps/averages on every face -- unlike the oracle's per-cell, mirrored-ghost
formulation -- so that agreement pins both.  Form of deal.II's
LocalIntegrators::Elasticity / GradDiv as called by the reference
(apps/linelasticity_integrator.h:31-124):
  cell      2 mu eps(u):eps(v) + lambda div u div v
  interior  2 mu (sigma [u].[v] - {eps(u) n}.[v] - {eps(v) n}.[u])
            + lambda (sigma [u.n][v.n] - {div u}[v.n] - {div v}[u.n])
  boundary  2 mu (2 sigma u.v - eps(u) n.v - eps(v) n.u)
            + lambda (2 sigma (u.n)(v.n) - div u v.n - div v u.n)
with sigma = ip * gamma / h, gamma = p (p + 1), on the Cartesian unit cube.
DoFs component-blocked: index c * n_cells * ndc + cell * ndc + local.
"""
import itertools

import numpy as np
from numpy.polynomial import legendre as L

from dense_sipg import eval_basis, gll_nodes, lagrange_basis


def assemble(dim, p, n, ip, mu, lam, with_mass=False):
    h = 1.0 / n
    C = lagrange_basis(gll_nodes(p))
    gx, gw = L.leggauss(p + 1)
    qx, qw = 0.5 * (gx + 1.0), 0.5 * gw
    nb = p + 1
    ndc = nb ** dim
    ncell = n ** dim
    Ns = ncell * ndc
    N = dim * Ns
    A = np.zeros((N, N))
    Mm = np.zeros((N, N))
    gamma = p * (p + 1) if p > 0 else 1
    sigma = ip * gamma / h
    loc = [[(i // nb ** e) % nb for e in range(dim)] for i in range(ndc)]

    def shape(xi):
        v1, d1 = zip(*[eval_basis(C, xi[e]) for e in range(dim)])
        val = np.array([np.prod([v1[e][0, li[e]] for e in range(dim)]) for li in loc])
        grad = np.zeros((ndc, dim))
        for a in range(dim):
            grad[:, a] = [np.prod([(d1[e][0, li[e]] / h) if e == a else v1[e][0, li[e]]
                                   for e in range(dim)]) for li in loc]
        return val, grad

    def vec_shape(xi):
        """values (dim*ndc, dim), gradients (dim*ndc, dim, dim) [i, comp, deriv]."""
        val, grad = shape(xi)
        V = np.zeros((dim * ndc, dim))
        G = np.zeros((dim * ndc, dim, dim))
        for c in range(dim):
            V[c * ndc:(c + 1) * ndc, c] = val
            G[c * ndc:(c + 1) * ndc, c, :] = grad
        return V, G

    def eps(G):
        return 0.5 * (G + G.transpose(0, 2, 1))

    def div(G):
        return np.trace(G, axis1=1, axis2=2)

    def dofs(cell):
        return np.concatenate([c * Ns + cell * ndc + np.arange(ndc) for c in range(dim)])

    def cell_index(cc):
        return sum(cc[e] * n ** e for e in range(dim))

    Ke = np.zeros((dim * ndc, dim * ndc))
    Me = np.zeros((dim * ndc, dim * ndc))
    for q in itertools.product(range(p + 1), repeat=dim):
        xi = [qx[q[e]] for e in range(dim)]
        w = np.prod([qw[q[e]] for e in range(dim)]) * h ** dim
        V, G = vec_shape(xi)
        E, D = eps(G), div(G)
        Ke += w * (2 * mu * np.einsum("icd,jcd->ij", E, E) + lam * np.outer(D, D))
        Me += w * V @ V.T
    for c in range(ncell):
        I = dofs(c)
        A[np.ix_(I, I)] += Ke
        Mm[np.ix_(I, I)] += Me

    for cc in itertools.product(range(n), repeat=dim):
        cc = list(cc)
        c = cell_index(cc)
        for d in range(dim):
            for side in (0, 1):
                nrm = np.zeros(dim)
                nrm[d] = 1.0 if side == 1 else -1.0
                nbr = cc[d] + (1 if side == 1 else -1)
                is_bdry = nbr < 0 or nbr >= n
                if not is_bdry and side == 0:
                    continue
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
                    Vm, Gm = vec_shape(xi_m)
                    En = np.einsum("icd,d->ic", eps(Gm), nrm)  # eps(phi) n
                    Dm = div(Gm)
                    Vn = Vm @ nrm
                    Im = dofs(c)
                    if is_bdry:
                        blk = (2 * mu * (2 * sigma * Vm @ Vm.T - Vm @ En.T - En @ Vm.T)
                               + lam * (2 * sigma * np.outer(Vn, Vn) - np.outer(Vn, Dm)
                                        - np.outer(Dm, Vn)))
                        A[np.ix_(Im, Im)] += w * blk
                        continue
                    cn = list(cc)
                    cn[d] = nbr
                    Ip = dofs(cell_index(cn))
                    Vp, Gp = vec_shape(xi_p)
                    Ep = np.einsum("icd,d->ic", eps(Gp), nrm)
                    Dp = div(Gp)
                    jump = {0: Vm, 1: -Vp}               # [phi]
                    avg_e = {0: 0.5 * En, 1: 0.5 * Ep}   # {eps(phi) n}
                    avg_d = {0: 0.5 * Dm, 1: 0.5 * Dp}   # {div phi}
                    idx = {0: Im, 1: Ip}
                    for a in (0, 1):       # test side
                        for b in (0, 1):   # trial side
                            ja, jb = jump[a], jump[b]
                            jna, jnb = ja @ nrm, jb @ nrm
                            blk = (2 * mu * (sigma * ja @ jb.T - ja @ avg_e[b].T - avg_e[a] @ jb.T)
                                   + lam * (sigma * np.outer(jna, jnb) - np.outer(jna, avg_d[b])
                                            - np.outer(avg_d[a], jnb)))
                            A[np.ix_(idx[a], idx[b])] += w * blk
    if with_mass:
        return A, Mm
    return A
