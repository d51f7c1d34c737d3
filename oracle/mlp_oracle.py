"""FP64 numpy oracle of the learned-damping pieces (features, target, network).

TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels in tpss_amd/csrc/mlp.hip.
The reference has no learned damping (it applies one scalar relaxation,
schwarz_preconditioner.templates.h:631-633); the definitions restated here
are the build's own and are pinned by finite differences, closed forms
(Cartesian meshes give exact local solvers, hence damping 1) and brute force.
"""
import numpy as np


def param_layout(n_in, h):
    sizes = [("W1", (h, n_in)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)), ("W3", (h,)),
             ("b3", (1,))]
    out, off = {}, 0
    for name, shp in sizes:
        n = int(np.prod(shp))
        out[name] = (off, shp)
        off += n
    return out, off


def unpack(params, n_in, h):
    lay, _ = param_layout(n_in, h)
    return {k: np.asarray(params[o:o + int(np.prod(s))], dtype=np.float64).reshape(s)
            for k, (o, s) in lay.items()}


def bf16_round(x):
    """Round-to-nearest-even to bfloat16 (as float64)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params, X, n_in, h, lo=0.2, hi=1.2):
    p = unpack(params, n_in, h)
    a1 = np.tanh(X @ p["W1"].T + p["b1"])
    a2 = np.tanh(a1 @ p["W2"].T + p["b2"])
    z = a2 @ p["W3"] + p["b3"][0]
    return lo + (hi - lo) / (1.0 + np.exp(-z)), (a1, a2, z)


def loss(params, X, t, n_in, h, lo=0.2, hi=1.2):
    w, _ = forward(params, X, n_in, h, lo, hi)
    return 0.5 * np.mean((w - t) ** 2)


def grad(params, X, t, n_in, h, lo=0.2, hi=1.2):
    """Analytic gradient of loss() (back-propagation)."""
    p = unpack(params, n_in, h)
    B = X.shape[0]
    w, (a1, a2, z) = forward(params, X, n_in, h, lo, hi)
    sg = 1.0 / (1.0 + np.exp(-z))
    dz = (w - t) / B * (hi - lo) * sg * (1 - sg)
    d2 = dz[:, None] * p["W3"][None, :] * (1 - a2 ** 2)
    d1 = (d2 @ p["W2"]) * (1 - a1 ** 2)
    g = {"W1": d1.T @ X, "b1": d1.sum(0), "W2": d2.T @ a1, "b2": d2.sum(0),
         "W3": a2.T @ dz, "b3": np.array([dz.sum()])}
    lay, n = param_layout(n_in, h)
    out = np.zeros(n)
    for k, (o, s) in lay.items():
        out[o:o + int(np.prod(s))] = g[k].ravel()
    return out


def patch_features(dim, n, vertices, patch_vertices):
    """Singular values (descending, times n) of the Jacobian at the centre of
    every cell of each patch, cells in lexicographic patch order."""
    V = None if vertices is None else np.asarray(vertices).reshape(-1, dim)
    feats = []
    for v in patch_vertices:
        row = []
        for cb in range(2 ** dim):
            cell = [v[e] - 1 + ((cb >> e) & 1) for e in range(dim)]
            if V is None:
                J = np.eye(dim) / n
            else:
                J = np.zeros((dim, dim))
                for c in range(2 ** dim):
                    idx = sum((cell[e] + ((c >> e) & 1)) * (n + 1) ** e for e in range(dim))
                    for b in range(dim):
                        wgt = np.prod([((1.0 if (c >> e) & 1 else -1.0) if e == b else 0.5)
                                       for e in range(dim)])
                        J[:, b] += wgt * V[idx]
            row.extend(np.sort(np.linalg.svd(J, compute_uv=False))[::-1] * n)
        feats.append(row)
    return np.array(feats)


def optimal_damping(smoother, n_samples, seed, lo=0.2, hi=1.2):
    """omega_j = sum_e d_j^T r / sum_e d_j^T A d_j with r = A e, e = U[-1,1] keyed
    by (seed + s, index), d_j = R_j^T A_j^{-1} R_j r (brute force per patch)."""
    import oracle as o
    idx = smoother.patch_dof_indices()
    num = np.zeros(smoother.P)
    den = np.zeros(smoother.P)
    for s in range(n_samples):
        e = o.splitmix_uniform(seed + s, smoother.N)
        r = smoother.operator(e)
        for j in range(smoother.P):
            u = smoother.local_apply(j, r[idx[j]])
            d = np.zeros(smoother.N)
            d[idx[j]] = u
            num[j] += u @ r[idx[j]]
            den[j] += d @ smoother.operator(d)
    w = np.where(den > 0, num / np.where(den > 0, den, 1.0), 1.0)
    return np.clip(w, lo, hi)
