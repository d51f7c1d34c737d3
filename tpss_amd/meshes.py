"""Vertex arrays of distorted unit cubes ((n+1)^dim points, x fastest), the two
constructions the distorted-mesh experiments use.

* distorted_vertices: every interior vertex of the n^dim mesh moved by
  amplitude * h * U[-1, 1]^dim (U from SplitMix64 keyed by (seed, vertex*dim + d)).
  The coarse levels of a GMG hierarchy built from it (every second vertex) are
  not nested in it.
* nested_distorted_vertices: the reference's construction
  (create_distorted_cube, apps/mesh.h:153-204): the interior vertices of a root
  mesh of n_root^dim cells are moved (here by amplitude * h_root * U[-1, 1]^dim,
  the same noise as above on the root mesh), then the root cells are refined
  uniformly; refinement of multilinear cells places every new vertex on the
  parent's multilinear map, so every GMG level down to the root mesh is nested.
"""
import numpy as np


def distorted_vertices(dim, n, amplitude, seed, graded=False, pattern="uniform"):
    """Interior vertices moved by (amplitude/n) * U[-1,1]^dim; boundary vertices
    fixed.  pattern "graded" (or graded=True) scales the amplitude by the vertex's
    x coordinate (no distortion at x = 0, the full amplitude at x = 1); "half"
    moves only the vertices with x > 1/2 (the other half of the mesh stays
    Cartesian)."""
    if pattern not in ("uniform", "graded", "half"):
        raise ValueError("pattern must be uniform, graded or half")
    graded = graded or pattern == "graded"
    g = np.linspace(0.0, 1.0, n + 1)
    mesh = np.meshgrid(*([g] * dim), indexing="ij")
    v = np.stack([m.transpose(tuple(range(dim))[::-1]).ravel() for m in mesh], axis=-1)
    nv = v.shape[0]
    i = np.arange(nv * dim, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * np.uint64(0x100000000) + i + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    noise = (2.0 * ((z >> np.uint64(11)).astype(np.float64) / 9007199254740992.0) - 1.0).reshape(nv, dim)
    idx = np.arange(nv)
    interior = np.ones(nv, dtype=bool)
    t = idx.copy()
    for _ in range(dim):
        c = t % (n + 1)
        interior &= (c > 0) & (c < n)
        t //= n + 1
    amp = amplitude * (v[:, :1] if graded else 1.0)
    if pattern == "half":
        amp = amplitude * (v[:, :1] > 0.5 + 1e-12)
    v[interior] += ((amp / n) * noise)[interior]
    return v


def nested_distorted_vertices(dim, n_root, n, amplitude, seed):
    """Root mesh of n_root cells per direction with interior vertices moved by
    amplitude * h_root * U[-1,1]^dim, refined uniformly to n cells per direction
    (n = n_root * 2^k): every fine vertex is the multilinear interpolant of the
    corners of its root cell."""
    if n % n_root != 0 or (n // n_root) & (n // n_root - 1):
        raise ValueError("n must be n_root times a power of two")
    vr = distorted_vertices(dim, n_root, amplitude, seed).reshape((n_root + 1,) * dim + (dim,))
    g = np.arange(n + 1) * (n_root / n)
    c = np.minimum(np.floor(g).astype(np.int64), n_root - 1)
    f = g - c
    out = np.zeros((n + 1,) * dim + (dim,))
    # array axes run (z, y, x) / (y, x): the last axis before the coordinates is x
    for corner in range(1 << dim):
        w = np.ones((n + 1,) * dim)
        ix = []
        for axis in range(dim):
            bit = (corner >> (dim - 1 - axis)) & 1
            shape = [1] * dim
            shape[axis] = n + 1
            w = w * (f if bit else 1.0 - f).reshape(shape)
            ix.append((c + bit).reshape(shape))
        out += w[..., None] * vr[tuple(ix)]
    return out.reshape(-1, dim)
