"""The distorted-mesh vertex generators (tpss_amd/meshes.py).  The nested
construction follows the reference's create_distorted_cube
(apps/mesh.h:153-204): distort the root mesh, then refine uniformly."""
import numpy as np
import pytest

from tpss_amd import meshes


def coarsen(v, dim, n):
    """Every second vertex: the coarse level of a GMG hierarchy."""
    g = v.reshape((n + 1,) * dim + (dim,))
    return g[(slice(None, None, 2),) * dim].reshape(-1, dim)


@pytest.mark.parametrize("dim,n_root,n", [(2, 3, 12), (3, 2, 8), (3, 3, 12)])
def test_nested_hierarchy(dim, n_root, n):
    v = meshes.nested_distorted_vertices(dim, n_root, n, 0.15, seed=3)
    root = meshes.distorted_vertices(dim, n_root, 0.15, seed=3)
    # every level down to the root mesh is the same construction at its size
    m = n
    while m > n_root:
        v = coarsen(v, dim, m)
        m //= 2
        np.testing.assert_allclose(v, meshes.nested_distorted_vertices(dim, n_root, m, 0.15, 3),
                                   rtol=0, atol=1e-15)
    np.testing.assert_array_equal(v, root)


def test_nested_refinement_is_multilinear_and_keeps_the_boundary():
    dim, n_root, n = 3, 2, 4
    v = meshes.nested_distorted_vertices(dim, n_root, n, 0.2, seed=5).reshape(5, 5, 5, 3)
    r = meshes.distorted_vertices(dim, n_root, 0.2, seed=5).reshape(3, 3, 3, 3)
    # a fine vertex at the centre of root cell (0,0,0) is the mean of its 8 corners
    np.testing.assert_allclose(v[1, 1, 1], r[:2, :2, :2].reshape(-1, 3).mean(0), atol=1e-15)
    # edge midpoint = mean of the two end points
    np.testing.assert_allclose(v[0, 0, 1], 0.5 * (r[0, 0, 0] + r[0, 0, 1]), atol=1e-15)
    # boundary faces stay on the unit cube
    g = np.linspace(0, 1, n + 1)
    np.testing.assert_allclose(v[:, :, 0, 0], 0.0, atol=1e-15)
    np.testing.assert_allclose(v[:, :, -1, 0], 1.0, atol=1e-15)
    np.testing.assert_allclose(v[0, 0, :, 0], g, atol=1e-15)


def test_nested_rejects_non_dyadic_sizes():
    with pytest.raises(ValueError):
        meshes.nested_distorted_vertices(3, 3, 10, 0.1, seed=1)


def test_fine_level_distortion_matches_training_meshes():
    from tpss_amd import learned_damping as ld
    np.testing.assert_array_equal(ld.distorted_vertices(3, 6, 0.2, 101),
                                  meshes.distorted_vertices(3, 6, 0.2, 101))
