"""Cell renumbering of a mesh (ins_mesh_renumber): a permutation of the same cells whose
operator is the permuted operator; cells of one shape class are grouped in their old order."""
import numpy as np
import pytest


@pytest.mark.parametrize("dim", [2, 3])
def test_shape_class_renumbering(ip, oracle, dim):
    mesh = ip.Mesh.square(5, jitter=0.1, seed=4) if dim == 2 else ip.Mesh.cube(3, 0.1, 4)
    coords, c0 = mesh.arrays()
    nf0, nb0 = mesh.n_faces, mesh.n_bfaces
    perm = mesh.renumber()
    _, c1 = mesh.arrays()
    assert np.array_equal(np.sort(perm), np.arange(len(perm)))
    assert np.array_equal(c1, c0[perm])            # same cells, same vertex order
    assert (mesh.n_faces, mesh.n_bfaces) == (nf0, nb0)
    # the operator of the renumbered mesh is the permuted operator
    N = 2
    o0 = oracle.Oracle(coords, c0, N)
    o1 = oracle.Oracle(coords, c1, N)
    x = np.random.default_rng(dim).standard_normal((len(perm), o0.Np))
    assert np.abs(o1.mult(x[perm]) - o0.mult(x)[perm]).max() < 1e-12 * np.abs(o0.mult(x)).max()
    # a structured subdivision has 2 (2D) or 6 (3D) classes, each kept in its old order
    assert np.sum(np.diff(perm) < 0) == (1 if dim == 2 else 5)
