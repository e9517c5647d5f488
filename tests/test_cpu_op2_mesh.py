"""Meshes built from a caller's OP2 face sets (ins_mesh_create_op2): the caller's face and
bface order is kept, and inconsistent maps or negatively oriented cells are rejected
(src/ins2d.cpp:192 reads these sets from the mesh file; the bface order indexes bc_types and
bc_data, src/solvers/2d/ins_solver.cpp:205,324)."""
import numpy as np
import pytest


def _maps(m):
    coords, c2n = m.arrays()
    f2c, fn, b2c, bn = m.faces()
    return coords, c2n, f2c, fn, b2c, bn


@pytest.mark.parametrize("kind", ["square", "cube"])
def test_caller_face_order_is_kept(ip, kind):
    m = ip.Mesh.square(5, jitter=0.1, seed=2) if kind == "square" else ip.Mesh.cube(3, 0.1, 2)
    coords, c2n, f2c, fn, b2c, bn = _maps(m)
    rng = np.random.default_rng(1)
    fo = rng.permutation(len(f2c))
    f2c, fn = f2c[fo].copy(), fn[fo].copy()
    sw = rng.random(len(f2c)) < 0.5  # either side first
    f2c[sw] = f2c[sw][:, ::-1]
    fn[sw] = fn[sw][:, ::-1]
    b2c, bn = b2c[::-1].copy(), bn[::-1].copy()
    m2 = ip.Mesh.from_op2(coords, c2n, f2c, fn, b2c, bn)
    g = m2.faces()
    for a, b in zip(g, (f2c, fn, b2c, bn)):
        assert np.array_equal(a, b)


def test_inconsistent_maps_are_rejected(ip):
    m = ip.Mesh.square(4, jitter=0.1, seed=3)
    coords, c2n, f2c, fn, b2c, bn = _maps(m)
    bad = [
        (f2c, (fn + 1) % 3, b2c, bn),          # wrong local faces
        (f2c, fn, b2c[:-1], bn[:-1]),          # a bface missing
        (f2c, fn, np.r_[b2c[:-1], b2c[0]], np.r_[bn[:-1], bn[0]]),  # a bface twice
        (f2c, fn, b2c, (bn + 1) % 3),          # bfaces not on the boundary
        (f2c[:-1], fn[:-1], b2c, bn),          # an interior face missing
    ]
    for a in bad:
        with pytest.raises(ip.InsPoissonError) as e:
            ip.Mesh.from_op2(coords, c2n, *a)
        assert e.value.code == ip.ERR_INVALID


@pytest.mark.parametrize("kind", ["square", "cube"])
def test_negatively_oriented_cell_is_rejected(ip, kind):
    m = ip.Mesh.square(3, jitter=0.1, seed=4) if kind == "square" else ip.Mesh.cube(2, 0.1, 4)
    coords, c2n, f2c, fn, b2c, bn = _maps(m)
    c2n = c2n.copy()
    d = m.dim
    c2n[5, [d - 1, d]] = c2n[5, [d, d - 1]]
    with pytest.raises(ip.InsPoissonError) as e:
        ip.Mesh.from_op2(coords, c2n, f2c, fn, b2c, bn)
    assert e.value.code == ip.ERR_INVALID
    assert "cell 5" in str(e.value)
    # ins_mesh_create re-orients the cell instead, and reports the order it uses
    m3 = ip.Mesh.from_arrays(coords, c2n)
    assert np.array_equal(m3.arrays()[1][5], m.arrays()[1][5])
