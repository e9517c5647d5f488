"""GPU parity of the matrix-core element pass (3D, orders 2-4, FP32) against the FP64 oracle.

With ELEMENT_PASS_MATRIX the FP32 operator apply of a 3D order-2..4 level runs on
v_mfma_f32_16x16x4f32 / v_mfma_f32_4x4x1f32 (sipg_mfma.hip); by default only order 4 does.
Its result is compared with the oracle's FP64 operator on the same seeded meshes:
within 2e-5 relative (max norm), the bound of the FP32 vector kernels. Cases cover every
boundary mode (Dirichlet, Neumann, mixed), the variable coefficient, cell counts that
leave partial 16-cell tiles and partial 256-cell blocks, and meshes of several blocks.
"""
import numpy as np
import pytest

from helpers import density_field

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def pair(ip, oracle, N, n, bc, factor, seed=11):
    mesh = ip.Mesh.cube(n, 0.1, seed)
    coords, c2n = mesh.arrays()
    s = ip.PoissonSolver(mesh, N)
    s.set_element_pass(ip.ELEMENT_PASS_MATRIX)
    o = oracle.Oracle(coords, c2n, N)
    nb = s.n_bfaces
    bt = {"dirichlet": np.zeros(nb, np.int32), "neumann": np.ones(nb, np.int32)}.get(bc)
    if bt is None:
        bt = np.zeros(nb, np.int32)
        bt[1::3] = 1
    s.set_bc_types(bt)
    o.set_bc_types(bt)
    if factor:
        k = 1.0 / density_field(o.node_coords(), 1.0 / n, rho_in=100.0)
        s.set_factor(k)
        o.set_factor(k)
    return s, o


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "mixed"])
@pytest.mark.parametrize("factor", [False, True])
def test_mfma_apply_matches_oracle(ip, oracle, N, bc, factor):
    # 5^3 x 6 = 750 cells: 2 full blocks of 256 and a tail with a partial 16-cell tile
    print(f"mfma apply N={N} bc={bc} factor={factor} cube 5", flush=True)
    s, o = pair(ip, oracle, N, 5, bc, factor)
    x = np.random.default_rng(100 + N).standard_normal((s.n_cells, s.Np))
    assert rel(s.mult_fp32(x), o.mult(x)) < 2e-5


@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_mfma_apply_sizes(ip, oracle, n):
    # 6, 48, 162 and 2058 cells: a single partial tile up to several blocks
    print(f"mfma apply N=3 cube {n}", flush=True)
    s, o = pair(ip, oracle, 3, n, "mixed", True, seed=n)
    x = np.random.default_rng(n).standard_normal((s.n_cells, s.Np))
    assert rel(s.mult_fp32(x), o.mult(x)) < 2e-5


def test_mfma_apply_is_linear_and_deterministic(ip, oracle):
    print("mfma linearity N=4 cube 4", flush=True)
    s, _ = pair(ip, oracle, 4, 4, "dirichlet", True)
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal((2, s.n_cells, s.Np))
    ax, ay, axy = s.mult_fp32(x), s.mult_fp32(y), s.mult_fp32(2.0 * x - y)
    assert rel(axy, 2.0 * ax - ay) < 1e-5
    assert np.array_equal(s.mult_fp32(x), ax)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_element_pass_choice(ip, oracle, N):
    """AUTO takes the matrix cores at order 4 and the vector kernel below; both kernels agree."""
    print(f"element pass choice N={N} cube 4", flush=True)
    s, o = pair(ip, oracle, N, 4, "mixed", True)
    x = np.random.default_rng(N).standard_normal((s.n_cells, s.Np))
    ym = s.mult_fp32(x)
    s.set_element_pass(ip.ELEMENT_PASS_VECTOR)
    yv = s.mult_fp32(x)
    s.set_element_pass(ip.ELEMENT_PASS_AUTO)
    ya = s.mult_fp32(x)
    assert rel(ym, yv) < 4e-6
    assert not np.array_equal(ym, yv)  # two different kernels ran
    assert np.array_equal(ya, ym if N == 4 else yv)
    assert rel(ya, o.mult(x)) < 2e-5
