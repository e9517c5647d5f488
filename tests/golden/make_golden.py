"""Writes tests/golden/oracle_golden.npz from the FP64 oracle.

The reference ships no fixtures for the Poisson path (its solver lives in the
un-vendored OP2-DG-Toolkit), so these fixtures pin the oracle's own results on
the config-0 workload (2D, 32x32x2 jittered triangles, N=3, Jacobi-PCG to
1e-10) and one small dense operator, for regression.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "ins-solver_amd"), os.path.join(ROOT, "tests")]

import ins_poisson as ip  # noqa: E402  (mesh generator only: host code, no device)
import oracle as O  # noqa: E402
from helpers import manufactured_sin  # noqa: E402

SEED = 20261202
mesh = ip.Mesh.square(32, jitter=0.1, seed=SEED)
coords, c2n = mesh.arrays()
o = O.Oracle(coords, c2n, 3)
xyz = o.node_coords()
u, f = manufactured_sin(xyz)
x, rc, it, rr = o.solve_jacobi(o.mass(f), rtol=1e-10)
assert rc == 0
err = o.mass_norm(x - u)
err_ref = np.sqrt(o.l2sq(x - u))
m2 = ip.Mesh.square(2, jitter=0.1, seed=1)
o2 = O.Oracle(*m2.arrays(), 2)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_golden.npz")
np.savez_compressed(out, cfg0_seed=SEED, cfg0_coords=coords, cfg0_solution=x, cfg0_l2err=err, cfg0_l2err_ref_pattern=err_ref,
                    cfg0_iters=it, dense_2d_n2=o2.dense())
print(out, "l2err", err, "iters", it)
