"""Shared problem set-ups for the CPU and GPU tests (manufactured solutions)."""
import numpy as np


def manufactured_sin(xyz, k=(1, 1, 1)):
    """u = prod sin(k_i pi x_i), f = -lap u."""
    dim = xyz.shape[0]
    u = np.ones_like(xyz[0])
    for d in range(dim):
        u = u * np.sin(k[d] * np.pi * xyz[d])
    f = sum((k[d] * np.pi) ** 2 for d in range(dim)) * u
    return u, f


def poly_solution(xyz, N):
    """A global polynomial of degree <= N and -lap of it."""
    dim = xyz.shape[0]
    x, y = xyz[0], xyz[1]
    if dim == 2:
        if N == 1:
            return 1 + 2 * x - 3 * y, 0 * x
        u = 1 + x * x - 0.5 * x * y + 2 * y * y + x
        f = -(2 + 4) * np.ones_like(x)
        if N >= 3:
            u = u + x ** 3 - x * y * y
            f = f - (6 * x - 2 * x)
        return u, f
    z = xyz[2]
    if N == 1:
        return 1 + 2 * x - 3 * y + z, 0 * x
    u = 1 + x * x - 0.5 * x * y + 2 * y * y + z * z - y * z + x
    f = -(2 + 4 + 2) * np.ones_like(x)
    if N >= 3:
        u = u + x ** 3 - x * z * z
        f = f - (6 * x - 2 * x)
    return u, f


def bc_values(xyz, bface2cells, bfaceNum, fmask, values):
    """Gather nodal values of the owner cell at each bface's face nodes (FMASK order)."""
    return np.ascontiguousarray(values[bface2cells[:, None], fmask[bfaceNum]])


def density_field(xyz, h, rho_in=1000.0, rho_out=1.0, radius=0.25, alpha_h=2.0):
    """rho from a sphere level set, smoothed as in src/kernels/ls_step.h:5-6."""
    dim = xyz.shape[0]
    c = 0.5
    s = np.sqrt(sum((xyz[d] - c) ** 2 for d in range(dim))) - radius
    alpha = alpha_h * h
    step = np.tanh(np.pi * s / alpha)
    return 0.5 * rho_out * (1 + step) + 0.5 * rho_in * (1 - step)
