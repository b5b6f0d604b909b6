"""Shared seeded inputs for the CPU and GPU tests."""
import numpy as np

from lbfs_amd import abi
from oracle import oracle as orc

LAT_PAIRS = [(abi.D3Q15, orc.OR_D3Q15), (abi.D3Q19, orc.OR_D3Q19), (abi.D2Q9, orc.OR_D2Q9)]


def centroids(mesh, closed=False):
    g = mesh.geometry(flags=abi.FLAG_CLOSED_FACES if closed else 0)
    return g["centroid"][: 3 * mesh.n_cells].reshape(-1, 3), g


def smooth_state(mesh, seed=0, amp=0.01, drho=1e-3, rho0=1.0):
    """A smooth, seeded, non-trivial state (sums of low Fourier modes over
    the box, so it is periodic-consistent on periodic axes)."""
    c, _ = centroids(mesh)
    L = mesh.lengths
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * (c - mesh.origin) / L
    out = []
    for var in range(4):
        f = np.zeros(mesh.n_cells)
        for _ in range(3):
            kx, ky, kz = rng.integers(0, 3, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            f += rng.uniform(-1, 1) * np.cos(kx * x[:, 0] + ph[0]) * np.cos(ky * x[:, 1] + ph[1]) * np.cos(
                kz * x[:, 2] + ph[2])
        out.append(f)
    rho = rho0 + drho * out[0]
    u, v, w = amp * out[1], amp * out[2], amp * out[3]
    if mesh.nz == 1:
        w = np.zeros_like(w)
    return rho, u, v, w


def tgv_state(mesh, U0=0.01, rho0=1.0):
    """Taylor-Green vortex initial condition (Solution.cpp:101-110, 158-171)
    with L the domain length (the reference reads unset structured sizes)."""
    c, _ = centroids(mesh)
    L = mesh.lengths[0]
    k = 2 * np.pi / L
    x, y = c[:, 0] - mesh.origin[0], c[:, 1] - mesh.origin[1]
    u = -U0 * np.cos(k * x) * np.sin(k * y)
    v = U0 * np.sin(k * x) * np.cos(k * y)
    rho = rho0 - rho0 * U0 ** 2 / 4.0 * 3.0 * (np.cos(2 * k * x) + np.cos(2 * k * y))
    return rho, u, v, np.zeros_like(u)


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def channel_mesh(nx, ny, nz, jitter=0.1, seed=2):
    """x-periodic box with no-slip walls (Dirichlet u = 0, Neumann rho) on y and z."""
    return abi.HexMesh(nx, ny, nz, periodic=(1, 0, 0), jitter=jitter, seed=seed)
