"""Mesh set-up on the GPU: the FP64 geometry (face areas, normals, centroids
and streaming steps; cell volumes, centroids and stencils; ghost centroids)
and the least-squares weights, computed by the face and cell kernels, equal
the host set-up bit for bit, so solvers built either way step alike bit for
bit (every quantity the kernels read derives from them)."""
import numpy as np
import pytest

import cases
from lbfs_amd import abi

pytestmark = pytest.mark.gpu


def run(mesh, params, membranes, steps, host):
    p = abi.make_params(**dict(params, flags=params.get("flags", 0) | (abi.FLAG_HOST_GEOMETRY if host else 0)))
    s = abi.Solver(mesh, p)
    if isinstance(mesh, abi.Ensemble):
        parts = [cases.smooth_state(m, seed=9 + k, amp=0.005) for k, m in enumerate(mesh.members)]
        s.set_state(*[np.concatenate([x[v] for x in parts]) for v in range(4)])
    else:
        s.set_state(*cases.smooth_state(mesh, seed=9, amp=0.005))
    if membranes:
        s.add_membranes(membranes, ibm_kernel=abi.IBM_COS4)
    s.step(steps)
    out = [np.concatenate(s.get_state()), s.cell_force()]
    if membranes:
        out += list(s.membrane_nodes())
    s.close()
    return out


def cell(center, d=6.0):
    x, t = abi.rbc(3, d, center=center)
    return abi.Membrane(x, t, **abi.RBC_PARAMS)


CASES = {
    "walled channel, hybrid gradients, one cell": (
        lambda: cases.channel_mesh(20, 14, 12, seed=31), dict(lattice=abi.D3Q19, visc=1 / 6, dt=0.25), True),
    "periodic box, least squares, determinant as written": (
        lambda: abi.HexMesh(12, 10, 8, periodic=(1, 1, 1), jitter=0.12, seed=32),
        dict(lattice=abi.D3Q15, gradient_type=abi.GRAD_LS, flags=abi.FLAG_LS_DET_COMPAT), False),
    "closed faces, Green-Gauss, capped face steps": (
        lambda: cases.channel_mesh(16, 12, 10, seed=33),
        dict(lattice=abi.D3Q19, gradient_type=abi.GRAD_GG, flags=abi.FLAG_CLOSED_FACES, fneq_min=0.05), False),
    "one-layer periodic mesh (each cell its own neighbour)": (
        lambda: abi.HexMesh(24, 20, 1, periodic=(1, 1, 1), jitter=0.1, seed=34), dict(lattice=abi.D2Q9), False),
    "ensemble of three boxes": (
        lambda: abi.Ensemble([cases.channel_mesh(12, 10, 10, seed=35), cases.channel_mesh(16, 8, 8, seed=36),
                              abi.HexMesh(8, 8, 8, periodic=(1, 1, 1), jitter=0.1, seed=37)]),
        dict(lattice=abi.D3Q19), False),
}


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_setup_equals_host_setup(name):
    make, params, with_cell = CASES[name]
    mesh = make()
    mems = [cell((10.0, 7.0, 6.0))] if with_cell else []
    a = run(mesh, params, mems, 6, host=False)
    b = run(mesh, params, mems, 6, host=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y), name


def test_gpu_setup_reports_bad_meshes():
    """An owner index out of range is refused with the host path's message."""
    mesh = cases.channel_mesh(6, 5, 4, seed=38)
    mesh.owner[7] = mesh.n_cells + 3
    with pytest.raises(abi.LbfsError, match="face 7 has an invalid owner/neighbour"):
        abi.Solver(mesh, abi.make_params())


def test_host_layout_independent_of_threads(monkeypatch):
    """The tile layout is built on several host threads (tiles are
    independent; per-cell sums keep the face order): one thread and seven
    give solvers that step alike bit for bit."""
    mesh = cases.channel_mesh(37, 21, 18, seed=39)
    params = dict(lattice=abi.D3Q19, visc=1 / 6, dt=0.25, tile=(8, 8, 4))
    outs = []
    for n in ("1", "7"):
        monkeypatch.setenv("LBFS_HOST_THREADS", n)
        outs.append(run(mesh, params, [cell((18.0, 10.0, 9.0))], 5, host=False))
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
