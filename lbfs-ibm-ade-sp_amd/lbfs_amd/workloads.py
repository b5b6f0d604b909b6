"""Seeded synthetic workloads of the benchmark configurations.

  channel_rbc     128x64x64 jittered hex channel (+-10 %, seed 2), x-periodic,
                  no-slip walls in y and z, D3Q19 FP32, body-force Poiseuille
                  flow, one biconcave RBC of 2562 nodes / 5120 triangles
  tgv_2d          64x64 periodic quads (+-5 %, seed 1), one z layer, D2Q9,
                  Taylor-Green vortex at Re = 100, U0 = 0.01
"""
import numpy as np

from . import abi

# duct flow: nu = 1/6 (relaxation factor 1 at dt_f = 0.5), dt = 0.25
CHANNEL_VISC = 1.0 / 6.0
CHANNEL_DT = 0.25


def duct_body_force(u_max, ny, nz, visc):
    """Body force giving centre-line velocity u_max in a rectangular duct of
    ny x nz cells (series solution of the Poiseuille problem, 50 odd terms)."""
    a, b = ny / 2.0, nz / 2.0
    s = 0.0
    for n in range(1, 100, 2):
        s += (-1) ** ((n - 1) // 2) / n ** 3 * (1.0 - 1.0 / np.cosh(n * np.pi * b / (2 * a)))
    # u_c = 16 a^2 G / (mu pi^3) * s  with G the pressure gradient per unit volume
    return u_max * visc * np.pi ** 3 / (16.0 * a * a * s)


def channel(nx=128, ny=64, nz=64, jitter=0.1, seed=2, lattice=abi.D3Q19, u_max=0.01, flags=0):
    mesh = abi.HexMesh(nx, ny, nz, periodic=(1, 0, 0), jitter=jitter, seed=seed)
    f = duct_body_force(u_max, ny, nz, CHANNEL_VISC)
    params = abi.make_params(lattice=lattice, visc=CHANNEL_VISC, dt=CHANNEL_DT, rho0=1.0, flags=flags,
                             body_force=(f, 0.0, 0.0))
    n = mesh.n_cells
    state = (np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n))
    return mesh, params, state


def single_rbc(mesh, frequency=16, d0=20.0, params=None):
    """One Evans-Fung cell (2562 nodes / 5120 triangles at frequency 16) at the
    channel centre, symmetry axis along z."""
    c = mesh.origin + 0.5 * mesh.lengths
    x, tri = abi.rbc(frequency, d0, center=c)
    return [abi.Membrane(x, tri, **(params or {}))]


# geodesic frequencies of the three resolutions: 642 (f=8), 1212 (f=11) and
# 2562 (f=16) nodes; 1280 is not a geodesic node count, 1212 is the nearest
SUSPENSION_FREQS = (8, 11, 16)


def random_rotation(rng):
    """Uniformly distributed rotation matrix (from a random unit quaternion)."""
    q = rng.standard_normal(4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _ef_half_thickness(rho, d0):
    """Half thickness of the Evans-Fung surface abi.evans_fung builds, at
    radius rho = 2 r / d0 (0 <= rho <= 1)."""
    r2 = rho * rho
    return d0 * np.sqrt(np.maximum(0.0, 1.0 - r2)) * (0.0518 + 2.0026 * r2 / 4.0 - 4.491 * r2 * r2 / 16.0)


def _near_cell(points, centre, rot, d0, gap):
    """True if any point lies inside the cell (centre, rot, diameter d0)
    inflated by `gap` (radially and in thickness)."""
    q = (points - centre) @ rot
    rho = 2.0 * np.hypot(q[:, 0], q[:, 1]) / d0
    inside = (rho <= 1.0 + 2.0 * gap / d0) & (np.abs(q[:, 2]) <= _ef_half_thickness(np.minimum(rho, 1.0), d0) + gap)
    return bool(np.any(inside))


def random_placement(rng, probe, n_cells, lengths, d0, periodic=(1, 0, 0), wall_gap=2.0, gap=1.0, attempts=20000):
    """Seeded random sequential placement of n_cells cells of diameter d0
    without overlap: each takes a uniformly random rotation and a uniformly
    random centre, kept `wall_gap` from the walls of the non-periodic axes; a
    candidate is accepted when no node of its probe shape (`probe`: nodes of
    the cell centred at the origin, symmetry axis z) lies inside an earlier
    cell inflated by `gap` and no probe node of an earlier cell lies inside
    it (periodic axes by the nearest image).  Returns (centres, rotations)."""
    L = np.asarray(lengths, dtype=np.float64)
    per = np.asarray(periodic, dtype=bool)
    centres, rots, placed = [], [], []
    reach = d0 + 2.0 * gap
    for _ in range(n_cells):
        for _ in range(attempts):
            rot = random_rotation(rng)
            x = probe @ rot.T
            lo, hi = x.min(0), x.max(0)
            c = np.where(per, rng.uniform(0.0, 1.0, 3) * L,
                         rng.uniform(wall_gap - lo, np.maximum(wall_gap - lo, L - wall_gap - hi)))
            if np.any(~per & ((c + lo < wall_gap) | (c + hi > L - wall_gap))):
                continue
            ok = True
            if centres:
                d = c - np.asarray(centres)
                d = np.where(per, d - L * np.round(d / L), d)
                for j in np.nonzero(np.einsum("ij,ij->i", d, d) < reach * reach)[0]:
                    # the candidate's image nearest cell j
                    cj = centres[j] + d[j]
                    if _near_cell(x + cj, centres[j], rots[j], d0, gap) or _near_cell(placed[j], cj, rot, d0, gap):
                        ok = False
                        break
            if ok:
                centres.append(c)
                rots.append(rot)
                placed.append(x + c)
                break
        else:
            raise RuntimeError("random placement: no room for cell %d after %d attempts" % (len(placed), attempts))
    return centres, rots


def suspension(nx=256, ny=128, nz=128, n_cells=256, hematocrit=0.2, seed=3, jitter=0.1, lattice=abi.D3Q19,
               u_max=0.01, params=None):
    """Batched suspension: n_cells Evans-Fung cells of seeded mixed resolution
    in a jittered channel at the given hematocrit, at seeded random positions
    and orientations without overlap (random_placement: at least one cell
    between membranes, two between a membrane and a wall)."""
    mesh = abi.HexMesh(nx, ny, nz, periodic=(1, 0, 0), jitter=jitter, seed=seed)
    f = duct_body_force(u_max, ny, nz, CHANNEL_VISC)
    pr = abi.make_params(lattice=lattice, visc=CHANNEL_VISC, dt=CHANNEL_DT, rho0=1.0, body_force=(f, 0.0, 0.0))
    # diameter giving the hematocrit: the Evans-Fung volume scales as d0^3
    u, t = abi.icosphere(16)
    x20 = abi.evans_fung(u, 20.0)
    tri = t
    N = np.cross(x20[tri[:, 1]] - x20[tri[:, 0]], x20[tri[:, 2]] - x20[tri[:, 0]])
    v20 = np.sum(np.einsum("ij,ij->i", N, x20[tri].mean(1))) / 6.0
    d0 = 20.0 * (hematocrit * nx * ny * nz / n_cells / v20) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    kinds = [SUSPENSION_FREQS[rng.integers(0, 3)] for _ in range(n_cells)]
    shapes, tris = {}, {}
    for f in SUSPENSION_FREQS:
        shapes[f], tris[f] = abi.rbc(f, d0)
    # placed with the coarsest resolution as the probe, in the box [0, L)
    # (the mesh origin is added after)
    centres, rots = random_placement(rng, shapes[min(SUSPENSION_FREQS)], n_cells, mesh.lengths, d0,
                                     periodic=(1, 0, 0))
    mems = [abi.Membrane(shapes[f] @ r.T + c + mesh.origin, tris[f], **(params or {}))
            for f, c, r in zip(kinds, centres, rots)]
    n = mesh.n_cells
    state = (np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n))
    return mesh, pr, state, mems, d0


def tgv_2d(n=64, jitter=0.05, seed=1, u0=0.01, re=100.0, lattice=abi.D2Q9, dt=0.25):
    mesh = abi.HexMesh(n, n, 1, periodic=(1, 1, 1), jitter=jitter, seed=seed)
    visc = u0 * n / re
    params = abi.make_params(lattice=lattice, visc=visc, dt=dt, rho0=1.0)
    g = mesh.geometry()
    c = g["centroid"][: 3 * mesh.n_cells].reshape(-1, 3)
    k = 2 * np.pi / n
    x, y = c[:, 0], c[:, 1]
    u = -u0 * np.cos(k * x) * np.sin(k * y)
    v = u0 * np.sin(k * x) * np.cos(k * y)
    rho = 1.0 - 0.75 * u0 * u0 * (np.cos(2 * k * x) + np.cos(2 * k * y))
    return mesh, params, (rho, u, v, np.zeros_like(u))


ENSEMBLE_EDGES = (32, 48, 64, 96)


def ensemble(n_instances=64, seed=5000, edges=ENSEMBLE_EDGES, max_rbcs=8, frequency=8, jitter=0.1,
             lattice=abi.D3Q19, u_max=0.01, params=None):
    """Ensemble of independent channel boxes stepped as one mesh: instance i
    has edges drawn from `edges` per axis and 1..max_rbcs cells of 10 f^2 + 2
    nodes (642 for f = 8), all seeded with seed + i.  Cells sit on slots along
    the box's x axis, diameter 0.45 of the smaller cross-section edge.  Each
    box has its own duct body force (centre-line velocity u_max)."""
    boxes, mems = [], []
    forces = []
    for i in range(n_instances):
        rng = np.random.default_rng(seed + i)
        nx, ny, nz = (int(e) for e in rng.choice(edges, 3))
        box = abi.HexMesh(nx, ny, nz, periodic=(1, 0, 0), jitter=jitter, seed=seed + i)
        boxes.append(box)
        forces.append(duct_body_force(u_max, ny, nz, CHANNEL_VISC))
        k = int(rng.integers(1, max_rbcs + 1))
        d0 = 0.45 * min(ny, nz)
        k = max(1, min(k, int(nx // (0.45 * d0 + 3))))
        for c in range(k):
            rot = abi.rotation((0, 0, 1), rng.uniform(0, 2 * np.pi)) @ abi.rotation((1, 0, 0), rng.uniform(-0.2, 0.2))
            x, tri = abi.rbc(frequency, d0, rot=rot)
            centre = np.array([(c + 0.5) * nx / k, 0.5 * ny, 0.5 * nz])
            mems.append(abi.Membrane(x + centre, tri, block=i, **(params or {})))
    mesh = abi.Ensemble(boxes, body_forces=[(f, 0.0, 0.0) for f in forces])
    pr = abi.make_params(lattice=lattice, visc=CHANNEL_VISC, dt=CHANNEL_DT, rho0=1.0, body_force=(0.0, 0.0, 0.0))
    n = mesh.n_cells
    state = (np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n))
    return mesh, pr, state, mems
