"""Pins of the oracle's wall (ghost-side) face flux and of its viscous and
preconditioned (gamma) terms by closed forms (CPU only).

Plane Poiseuille flow between no-slip walls y = 0 and y = H = N h on an
orthogonal mesh (periodic in x and z, body force f along x).  With u = u(y)
only, the lattice sums of the reference's face flux reduce by hand to a 1-D
scheme (derivation below); its steady state is the parabola of the
continuous problem shifted by a constant,

    u_j = f / (2 nu) y_j (H - y_j) + delta,     y_j = (j + 1/2) h,

    delta / U = -3 / (2 N) + 10 / (9 N^2)   owner gradient on the owner side of
                                            a wall face (Solver.cpp:918-1000)
    delta / U = -3 / (2 N) + 1 / N^2        ghost gradient at every point of a
                                            wall face (LBFS.cu:391-394)

with U = f H^2 / (8 nu), when the face streaming step is dt_f = h / 2
(unstructured_mesh.cpp:1435-1443 with the fneq cap off).

Derivation (linear in u; the neglected terms are O(u^2) relative):
  * least squares (gradients.cpp:149-313, weights (2 l_n / L_n)^2 S / L):
    interior cells g_j = (u_{j+1} - u_{j-1}) / (2 h); the wall cell's ghost
    sits halfway between its centroid and the face (unstructured_mesh.cpp:
    1982-1992), at h/4 with weight 64 S / h and the mirror value -u_0
    (Solution.cpp:759-833), so g_0 = (u_1 + 31 u_0) / (5 h);
  * Dirichlet ghost gradient 2 (0 - u_0) / (h/4) along the outward normal
    (gradients.cpp:592-648): 8 u_0 / h in y at the bottom wall;
  * a face with normal +y: the points x_f - e dt_f with e_y = +1 take the
    owner extrapolation p_o - g_o dt_f, those with e_y = -1 the neighbour's
    p_n + g_n dt_f (Solver.cpp:918-1000); the x-momentum flux through it is
    3 tau rho sum_k w_k e_x^2 e_y u_k with tau = 3 nu / dt_f (LBFS.cu:1488,
    1573-1647), sum over e_x^2 e_y = +-1 of w = 2/36 on D2Q9, D3Q15 and D3Q19:
        Pi = nu / (2 dt_f) (p_o - p_n) - nu / 2 (g_o + g_n),
    p_o = u_j + h/2 g_j, p_n = u_{j+1} - h/2 g_{j+1};
  * the bottom wall (face value 0): Pi_w = -nu / 2 (8 u_0 / h + g_0), or
    -nu / 2 (16 u_0 / h) when every point takes the ghost gradient;
  * steady cell balance: Pi_{j+1/2} - Pi_{j-1/2} = f h.
Substituting the shifted parabola satisfies every interior cell identically
and the wall cells fix delta.  The preconditioned flux (LBFS.cu:1041-1354,
1488) divides tau by gamma and leaves the linear part of the equilibrium
alone, so nu becomes nu / gamma.

A dropped factor in the ghost placement, its weight, the mirror value, the
ghost gradient or the wall upwinding moves delta and fails these tests.
"""
import numpy as np
import pytest

import cases
from lbfs_amd import abi
from oracle import oracle as orc

NU = 1.0 / 6.0


def poiseuille_1d(N, h, nu, dt_f, f, ghost_grad_all):
    """Steady state of the hand-reduced 1-D scheme (module docstring),
    assembled as a linear system and solved."""
    def residual(u):
        g = np.zeros(N)
        g[1:-1] = (u[2:] - u[:-2]) / (2 * h)
        g[0] = (u[1] + 31 * u[0]) / (5 * h)
        g[-1] = -(u[-2] + 31 * u[-1]) / (5 * h)
        Pi = np.zeros(N + 1)
        for j in range(N - 1):
            po = u[j] + 0.5 * h * g[j]
            pn = u[j + 1] - 0.5 * h * g[j + 1]
            Pi[j + 1] = nu / (2 * dt_f) * (po - pn) - 0.5 * nu * (g[j] + g[j + 1])
        gw0, gwN = 8 * u[0] / h, -8 * u[-1] / h
        if ghost_grad_all:
            Pi[0], Pi[N] = -nu * gw0, -nu * gwN
        else:
            Pi[0], Pi[N] = -0.5 * nu * (gw0 + g[0]), -0.5 * nu * (gwN + g[-1])
        return -(Pi[1:] - Pi[:-1]) / h + f
    r0 = residual(np.zeros(N))
    A = np.stack([residual(np.eye(N)[k]) - r0 for k in range(N)], axis=1)
    return np.linalg.solve(A, -r0)


def poiseuille_closed_form(N, h, nu, f, ghost_grad_all):
    H = N * h
    y = (np.arange(N) + 0.5) * h
    U = f * H * H / (8 * nu)
    delta = U * (-1.5 / N + (1.0 if ghost_grad_all else 10.0 / 9.0) / (N * N))
    return f / (2 * nu) * y * (H - y) + delta


def _channel(N, lattice):
    nz = 1 if lattice == orc.OR_D2Q9 else 2
    return abi.HexMesh(2, N, nz, periodic=(1, 0, 1), jitter=0.0)


def _rows(mesh, N):
    c = mesh.geometry()["centroid"][: 3 * mesh.n_cells].reshape(-1, 3)
    return np.round(c[:, 1] - 0.5).astype(int)


def _flags(form):
    return {"cpu": 0, "ghost_grad": orc.FLAG_GHOST_FACE_GRAD,
            "gpu_reference": orc.FLAG_UPWIND_BAND | orc.FLAG_GHOST_DIR_GPU | orc.FLAG_GHOST_FACE_GRAD}[form]


@pytest.mark.parametrize("N", [8, 12, 32])
@pytest.mark.parametrize("ghost_grad_all", [False, True])
def test_reduced_scheme_is_the_shifted_parabola(N, ghost_grad_all):
    """The hand-reduced 1-D scheme's steady state is exactly the shifted
    parabola (so the closed form is the scheme's, at any resolution)."""
    f = 1e-6
    u = poiseuille_1d(N, 1.0, NU, 0.5, f, ghost_grad_all)
    np.testing.assert_allclose(u, poiseuille_closed_form(N, 1.0, NU, f, ghost_grad_all), rtol=1e-11)


@pytest.mark.parametrize("lattice", [orc.OR_D3Q15, orc.OR_D3Q19, orc.OR_D2Q9])
@pytest.mark.parametrize("form", ["cpu", "ghost_grad", "gpu_reference"])
@pytest.mark.parametrize("N", [8, 12])
def test_poiseuille_profile_is_a_discrete_equilibrium(lattice, form, N):
    """The closed-form profile is a steady state of the oracle's full step:
    its face-flux residual cancels the body force in every cell (the O(U^2)
    terms the reduction drops cancel by lattice isotropy for a parallel
    flow), and the mass and transverse momentum residuals vanish."""
    m = _channel(N, lattice)
    f = 8 * NU * 1e-5 / N ** 2
    ghost_all = form != "cpu"
    u_row = poiseuille_closed_form(N, 1.0, NU, f, ghost_all)
    rows = _rows(m, N)
    n = m.n_cells
    o = orc.OracleFluid(m, lattice, visc=NU, dt=0.25, flags=_flags(form))
    o.set_state(np.ones(n), u_row[rows], np.zeros(n), np.zeros(n))
    res = o.residual().reshape(-1, 4)
    assert np.abs(res[:, 1] + f).max() < 1e-7 * f
    # mass and the transverse momenta balance to the round-off of the rho/3 n A terms
    assert np.abs(res[:, [0, 2, 3]]).max() < 1e-15
    # and the other wall form's profile is not (the two deltas differ by U / (9 N^2))
    other = poiseuille_closed_form(N, 1.0, NU, f, not ghost_all)
    o.set_state(np.ones(n), other[rows], np.zeros(n), np.zeros(n))
    res = o.residual().reshape(-1, 4)
    assert np.abs(res[:, 1] + f).max() > 1e-3 * f


def test_poiseuille_reached_from_rest():
    """Run from rest, the oracle settles on the closed-form profile (it is
    the attractor, not only a fixed point)."""
    N = 8
    m = _channel(N, orc.OR_D3Q19)
    f = 8 * NU * 1e-5 / N ** 2
    n = m.n_cells
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=NU, dt=0.25)
    o.set_state(np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n))
    o.step(6000, force=np.tile([f, 0.0, 0.0], n))
    _, u, v, w = o.state()
    rows = _rows(m, N)
    np.testing.assert_allclose(u, poiseuille_closed_form(N, 1.0, NU, f, False)[rows], rtol=1e-7)
    assert np.abs(v).max() < 1e-14 and np.abs(w).max() < 1e-14


@pytest.mark.parametrize("gamma", [0.5, 0.2])
def test_preconditioned_viscous_flux(gamma):
    """gamma != 1 (LBFS.cu:1041-1354, 1488): the steady profile is the
    closed form with nu / gamma (a profile built with nu is not steady)."""
    N = 8
    m = _channel(N, orc.OR_D3Q19)
    f = 8 * NU * 1e-5 / N ** 2
    rows = _rows(m, N)
    n = m.n_cells
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=NU, dt=0.25, gamma=gamma)
    o.set_state(np.ones(n), poiseuille_closed_form(N, 1.0, NU / gamma, f, False)[rows], np.zeros(n), np.zeros(n))
    res = o.residual().reshape(-1, 4)
    assert np.abs(res[:, 1] + f).max() < 1e-9 * f
    o.set_state(np.ones(n), poiseuille_closed_form(N, 1.0, NU, f, False)[rows], np.zeros(n), np.zeros(n))
    assert np.abs(o.residual().reshape(-1, 4)[:, 1] + f).max() > 0.1 * f


def test_face_dt_cap_moves_the_wall_profile():
    """fneq_min > 0 caps the face streaming step at 3 nu / (gamma fneq_min)
    (unstructured_mesh.cpp:1435-1443).  At dt_f = 0.3 h the steady state is
    no longer the shifted parabola; it is the reduced scheme's solution with
    that dt_f."""
    N = 8
    m = _channel(N, orc.OR_D3Q19)
    f = 8 * NU * 1e-5 / N ** 2
    rows = _rows(m, N)
    n = m.n_cells
    fneq_min = 3 * NU / 0.3
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=NU, dt=0.25, fneq_min=fneq_min)
    assert np.allclose(o.fdt, 0.3)
    u1 = poiseuille_1d(N, 1.0, NU, 0.3, f, False)
    o.set_state(np.ones(n), u1[rows], np.zeros(n), np.zeros(n))
    assert np.abs(o.residual().reshape(-1, 4)[:, 1] + f).max() < 1e-9 * f
    u0 = poiseuille_closed_form(N, 1.0, NU, f, False)
    assert np.abs(u1 - u0).max() > 1e-3 * np.abs(u0).max()


@pytest.mark.parametrize("kind", [orc.OR_D3Q15, orc.OR_D3Q19, orc.OR_D2Q9])
@pytest.mark.parametrize("gamma", [1.0, 0.35])
def test_uniform_state_flux_is_the_euler_flux(kind, gamma):
    """For a uniform state every lattice point carries the same equilibrium,
    the non-equilibrium part vanishes and the face flux is the closed-form
    Euler flux of the preconditioned equilibrium:
    (rho u.n, rho/3 n + rho u (u.n) / gamma) A, on every face of a jittered
    periodic mesh."""
    nz = 1 if kind == orc.OR_D2Q9 else 5
    m = abi.HexMesh(6, 5, nz, periodic=(1, 1, 1), jitter=0.1, seed=7)
    n = m.n_cells
    rho, u = 1.02, np.array([0.03, -0.02, 0.0 if kind == orc.OR_D2Q9 else 0.015])
    o = orc.OracleFluid(m, kind, visc=0.05, dt=0.25, gamma=gamma)
    o.set_state(np.full(n, rho), np.full(n, u[0]), np.full(n, u[1]), np.full(n, u[2]))
    fl = o.face_flux().reshape(-1, 4)
    nrm = o.normal.reshape(-1, 3)
    A = o.area
    un = nrm @ u
    np.testing.assert_allclose(fl[:, 0], rho * un * A, atol=1e-15)
    dims = 2 if kind == orc.OR_D2Q9 else 3
    for c in range(dims):
        exp = (rho / 3.0 * nrm[:, c] + rho * u[c] * un / gamma) * A
        np.testing.assert_allclose(fl[:, 1 + c], exp, atol=1e-15)


def _mass_change(mesh, flags, force, steps=1000, closed=False):
    n = mesh.n_cells
    o = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=NU, dt=0.25, flags=flags, closed=closed)
    o.set_state(np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n))
    m0 = np.dot(o.state()[0], o.vol)
    o.step(steps, force=np.tile([force, 0.0, 0.0], n))
    return np.dot(o.state()[0], o.vol) / m0 - 1.0


def test_wall_mass_flux_of_the_two_wall_forms():
    """Mass through no-slip walls of a duct (walls in y and z).
    Orthogonal mesh: with the owner's gradient on the owner side of a wall
    face the mass change is exactly quadratic in the flow (doubling the
    force quadruples it: the O(u^2) equilibrium terms of points extrapolated
    with different tangential gradients on the two sides); with the ghost
    gradient at every point (LBFS.cu:391-394) those terms cancel and mass is
    conserved to round-off.  Jittered mesh with closed faces: the owner-side
    form's wall mass flux is linear in the flow (the third lattice moment
    does not cancel when the two sides' gradients differ)."""
    m = abi.HexMesh(8, 12, 12, periodic=(1, 0, 0), jitter=0.0)
    a, b = _mass_change(m, 0, 1e-5), _mass_change(m, 0, 2e-5)
    assert a > 1e-10 and abs(b / a - 4.0) < 1e-4
    assert abs(_mass_change(m, orc.FLAG_GHOST_FACE_GRAD, 2e-5)) < 1e-14
    j = abi.HexMesh(8, 12, 12, periodic=(1, 0, 0), jitter=0.1, seed=2)
    a, b = _mass_change(j, 0, 1e-5, closed=True), _mass_change(j, 0, 2e-5, closed=True)
    assert a > 1e-7 and abs(b / a - 2.0) < 1e-3
    assert _mass_change(j, 0, 0.0, closed=True) == 0.0


def test_reference_face_areas_drive_the_rest_state_through_the_walls():
    """On the jittered channel the reference's face areas (the sum of the
    triangle areas, unstructured_mesh.cpp:1193-1196) do not close the
    cells: a uniform rest state is not an equilibrium, the spurious flow it
    drives carries mass through the walls, about 0.9 % in 1000 steps of this
    channel with no body force at all (the same with a force); with closed
    faces the rest state stays at rest."""
    j = abi.HexMesh(8, 12, 12, periodic=(1, 0, 0), jitter=0.1, seed=2)
    d0 = _mass_change(j, 0, 0.0)
    assert 5e-3 < d0 < 2e-2
    assert abs(_mass_change(j, 0, 2e-5) - d0) < 1e-3 * d0
