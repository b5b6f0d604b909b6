"""Pins of the FP64 membrane and IBM oracle (CPU only).

The reference has no tests for spring_network.cu or the IBM kernels; the
oracle is pinned by translation/rotation invariance, the zero-force reference
state, forces equal to minus the gradient of the energies they derive from
(finite differences), an independent dihedral-angle evaluation for the
bending terms, the closed form of the area difference, frozen-coefficient
gradients for the terms that are not energy gradients (curvature, local
bending, area difference), a dissipation function for the membrane
viscosity, and partition of unity / conservation / adjointness of the IBM
weights.
"""
import ctypes as C
import numpy as np
import pytest

import cases
from lbfs_amd import abi
from oracle import oracle as orc

NO_BEND = dict(abi.RBC_PARAMS, global_bending_modulus=0.0, local_bending_modulus=0.0, viscosity=0.0)


def small_rbc(freq=4, d0=8.0, seed=0, noise=0.0):
    x, t = abi.rbc(freq, d0)
    if noise:
        x = x + noise * np.random.default_rng(seed).standard_normal(x.shape)
    return x, t


def test_springs_match_and_cover_every_edge_once():
    for freq in (2, 4, 8):
        x, t = abi.icosphere(freq)
        sp = abi.springs(len(x), t)
        assert np.array_equal(sp, orc.build_springs(len(x), t))
        assert len(sp) * 2 == 3 * len(t)
        edges = {tuple(sorted((a, c))) for a, b, c, d in sp}
        assert len(edges) == len(sp)
        tris = {tuple(np.roll(tt, -np.argmin(tt))) for tt in t}
        for a, b, c, d in sp[:50]:
            for tri in ((a, b, c), (a, c, d)):
                assert tuple(np.roll(tri, -np.argmin(tri))) in tris


def test_geodesic_counts_and_outward_winding():
    for freq in (1, 3, 8, 11, 16):
        x, t = abi.icosphere(freq)
        assert len(x) == 10 * freq * freq + 2 and len(t) == 20 * freq * freq
        assert np.allclose(np.linalg.norm(x, axis=1), 1.0)
        N = np.cross(x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]])
        assert np.all(np.einsum("ij,ij->i", N, x[t].mean(1)) > 0)


def test_forces_are_translation_and_rotation_free():
    x0, t = small_rbc()
    params = dict(abi.RBC_PARAMS, local_bending_modulus=2e-3, viscosity=0.05)
    m = orc.OracleMembrane(x0, t, params)
    rng = np.random.default_rng(1)
    X = x0 + 0.05 * rng.standard_normal(x0.shape)
    V = 0.01 * rng.standard_normal(x0.shape)
    F = m.forces(X, V)
    assert np.abs(F.sum(0)).max() < 1e-13
    assert np.abs(np.cross(X, F).sum(0)).max() < 1e-12


def test_reference_state_is_force_free_without_bending():
    x0, t = small_rbc()
    m = orc.OracleMembrane(x0, t, NO_BEND)
    assert np.abs(m.forces()).max() < 1e-14


def _energy(m, X, params, n0):
    """WLC + local area + global area + volume energies whose negative
    gradients are the force terms of spring_network.cu:613-971."""
    X = X.reshape(-1, 3)
    sp = m.spring.reshape(-1, 4)
    L = np.linalg.norm(X[sp[:, 2]] - X[sp[:, 0]], axis=1)
    r = params["wlc_ratio"]
    Cc = 0.25 * r * r / (r - 1) ** 2 - 0.25 + 1 / r
    Lm, ks = m.arr["contour_length"], m.arr["spring_constant"]
    x = L / Lm
    e_wlc = np.sum(ks * Lm * (0.25 / (1 - x) - x / 4 + x * x / 2 - Cc * x))
    t = m.tri.reshape(-1, 3)
    N = np.cross(X[t[:, 1]] - X[t[:, 0]], X[t[:, 2]] - X[t[:, 0]])
    n = np.linalg.norm(N, axis=1)
    e_loc = np.sum(params["local_area_modulus"] * (n - n0) ** 2 / (2 * n0))
    A, A0 = 0.5 * n.sum(), m.mb.total_area_0
    e_glob = params["global_area_modulus"] * (A - A0) ** 2 / A0
    V = np.sum(np.einsum("ij,ij->i", N, X[t].mean(1))) / 6.0
    V0 = m.mb.total_volume_0
    e_vol = params["volume_modulus"] * (V - V0) ** 2 / (12.0 * V0)
    return e_wlc + e_loc + e_glob + e_vol


def test_conservative_forces_are_energy_gradients():
    x0, t = small_rbc(freq=3, d0=6.0)
    params = dict(NO_BEND, volume_ratio=0.9)
    m = orc.OracleMembrane(x0, t, params)
    tt = t.reshape(-1, 3)
    n0 = np.linalg.norm(np.cross(x0[tt[:, 1]] - x0[tt[:, 0]], x0[tt[:, 2]] - x0[tt[:, 0]]), axis=1)
    X = x0 + 0.03 * np.random.default_rng(2).standard_normal(x0.shape)
    F = m.forces(X).reshape(-1)
    h = 1e-6
    Xf = X.reshape(-1).copy()
    for j in np.random.default_rng(3).choice(len(Xf), 25, replace=False):
        xp, xm = Xf.copy(), Xf.copy()
        xp[j] += h
        xm[j] -= h
        g = (_energy(m, xp, params, n0) - _energy(m, xm, params, n0)) / (2 * h)
        assert abs(-g - F[j]) < 1e-7 + 1e-5 * abs(F[j]), (j, -g, F[j])


def _dihedral(X, sp):
    """Signed dihedral angle of each spring with the convention of
    spring_network.cu:752-773 (convex positive)."""
    A, B, Cn, D = (X[sp[:, k]] for k in range(4))
    N1 = np.cross(B - A, Cn - A)
    N2 = np.cross(Cn - A, D - A)
    cs = np.clip(np.einsum("ij,ij->i", N1, N2) / (np.linalg.norm(N1, axis=1) * np.linalg.norm(N2, axis=1)), -1, 1)
    d = np.einsum("ij,ij->i", N1, (A + Cn + D) / 3 - (A + B + Cn) / 3)
    return np.arccos(cs) * ((d < 0).astype(float) - (d > 0).astype(float))


def test_bending_force_is_gradient_of_theta_length_with_frozen_coefficients():
    """F = -sum_e c_e d(theta_e L_e)/dX with c_e = k (kappa_e - kappa0_e) taken
    at the evaluation point (the curvature term of spring_network.cu:899-916)."""
    x0, t = small_rbc(freq=3, d0=6.0)
    k = 2e-3
    params = dict(NO_BEND, global_bending_modulus=k, shear_modulus=0.0, volume_modulus=0.0,
                  local_area_modulus=0.0, global_area_modulus=0.0)
    m = orc.OracleMembrane(x0, t, params)
    X = x0 + 0.02 * np.random.default_rng(4).standard_normal(x0.shape)
    F = m.forces(X).reshape(-1)
    c = k * (m.arr["curvature"] - m.arr["curvature_0"])
    sp = m.spring.reshape(-1, 4)

    def G(xf):
        Xr = xf.reshape(-1, 3)
        L = np.linalg.norm(Xr[sp[:, 2]] - Xr[sp[:, 0]], axis=1)
        return np.sum(c * _dihedral(Xr, sp) * L)

    Xf = X.reshape(-1).copy()
    h = 1e-6
    for j in np.random.default_rng(5).choice(len(Xf), 25, replace=False):
        xp, xm = Xf.copy(), Xf.copy()
        xp[j] += h
        xm[j] -= h
        g = (G(xp) - G(xm)) / (2 * h)
        assert abs(-g - F[j]) < 1e-9 + 1e-5 * abs(F[j]), (j, -g, F[j])


def _theta_length(X, sp):
    L = np.linalg.norm(X[sp[:, 2]] - X[sp[:, 0]], axis=1)
    return _dihedral(X, sp) * L


def _spring_norms(X, sp):
    """|N1|, |N2|: twice the areas of the spring's triangles (a, b, c) and (a, c, d)."""
    A, B, Cn, D = (X[sp[:, k]] for k in range(4))
    return (np.linalg.norm(np.cross(B - A, Cn - A), axis=1), np.linalg.norm(np.cross(Cn - A, D - A), axis=1))


def _surface_area(X, t):
    tt = t.reshape(-1, 3)
    return 0.5 * np.sum(np.linalg.norm(np.cross(X[tt[:, 1]] - X[tt[:, 0]], X[tt[:, 2]] - X[tt[:, 0]]), axis=1))


def test_mad_term_uses_the_previous_call_and_the_closed_form_mad():
    """The area-difference term of spring_network.cu:645 and :899-916: a second
    force call adds mad_coef d(sum theta L)/dX with mad_coef = -2 k (MAD - MAD0) / A0
    taken from the previous call's MAD, and MAD = sum kappa (|N1| + |N2|) / 3 / A
    reduces to 2 sum theta L / A (kappa = theta L / ((|N1| + |N2|) / 6))."""
    x0, t = small_rbc(freq=3, d0=6.0)
    k = 2e-3
    params = dict(NO_BEND, global_bending_modulus=k, shear_modulus=0.0, volume_modulus=0.0,
                  local_area_modulus=0.0, global_area_modulus=0.0)
    m = orc.OracleMembrane(x0, t, params)
    sp = m.spring.reshape(-1, 4)
    mad = lambda X: 2.0 * np.sum(_theta_length(X, sp)) / _surface_area(X, t)
    assert abs(m.globals()[5] - mad(x0)) < 1e-12 * abs(mad(x0))
    rng = np.random.default_rng(6)
    X1 = 1.3 * x0 + 0.05 * rng.standard_normal(x0.shape)  # MAD scales as 1/size
    m.forces(X1)
    assert abs(m.globals()[2] - mad(X1)) < 1e-12 * abs(mad(X1))
    X = x0 + 0.02 * rng.standard_normal(x0.shape)
    F = m.forces(X).reshape(-1)
    mad_coef = -2.0 * k * (mad(X1) - mad(x0)) / _surface_area(x0, t)
    assert abs(mad_coef) > 0.1 * k * abs(mad(x0)) / _surface_area(x0, t)
    c = k * (m.arr["curvature"] - m.arr["curvature_0"]) - mad_coef
    G = lambda xf: np.sum(c * _theta_length(xf.reshape(-1, 3), sp))
    Xf = X.reshape(-1).copy()
    h = 1e-6
    for j in np.random.default_rng(7).choice(len(Xf), 25, replace=False):
        xp, xm = Xf.copy(), Xf.copy()
        xp[j] += h
        xm[j] -= h
        g = (G(xp) - G(xm)) / (2 * h)
        assert abs(-g - F[j]) < 1e-9 + 1e-5 * abs(F[j]), (j, -g, F[j])
    # the term alone: a fresh membrane's first call at X has mad_coef = 0 and
    # the same curvature coefficients, so the difference is
    # mad_coef d(sum theta L)/dX (this pins the factor 2, A0 rather than A and
    # the previous call's MAD rather than the current one)
    m2 = orc.OracleMembrane(x0, t, params)
    dF = F - m2.forces(X).reshape(-1)
    S = lambda xf: np.sum(_theta_length(xf.reshape(-1, 3), sp))
    for j in np.random.default_rng(12).choice(len(Xf), 25, replace=False):
        xp, xm = Xf.copy(), Xf.copy()
        xp[j] += h
        xm[j] -= h
        g = mad_coef * (S(xp) - S(xm)) / (2 * h)
        assert abs(g - dF[j]) < 1e-6 * np.max(np.abs(dF)), (j, g, dF[j])


def test_local_bending_force_is_gradient_of_spring_area_with_frozen_curvature():
    """The local bending term of spring_network.cu:870-896:
    F = -d/dX sum_e 0.5 k_b (kappa0_e - kappa_e)^2 (|N1| + |N2|) / 6 with kappa
    held at the evaluation point."""
    x0, t = small_rbc(freq=3, d0=6.0)
    kb = 5e-3
    params = dict(NO_BEND, local_bending_modulus=kb, shear_modulus=0.0, volume_modulus=0.0,
                  local_area_modulus=0.0, global_area_modulus=0.0)
    m = orc.OracleMembrane(x0, t, params)
    sp = m.spring.reshape(-1, 4)
    X = x0 + 0.02 * np.random.default_rng(8).standard_normal(x0.shape)
    F = m.forces(X).reshape(-1)
    assert np.max(np.abs(F)) > 0
    w = 0.5 * kb * (m.arr["curvature_0"] - m.arr["curvature"]) ** 2

    def G(xf):
        n1, n2 = _spring_norms(xf.reshape(-1, 3), sp)
        return np.sum(w * (n1 + n2) / 6.0)

    Xf = X.reshape(-1).copy()
    h = 1e-6
    for j in np.random.default_rng(9).choice(len(Xf), 25, replace=False):
        xp, xm = Xf.copy(), Xf.copy()
        xp[j] += h
        xm[j] -= h
        g = (G(xp) - G(xm)) / (2 * h)
        assert abs(-g - F[j]) < 1e-12 + 1e-5 * abs(F[j]), (j, -g, F[j])
    # the curvature the term uses is theta L / ((|N1| + |N2|) / 6)
    n1, n2 = _spring_norms(X, sp)
    kappa = _theta_length(X, sp) / ((n1 + n2) / 6.0)
    np.testing.assert_allclose(m.arr["curvature"], kappa, rtol=1e-10, atol=1e-12)


def test_osay_viscous_force_is_minus_the_gradient_of_a_dissipation_function():
    """The viscous term of spring_network.cu:851-866 is linear in the node
    velocities and equals -dR/dV with R = (2 / sqrt 3) eta sum_e (e_hat . (v_c - v_a))^2;
    it conserves momentum, vanishes for rigid translation and rotation and
    dissipates (V . F <= 0)."""
    x0, t = small_rbc(freq=3, d0=6.0)
    eta = 0.7
    params = dict(abi.RBC_PARAMS, viscosity=eta)
    m = orc.OracleMembrane(x0, t, params)
    sp = m.spring.reshape(-1, 4)
    rng = np.random.default_rng(10)
    X = x0 + 0.02 * rng.standard_normal(x0.shape)
    m.forces(X)  # the area-difference term uses the previous call's MAD: make it MAD(X)
    F0 = m.forces(X)
    visc = lambda V: (m.forces(X, V) - F0).reshape(-1)

    d = X[sp[:, 2]] - X[sp[:, 0]]
    e = d / np.linalg.norm(d, axis=1)[:, None]

    def R(vf):
        V = vf.reshape(-1, 3)
        return 2.0 / np.sqrt(3.0) * eta * np.sum(np.einsum("ij,ij->i", e, V[sp[:, 2]] - V[sp[:, 0]]) ** 2)

    V = rng.standard_normal(X.shape)
    Fv = visc(V)
    Vf = V.reshape(-1).copy()
    h = 1e-3  # R is quadratic: central differences are exact up to round-off
    for j in np.random.default_rng(11).choice(len(Vf), 30, replace=False):
        vp, vm = Vf.copy(), Vf.copy()
        vp[j] += h
        vm[j] -= h
        g = (R(vp) - R(vm)) / (2 * h)
        assert abs(-g - Fv[j]) < 1e-10 + 1e-8 * abs(Fv[j]), (j, -g, Fv[j])
    W = rng.standard_normal(X.shape)
    np.testing.assert_allclose(visc(2.0 * V - 3.0 * W), 2.0 * Fv - 3.0 * visc(W), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(Fv.reshape(-1, 3).sum(axis=0), 0.0, atol=1e-10 * np.abs(Fv).max())
    assert np.dot(Vf, Fv) < 0
    rigid = np.array([0.3, -0.2, 0.1]) + np.cross(np.array([0.05, 0.02, -0.04]), X - X.mean(axis=0))
    assert np.max(np.abs(visc(rigid))) < 1e-12 * np.max(np.abs(Fv))


def test_spring_constant_formula():
    """k_s = 4 r L0 mu / (sqrt(3) (0.5 (1 - 1/r)^-3 + 1)), L_max = r L0
    (spring_network.cu:516-604, 1600-1615)."""
    x0, t = small_rbc()
    p = abi.RBC_PARAMS
    m = orc.OracleMembrane(x0, t, p)
    sp = m.spring.reshape(-1, 4)
    L0 = np.linalg.norm(x0[sp[:, 2]] - x0[sp[:, 0]], axis=1)
    r = p["wlc_ratio"]
    ks = 4 * r * L0 * p["shear_modulus"] / (np.sqrt(3) * (0.5 * (1 - 1 / r) ** -3 + 1))
    np.testing.assert_allclose(m.arr["spring_constant"], ks, rtol=1e-14)
    np.testing.assert_allclose(m.arr["contour_length"], r * L0, rtol=1e-14)


def _ibm_setup(kernel, periodic=(1, 0, 0)):
    mesh = abi.HexMesh(12, 10, 10, periodic=periodic, jitter=0.1, seed=8)
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=1 / 6, dt=0.25)
    return mesh, fl, orc.OracleCoupled(fl, mesh, [], kernel)


@pytest.mark.parametrize("kernel", [abi.IBM_LINEAR2, abi.IBM_COS4])
def test_ibm_partition_of_unity_and_linear_exactness(kernel):
    mesh, fl, cp = _ibm_setup(kernel)
    rng = np.random.default_rng(6)
    X = np.column_stack([rng.uniform(0, 12, 30), rng.uniform(3, 7, 30), rng.uniform(3, 7, 30)]).reshape(-1)
    n = mesh.n_cells
    # constant field: interpolated exactly by both kernels
    fl.st[:] = 0.0
    fl.st[1:4 * n:4] = 0.7
    U = cp.interpolate(X).reshape(-1, 3)
    np.testing.assert_allclose(U[:, 0], 0.7, rtol=1e-14)
    if kernel == abi.IBM_LINEAR2:
        # trilinear reproduces linear functions of the logical cell centres exactly
        i = np.arange(n)
        cx = (i % 12) + 0.5
        fl.st[1:4 * n:4] = 0.3 + 0.1 * cx
        U = cp.interpolate(X).reshape(-1, 3)
        Xr = X.reshape(-1, 3)
        inner = (Xr[:, 0] > 1.0) & (Xr[:, 0] < 11.0)
        np.testing.assert_allclose(U[inner, 0], 0.3 + 0.1 * Xr[inner, 0], rtol=1e-13)


def test_ibm_cos4_conserves_momentum_and_is_adjoint():
    mesh, fl, cp = _ibm_setup(abi.IBM_COS4)
    rng = np.random.default_rng(7)
    X = np.column_stack([rng.uniform(0, 12, 20), rng.uniform(3, 7, 20), rng.uniform(3, 7, 20)]).reshape(-1)
    F = rng.standard_normal(60)
    f = np.zeros(3 * mesh.n_cells)
    cp.spread(X, F, f)
    np.testing.assert_allclose((f.reshape(-1, 3) * fl.vol[:, None]).sum(0), F.reshape(-1, 3).sum(0), rtol=1e-12)
    # adjointness: sum_n U_n . F_n = sum_c u_c . f_c V_c
    n = mesh.n_cells
    u = rng.standard_normal((n, 3))
    fl.st[:] = 0.0
    for a in range(3):
        fl.st[1 + a:4 * n:4] = u[:, a]
    U = cp.interpolate(X)
    lhs = np.dot(U, F)
    rhs = np.sum(u * f.reshape(-1, 3) * fl.vol[:, None])
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_ibm_periodic_wrap_is_seamless():
    """A node near x = 0 and its periodic image near x = L interpolate alike."""
    mesh, fl, cp = _ibm_setup(abi.IBM_COS4)
    n = mesh.n_cells
    c, _ = cases.centroids(mesh)
    fl.st[:] = 0.0
    fl.st[1:4 * n:4] = np.cos(2 * np.pi * c[:, 0] / 12.0)
    X1 = np.array([0.3, 5.0, 5.0])
    X2 = np.array([12.3, 5.0, 5.0])
    assert abs(cp.interpolate(X1)[0] - cp.interpolate(X2)[0]) < 1e-14


def test_internal_viscosity_marks_interior():
    """Plucker ray parity (immersed_boundary_method.cu:389-527): cells well
    inside a closed membrane get the ratio, cells in the box but outside get 1."""
    mesh = abi.HexMesh(16, 16, 16, periodic=(0, 0, 0), jitter=0.1, seed=9)
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=1 / 6, dt=0.25)
    u, t = abi.icosphere(6)
    x = 5.0 * u + 8.0
    m = orc.OracleMembrane(x, t, dict(abi.RBC_PARAMS, internal_viscosity_ratio=5.0))
    cp = orc.OracleCoupled(fl, mesh, [m], abi.IBM_COS4)
    vf = cp.internal_viscosity()
    c = fl.cc[: 3 * mesh.n_cells].reshape(-1, 3)
    r = np.linalg.norm(c - 8.0, axis=1)
    assert np.all(vf[r < 4.0] == 5.0)
    assert np.all(vf[r > 6.0] == 1.0)
    assert np.all((vf >= 1.0) & (vf <= 5.0))


def test_energies_vanish_at_rest_and_spring_energy_matches_wlc_force():
    """Output-step energies (Solver_gpu.cu:1060-1139): every term is zero in
    the stress-free reference state; with only WLC springs active, the
    spring energy's gradient is half the negative WLC force (the reference
    halves the sum, counting each spring twice), checked by central
    differences against the force routine; the global terms equal their
    closed forms under a uniform scaling."""
    x0, t = small_rbc(freq=3, d0=6.0)
    params = dict(NO_BEND, volume_ratio=1.0)
    m = orc.OracleMembrane(x0, t, params)
    e = m.energies()
    assert np.abs(e).max() < 1e-14
    wlc_only = dict(params, volume_modulus=0.0, local_area_modulus=0.0, global_area_modulus=0.0)
    w = orc.OracleMembrane(x0, t, wlc_only)
    X = x0 + 0.05 * np.random.default_rng(7).standard_normal(x0.shape)
    F = w.forces(X).reshape(-1)
    Xf = X.reshape(-1)
    h = 1e-6
    for j in np.random.default_rng(8).choice(len(Xf), 12, replace=False):
        xp, xm = Xf.copy(), Xf.copy()
        xp[j] += h
        xm[j] -= h
        w.forces(xp)
        ep = w.energies(xp)[0]
        w.forces(xm)
        em = w.energies(xm)[0]
        g = (ep - em) / (2 * h)
        assert abs(g + 0.5 * F[j]) < 1e-8 + 1e-5 * abs(F[j]), (j, g, F[j])
    # closed-form global terms after a uniform scaling
    sc = 1.05
    Xs = x0 * sc
    m.forces(Xs)
    e = m.energies(Xs)
    g = m.globals()
    V, A, V0, A0 = g[0], g[1], g[3], g[4]
    assert abs(V / V0 - sc ** 3) < 1e-12 and abs(A / A0 - sc ** 2) < 1e-12
    assert abs(e[4] - 0.5 * params["volume_modulus"] * (V - V0) ** 2 / V0) < 1e-15
    assert abs(e[3] - 0.5 * params["global_area_modulus"] * (A - A0) ** 2 / A0) < 1e-15
    # local area term: every |N| scales by sc^2
    n0 = m.arr["spring_area_0"]
    exp_la = params["global_area_modulus"] * 0.5 * np.sum((sc * sc - 1) ** 2 * n0)
    assert abs(e[1] - exp_la) < 1e-12 * exp_la
    assert abs(e[6] - e[:6].sum()) < 1e-15


def _stretched_mesh(nx=20, ny=16, nz=16, amp=0.25, periodic=(1, 0, 0), seed=1):
    """Tensor-product stretched box: vertex coordinates x -> x + amp L/(2 pi) sin(2 pi x / L)
    per axis (no jitter), so the vertex planes and cell lines are exact."""
    m = abi.HexMesh(nx, ny, nz, periodic=periodic, jitter=0.0, seed=seed)
    P = m.points.reshape(-1, 3)
    for a, n in enumerate((nx, ny, nz)):
        L = float(n)
        P[:, a] = P[:, a] + amp * L / (2 * np.pi) * np.sin(2 * np.pi * P[:, a] / L)
    return m


def lagrange_grid_arrays(m):
    """Vertex planes and first-line cell centroids per axis from the points and the oracle geometry."""
    P = m.points.reshape(-1, 3)
    planes = [np.unique(np.round(P[:, a], 12)) for a in range(3)]
    cc = orc.OracleFluid(m, orc.OR_D3Q19, visc=1 / 6, dt=0.25).cc.reshape(-1, 3)
    nx, ny = m.nx, m.ny
    lines = [cc[np.arange(m.nx), 0], cc[np.arange(m.ny) * nx, 1], cc[np.arange(m.nz) * nx * ny, 2]]
    return planes, lines


def test_lagrange4_reproduces_cubics_on_stretched_grid():
    """OR_IBM_LAGRANGE4 (immersed_boundary_method.cu:588-872): on a
    tensor-product stretched grid the weights are the cubic Lagrange basis of
    the 4 nearest cell lines per axis, so any field cubic in each coordinate
    is interpolated exactly, the weights sum to 1 and reproduce the node
    position; spreading conserves momentum (sum f V = F)."""
    m = _stretched_mesh()
    planes, lines = lagrange_grid_arrays(m)
    assert [len(p) for p in planes] == [m.nx + 1, m.ny + 1, m.nz + 1]
    fl = orc.OracleFluid(m, orc.OR_D3Q19, visc=1 / 6, dt=0.25)
    cc = fl.cc.reshape(-1, 3)[: m.n_cells]
    cp = orc.OracleCoupled(fl, m, [], abi.IBM_LAGRANGE4, planes=planes, lines=lines)
    rng = np.random.default_rng(4)
    X = np.column_stack([rng.uniform(3, m.nx - 3, 40), rng.uniform(3, m.ny - 3, 40), rng.uniform(3, m.nz - 3, 40)])
    def field(p):
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        return 0.01 + 1e-4 * x ** 3 - 2e-4 * x * y ** 2 + 3e-5 * z ** 3 * y - 1e-3 * z
    n = m.n_cells
    st = np.zeros(4 * n)
    st[1::4] = field(cc)
    U = np.zeros(3 * len(X))
    orc.load().or_ibm_interpolate(abi.IBM_LAGRANGE4, C.byref(cp.g), len(X), orc.ptr(np.ascontiguousarray(X.reshape(-1))),
                                  orc.ptr(st), orc.ptr(U))
    np.testing.assert_allclose(U[0::3], field(X), rtol=0, atol=1e-12)
    for comp, vals in ((2, np.ones(n)), (3, cc[:, 0])):
        st2 = np.zeros(4 * n)
        st2[comp::4] = vals
        U2 = np.zeros(3 * len(X))
        orc.load().or_ibm_interpolate(abi.IBM_LAGRANGE4, C.byref(cp.g), len(X),
                                      orc.ptr(np.ascontiguousarray(X.reshape(-1))), orc.ptr(st2), orc.ptr(U2))
        expect = np.ones(len(X)) if comp == 2 else X[:, 0]
        np.testing.assert_allclose(U2[comp - 1::3], expect, rtol=0, atol=1e-12)
    F = rng.standard_normal(X.shape)
    f = np.zeros(3 * n)
    orc.load().or_ibm_spread(abi.IBM_LAGRANGE4, C.byref(cp.g), len(X), orc.ptr(np.ascontiguousarray(X.reshape(-1))),
                             orc.ptr(np.ascontiguousarray(F.reshape(-1))), orc.ptr(f))
    np.testing.assert_allclose((f.reshape(-1, 3) * fl.vol[:, None]).sum(0), F.sum(0), atol=1e-12)


def test_mad_compat_sums_the_first_n_nodes_springs():
    """mad_compat (Solver_gpu.cu:2355: the MAD block sum runs over the first
    n_nodes entries of the per-spring array): MAD = 2 sum_{e < n_nodes}
    theta_e L_e / A, against the full sum 2 sum_e theta_e L_e / A without it."""
    x0, t = small_rbc(freq=3, d0=6.0)
    X = x0 + 0.03 * np.random.default_rng(13).standard_normal(x0.shape)
    params = dict(NO_BEND, global_bending_modulus=1e-3)
    full = orc.OracleMembrane(X, t, params)
    part = orc.OracleMembrane(X, t, dict(params, mad_compat=1))
    sp = full.spring.reshape(-1, 4)
    n = full.n
    assert n < len(sp)
    tl = _theta_length(X, sp)
    A = _surface_area(X, t)
    assert abs(full.globals()[5] - 2.0 * np.sum(tl) / A) < 1e-12 * abs(full.globals()[5])
    assert abs(part.globals()[5] - 2.0 * np.sum(tl[:n]) / A) < 1e-12 * abs(part.globals()[5])
    assert abs(part.globals()[5] - full.globals()[5]) > 1e-3 * abs(full.globals()[5])


def test_sphere_reference_bending_uses_the_spherical_curvature():
    """sphere_reference (spring_network.cu:493-495): every spring's
    spontaneous curvature is curvature_ratio / 3.265483301, so the curvature
    and local bending terms are the frozen-coefficient gradients with that
    constant kappa0 (kappa = theta L / ((|N1| + |N2|) / 6) computed here)."""
    x0, t = small_rbc(freq=3, d0=6.0)
    k, kb, ratio = 2e-3, 3e-3, 1.7
    params = dict(NO_BEND, global_bending_modulus=k, local_bending_modulus=kb, shear_modulus=0.0,
                  volume_modulus=0.0, local_area_modulus=0.0, global_area_modulus=0.0,
                  sphere_reference=1, curvature_ratio=ratio)
    m = orc.OracleMembrane(x0, t, params)
    kappa0 = ratio / 3.265483301
    sp = m.spring.reshape(-1, 4)
    X = x0 + 0.02 * np.random.default_rng(14).standard_normal(x0.shape)
    F = m.forces(X).reshape(-1)
    n1, n2 = _spring_norms(X, sp)
    kappa = _theta_length(X, sp) / ((n1 + n2) / 6.0)
    c = k * (kappa - kappa0)
    w = 0.5 * kb * (kappa0 - kappa) ** 2

    def G(xf):
        Xr = xf.reshape(-1, 3)
        a1, a2 = _spring_norms(Xr, sp)
        return np.sum(c * _theta_length(Xr, sp)) + np.sum(w * (a1 + a2) / 6.0)

    Xf = X.reshape(-1).copy()
    h = 1e-6
    for j in np.random.default_rng(15).choice(len(Xf), 25, replace=False):
        xp, xm = Xf.copy(), Xf.copy()
        xp[j] += h
        xm[j] -= h
        g = (G(xp) - G(xm)) / (2 * h)
        assert abs(-g - F[j]) < 1e-10 + 1e-5 * abs(F[j]), (j, -g, F[j])
    # without it the spontaneous curvature is zero and the force differs
    m0 = orc.OracleMembrane(x0, t, dict(params, sphere_reference=0))
    assert np.max(np.abs(m0.forces(X).reshape(-1) - F)) > 1e-3 * np.max(np.abs(F))


POW = dict(NO_BEND, shear_modulus=5e-3, wlc_ratio=2.2, volume_modulus=0.0, local_area_modulus=0.0,
           global_area_modulus=0.0, spring_law=1)


def test_pow_spring_constants_give_the_published_shear_modulus():
    """WLC-POW constants (get_spring_pow_constants, spring_network.cu:1655-1682):
    the force vanishes at the rest length, k_p < 0 (a repulsive POW term in
    the reference's sign), and the linear shear modulus of the network in the
    published WLC-POW form (Fedosov, Caswell & Karniadakis 2010, with
    k_B T / p = 4 k_s and exponent m = 2),
    mu0 = sqrt3 k_B T / (4 p L_max x0) (x0 / (2 (1-x0)^3) - 1 / (4 (1-x0)^2) + 1/4)
          + sqrt3 (m + 1) |k_p| / (4 L0^(m+1)),
    equals the membrane's shear modulus on every spring."""
    x0, t = small_rbc(freq=4, d0=8.0)
    m = orc.OracleMembrane(x0, t, POW)
    sp = m.spring.reshape(-1, 4)
    L0 = np.linalg.norm(x0[sp[:, 2]] - x0[sp[:, 0]], axis=1)
    ks, kp, Lm = m.arr["spring_constant"], m.arr["pow_constant"], m.arr["contour_length"]
    r = POW["wlc_ratio"]
    x = 1.0 / r
    np.testing.assert_allclose(Lm, r * L0, rtol=1e-14)
    assert np.all(kp < 0)
    kbt_p = 4.0 * ks
    mu0 = (np.sqrt(3.0) * kbt_p / (4.0 * Lm * x) * (x / (2 * (1 - x) ** 3) - 1 / (4 * (1 - x) ** 2) + 0.25)
           + np.sqrt(3.0) * 3.0 * np.abs(kp) / (4.0 * L0 ** 3))
    np.testing.assert_allclose(mu0, POW["shear_modulus"], rtol=1e-12)
    # no force at the rest length: the stress-free shape is force free
    F = m.forces(x0)
    assert np.abs(F).max() < 1e-12 * np.abs(ks).max()


def test_pow_spring_force_is_minus_the_energy_gradient():
    """The WLC-POW force -(k_s (1/(1-x)^2 - 1 + 4x) + k_p / L^2) along the
    spring (spring_network.cu:221-223) is -dE/dX of its energy, with
    E = k_s L_max (g(x) - g(x0)) - k_p (1/L - 1/L0), g(x) = 1/(1-x) - x + 2x^2,
    by central differences; and it differs from the WLC law."""
    x0, t = small_rbc(freq=3, d0=6.0)
    m = orc.OracleMembrane(x0, t, POW)
    X = x0 + 0.03 * np.random.default_rng(16).standard_normal(x0.shape)
    F = m.forces(X).reshape(-1)
    sp = m.spring.reshape(-1, 4)
    ks, kp, Lm = m.arr["spring_constant"], m.arr["pow_constant"], m.arr["contour_length"]
    L0 = Lm / POW["wlc_ratio"]

    def E(xf):
        Xr = xf.reshape(-1, 3)
        L = np.linalg.norm(Xr[sp[:, 2]] - Xr[sp[:, 0]], axis=1)
        g = lambda s: 1 / (1 - s) - s + 2 * s * s  # noqa: E731
        return np.sum(ks * Lm * (g(L / Lm) - g(L0 / Lm)) - kp * (1 / L - 1 / L0))

    Xf = X.reshape(-1).copy()
    h = 1e-6
    for j in np.random.default_rng(17).choice(len(Xf), 25, replace=False):
        xp, xm = Xf.copy(), Xf.copy()
        xp[j] += h
        xm[j] -= h
        g = (E(xp) - E(xm)) / (2 * h)
        assert abs(-g - F[j]) < 1e-11 + 1e-6 * np.abs(F).max(), (j, -g, F[j])
    # the energy routine reports the same energy
    m.forces(X)
    np.testing.assert_allclose(m.energies(X)[0], E(Xf), rtol=1e-10)
    w = orc.OracleMembrane(x0, t, dict(POW, spring_law=0))
    assert np.abs(w.forces(X).reshape(-1) - F).max() > 0.05 * np.abs(F).max()
