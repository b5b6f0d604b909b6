"""GPU discrete adjoint of the coupled step against the FP64 oracle.

The reference gradient is the central difference of the FP64 oracle's run
(tests/ only), so these tests pin the hand-written adjoint against the
derivative of the reference algorithm itself.  Loss functionals:
fluid  L = sum_i c_i . st_i(N)  (c seeded, 4 per cell);
coupled L = sum_n |X_n(N) - X_target|^2 (the stress-free shape).
"""
import numpy as np
import pytest

import cases
from lbfs_amd import abi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

VISC, DT = 1.0 / 6.0, 0.25
BODY = (2e-5, 0.0, 0.0)
PARAMS = dict(abi.RBC_PARAMS, local_bending_modulus=2e-3, global_bending_modulus=1e-3, viscosity=0.02,
              volume_ratio=0.95)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# ---------------------------------------------------------------------------
# fluid only
# ---------------------------------------------------------------------------
def fluid_case():
    mesh = cases.channel_mesh(16, 12, 12, seed=5)
    st = cases.smooth_state(mesh, seed=5, amp=0.005)
    rng = np.random.default_rng(7)
    c = rng.standard_normal((mesh.n_cells, 4))
    return mesh, st, c


def oracle_fluid_loss(mesh, st, c, steps, fx):
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=VISC, dt=DT)
    fl.set_state(*st)
    fl.step(steps, force=np.tile([fx, 0.0, 0.0], mesh.n_cells))
    r, u, v, w = fl.state()
    return float(np.sum(c * np.stack([r, u, v, w], axis=1)))


def gpu_fluid_adjoint(mesh, st, c, steps, every, fused=True):
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=BODY))
    s.set_fused(fused)
    s.set_state(*st)
    s.adjoint_forward(steps, every)
    g = s.adjoint_backward(lam_state=c)
    s.close()
    return g


def test_fluid_adjoint_matches_oracle_differences():
    """dL/d(body force x) and the initial-state adjoint along a seeded
    direction against central differences of the FP64 oracle."""
    mesh, st, c = fluid_case()
    steps = 8
    g = gpu_fluid_adjoint(mesh, st, c, steps, 3)
    h = 1e-6
    fd_f = (oracle_fluid_loss(mesh, st, c, steps, BODY[0] + h) - oracle_fluid_loss(mesh, st, c, steps, BODY[0] - h)) / (2 * h)
    e_f = rel(g["body_force"][0], fd_f)
    rng = np.random.default_rng(11)
    v = [1e-3 * rng.standard_normal(mesh.n_cells) for _ in range(4)]
    eps = 1e-3
    plus = [st[k] + eps * v[k] for k in range(4)]
    minus = [st[k] - eps * v[k] for k in range(4)]
    fd_s = (oracle_fluid_loss(mesh, plus, c, steps, BODY[0]) - oracle_fluid_loss(mesh, minus, c, steps, BODY[0])) / (2 * eps)
    ad_s = float(np.sum(g["state0"] * np.stack(v, axis=1)))
    e_s = rel(ad_s, fd_s)
    print("fluid adjoint: dL/dfx %.9e vs %.9e (rel %.2e); state direction %.9e vs %.9e (rel %.2e)"
          % (g["body_force"][0], fd_f, e_f, ad_s, fd_s, e_s))
    assert e_f < 1e-4, e_f
    assert e_s < 1e-4, e_s


def test_adjoint_independent_of_checkpoint_interval():
    """Replays recompute the recorded trajectory bit for bit, so the
    gradients do not depend on the checkpoint interval, nor on whether the
    steps and replays run the fused fluid pass (which then also writes the
    recorded gradients) or the gradient and flux passes."""
    mesh, st, c = fluid_case()
    outs = [gpu_fluid_adjoint(mesh, st, c, 7, k) for k in (1, 3, 7)]
    outs.append(gpu_fluid_adjoint(mesh, st, c, 7, 3, fused=False))
    for o in outs[1:]:
        assert np.array_equal(o["body_force"], outs[0]["body_force"])
        assert np.array_equal(o["state0"], outs[0]["state0"])


# ---------------------------------------------------------------------------
# coupled: fluid + IBM + membrane
# ---------------------------------------------------------------------------
CENTER = np.array([12.0, 8.0, 8.0])


def shapes():
    xs, tri = abi.rbc(4, 8.0, center=tuple(CENTER), rot=abi.rotation((1, 1, 0), 0.4))
    xs = xs.reshape(-1, 3)
    # start stretched along x and squeezed across: every force term is active
    x0 = CENTER + (xs - CENTER) * np.array([1.06, 0.97, 0.97])
    return x0, xs, tri


def coupled_inputs():
    mesh = cases.channel_mesh(24, 16, 16, seed=21)
    st = cases.smooth_state(mesh, seed=21, amp=0.005)
    x0, xs, tri = shapes()
    return mesh, st, x0, xs, tri


def oracle_coupled_loss(mesh, st, x0, xs, tri, steps, shear, lb, gb, fx, dx=None):
    p = dict(PARAMS, shear_modulus=shear, local_bending_modulus=lb, global_bending_modulus=gb)
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=VISC, dt=DT)
    fl.set_state(*st)
    mo = orc.OracleMembrane(x0, tri, p, nodes_stress_free=xs, nodes_curvature=xs)
    if dx is not None:
        # move the nodes after the set-up: reference areas, volume and MAD stay
        mo.X += np.ascontiguousarray(dx, dtype=np.float64).reshape(-1)
    cp = orc.OracleCoupled(fl, mesh, [mo], abi.IBM_COS4, body_force=(fx, 0.0, 0.0))
    cp.step(steps)
    return float(np.sum((mo.X.reshape(-1, 3) - xs) ** 2))


def gpu_coupled_solver(mesh, st, x0, xs, tri):
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=BODY))
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, nodes_stress_free=xs, nodes_curvature=xs, **PARAMS)],
                    ibm_kernel=abi.IBM_COS4)
    return s


def test_coupled_adjoint_matches_oracle_differences():
    """Gradients of the node loss with respect to the shear modulus, both
    bending moduli, the body force and the initial node positions (seeded
    direction, reference quantities of the set-up held) against central
    differences of the FP64 coupled oracle."""
    mesh, st, x0, xs, tri = coupled_inputs()
    steps = 12
    s = gpu_coupled_solver(mesh, st, x0, xs, tri)
    s.adjoint_forward(steps, 4)
    X, _, _ = s.membrane_nodes()
    loss = float(np.sum((X - xs) ** 2))
    g = s.adjoint_backward(lam_nodes=2.0 * (X - xs))
    s.close()
    base = dict(shear=PARAMS["shear_modulus"], lb=PARAMS["local_bending_modulus"], gb=PARAMS["global_bending_modulus"],
                fx=BODY[0])
    lo = oracle_coupled_loss(mesh, st, x0, xs, tri, steps, **base)
    assert rel(loss, lo) < 1e-5, (loss, lo)
    ad = {"shear": g["moduli"][0, 0], "lb": g["moduli"][0, 1], "gb": g["moduli"][0, 2], "fx": g["body_force"][0]}
    errs = {}
    for k, v0 in base.items():
        h = 1e-3 * abs(v0)
        up, dn = dict(base), dict(base)
        up[k] = v0 + h
        dn[k] = v0 - h
        fd = (oracle_coupled_loss(mesh, st, x0, xs, tri, steps, **up) -
              oracle_coupled_loss(mesh, st, x0, xs, tri, steps, **dn)) / (2 * h)
        errs[k] = (ad[k], fd, rel(ad[k], fd))
    rng = np.random.default_rng(3)
    dx = 1e-2 * rng.standard_normal(x0.shape)
    eps = 1e-3
    fd_x = (oracle_coupled_loss(mesh, st, x0, xs, tri, steps, dx=eps * dx, **base) -
            oracle_coupled_loss(mesh, st, x0, xs, tri, steps, dx=-eps * dx, **base)) / (2 * eps)
    ad_x = float(np.sum(g["nodes0"] * dx))
    errs["nodes0"] = (ad_x, fd_x, rel(ad_x, fd_x))
    for k, (a, f, e) in errs.items():
        print("coupled adjoint %-7s %.9e vs FD %.9e (rel %.2e)" % (k, a, f, e))
    for k, (a, f, e) in errs.items():
        assert e < 1e-4, (k, a, f, e)


def test_torch_function_matches_adjoint():
    """The autograd Function returns the ABI adjoint's gradients for
    L = sum |X - target|^2, and a second call restarts from the same state."""
    torch = pytest.importorskip("torch")
    from lbfs_amd.adjoint import DeformationRun
    mesh, st, x0, xs, tri = coupled_inputs()
    s = gpu_coupled_solver(mesh, st, x0, xs, tri)
    run = DeformationRun(s, 6, 2)
    p = [torch.tensor(v, dtype=torch.float64, requires_grad=True)
         for v in (PARAMS["shear_modulus"], PARAMS["local_bending_modulus"], BODY[0])]
    target = torch.from_numpy(xs)
    grads = []
    for _ in range(2):
        for t in p:
            t.grad = None
        L = ((run(*p) - target) ** 2).sum()
        L.backward()
        grads.append([float(t.grad) for t in p])
    assert grads[0] == grads[1]
    s.adjoint_rewind()
    loss, gb, gm = s.adjoint_run(6, 2, xs)
    assert np.isclose(loss, float(L.detach()), rtol=1e-12)
    np.testing.assert_allclose(grads[0], [gm[0, 0], gm[0, 1], gb[0]], rtol=1e-12)
    s.close()


def test_coupled_adjoint_pow_springs():
    """The adjoint differentiates the WLC-POW law through the same edge
    template: dL/d(shear modulus) (k_s and k_p both scale with it) against
    central differences of the FP64 oracle."""
    mesh, st, x0, xs, tri = coupled_inputs()
    steps = 8
    p = dict(PARAMS, spring_law=abi.SPRING_WLC_POW)
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=BODY))
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, nodes_stress_free=xs, nodes_curvature=xs, **p)], ibm_kernel=abi.IBM_COS4)
    s.adjoint_forward(steps, 3)
    X, _, _ = s.membrane_nodes()
    g = s.adjoint_backward(lam_nodes=2.0 * (X - xs))
    s.close()

    def loss(shear):
        q = dict(p, shear_modulus=shear)
        fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=VISC, dt=DT)
        fl.set_state(*st)
        mo = orc.OracleMembrane(x0, tri, q, nodes_stress_free=xs, nodes_curvature=xs)
        orc.OracleCoupled(fl, mesh, [mo], abi.IBM_COS4, body_force=BODY).step(steps)
        return float(np.sum((mo.X.reshape(-1, 3) - xs) ** 2))

    mu = p["shear_modulus"]
    h = 1e-3 * mu
    fd = (loss(mu + h) - loss(mu - h)) / (2 * h)
    e = rel(g["moduli"][0, 0], fd)
    print("POW adjoint dL/dmu %.9e vs FD %.9e (rel %.2e)" % (g["moduli"][0, 0], fd, e))
    assert e < 1e-4, e


def test_coupled_adjoint_two_membranes():
    """Two cells with their own moduli in one run: each cell's modulus
    gradients and the body-force gradient against central differences of the
    FP64 oracle stepping both cells (the per-membrane sums of the adjoint are
    split between workgroups that hold edges of both cells)."""
    mesh = cases.channel_mesh(24, 16, 16, seed=23)
    st = cases.smooth_state(mesh, seed=23, amp=0.005)
    cells = []
    for k, (c, ax) in enumerate((((7.0, 8.0, 8.0), (1, 1, 0)), ((17.0, 8.5, 7.5), (0, 1, 1)))):
        xs, tri = abi.rbc(4, 6.0, center=c, rot=abi.rotation(ax, 0.3 + 0.2 * k))
        xs = xs.reshape(-1, 3)
        x0 = np.array(c) + (xs - np.array(c)) * np.array([1.05, 0.97, 0.98])
        cells.append((x0, xs, tri))
    pars = [dict(PARAMS), dict(PARAMS, shear_modulus=1.5 * PARAMS["shear_modulus"],
                                global_bending_modulus=2.0 * PARAMS["global_bending_modulus"])]
    steps = 8
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=BODY))
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, nodes_stress_free=xs, nodes_curvature=xs, **p)
                     for (x0, xs, tri), p in zip(cells, pars)], ibm_kernel=abi.IBM_COS4)
    s.adjoint_forward(steps, 3)
    X, _, _ = s.membrane_nodes()
    target = np.concatenate([xs for _, xs, _ in cells])
    g = s.adjoint_backward(lam_nodes=2.0 * (X - target))
    s.close()

    def loss(ps, fx=BODY[0]):
        fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=VISC, dt=DT)
        fl.set_state(*st)
        mos = [orc.OracleMembrane(x0, tri, q, nodes_stress_free=xs, nodes_curvature=xs)
               for (x0, xs, tri), q in zip(cells, ps)]
        orc.OracleCoupled(fl, mesh, mos, abi.IBM_COS4, body_force=(fx, 0.0, 0.0)).step(steps)
        return sum(float(np.sum((mo.X.reshape(-1, 3) - xs) ** 2)) for mo, (_, xs, _) in zip(mos, cells))

    errs = {}
    for k in range(2):
        for col, key in ((0, "shear_modulus"), (2, "global_bending_modulus")):
            v0 = pars[k][key]
            h = 1e-3 * v0
            up = [dict(p) for p in pars]
            dn = [dict(p) for p in pars]
            up[k][key] = v0 + h
            dn[k][key] = v0 - h
            fd = (loss(up) - loss(dn)) / (2 * h)
            errs["cell%d %s" % (k, key)] = (g["moduli"][k, col], fd)
    h = 1e-3 * BODY[0]
    errs["body force x"] = (g["body_force"][0], (loss(pars, BODY[0] + h) - loss(pars, BODY[0] - h)) / (2 * h))
    for k, (a, f) in errs.items():
        print("two-cell adjoint %-32s %.9e vs FD %.9e (rel %.2e)" % (k, a, f, rel(a, f)))
    for k, (a, f) in errs.items():
        assert rel(a, f) < 1e-4, (k, a, f)
