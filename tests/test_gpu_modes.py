"""GPU parity, through the C ABI, for the reference modes a drop-in binds
(the GPU solver's semantics flags, preconditioning gamma, the gradient
types, the face-step cap, the membrane's MAD partial sum and spherical
reference curvature) against the FP64 oracle given the same modes."""
import numpy as np
import pytest

import cases
from lbfs_amd import abi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

BODY = (2e-5, 0.0, 0.0)
ORACLE_FLAGS = abi.FLAG_UPWIND_BAND | abi.FLAG_NEUMANN_SIGNED | abi.FLAG_GHOST_DIR_GPU | abi.FLAG_GHOST_FACE_GRAD


def step_pair(mesh, steps, lattice=(abi.D3Q19, orc.OR_D3Q19), flags=0, visc=1.0 / 6.0, dt=0.25, gamma=1.0,
              fneq_min=0.0, grad_type=abi.GRAD_HYBRID, seed=14):
    closed = bool(flags & abi.FLAG_CLOSED_FACES)
    s = abi.Solver(mesh, abi.make_params(lattice=lattice[0], visc=visc, dt=dt, flags=flags, body_force=BODY,
                                         gamma=gamma, fneq_min=fneq_min, gradient_type=grad_type))
    o = orc.OracleFluid(mesh, lattice[1], visc=visc, dt=dt, gamma=gamma, fneq_min=fneq_min, flags=flags & ORACLE_FLAGS,
                        closed=closed, det_compat=bool(flags & abi.FLAG_LS_DET_COMPAT), grad_type=grad_type)
    st = cases.smooth_state(mesh, seed=seed)
    s.set_state(*st)
    o.set_state(*st)
    s.step(steps)
    o.step(steps, force=np.tile(np.array(BODY), mesh.n_cells))
    rs, us, vs, ws = s.get_state()
    ro, uo, vo, wo = o.state()
    vel = cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo]))
    rho = cases.rel_l2(rs - 1.0, ro - 1.0)
    return vel, rho


@pytest.mark.parametrize("flags", [abi.FLAG_GHOST_FACE_GRAD, abi.FLAGS_REFERENCE_GPU,
                                   abi.FLAGS_REFERENCE_GPU | abi.FLAG_NEUMANN_SIGNED | abi.FLAG_LS_DET_COMPAT])
@pytest.mark.parametrize("lat", [(abi.D3Q15, orc.OR_D3Q15), (abi.D3Q19, orc.OR_D3Q19)])
def test_reference_gpu_semantics_walled_channel(flags, lat):
    """The reference GPU solver's face semantics (tie band, ghost gradient
    along owner -> ghost, the ghost gradient at every point of a wall face,
    LBFS.cu:349-401, 1965-1985) on the walled channel, 100 steps.  With the
    ghost gradient at every wall-face point the scheme is unstable at
    nu = 1/6, dt = 0.25 on this channel (the FP64 oracle's largest speed
    grows 7x per 40 steps and then diverges), so the step is 0.1 here."""
    vel, rho = step_pair(cases.channel_mesh(16, 12, 10), 100, lattice=lat, flags=flags, dt=0.1)
    assert vel < 1e-5 and rho < 1e-4, (vel, rho)


@pytest.mark.parametrize("gamma", [0.5, 0.2])
def test_preconditioned_step_parity(gamma):
    """Preconditioned equilibria and viscous term (gamma != 1,
    LBFS.cu:1041-1354, 1488; gamma = 2 sqrt(3) Mach from the case file)."""
    vel, rho = step_pair(cases.channel_mesh(16, 12, 10), 100, visc=0.02, gamma=gamma)
    assert vel < 1e-5 and rho < 1e-4, (vel, rho)


@pytest.mark.parametrize("grad_type", [abi.GRAD_LS, abi.GRAD_GG])
def test_gradient_type_parity(grad_type):
    """Least-squares and Green-Gauss weights (gradients.cpp:255-265):
    gradients, then 100 steps."""
    m = cases.channel_mesh(16, 12, 10)
    s = abi.Solver(m, abi.make_params(lattice=abi.D3Q19, visc=1.0 / 6.0, dt=0.25, gradient_type=grad_type))
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=1.0 / 6.0, dt=0.25, grad_type=grad_type)
    st = cases.smooth_state(m, seed=11)
    s.set_state(*st)
    o.set_state(*st)
    gs, go = s.gradients().reshape(-1, 4, 3), o.gradients().reshape(-1, 4, 3)
    for v in range(4):
        assert cases.rel_l2(gs[:, v], go[:, v]) < 2e-5, v
    # the gradient types differ from the hybrid default on this mesh
    h = orc.OracleFluid(m, orc.OR_D3Q19, visc=1.0 / 6.0, dt=0.25)
    h.set_state(*st)
    assert cases.rel_l2(go, h.gradients().reshape(-1, 4, 3)) > 1e-3
    vel, rho = step_pair(m, 100, grad_type=grad_type)
    assert vel < 1e-5 and rho < 1e-4, (vel, rho)


def test_face_step_cap_parity():
    """fneq_min > 0: dt_f = min(sqrt(A)/2, 3 nu / (gamma fneq_min))
    (unstructured_mesh.cpp:1435-1443) with the cap active on part of the
    faces of the jittered mesh."""
    m = cases.channel_mesh(16, 12, 10)
    visc = 1.0 / 6.0
    fneq_min = 3 * visc / 0.49
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=visc, dt=0.25, fneq_min=fneq_min)
    capped = np.mean(np.isclose(o.fdt, 0.49))
    assert 0.1 < capped < 0.9, capped
    vel, rho = step_pair(m, 100, fneq_min=fneq_min)
    assert vel < 1e-5 and rho < 1e-4, (vel, rho)


PARAMS = dict(abi.RBC_PARAMS, local_bending_modulus=2e-3, global_bending_modulus=2e-3, viscosity=0.02,
              volume_ratio=0.95)


@pytest.mark.parametrize("mode", [dict(mad_compat=1), dict(sphere_reference=1, curvature_ratio=1.5),
                                  dict(mad_compat=1, sphere_reference=1, curvature_ratio=1.0)])
def test_membrane_reference_modes_coupled_parity(mode):
    """The MAD summed over the first n_nodes springs (Solver_gpu.cu:2355,
    the reference's partial sum) and the spherical spontaneous curvature
    curvature_ratio / 3.265483301 (spring_network.cu:493-495): node forces
    after one step within 1e-4, 40 coupled steps within 1e-5 (velocity,
    node positions) of the oracle, MAD0 within 1e-9."""
    p = dict(PARAMS, **mode)
    mesh = cases.channel_mesh(24, 16, 16, seed=21)
    x0, tri = abi.rbc(4, 8.0, center=(12.0, 8.0, 8.0), rot=abi.rotation((1, 1, 0), 0.4))
    st = cases.smooth_state(mesh, seed=21, amp=0.005)
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=1.0 / 6.0, dt=0.25, body_force=BODY))
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, **p)], ibm_kernel=abi.IBM_COS4)
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=1.0 / 6.0, dt=0.25)
    fl.set_state(*st)
    mo = orc.OracleMembrane(x0, tri, p)
    cp = orc.OracleCoupled(fl, mesh, [mo], abi.IBM_COS4, body_force=BODY)
    g = s.membrane_globals()[0]
    np.testing.assert_allclose(g[5], mo.globals()[5], rtol=1e-9)
    if mode.get("mad_compat"):
        full = orc.OracleMembrane(x0, tri, dict(p, mad_compat=0)).globals()[5]
        assert abs(full - mo.globals()[5]) > 1e-3 * abs(full)
    s.step(1)
    cp.step(1)
    _, _, F = s.membrane_nodes()
    assert cases.rel_l2(F, mo.F.reshape(-1, 3)) < 1e-4
    s.step(39)
    cp.step(39)
    X, _, _ = s.membrane_nodes()
    _, us, vs, ws = s.get_state()
    _, uo, vo, wo = fl.state()
    assert cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo])) < 1e-5
    assert cases.rel_l2(X, mo.X.reshape(-1, 3)) < 1e-5
    assert cases.rel_l2(X - x0, mo.X.reshape(-1, 3) - x0) < 1e-4
