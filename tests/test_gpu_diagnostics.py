"""GPU residual norms, convergence exit and error reporting against the FP64
oracle (residuals.cpp:114-132, Solver_gpu.cu:1032-1046, 1206), through the
C ABI."""
import numpy as np
import pytest

import cases
from lbfs_amd import abi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

VISC, DT = 1.0 / 6.0, 0.25
BODY = (2e-5, 0.0, 0.0)
PARAMS = dict(abi.RBC_PARAMS, local_bending_modulus=1e-3, viscosity=0.02, volume_ratio=0.95)


def fluid_pair(closed=False):
    m = cases.channel_mesh(12, 10, 8)
    flags = abi.FLAG_CLOSED_FACES if closed else 0
    s = abi.Solver(m, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=BODY, flags=flags))
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=VISC, dt=DT, closed=closed)
    st = cases.smooth_state(m, seed=13)
    s.set_state(*st)
    o.set_state(*st)
    return m, s, o


def rms4(d):
    return np.array([d["rho_rms"], d["u_rms"], d["v_rms"], d["w_rms"]])


@pytest.mark.parametrize("closed", [False, True])
def test_step_residual_rms_matches_oracle(closed):
    """The residual RMS of a step (flux residual + body force, squared and
    summed in FP64) equals or_residual_rms of the oracle's step to 1e-4 (the
    bound of the per-cell residual parity), and the pre-step reduction gives
    the same numbers for the plain fluid."""
    m, s, o = fluid_pair(closed)
    pre = rms4(s.diagnostics(ref_rho=1.0, ref_u=0.01))
    d = s.step_residual(ref_rho=1.0, ref_u=0.01)
    o.step(1, force=np.tile(np.array(BODY), m.n_cells))
    ro = o.residual_rms(1.0, 0.01)
    np.testing.assert_allclose(rms4(d), ro, rtol=1e-4)
    np.testing.assert_allclose(pre, ro, rtol=1e-4)
    # the state after the residual step is the state after one plain step
    for a, b in zip(s.get_state()[1:], o.state()[1:]):
        assert cases.rel_l2(a, b) < 1e-5
    assert d["finite"] == 1


def test_step_residual_with_membrane_matches_oracle():
    """With a membrane the step's force density includes the force it spread
    (Solver_gpu.cu:899-915, 1008-1010)."""
    mesh = cases.channel_mesh(24, 16, 16, seed=21)
    x0, tri = abi.rbc(4, 8.0, center=(12.0, 8.0, 8.0), rot=abi.rotation((1, 1, 0), 0.4))
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=BODY))
    st = cases.smooth_state(mesh, seed=21, amp=0.005)
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, **PARAMS)], ibm_kernel=abi.IBM_COS4)
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=VISC, dt=DT)
    fl.set_state(*st)
    mo = orc.OracleMembrane(x0, tri, PARAMS)
    cp = orc.OracleCoupled(fl, mesh, [mo], abi.IBM_COS4, body_force=BODY)
    for k in range(3):
        d = s.step_residual(1.0, 0.01)
        cp.step(1)
        np.testing.assert_allclose(rms4(d), fl.residual_rms(1.0, 0.01), rtol=1e-4, err_msg="step %d" % k)
    # the residual steps continue the same trajectory as plain steps
    s.step(2)
    cp.step(2)
    X, _, _ = s.membrane_nodes()
    assert cases.rel_l2(X - x0, mo.X.reshape(-1, 3) - x0) < 1e-4


def test_run_to_tolerance_stops_at_the_oracle_step():
    """A decaying state without forcing: the loop stops at the first step
    whose momentum RMS is below the tolerance, the same step as the FP64
    oracle (tolerance set between the oracle's RMS of that step and the
    smallest RMS before it, so FP32 differences cannot move the stop)."""
    m = abi.HexMesh(12, 10, 8, periodic=(1, 1, 1), jitter=0.1, seed=7)
    st = cases.smooth_state(m, seed=3)
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=VISC, dt=DT)
    o.set_state(*st)
    err = []
    for _ in range(12):
        o.step(1)
        err.append(max(o.residual_rms(1.0, 0.01)[1:]))
    ks = [k for k in range(2, 12) if err[k] < min(err[:k]) * 0.999]
    assert ks, err
    k = ks[0]
    tol = np.sqrt(err[k] * min(err[:k]))

    def fresh():
        s = abi.Solver(m, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT))
        s.set_state(*st)
        return s

    steps, conv, d = fresh().run_to_tolerance(100, tol, check_every=1)
    assert conv and steps == k + 1, (steps, k + 1)
    np.testing.assert_allclose(max(d["u_rms"], d["v_rms"], d["w_rms"]), err[k], rtol=1e-4)
    # checking every 4th step: stops at the first checked step below it
    first4 = [q for q in (4, 8, 12) if err[q - 1] < tol]
    if first4:
        steps, conv, _ = fresh().run_to_tolerance(100, tol, check_every=4)
        assert conv and steps == first4[0]
    # out of steps: not converged, every step taken, same state as plain steps
    a = fresh()
    steps, conv, _ = a.run_to_tolerance(5, 0.0, check_every=2)
    assert not conv and steps == 5
    b = fresh()
    b.step(5)
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)


def test_fixed_point_range_is_reported():
    """A membrane force far outside the fixed-point range of the spreading
    (shear modulus 1e9, node forces ~1e7) is reported as LBFS_E_NONFINITE
    by the next synchronising call instead of wrapping the integer sums."""
    mesh = cases.channel_mesh(24, 16, 16, seed=21)
    x0, tri = abi.rbc(4, 8.0, center=(12.0, 8.0, 8.0))
    # stretch the cell by 30 % so the springs carry force from the first step
    x1 = (x0 - x0.mean(0)) * 1.3 + x0.mean(0)
    p = dict(PARAMS, shear_modulus=1e9)
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT))
    s.add_membranes([abi.Membrane(x1, tri, nodes_stress_free=x0, **p)], ibm_kernel=abi.IBM_COS4)
    s.step(1)
    with pytest.raises(abi.LbfsError) as e:
        s.sync()
    assert e.value.code == abi.E_NONFINITE
    assert "fixed-point" in str(e.value)
    with pytest.raises(abi.LbfsError):
        s.diagnostics()
