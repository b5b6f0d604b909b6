"""GPU parity of the fluid path: HIP kernels (FP32, through the C ABI) against
the FP64 oracle on the same seeded meshes and states.

Tolerances (stated per test) are relative L2 norms; the north-star bar for
the full step is 1e-5 on velocity after N steps.
"""
import numpy as np
import pytest

import cases
from lbfs_amd import abi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

VISC, DT = 1.0 / 6.0, 0.25


def pair(mesh, lat_a, lat_o, flags=0, visc=VISC, dt=DT, body=(0.0, 0.0, 0.0), grad_type=0):
    closed = bool(flags & abi.FLAG_CLOSED_FACES)
    det = bool(flags & abi.FLAG_LS_DET_COMPAT)
    oflags = flags & (abi.FLAG_UPWIND_BAND | abi.FLAG_NEUMANN_SIGNED | abi.FLAG_GHOST_DIR_GPU)
    s = abi.Solver(mesh, abi.make_params(lattice=lat_a, visc=visc, dt=dt, flags=flags, body_force=body,
                                         gradient_type=grad_type))
    o = orc.OracleFluid(mesh, lat_o, visc=visc, dt=dt, flags=oflags, closed=closed, det_compat=det,
                        grad_type=grad_type)
    return s, o


def test_gradients_parity():
    m = cases.channel_mesh(16, 12, 10)
    s, o = pair(m, abi.D3Q19, orc.OR_D3Q19)
    st = cases.smooth_state(m, seed=11)
    s.set_state(*st)
    o.set_state(*st)
    gs, go = s.gradients(), o.gradients()
    g_s, g_o = gs.reshape(-1, 4, 3), go.reshape(-1, 4, 3)
    for v in range(4):
        assert cases.rel_l2(g_s[:, v], g_o[:, v]) < 2e-5, v


@pytest.mark.parametrize("lat_a,lat_o", cases.LAT_PAIRS)
@pytest.mark.parametrize("flags", [0, abi.FLAG_UPWIND_BAND | abi.FLAG_GHOST_DIR_GPU, abi.FLAG_NEUMANN_SIGNED])
def test_face_flux_parity(lat_a, lat_o, flags):
    m = cases.channel_mesh(12, 10, 8)
    s, o = pair(m, lat_a, lat_o, flags=flags)
    st = cases.smooth_state(m, seed=12)
    s.set_state(*st)
    o.set_state(*st)
    fs, fo = s.face_flux().reshape(-1, 4), o.face_flux().reshape(-1, 4)
    # mass flux compared to its own scale; momentum flux carries rho0/3 n A
    assert cases.rel_l2(fs[:, 0], fo[:, 0]) < 1e-5
    for v in range(1, 4):
        d = fs[:, v] - fo[:, v]
        assert np.sqrt(np.mean(d * d)) < 2e-7, (v, np.abs(d).max())


@pytest.mark.parametrize("closed", [False, True])
def test_residual_parity(closed):
    m = cases.channel_mesh(12, 10, 8)
    s, o = pair(m, abi.D3Q19, orc.OR_D3Q19, flags=abi.FLAG_CLOSED_FACES if closed else 0)
    st = cases.smooth_state(m, seed=13)
    s.set_state(*st)
    o.set_state(*st)
    rs, ro = s.residual().reshape(-1, 4), o.residual().reshape(-1, 4)
    for v in range(4):
        assert cases.rel_l2(rs[:, v], ro[:, v]) < 1e-4, v


@pytest.mark.parametrize("lat_a,lat_o", cases.LAT_PAIRS)
@pytest.mark.parametrize("flags", [0, abi.FLAG_CLOSED_FACES | abi.FLAG_UPWIND_BAND, abi.FLAG_LS_DET_COMPAT])
def test_channel_step_parity(lat_a, lat_o, flags):
    """Body-force channel flow from a perturbed state, 100 steps: relative L2
    of (u, v, w) <= 1e-5 and of rho - rho0 <= 1e-4 against FP64."""
    m = cases.channel_mesh(16, 12, 10)
    body = (2e-5, 0.0, 0.0)
    s, o = pair(m, lat_a, lat_o, flags=flags, body=body)
    st = cases.smooth_state(m, seed=14)
    s.set_state(*st)
    o.set_state(*st)
    n = m.n_cells
    force = np.tile(np.array(body), n)
    steps = 100
    s.step(steps)
    o.step(steps, force=force)
    rs, us, vs, ws = s.get_state()
    ro, uo, vo, wo = o.state()
    vel_err = cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo]))
    rho_err = cases.rel_l2(rs - 1.0, ro - 1.0)
    assert vel_err < 1e-5, vel_err
    assert rho_err < 1e-4, rho_err


def test_taylor_green_2d_parity_and_decay():
    """Taylor-Green vortex: D2Q9 on a jittered 2D periodic mesh, 400 steps,
    against the FP64 oracle and the analytic decay."""
    N = 32
    m = abi.HexMesh(N, N, 1, periodic=(1, 1, 1), jitter=0.05, seed=1)
    nu = 0.01 * N / 100.0
    s, o = pair(m, abi.D2Q9, orc.OR_D2Q9, visc=nu)
    st = cases.tgv_state(m, 0.01)
    s.set_state(*st)
    o.set_state(*st)
    s.step(400)
    o.step(400)
    rs, us, vs, _ = s.get_state()
    ro, uo, vo, _ = o.state()
    assert cases.rel_l2(np.concatenate([us, vs]), np.concatenate([uo, vo])) < 1e-5
    k = 2 * np.pi / N
    decay = np.exp(-2 * nu * k * k * DT * 400)
    u0 = st[1]
    assert abs(np.dot(us, u0) / np.dot(u0, u0) / decay - 1.0) < 0.02


def test_taylor_green_64_for_1000_steps():
    """The Taylor-Green verification case at its full size: 64x64 D2Q9 on a
    jittered periodic mesh (jitter 0.05, seed 1), Re 100, U0 0.01, 1000 steps:
    velocity within 1e-5 relative L2 of the FP64 oracle and the amplitude
    within 0.2 % of exp(-2 nu k^2 t)."""
    N = 64
    m = abi.HexMesh(N, N, 1, periodic=(1, 1, 1), jitter=0.05, seed=1)
    nu = 0.01 * N / 100.0
    s, o = pair(m, abi.D2Q9, orc.OR_D2Q9, visc=nu)
    st = cases.tgv_state(m, 0.01)
    s.set_state(*st)
    o.set_state(*st)
    s.step(1000)
    o.step(1000)
    _, us, vs, _ = s.get_state()
    _, uo, vo, _ = o.state()
    err = cases.rel_l2(np.concatenate([us, vs]), np.concatenate([uo, vo]))
    k = 2 * np.pi / N
    decay = np.exp(-2 * nu * k * k * DT * 1000)
    q0 = np.concatenate([st[1], st[2]])
    amp = np.dot(np.concatenate([us, vs]), q0) / np.dot(q0, q0)
    print("Taylor-Green 64x64, 1000 steps: rel L2 vs oracle %.2e, amplitude %.6f vs analytic %.6f" % (err, amp, decay))
    assert err < 1e-5, err
    assert abs(amp / decay - 1.0) < 2e-3
    s.close()


def test_bitwise_repeatable():
    m = cases.channel_mesh(16, 12, 10)
    outs = []
    for _ in range(2):
        s = abi.Solver(m, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=(1e-5, 0, 0)))
        s.set_state(*cases.smooth_state(m, seed=15))
        s.step(20)
        outs.append(np.concatenate(s.get_state()))
        s.close()
    assert np.array_equal(outs[0], outs[1])


def test_mass_conserved_periodic():
    m = abi.HexMesh(16, 16, 16, periodic=(1, 1, 1), jitter=0.1, seed=21)
    s = abi.Solver(m, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT))
    s.set_state(*cases.smooth_state(m, seed=16))
    d0 = s.diagnostics()
    s.step(200)
    d1 = s.diagnostics()
    assert d1["finite"] == 1
    assert abs(d1["total_mass"] - d0["total_mass"]) / d0["total_mass"] < 1e-6


def test_invalid_arguments_return_codes():
    m = cases.channel_mesh(4, 4, 4)
    with pytest.raises(abi.LbfsError):
        abi.Solver(m, abi.make_params(lattice=7))
    s = abi.Solver(m, abi.make_params())
    bad = np.full(m.n_cells, np.nan)
    with pytest.raises(abi.LbfsError, match="non-finite"):
        s.set_state(bad, bad, bad, bad)


def test_profile_and_empty_boundaries_parity():
    """Parabolic duct inlet (9) and shear (10) velocity BCs on a jittered duct,
    and a one-layer mesh with empty front/back patches: 60 steps, relative L2
    of velocity <= 1e-5 against FP64, face fluxes of empty faces exactly 0."""
    H = 8.0
    m = abi.HexMesh(12, 8, 8, lengths=(12.0, H, H), origin=(0.0, -H / 2, -H / 2), periodic=(0, 0, 0),
                    jitter=0.1, seed=6)
    m.set_patch_bc(0, abi.BC_NEUMANN, abi.BC_PARABOLIC, (1.0, 0.01, 0.0, 0.0))
    m.set_patch_bc(1, abi.BC_DIRICHLET, abi.BC_SHEAR, (1.0, 0.001, 0.0, 0.0))
    s = abi.Solver(m, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, channel_diameter=H))
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=VISC, dt=DT, channel_diameter=H)
    st = cases.smooth_state(m, seed=15)
    s.set_state(*st)
    o.set_state(*st)
    np.testing.assert_allclose(s.face_flux().reshape(-1, 4), o.face_flux().reshape(-1, 4), atol=2e-6)
    s.step(60)
    o.step(60)
    _, us, vs, ws = s.get_state()
    _, uo, vo, wo = o.state()
    assert cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo])) < 1e-5

    m2 = abi.HexMesh(16, 16, 1, periodic=(1, 1, 0), jitter=0.05, seed=1)
    for patch in (4, 5):
        m2.set_patch_bc(patch, abi.BC_EMPTY, abi.BC_EMPTY, (1.0, 0.0, 0.0, 0.0))
    s2 = abi.Solver(m2, abi.make_params(lattice=abi.D3Q15, visc=0.0024, dt=DT))
    o2 = orc.OracleFluid(m2, orc.OR_D3Q15, visc=0.0024, dt=DT)
    st2 = cases.tgv_state(m2, 0.01)
    s2.set_state(*st2)
    o2.set_state(*st2)
    fl = s2.face_flux().reshape(-1, 4)
    assert np.all(fl[m2.n_internal:] == 0.0)
    s2.step(60)
    o2.step(60)
    _, us, vs, ws = s2.get_state()
    _, uo, vo, wo = o2.state()
    assert cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo])) < 1e-5


def test_unsupported_bc_types_rejected():
    m = cases.channel_mesh(6, 5, 4)
    m.set_patch_bc(2, abi.BC_NEUMANN, 4, (1.0, 0.0, 0.0, 0.0))   # parabolic-N: not on the GPU path
    with pytest.raises(abi.LbfsError):
        abi.Solver(m, abi.make_params())
    m = cases.channel_mesh(6, 5, 4)
    m.set_patch_bc(2, abi.BC_NEUMANN, abi.BC_PARABOLIC, (1.0, 0.01, 0.0, 0.0))
    with pytest.raises(abi.LbfsError):
        abi.Solver(m, abi.make_params())          # parabolic without channel_diameter


@pytest.mark.parametrize("ts", [abi.TS_LOCAL, abi.TS_MIN])
def test_local_and_min_time_stepping_parity(ts):
    """gpu_time_stepping 1 (local) and 3 (min): per-cell CFL steps
    (LBFS.cu:2871-2936) and their minimum; 80 steps against FP64, relative L2
    of velocity <= 1e-5."""
    m = cases.channel_mesh(16, 12, 10)
    body = (2e-5, 0.0, 0.0)
    cfl, U = 0.5, 0.02
    s = abi.Solver(m, abi.make_params(lattice=abi.D3Q19, visc=0.05, dt=cfl, body_force=body, time_stepping=ts,
                                      max_velocity=U))
    o = orc.OracleFluid(m, orc.OR_D3Q19, visc=0.05, dt=cfl, time_stepping=ts, max_velocity=U)
    st = cases.smooth_state(m, seed=16)
    s.set_state(*st)
    o.set_state(*st)
    s.step(80)
    o.step(80, force=np.tile(np.array(body), m.n_cells))
    _, us, vs, ws = s.get_state()
    _, uo, vo, wo = o.state()
    assert cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo])) < 1e-5


def test_tile_shape_and_workgroup_size_do_not_change_results():
    """The cell tiling and the flux/update workgroup size (the autotuner's
    knobs) change neither arithmetic nor summation order: 30 steps are
    bitwise equal across shapes, including ragged tiles."""
    m = cases.channel_mesh(20, 12, 10)
    body = (2e-5, 0.0, 0.0)
    st = cases.smooth_state(m, seed=17)
    out = []
    for tile, nt in (((8, 8, 4), 256), ((8, 8, 8), 512), ((16, 4, 4), 256), ((3, 5, 7), 512), ((1, 1, 1), 256)):
        s = abi.Solver(m, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=body, tile=tile,
                                          block_threads=nt))
        s.set_state(*st)
        s.step(30)
        out.append(np.concatenate(s.get_state()))
        s.close()
    for o in out[1:]:
        assert np.array_equal(o, out[0])


def test_autotuner_picks_a_valid_candidate():
    from lbfs_amd import autotune
    m = cases.channel_mesh(32, 16, 16)
    p = abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT)
    best, fused, table = autotune.tune(m, p, steps=4, warmup=2)
    assert len(table) >= 3
    assert (tuple(best.tile), best.block_threads, fused) in [(t, nt, f) for t, nt, f, _ in table]
    assert best.block_threads in (256, 512)
    assert any(f for _, _, f, _ in table) and not all(f for _, _, f, _ in table)


def test_restart_files_continue_the_run(tmp_path):
    """A run written to the reference's restart files (17 digits) and read
    into a new solver continues bit for bit like the uninterrupted run."""
    mesh = cases.channel_mesh(16, 12, 12, seed=8)
    params = abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0))
    a = abi.Solver(mesh, params)
    a.set_state(*cases.smooth_state(mesh, seed=8, amp=0.005))
    a.step(10)
    a.write_restart(tmp_path)
    b = abi.Solver(mesh, params)
    b.read_restart(tmp_path)
    assert np.array_equal(np.concatenate(a.get_state()), np.concatenate(b.get_state()))
    a.step(10)
    b.step(10)
    assert np.array_equal(np.concatenate(a.get_state()), np.concatenate(b.get_state()))
    a.close()
    b.close()


@pytest.mark.parametrize("lat_a,lat_o", cases.LAT_PAIRS)
@pytest.mark.parametrize("flags", [0, abi.FLAGS_REFERENCE_GPU | abi.FLAG_NEUMANN_SIGNED])
@pytest.mark.parametrize("tile,threads", [((8, 8, 8), 512), ((8, 8, 4), 256), ((4, 4, 16), 256)])
def test_fused_fluid_pass_equals_two_passes(lat_a, lat_o, flags, tile, threads):
    """Whole steps with the fused fluid pass (each tile's gradients in LDS,
    the cells across its border recomputed) equal the gradient pass + face
    pass bit for bit, with local time steps too."""
    if lat_a == abi.D2Q9:
        m = abi.HexMesh(20, 12, 1, periodic=(1, 0, 1), jitter=0.1, seed=2)
    else:
        m = cases.channel_mesh(20, 12, 10)
    st = cases.smooth_state(m, seed=15)
    outs = []
    for fused in (True, False):
        for ts in (0, abi.TS_LOCAL):
            p = abi.make_params(lattice=lat_a, visc=VISC, dt=DT if ts == 0 else 0.4, flags=flags,
                                body_force=(2e-5, 0.0, 0.0), time_stepping=ts, max_velocity=0.02,
                                tile=tile, block_threads=threads)
            s = abi.Solver(m, p)
            s.set_fused(fused)
            s.set_state(*st)
            s.step(7)
            outs.append(np.concatenate(s.get_state()))
            s.close()
    assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[1], outs[3])
    assert np.all(np.isfinite(outs[0])) and not np.array_equal(outs[0], outs[1])
