"""GPU parity of the coupled step (IBM interpolation/advection, spring-network
membrane forces, spreading, fluid) against the FP64 oracle, through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

import cases
from lbfs_amd import abi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

VISC, DT = 1.0 / 6.0, 0.25
PARAMS = dict(abi.RBC_PARAMS, local_bending_modulus=1e-3, viscosity=0.02, volume_ratio=0.95)


def coupled_pair(kernel, steps, freq=4, d0=8.0, body=(2e-5, 0.0, 0.0), seed=21, lattice=(abi.D3Q19, orc.OR_D3Q19),
                 params=PARAMS):
    mesh = cases.channel_mesh(24, 16, 16, seed=seed)
    x0, tri = abi.rbc(freq, d0, center=(12.0, 8.0, 8.0), rot=abi.rotation((1, 1, 0), 0.4))
    s = abi.Solver(mesh, abi.make_params(lattice=lattice[0], visc=VISC, dt=DT, body_force=body))
    st = cases.smooth_state(mesh, seed=seed, amp=0.005)
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, **params)], ibm_kernel=kernel)
    fl = orc.OracleFluid(mesh, lattice[1], visc=VISC, dt=DT)
    fl.set_state(*st)
    mo = orc.OracleMembrane(x0, tri, params)
    cp = orc.OracleCoupled(fl, mesh, [mo], kernel, body_force=body)
    s.step(steps)
    cp.step(steps)
    return s, cp, mo, x0


@pytest.mark.parametrize("kernel", [abi.IBM_LINEAR2, abi.IBM_COS4])
def test_one_step_forces_and_spreading(kernel):
    s, cp, mo, x0 = coupled_pair(kernel, 1)
    X, U, F = s.membrane_nodes()
    Fo = mo.F.reshape(-1, 3)
    assert cases.rel_l2(U, mo.U.reshape(-1, 3)) < 1e-5
    assert cases.rel_l2(F, Fo) < 1e-4
    fc = s.cell_force().reshape(-1, 3)
    assert cases.rel_l2(fc, cp.last_force.reshape(-1, 3)) < 1e-4
    g = s.membrane_globals()[0]
    o = mo.globals()
    np.testing.assert_allclose(g[:2], o[:2], rtol=1e-6)
    np.testing.assert_allclose(g[3:5], o[3:5], rtol=1e-9)


@pytest.mark.parametrize("kernel", [abi.IBM_LINEAR2, abi.IBM_COS4])
def test_coupled_step_parity(kernel):
    """60 coupled steps: fluid velocity and node displacement within 1e-5 /
    1e-4 relative L2 of the FP64 oracle."""
    s, cp, mo, x0 = coupled_pair(kernel, 60)
    X, U, F = s.membrane_nodes()
    Xo = mo.X.reshape(-1, 3)
    disp_err = cases.rel_l2(X - x0, Xo - x0)
    pos_err = cases.rel_l2(X, Xo)
    rs, us, vs, ws = s.get_state()
    ro, uo, vo, wo = cp.fluid.state()
    vel_err = cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo]))
    assert pos_err < 1e-5, pos_err
    assert disp_err < 1e-4, disp_err
    assert vel_err < 1e-5, vel_err


def test_internal_viscosity_field_and_step():
    """Viscosity ratio 5 inside the cell: the field built at t = 1
    (Solver_gpu.cu:884-889) and the following steps match the oracle."""
    p = dict(PARAMS, internal_viscosity_ratio=5.0)
    s, cp, mo, x0 = coupled_pair(abi.IBM_COS4, 6, params=p)
    assert cp.visc is not None and np.any(cp.visc > 1.0)
    rs, us, vs, ws = s.get_state()
    ro, uo, vo, wo = cp.fluid.state()
    assert cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo])) < 1e-5
    X, _, _ = s.membrane_nodes()
    assert cases.rel_l2(X - x0, mo.X.reshape(-1, 3) - x0) < 1e-4


def test_coupled_step_bitwise_repeatable():
    outs = []
    for _ in range(2):
        s, _, _, _ = coupled_pair(abi.IBM_COS4, 10)
        X, U, F = s.membrane_nodes()
        outs.append(np.concatenate([X.ravel(), F.ravel(), np.concatenate(s.get_state())]))
        s.close()
    assert np.array_equal(outs[0], outs[1])


def test_spreading_conserves_momentum():
    s, _, _, _ = coupled_pair(abi.IBM_COS4, 3)
    X, U, F = s.membrane_nodes()
    f = s.cell_force().reshape(-1, 3)
    g = s.mesh.geometry()
    vol = g["volume"]
    body = np.array([2e-5, 0.0, 0.0])
    spread = ((f - body) * vol[:, None]).sum(0)
    np.testing.assert_allclose(spread, F.sum(0), atol=1e-6 * np.abs(F).sum())


def test_many_membranes_match_single():
    """Two membranes with different resolutions in one launch behave like
    each run alone when they are far apart (per-membrane parameters and
    offsets)."""
    mesh = abi.HexMesh(40, 16, 16, periodic=(1, 0, 0), jitter=0.1, seed=5)
    xa, ta = abi.rbc(3, 7.0, center=(10.0, 8.0, 8.0))
    xb, tb = abi.rbc(4, 8.0, center=(30.0, 8.0, 8.0))
    pa = dict(PARAMS, shear_modulus=4e-3)
    pb = dict(PARAMS, shear_modulus=6e-3)
    both = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT))
    both.add_membranes([abi.Membrane(xa, ta, **pa), abi.Membrane(xb, tb, **pb)])
    both.step(5)
    Xb, _, Fb = both.membrane_nodes()
    one = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT))
    one.add_membranes([abi.Membrane(xa, ta, **pa)])
    one.step(5)
    X1, _, F1 = one.membrane_nodes()
    na = len(xa)
    assert cases.rel_l2(Xb[:na], X1) < 1e-6
    assert cases.rel_l2(Fb[:na], F1) < 1e-3


def test_graph_replays_match_eager_launches():
    """Steps replayed from the captured hipGraph equal eager launches bitwise,
    replay after replay, with eager work and synchronisation in between."""
    def build(graph):
        mesh = cases.channel_mesh(24, 16, 16, seed=3)
        x0, tri = abi.rbc(4, 8.0, center=(12.0, 8.0, 8.0))
        s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
        s.set_state(*cases.smooth_state(mesh, seed=4, amp=0.005))
        s.add_membranes([abi.Membrane(x0, tri, **PARAMS)])
        s.set_graph(graph)
        return s

    g, e = build(True), build(False)
    for rep in range(4):
        g.step(6)       # three graph replays
        e.step(6)
        g.step(1)       # an eager step between replays (odd count)
        e.step(1)
        g.sync()
        sg, se = np.concatenate(g.get_state()), np.concatenate(e.get_state())
        assert np.array_equal(sg, se), rep
        assert np.array_equal(g.membrane_nodes()[0], e.membrane_nodes()[0]), rep


def test_membrane_energies_match_oracle():
    """Output-step energies after 20 coupled steps (Solver_gpu.cu:1060-1139)
    against the oracle's, per term (FP32 state, FP64 sums)."""
    p = dict(PARAMS, local_bending_modulus=2e-3, global_bending_modulus=5e-4)
    s, cp, mo, x0 = coupled_pair(abi.IBM_COS4, 20, params=p)
    e = s.membrane_energies()[0]
    eo = mo.energies()
    assert np.all(eo[[0, 1, 6]] > 0)
    np.testing.assert_allclose(e, eo, rtol=2e-3, atol=2e-4 * eo[6])


def test_springs_from_mz_file_parity(tmp_path):
    """Springs read from an MZ file (mesh_read, lagrangian_object.cpp:495-587)
    in another valid orientation (c, d, a, b) drive the same coupled step as
    the oracle given those springs."""
    x0, tri = abi.rbc(4, 8.0, center=(12.0, 8.0, 8.0))
    sp = abi.springs(len(x0), tri)[:, [2, 3, 0, 1]]
    path = tmp_path / "cell.mz"
    abi.write_membrane_mesh(path, x0, tri, fmt=abi.MESH_MZ, springs_=sp)
    xr, tr, spr = abi.read_membrane_mesh(path, fmt=abi.MESH_MZ)
    mesh = cases.channel_mesh(24, 16, 16, seed=8)
    s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
    st = cases.smooth_state(mesh, seed=8, amp=0.005)
    s.set_state(*st)
    s.add_membranes([abi.Membrane(xr, tr, springs=spr, **PARAMS)])
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=VISC, dt=DT)
    fl.set_state(*st)
    mo = orc.OracleMembrane(xr, tr, PARAMS, springs=spr)
    cp = orc.OracleCoupled(fl, mesh, [mo], abi.IBM_COS4, body_force=(2e-5, 0.0, 0.0))
    s.step(20)
    cp.step(20)
    X, _, F = s.membrane_nodes()
    assert cases.rel_l2(X - xr, mo.X.reshape(-1, 3) - xr) < 1e-4
    assert cases.rel_l2(F, mo.F.reshape(-1, 3)) < 1e-3


def test_named_stages_equal_whole_steps():
    """Running the named stages of each step in order (the reference's kernel
    groups) gives bitwise the same state as lbfs_step; the C++ host mirror's
    check program does the same through gpu_solver::General_Purpose_Solver_mk_i
    and the reference-named group wrappers."""
    def build():
        mesh = cases.channel_mesh(24, 16, 16, seed=9)
        x0, tri = abi.rbc(4, 8.0, center=(12.0, 8.0, 8.0))
        s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
        s.set_state(*cases.smooth_state(mesh, seed=9, amp=0.005))
        s.add_membranes([abi.Membrane(x0, tri, **PARAMS)])
        return s

    a, b = build(), build()
    a.step(7)
    names = b.stage_names()
    assert names[0] == "ibm_interp" and names[-2:] == ["flux_update", "ibm_apply"]
    for _ in range(7):
        for n in b.stage_names():
            b.stage_run(n)
        b.stage_end_step()
    assert np.array_equal(np.concatenate(a.get_state()), np.concatenate(b.get_state()))
    assert np.array_equal(a.membrane_nodes()[0], b.membrane_nodes()[0])
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lbfs-ibm-ade-sp_amd", "host",
                       "gpu_solver_check")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "groups equal whole steps" in r.stdout


def test_reference_signature_entry_equals_direct_abi_run(tmp_path):
    """General_Purpose_Solver_mk_i with the reference's parameter list
    (host/reference_entry.hpp), called on stand-ins of the reference's classes
    with a polyMesh case, a cell read from its Osay mesh file and the
    pressure-gradient body force, gives bit for bit the run made through the C
    ABI directly, calls the output hook every output_step steps, and stops
    after one step when the tolerance is above the first residual."""
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lbfs-ibm-ade-sp_amd", "host",
                       "reference_entry_check")
    r = subprocess.run([exe, str(tmp_path / "case")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "direct ABI run equal" in r.stdout and "stop after 1 step" in r.stdout


@pytest.mark.parametrize("per_box_force", [False, True])
def test_ensemble_equals_members_run_alone(per_box_force):
    """Independent boxes stepped together as one mesh (one launch per stage
    over all boxes, membranes in their own box) give bit for bit what each box
    gives run alone, with one body force for all boxes or one per box."""
    members = [abi.HexMesh(16, 10, 10, periodic=(1, 0, 0), jitter=0.1, seed=31),
               abi.HexMesh(20, 12, 10, periodic=(1, 0, 0), jitter=0.1, seed=32),
               abi.HexMesh(12, 10, 12, periodic=(0, 0, 0), jitter=0.1, seed=33)]
    cells = {0: [(8.0, 5.0, 5.0)], 1: [(6.0, 6.0, 5.0), (15.0, 6.0, 5.0)], 2: []}
    body = (2e-5, 0.0, 0.0)
    states = [cases.smooth_state(m, seed=50 + k, amp=0.004) for k, m in enumerate(members)]

    def mems(k, block):
        out = []
        for c in cells[k]:
            x, t = abi.rbc(3, 5.0, center=c)
            out.append(abi.Membrane(x, t, block=block, **PARAMS))
        return out

    forces = [(2e-5 * (k + 1), 1e-6 * k, -1e-6) if per_box_force else body for k in range(len(members))]
    ens = abi.Ensemble(members, body_forces=forces if per_box_force else None)
    s = abi.Solver(ens, abi.make_params(visc=VISC, dt=DT, body_force=(0.0, 0.0, 0.0) if per_box_force else body))
    s.set_state(*[np.concatenate([st[v] for st in states]) for v in range(4)])
    s.add_membranes([mm for k in range(3) for mm in mems(k, k)])
    s.step(12)
    ens_state = s.get_state()
    ens_nodes = s.membrane_nodes()[0]
    node0 = 0
    for k, m in enumerate(members):
        r = abi.Solver(m, abi.make_params(visc=VISC, dt=DT, body_force=forces[k]))
        r.set_state(*states[k])
        if cells[k]:
            r.add_membranes(mems(k, 0))
        r.step(12)
        for v in range(4):
            assert np.array_equal(ens.split(ens_state[v])[k], r.get_state()[v]), (k, v)
        if cells[k]:
            X = r.membrane_nodes()[0]
            assert np.array_equal(ens_nodes[node0: node0 + len(X)], X), k
            node0 += len(X)
        r.close()


def test_lagrange4_coupled_parity():
    """4-point Lagrange IBM kernel on a stretched tensor-product channel
    (mesh_type 4 of the reference): 30 coupled steps against the oracle."""
    from test_oracle_membrane import _stretched_mesh, lagrange_grid_arrays
    m = _stretched_mesh(24, 16, 16)
    planes, lines = lagrange_grid_arrays(m)
    x0, tri = abi.rbc(4, 7.0, center=(12.0, 8.0, 8.0))
    body = (2e-5, 0.0, 0.0)
    s = abi.Solver(m, abi.make_params(visc=VISC, dt=DT, body_force=body))
    st = cases.smooth_state(m, seed=61, amp=0.004)
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, **PARAMS)], ibm_kernel=abi.IBM_LAGRANGE4)
    fl = orc.OracleFluid(m, orc.OR_D3Q19, visc=VISC, dt=DT)
    fl.set_state(*st)
    mo = orc.OracleMembrane(x0, tri, PARAMS)
    cp = orc.OracleCoupled(fl, m, [mo], abi.IBM_LAGRANGE4, body_force=body, planes=planes, lines=lines)
    s.step(30)
    cp.step(30)
    X, U, F = s.membrane_nodes()
    assert cases.rel_l2(X - x0, mo.X.reshape(-1, 3) - x0) < 1e-4
    _, us, vs, ws = s.get_state()
    _, uo, vo, wo = cp.fluid.state()
    assert cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo])) < 1e-5


@pytest.mark.parametrize("kernel,tile", [(abi.IBM_COS4, (0, 0, 0)), (abi.IBM_LINEAR2, (0, 0, 0)),
                                         (abi.IBM_COS4, (3, 5, 2)), (abi.IBM_COS4, (16, 4, 4))])
def test_spreading_paths_are_bitwise_equal(kernel, tile):
    """Per-node spreading into integer cell accumulators in global memory,
    per-tile gathering (each fluid tile gathering the nodes listed around it,
    sums in LDS) and per-bin scattering (LDS windows added to the
    accumulators) add the same 64-bit fixed-point terms, so whole coupled
    steps agree bit for bit, node forces included (the bin scattering gathers
    them from the edge records itself); also with tiles narrower than the
    stencil (3x5x2) and ragged tiles (16 cells along a 24-cell axis)."""
    outs = []
    for mode in (0, 1, 2):
        mesh = cases.channel_mesh(24, 16, 16, seed=12)
        s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0), tile=tile))
        s.set_state(*cases.smooth_state(mesh, seed=12, amp=0.005))
        s.set_ibm_binned(mode)
        xa, ta = abi.rbc(4, 8.0, center=(8.0, 8.0, 8.0))
        xb, tb = abi.rbc(3, 6.0, center=(18.0, 8.0, 8.0))
        s.add_membranes([abi.Membrane(xa, ta, **PARAMS), abi.Membrane(xb, tb, **PARAMS)], ibm_kernel=kernel)
        # the bin scattering gathers the node forces itself (no node-force pass)
        assert ("mem_nodes" in s.stage_names()) == (mode != 2)
        s.step(9)
        X, U, F = s.membrane_nodes()
        outs.append((np.concatenate(s.get_state()), s.cell_force(), X, F))
        s.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_full_tiles_overflow_to_the_node_list():
    """A 2562-node cell 6 cells across puts more than a thousand nodes in one
    8x8x4 fluid tile (more than its 512 slots) and hundreds in 4x4x4 bins
    (128 slots): the rest go to the overflow list (gathered by every tile, or
    spread node by node), and the step still equals the per-node spreading
    bit for bit."""
    outs = []
    for mode in (0, 1, 2):
        mesh = cases.channel_mesh(24, 16, 16, seed=13)
        s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
        s.set_state(*cases.smooth_state(mesh, seed=13, amp=0.005))
        s.set_ibm_binned(mode)
        x, t = abi.rbc(16, 6.0, center=(12.0, 8.0, 8.0))
        s.add_membranes([abi.Membrane(x, t, **PARAMS)], ibm_kernel=abi.IBM_COS4)
        s.step(3)
        X, U, F = s.membrane_nodes()
        outs.append((np.concatenate(s.get_state()), s.cell_force(), X, F))
        s.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_step_schedules_are_bitwise_equal():
    """The step schedules (stream assignments of the same kernels) give the
    same step bit for bit."""
    outs = []
    for first in (0, 1, 2):
        mesh = cases.channel_mesh(24, 16, 16, seed=14)
        s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
        s.set_state(*cases.smooth_state(mesh, seed=14, amp=0.005))
        s.set_step_schedule(first)
        xa, ta = abi.rbc(4, 8.0, center=(8.0, 8.0, 8.0))
        xb, tb = abi.rbc(3, 6.0, center=(18.0, 8.0, 8.0))
        s.add_membranes([abi.Membrane(xa, ta, **PARAMS), abi.Membrane(xb, tb, **PARAMS)], ibm_kernel=abi.IBM_COS4)
        s.step(12)
        X, U, F = s.membrane_nodes()
        outs.append((np.concatenate(s.get_state()), X, F, s.membrane_globals()))
        s.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_membrane_zone_file_from_the_solver(tmp_path):
    """The zone file of a membrane after some coupled steps holds the solver's
    node positions, forces and velocities in the reference's columns."""
    s, cp, mo, x0 = coupled_pair(abi.IBM_COS4, 5)
    X, U, F = s.membrane_nodes()
    area, curv = s.write_membrane_dat(tmp_path / "cell.dat", time=5, precision=17)
    rows = (tmp_path / "cell.dat").read_text().split("\n")[4:4 + len(X)]
    vals = np.array([[float(v) for v in r.split("\t")[:12]] for r in rows])
    assert np.array_equal(vals[:, 0:3], X) and np.array_equal(vals[:, 3:6], F) and np.array_equal(vals[:, 6:9], U)
    assert np.array_equal(vals[:, 10], np.arange(len(X)))
    np.testing.assert_allclose(area.sum(), s.membrane_globals()[0][1], rtol=1e-6)
    s.close()


@pytest.mark.parametrize("nx", [22, 9])
def test_tile_spreading_across_a_ragged_periodic_axis(nx):
    """A cell straddling the periodic boundary of an axis whose length is not a
    multiple of the 8-cell fluid tiles or 4-cell bins (22 cells; 9 cells,
    where the last tile is one cell wide, narrower than the stencil, so nodes
    listed two tiles away reach it, and a 7-cell bin window covers cells
    twice): the tile gathering and the bin scattering reach the wrapped
    stencils and equal the per-node spreading bit for bit."""
    outs = []
    for mode in (0, 1, 2):
        mesh = cases.channel_mesh(nx, 16, 16, seed=16)
        s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
        s.set_state(*cases.smooth_state(mesh, seed=16, amp=0.005))
        s.set_ibm_binned(mode)
        x, t = abi.rbc(4, 8.0, center=(1.0, 8.0, 8.0))
        s.add_membranes([abi.Membrane(x, t, **PARAMS)], ibm_kernel=abi.IBM_COS4)
        s.step(6)
        X, U, F = s.membrane_nodes()
        outs.append((np.concatenate(s.get_state()), s.cell_force(), X, F))
        s.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_fused_fluid_pass_coupled_step_equals_two_passes():
    """The coupled step with the fused fluid pass (the default) equals the
    one with the gradient and face passes bit for bit, graph and eager."""
    outs = []
    for fused in (True, False):
        for graph in (True, False):
            mesh = cases.channel_mesh(24, 16, 16, seed=14)
            s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0), tile=(8, 8, 8),
                                                 block_threads=512))
            s.set_fused(fused)
            s.set_graph(graph)
            s.set_state(*cases.smooth_state(mesh, seed=14, amp=0.005))
            x, t = abi.rbc(4, 8.0, center=(12.0, 8.0, 8.0))
            s.add_membranes([abi.Membrane(x, t, **PARAMS)], ibm_kernel=abi.IBM_COS4)
            s.step(8)
            X, U, F = s.membrane_nodes()
            outs.append((np.concatenate(s.get_state()), s.cell_force(), X, F))
            s.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_pow_springs_coupled_parity_and_energy():
    """WLC-POW springs (spring_network.cu:221-223, 1655-1682): node forces
    after one step within 1e-4 and 40 coupled steps within 1e-5 (velocity,
    nodes) of the FP64 oracle; the spring energy within 1e-5."""
    p = dict(PARAMS, spring_law=abi.SPRING_WLC_POW)
    mesh = cases.channel_mesh(24, 16, 16, seed=22)
    xs, tri = abi.rbc(4, 8.0, center=(12.0, 8.0, 8.0), rot=abi.rotation((1, 1, 0), 0.4))
    c = np.array([12.0, 8.0, 8.0])
    x0 = c + (xs - c) * np.array([1.08, 0.96, 0.96])   # stretched: the springs carry force
    st = cases.smooth_state(mesh, seed=22, amp=0.005)
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, nodes_stress_free=xs, **p)], ibm_kernel=abi.IBM_COS4)
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=VISC, dt=DT)
    fl.set_state(*st)
    mo = orc.OracleMembrane(x0, tri, p, nodes_stress_free=xs)
    cp = orc.OracleCoupled(fl, mesh, [mo], abi.IBM_COS4, body_force=(2e-5, 0.0, 0.0))
    s.step(1)
    cp.step(1)
    X, U, F = s.membrane_nodes()
    assert cases.rel_l2(F, mo.F.reshape(-1, 3)) < 1e-4
    eg = s.membrane_energies()[0]
    eo = mo.energies()
    assert abs(eg[0] - eo[0]) < 1e-5 * abs(eo[0]), (eg[0], eo[0])
    s.step(39)
    cp.step(39)
    X, _, _ = s.membrane_nodes()
    _, us, vs, ws = s.get_state()
    _, uo, vo, wo = fl.state()
    assert cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo])) < 1e-5
    assert cases.rel_l2(X, mo.X.reshape(-1, 3)) < 1e-5
    assert cases.rel_l2(X - x0, mo.X.reshape(-1, 3) - x0) < 1e-4


@pytest.mark.parametrize("layout", ["three cells", "far from the origin", "ensemble"])
def test_internal_viscosity_search_equals_every_triangle(layout):
    """The internal-viscosity field from the binned search (triangles by the
    y-z bins their projections overlap, nodes by 3-D bins within 2 delta_h)
    equals the field from every triangle of each membrane bit for bit: three
    cells of different ratios and resolutions close together (one with a
    bounding box overlapping another's), two cells near (250, 120, 120)
    where the crossing test in absolute coordinates rounds by more than the
    bin margin, or two boxes of an ensemble."""
    outs = []
    for search in (1, 0):
        if layout != "ensemble":
            if layout == "three cells":
                mesh = cases.channel_mesh(28, 18, 16, seed=41)
                shapes = [abi.rbc(4, 8.0, center=(8.0, 9.0, 8.0), rot=abi.rotation((0, 1, 1), 0.5)),
                          abi.rbc(3, 6.0, center=(15.5, 8.0, 8.5), rot=abi.rotation((1, 0, 1), 1.1)),
                          abi.rbc(4, 7.0, center=(22.0, 9.5, 7.5))]
                ratios = (5.0, 3.0, 0.5)
            else:
                mesh = abi.HexMesh(26, 18, 18, origin=(238.0, 111.0, 111.0), periodic=(1, 0, 0), jitter=0.1, seed=45)
                shapes = [abi.rbc(4, 8.0, center=(246.0, 120.0, 119.5), rot=abi.rotation((0, 1, 1), 0.7)),
                          abi.rbc(4, 7.0, center=(256.5, 119.0, 120.5), rot=abi.rotation((1, 0, 1), 0.4))]
                ratios = (5.0, 3.0)
            mems = [abi.Membrane(x, t, **dict(PARAMS, internal_viscosity_ratio=r)) for (x, t), r in zip(shapes, ratios)]
            s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
            s.set_state(*cases.smooth_state(mesh, seed=41, amp=0.005))
        else:
            boxes = [cases.channel_mesh(16, 14, 14, seed=42), cases.channel_mesh(18, 16, 14, seed=43)]
            mesh = abi.Ensemble(boxes)
            mems = []
            for k, (c, r) in enumerate((((8.0, 7.0, 7.0), 5.0), ((9.0, 8.0, 7.0), 2.0))):
                x, t = abi.rbc(4, 7.0, center=c, rot=abi.rotation((1, 1, 0), 0.3 * k))
                mems.append(abi.Membrane(x, t, block=k, **dict(PARAMS, internal_viscosity_ratio=r)))
            s = abi.Solver(mesh, abi.make_params(visc=VISC, dt=DT, body_force=(2e-5, 0.0, 0.0)))
            parts = [cases.smooth_state(m, seed=44 + k, amp=0.005) for k, m in enumerate(boxes)]
            s.set_state(*[np.concatenate([p[v] for p in parts]) for v in range(4)])
        s.add_membranes(mems, ibm_kernel=abi.IBM_COS4)
        s.set_visc_search(search)
        s.step(5)  # the field is built at t = 1
        X, _, F = s.membrane_nodes()
        outs.append((np.concatenate(s.get_state()), X, F))
        s.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
