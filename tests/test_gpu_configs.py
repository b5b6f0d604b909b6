"""The benchmark workloads at their full sizes on the HIP path against the
FP64 oracle, through the C ABI: the 128x64x64 channel with its 2562-node cell,
three cell resolutions in one launch, the 256-cell suspension on the
256x128x128 channel, sampled instances of the 64-box ensemble (each run
alone through the single-instance oracle, as the reference holds one object,
Solver_gpu.cu:2342-2349), and the adjoint on the channel against central
differences of the oracle.  Short windows keep the oracle to seconds; the
longer runs check what does not need it (finite, repeatable, graph replay
equal to eager launches)."""
import numpy as np
import pytest

import cases
from lbfs_amd import abi, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

VISC, DT = workloads.CHANNEL_VISC, workloads.CHANNEL_DT


def oracle_membranes(mems):
    return [orc.OracleMembrane(m.nodes, m.tri, m.params, nodes_stress_free=m.sf, nodes_curvature=m.cv) for m in mems]


def coupled_run(mesh, params, st, mems, steps, kernel=abi.IBM_COS4, graph=True):
    s = abi.Solver(mesh, params)
    if not graph:
        s.set_graph(False)
    s.set_state(*st)
    s.add_membranes(mems, ibm_kernel=kernel)
    s.step(steps)
    return s


def oracle_run(mesh, params, st, mems, steps, kernel=abi.IBM_COS4, body_force=None):
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=params.visc, dt=params.dt)
    fl.set_state(*st)
    mos = oracle_membranes(mems)
    bf = tuple(params.body_force) if body_force is None else tuple(body_force)
    cp = orc.OracleCoupled(fl, mesh, mos, kernel, body_force=bf)
    cp.step(steps)
    return fl, mos


def errors(s, fl, mos, mems):
    """velocity, node-position and node-displacement relative L2 errors, and
    the largest node-position error in units of the FP32 spacing of the
    positions (stored relative to each membrane's centroid)"""
    _, us, vs, ws = s.get_state()
    _, uo, vo, wo = fl.state()
    vel = cases.rel_l2(np.concatenate([us, vs, ws]), np.concatenate([uo, vo, wo]))
    X, _, _ = s.membrane_nodes()
    Xo = np.concatenate([m.X.reshape(-1, 3) for m in mos])
    X0 = np.concatenate([m.nodes.reshape(-1, 3) for m in mems])
    R = max(np.abs(m.nodes.reshape(-1, 3) - m.nodes.reshape(-1, 3).mean(0)).max() for m in mems)
    ulps = np.abs(X - Xo).max() / float(np.spacing(np.float32(R)))
    return vel, cases.rel_l2(X, Xo), cases.rel_l2(X - X0, Xo - X0), ulps


def test_channel_rbc_full_size_parity():
    """The headline workload (128x64x64 jittered channel, D3Q19, body-force
    duct flow, one Evans-Fung cell of 2562 nodes / 5120 triangles, 4-point
    cosine IBM) from a perturbed state, 20 steps: velocity and node
    positions within 1e-5 relative L2 of FP64, displacements within 1e-4."""
    mesh, params, _ = workloads.channel()
    mems = workloads.single_rbc(mesh)
    st = cases.smooth_state(mesh, seed=2, amp=0.005)
    s = coupled_run(mesh, params, st, mems, 20)
    fl, mos = oracle_run(mesh, params, st, mems, 20)
    vel, pos, disp, ulps = errors(s, fl, mos, mems)
    print("channel 128x64x64 + 2562-node cell, 20 steps: velocity %.2e, nodes %.2e, displacement %.2e (%.1f ulp)"
          % (vel, pos, disp, ulps))
    assert vel < 1e-5 and pos < 1e-5 and disp < 1e-4, (vel, pos, disp)


def test_channel_rbc_full_size_long_run_properties():
    """600 steps from rest: finite, graph replay equal to eager launches bit
    for bit, and a second run equal to the first."""
    mesh, params, st = workloads.channel()
    mems = workloads.single_rbc(mesh)
    outs = []
    for graph in (True, False, True):
        s = coupled_run(mesh, params, st, mems, 600, graph=graph)
        X, U, F = s.membrane_nodes()
        outs.append(np.concatenate([np.concatenate(s.get_state()), X.ravel(), F.ravel()]))
        d = s.diagnostics()
        assert d["finite"] == 1 and d["max_speed"] > 0.0
        s.close()
    assert np.all(np.isfinite(outs[0]))
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


def test_three_resolutions_in_one_launch_parity():
    """642-, 1212- and 2562-node cells in one set of launches against the
    oracle given the same list, 15 steps from a perturbed state."""
    mesh = abi.HexMesh(96, 40, 40, periodic=(1, 0, 0), jitter=0.1, seed=17)
    f = workloads.duct_body_force(0.01, 40, 40, VISC)
    params = abi.make_params(lattice=abi.D3Q19, visc=VISC, dt=DT, body_force=(f, 0.0, 0.0))
    mems = []
    for freq, d0, cx, ang in ((8, 12.0, 16.0, 0.3), (11, 14.0, 48.0, 1.1), (16, 16.0, 80.0, 2.0)):
        x, tri = abi.rbc(freq, d0, center=(cx, 20.0, 20.0), rot=abi.rotation((1, 1, 0), ang))
        mems.append(abi.Membrane(x, tri, local_bending_modulus=1e-3, viscosity=0.02))
    assert [len(m.nodes) // 3 for m in mems] == [642, 1212, 2562]
    st = cases.smooth_state(mesh, seed=17, amp=0.005)
    s = coupled_run(mesh, params, st, mems, 15)
    fl, mos = oracle_run(mesh, params, st, mems, 15)
    vel, pos, disp, ulps = errors(s, fl, mos, mems)
    print("three resolutions, 15 steps: velocity %.2e, nodes %.2e, displacement %.2e (%.1f ulp)"
          % (vel, pos, disp, ulps))
    assert vel < 1e-5 and pos < 1e-5 and disp < 1e-4, (vel, pos, disp)
    # each membrane on its own
    X, _, _ = s.membrane_nodes()
    k = 0
    for m, mo in zip(mems, mos):
        n = len(m.nodes) // 3
        x0 = m.nodes.reshape(-1, 3)
        assert cases.rel_l2(X[k:k + n] - x0, mo.X.reshape(-1, 3) - x0) < 1e-4
        k += n


def test_suspension_full_size_parity_and_replay():
    """The 256-cell suspension (256x128x128, 380052 nodes of three
    resolutions at 20 % hematocrit) from a perturbed state: 2 steps against
    the oracle holding the same 256 membranes (velocity and node positions
    within 1e-5, node positions within 6 FP32 spacings), then 10 steps with
    graph replay equal to eager launches bit for bit."""
    mesh, params, _, mems, _ = workloads.suspension()
    st = cases.smooth_state(mesh, seed=3, amp=0.005)
    s = coupled_run(mesh, params, st, mems, 2)
    fl, mos = oracle_run(mesh, params, st, mems, 2)
    vel, pos, disp, ulps = errors(s, fl, mos, mems)
    print("suspension 256 cells, 2 steps: velocity %.2e, nodes %.2e, displacement %.2e (%.1f ulp)"
          % (vel, pos, disp, ulps))
    # two steps move the nodes by ~1e-3: their displacement is resolved to
    # the FP32 spacing of the positions, so it is bounded in those units
    assert vel < 1e-5 and pos < 1e-5 and ulps <= 2 + 4, (vel, pos, disp, ulps)
    del fl, mos
    s.step(8)
    e = coupled_run(mesh, params, st, mems, 10, graph=False)
    for a, b in zip(s.get_state(), e.get_state()):
        assert np.array_equal(a, b)
    Xa, _, Fa = s.membrane_nodes()
    Xb, _, Fb = e.membrane_nodes()
    assert np.array_equal(Xa, Xb) and np.array_equal(Fa, Fb) and np.all(np.isfinite(Xa))
    s.close()
    e.close()


def test_ensemble_full_size_sampled_instances():
    """The 64-instance ensemble (12.1 M cells, 204 cells of 642 nodes, each
    box driven by its own duct body force) stepped together for 10 steps from
    a perturbed state; three instances (the
    smallest, one with the most cells, the largest box) each run alone
    through the single-instance oracle agree within 1e-5 (velocity, nodes)
    and their node positions within 14 FP32 spacings (one rounding of the
    advection per step and the interpolated velocities' error)."""
    ens, params, _, mems = workloads.ensemble()
    members = ens.members
    states = [cases.smooth_state(m, seed=5000 + i, amp=0.005) for i, m in enumerate(members)]
    st = [np.concatenate([x[v] for x in states]) for v in range(4)]
    s = coupled_run(ens, params, st, mems, 10)
    _, us, vs, ws = s.get_state()
    X, _, _ = s.membrane_nodes()
    starts = np.cumsum([0] + [len(m.nodes) // 3 for m in mems])
    by_block = {}
    for k, m in enumerate(mems):
        by_block.setdefault(m.block, []).append(k)
    sizes = [m.n_cells for m in members]
    counts = [len(by_block.get(i, [])) for i in range(len(members))]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(counts)), int(np.argmax(sizes))})
    for i in picks:
        mi = members[i]
        mm = [mems[k] for k in by_block[i]]
        fl, mos = oracle_run(mi, params, states[i], mm, 10, body_force=ens.body_forces[3 * i: 3 * i + 3])
        _, uo, vo, wo = fl.state()
        vel = cases.rel_l2(np.concatenate([ens.split(us)[i], ens.split(vs)[i], ens.split(ws)[i]]),
                           np.concatenate([uo, vo, wo]))
        Xi = np.concatenate([X[starts[k]:starts[k + 1]] for k in by_block[i]])
        Xo = np.concatenate([mo.X.reshape(-1, 3) for mo in mos])
        X0 = np.concatenate([m.nodes.reshape(-1, 3) for m in mm])
        pos, disp = cases.rel_l2(Xi, Xo), cases.rel_l2(Xi - X0, Xo - X0)
        R = max(np.abs(m.nodes.reshape(-1, 3) - m.nodes.reshape(-1, 3).mean(0)).max() for m in mm)
        ulps = np.abs(Xi - Xo).max() / float(np.spacing(np.float32(R)))
        print("ensemble instance %d (%s, %d cells): velocity %.2e, nodes %.2e, displacement %.2e (%.1f ulp)"
              % (i, "x".join(str(v) for v in (mi.nx, mi.ny, mi.nz)), len(mm), vel, pos, disp, ulps))
        # ten steps: displacement resolved to a few FP32 spacings of the positions
        assert vel < 1e-5 and pos < 1e-5 and ulps <= 10 + 4, (i, vel, pos, disp, ulps)
    s.close()


def test_adjoint_channel_full_size_matches_oracle_differences():
    """Adjoint on the headline channel (128x64x64 + 2562-node cell started off
    its stress-free shape), 6 steps with a checkpoint every 2: gradients of
    L = sum |X - X_sf|^2 with respect to the shear modulus, both bending
    moduli and the body force against central differences of the FP64
    oracle, within 1e-4."""
    mesh, params, _ = workloads.channel()
    c = mesh.origin + 0.5 * mesh.lengths
    xs, tri = abi.rbc(16, 20.0, center=c)
    x0 = c + (xs - c) * np.array([1.04, 0.98, 0.98])
    p = dict(abi.RBC_PARAMS, local_bending_modulus=2e-3, global_bending_modulus=1e-3, viscosity=0.02,
             volume_ratio=0.95)
    st = cases.smooth_state(mesh, seed=4, amp=0.005)
    fx = params.body_force[0]
    steps, every = 6, 2
    s = abi.Solver(mesh, params)
    s.set_state(*st)
    s.add_membranes([abi.Membrane(x0, tri, nodes_stress_free=xs, nodes_curvature=xs, **p)], ibm_kernel=abi.IBM_COS4)
    s.adjoint_forward(steps, every)
    X, _, _ = s.membrane_nodes()
    loss = float(np.sum((X - xs) ** 2))
    g = s.adjoint_backward(lam_nodes=2.0 * (X - xs))
    s.close()

    def oracle_loss(shear, lb, gb, f):
        q = dict(p, shear_modulus=shear, local_bending_modulus=lb, global_bending_modulus=gb)
        fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=VISC, dt=DT)
        fl.set_state(*st)
        mo = orc.OracleMembrane(x0, tri, q, nodes_stress_free=xs, nodes_curvature=xs)
        orc.OracleCoupled(fl, mesh, [mo], abi.IBM_COS4, body_force=(f, 0.0, 0.0)).step(steps)
        return float(np.sum((mo.X.reshape(-1, 3) - xs) ** 2))

    base = dict(shear=p["shear_modulus"], lb=p["local_bending_modulus"], gb=p["global_bending_modulus"], f=fx)
    lo = oracle_loss(**base)
    assert abs(loss - lo) / lo < 1e-5, (loss, lo)
    ad = {"shear": g["moduli"][0, 0], "lb": g["moduli"][0, 1], "gb": g["moduli"][0, 2], "f": g["body_force"][0]}
    for k, v0 in base.items():
        h = 1e-3 * abs(v0)
        up, dn = dict(base), dict(base)
        up[k], dn[k] = v0 + h, v0 - h
        fd = (oracle_loss(**up) - oracle_loss(**dn)) / (2 * h)
        e = abs(ad[k] - fd) / abs(fd)
        print("channel adjoint %-5s %.9e vs FD %.9e (rel %.2e)" % (k, ad[k], fd, e))
        assert e < 1e-4, (k, ad[k], fd, e)
