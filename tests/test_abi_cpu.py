"""CPU-side checks of the C ABI library: it loads, exports every symbol the
headers declare, and its host-only mesh helpers behave (no device work)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import cases
from lbfs_amd import abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(lbfs_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_declared_symbols():
    lib = abi.load()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_device_fails_cleanly(has_gpu):
    if has_gpu:
        pytest.skip("a device is present")
    m = cases.channel_mesh(4, 4, 4)
    with pytest.raises(abi.LbfsError, match="no HIP device"):
        abi.Solver(m, abi.make_params())


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 0), (1, 1, 1)])
def test_hex_counts_and_topology(periodic):
    nx, ny, nz = 5, 4, 3
    m = abi.HexMesh(nx, ny, nz, periodic=periodic, jitter=0.1, seed=9)
    # every cell has exactly six faces
    cnt = np.zeros(m.n_cells, dtype=int)
    np.add.at(cnt, m.owner, 1)
    np.add.at(cnt, m.neighbour[: m.n_internal], 1)
    assert np.all(cnt == 6)
    # periodic wraps carry the box length, other faces no shift
    sh = m.face_shift.reshape(-1, 3)
    for a in range(3):
        n_wrap = np.sum(np.abs(sh[:, a] - m.lengths[a]) < 1e-12)
        assert n_wrap == (m.n_cells // [nx, ny, nz][a] if periodic[a] else 0)
    # boundary patches only on non-periodic axes
    pats = set(m.bc_patch[: m.n_bc].tolist()) if m.n_bc else set()
    for a in range(3):
        assert ((2 * a) in pats) == (not periodic[a])


def test_jitter_is_seeded_and_bounded():
    a = abi.HexMesh(6, 5, 4, jitter=0.1, seed=3)
    b = abi.HexMesh(6, 5, 4, jitter=0.1, seed=3)
    c = abi.HexMesh(6, 5, 4, jitter=0.1, seed=4)
    assert np.array_equal(a.points, b.points)
    assert not np.array_equal(a.points, c.points)
    p = a.points.reshape(-1, 3)
    ijk = np.stack(np.meshgrid(np.arange(7), np.arange(6), np.arange(5), indexing="ij"), -1)
    ijk = ijk.transpose(2, 1, 0, 3).reshape(-1, 3)
    assert np.abs(p - ijk).max() <= 0.1 + 1e-6
    # FP32-representable coordinates
    assert np.array_equal(p.astype(np.float32).astype(np.float64), p)


def test_polymesh_round_trip(tmp_path):
    """A generated mesh written as ASCII polyMesh files and read back is the
    same mesh: points to the last bit, faces, owner/neighbour, patches and
    periodic translations, and the same product geometry."""
    m = abi.HexMesh(6, 5, 4, periodic=(1, 0, 1), jitter=0.15, seed=9)
    m.write_polymesh(tmp_path)
    r = abi.HexMesh.read_polymesh(tmp_path, 6, 5, 4)
    assert (r.n_points, r.n_cells, r.n_faces, r.n_internal) == (m.n_points, m.n_cells, m.n_faces, m.n_internal)
    assert np.array_equal(r.points, m.points)
    assert np.array_equal(r.face_vertices, m.face_vertices)
    assert np.array_equal(r.owner, m.owner)
    assert np.array_equal(r.neighbour[: m.n_internal], m.neighbour[: m.n_internal])
    assert np.array_equal(r.face_shift, m.face_shift)
    assert np.array_equal(r.patch, m.patch)
    assert r.patch_names == abi.PATCH_NAMES and list(r.periodic) == [1, 0, 1]
    ga, gb = m.geometry(), r.geometry()
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k


def test_polymesh_rejects_non_quad(tmp_path):
    m = abi.HexMesh(2, 2, 2, periodic=(0, 0, 0), jitter=0.0)
    m.write_polymesh(tmp_path)
    txt = (tmp_path / "faces").read_text().replace("4(", "3(", 1)
    (tmp_path / "faces").write_text(txt)
    with pytest.raises(abi.LbfsError):
        abi.HexMesh.read_polymesh(tmp_path)


@pytest.mark.parametrize("fmt", [abi.MESH_OSAY, abi.MESH_MZ])
def test_membrane_mesh_files_round_trip(tmp_path, fmt):
    """Osay and MZ membrane files (lagrangian_object.cpp:238-302, 495-587):
    written and read back with the object centre added to every node; MZ
    carries the springs, which equal the ones built from the triangles."""
    x, t = abi.rbc(3, 7.0)
    path = tmp_path / "cell.dat"
    abi.write_membrane_mesh(path, x, t, fmt=fmt)
    c = (1.5, -2.0, 3.25)
    xr, tr, sp = abi.read_membrane_mesh(path, fmt=fmt, centre=c)
    np.testing.assert_array_equal(tr, t.reshape(-1, 3))
    np.testing.assert_allclose(xr, x.reshape(-1, 3) + np.array(c), rtol=0, atol=1e-12)
    if fmt == abi.MESH_MZ:
        np.testing.assert_array_equal(sp, abi.springs(len(x), t))
        # degree lines: every node lists its spring neighbours
        lines = path.read_text().split("\n")
        n, s_, tt = (int(v) for v in lines[0].split())
        deg = [int(l.split()[0]) for l in lines[1 + n + tt + s_: 1 + n + tt + s_ + n]]
        assert sum(deg) == 2 * s_
    else:
        assert sp is None


def test_membrane_mesh_rejects_bad_ids(tmp_path):
    path = tmp_path / "bad.dat"
    path.write_text("4 2\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n1 2 3\n1 3 9\n")
    with pytest.raises(abi.LbfsError):
        abi.read_membrane_mesh(path)


def test_case_xml_reader():
    """The reference's XML schema (preprocessor.cpp:87-857) with its derived
    values (global_variables.cpp:20-68): nu = mach_no ref_length / Re,
    preconditioned gamma = 2 sqrt(3) mach_no, BC velocities scaled by
    mach_no / scale, signed body force along pressure_direction, object
    fields carried over from the previous object, sphere_reference written
    into the periodic flag."""
    c = abi.read_case(os.path.join(REPO, "tests", "golden", "channel_case.xml"))
    assert abs(c.visc - 0.02 * 100 / 12) < 1e-15
    assert abs(c.gamma - 2 * np.sqrt(3) * 0.02) < 1e-15
    assert c.time_stepping == abi.TS_MIN and c.gradient_type == abi.GRAD_LS
    assert c.import_file == b"polyMesh" and c.output_step == 50 and c.mesh_type == 3
    assert list(c.body_force) == [2e-5, 0.0, 0.0]
    assert c.n_patches == 4
    p = {c.patches[k].name.decode(): c.patches[k] for k in range(4)}
    assert (p["ymin"].rho_type, p["ymin"].vel_type) == (abi.BC_NEUMANN, abi.BC_DIRICHLET)
    assert abs(p["ymax"].values[1] - 0.5 * 0.02) < 1e-17
    assert p["zmin"].vel_type == abi.BC_PARABOLIC and abs(p["zmin"].values[1] - 0.02) < 1e-17
    assert p["zmax"].rho_type == abi.BC_DIRICHLET and abs(p["zmax"].values[0] - 1.01) < 1e-15
    assert c.n_objects == 2
    a, b = c.objects[0], c.objects[1]
    assert a.name == b"cell_a" and a.type == 2 and a.num_nodes == 642 and a.mesh_file == b"cell.dat"
    assert a.sphere_reference == 1 and a.periodic == 0            # <sphere_reference>false</...> -> periodic
    assert abs(a.volume_ratio - 0.95) < 1e-15 and abs(a.curvature_ratio - 1.5) < 1e-15
    assert b.name == b"cell_b" and abs(b.shear_modulus - 0.006) < 1e-15
    assert b.num_nodes == 642 and abs(b.volume_modulus - 0.1) < 1e-15 and b.mesh_file == b"cell.dat"
    assert list(b.centre) == [4.0, 6.0, 5.0]
    params = abi.case_params(c)
    assert params.time_stepping == abi.TS_MIN and abs(params.channel_diameter - 12.0) < 1e-15
    m = abi.HexMesh(16, 12, 10)
    abi.apply_case_bcs(m, c)
    sel = m.bc_patch[: m.n_bc] == 4
    assert np.all(m.bc_vel_type[: m.n_bc][sel] == abi.BC_PARABOLIC)


def test_case_xml_errors(tmp_path):
    bad = tmp_path / "bad.xml"
    bad.write_text("<case><global_variables><mach_no>0.1</mach_no></global_variables>")
    with pytest.raises(abi.LbfsError):
        abi.read_case(bad)


def test_cpp_host_mirror_builds_and_fails_cleanly_without_device():
    """The C++ host mirror (host/gpu_solver.hpp) compiles against the C ABI;
    without a HIP device its check program reports LBFS_E_NO_DEVICE (exit 3)
    instead of crashing."""
    import subprocess
    exe = os.path.join(REPO, "lbfs-ibm-ade-sp_amd", "host", "gpu_solver_check")
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode in (0, 3), r.stdout + r.stderr


def test_reference_signature_entry_compiles_and_writes_its_case(tmp_path):
    """The reference-signature step entry (host/reference_entry.hpp) compiles
    against stand-in classes carrying the reference's member names; without a
    HIP device its check program writes the polyMesh case and the cell's mesh
    file, then stops at lbfs_create with LBFS_E_NO_DEVICE (exit 3)."""
    import subprocess
    exe = os.path.join(REPO, "lbfs-ibm-ade-sp_amd", "host", "reference_entry_check")
    assert os.path.exists(exe)
    case = tmp_path / "case"
    r = subprocess.run([exe, str(case)], capture_output=True, text=True, timeout=120)
    assert r.returncode in (0, 3), r.stdout + r.stderr
    for f in ("points", "faces", "owner", "neighbour", "boundary", "lbfsFaceShift", "cell.dat"):
        assert (case / f).exists(), f
    m = abi.HexMesh.read_polymesh(case, 16, 12, 12)
    assert m.n_cells == 16 * 12 * 12 and m.patch_names == ("bottom", "top", "back", "front")
    assert list(m.periodic) == [1, 0, 0]


def test_state_files_in_the_reference_format(tmp_path):
    """Restart files: one "i ,value" line per cell in caller order
    (Solution::output); 17 digits round-trip exactly, the default 6 digits
    equal the stream format the reference writes; files written in the
    reference's own layout read back."""
    rng = np.random.default_rng(4)
    n = 37
    st = [1.0 + 1e-3 * rng.standard_normal(n)] + [1e-2 * rng.standard_normal(n) for _ in range(3)]
    abi.write_state_txt(tmp_path, *st)
    back = abi.read_state_txt(tmp_path, n)
    for a, b in zip(st, back):
        assert np.array_equal(a, b)
    lines = (tmp_path / "u.txt").read_text().splitlines()
    assert len(lines) == n and lines[5].startswith("5 ,")
    abi.write_state_txt(tmp_path, *st, precision=0)
    lines = (tmp_path / "rho.txt").read_text().splitlines()
    assert lines[3] == "3 ,%g" % st[0][3]
    back6 = abi.read_state_txt(tmp_path, n)
    np.testing.assert_allclose(back6[0], st[0], rtol=1e-5)
    # a file as the reference writes it (stream default format, index, " ,", value)
    d = tmp_path / "ref"
    d.mkdir()
    for name, vals in zip(("rho", "u", "v", "w"), ([1.00001, 0.99998], [0.01, -2e-05], [0, 1e-07], [3.5, -1])):
        (d / (name + ".txt")).write_text("".join("%d ,%s\n" % (i, repr(x) if isinstance(x, float) else x)
                                                 for i, x in enumerate(vals)))
    r, u, v, w = abi.read_state_txt(d, 2)
    assert list(r) == [1.00001, 0.99998] and list(u) == [0.01, -2e-05] and list(v) == [0.0, 1e-07] and list(w) == [3.5, -1.0]
    with pytest.raises(abi.LbfsError):
        abi.read_state_txt(d, 3)  # too few lines


def test_membrane_dat_file(tmp_path):
    """Membrane zone file: the reference's header and columns, 1-based
    triangle ids one per line; nodal areas sum to the surface area, and the
    curvature sum 0.5 theta L over all springs of a sphere approaches the
    integral of twice the mean curvature, 8 pi R (convex positive)."""
    R = 5.0
    x, tri = abi.icosphere(12)
    x = R * x.reshape(-1, 3)
    n, t = len(x), len(tri)
    F = np.tile([1.0, 2.0, 3.0], (n, 1))
    U = np.tile([0.1, 0.2, 0.3], (n, 1))
    area, curv = abi.write_membrane_dat(tmp_path / "cell.dat", x, F, U, tri, time=40)
    text = (tmp_path / "cell.dat").read_text().split("\n")
    assert text[0].startswith('VARIABLES = "nodx" "nody" "nodz" "i_force"') and text[0].endswith('"curvature"')
    assert text[1] == "ZONE Nodes= %d Elements= %d F=FEPOINT ET=TRIANGLE " % (n, t // 3 if tri.ndim == 1 else t)
    assert text[2] == "SOLUTIONTIME = 40 "
    row = text[4].split("\t")
    assert len(row) == 13 and row[-1] == "" and row[10] == "0"
    assert float(row[3]) == 1.0 and float(row[8]) == 0.3
    tri_flat = np.asarray(tri).reshape(-1)
    ids = [int(s.strip()) for s in text[4 + n: 4 + n + len(tri_flat)]]
    assert ids == list(tri_flat + 1)
    a_tri = 0.0
    for a, b, c in tri_flat.reshape(-1, 3):
        a_tri += 0.5 * np.linalg.norm(np.cross(x[b] - x[a], x[c] - x[a]))
    assert abs(area.sum() - a_tri) < 1e-9 * a_tri
    assert abs(curv.sum() - 8 * np.pi * R) < 0.01 * 8 * np.pi * R
    assert np.all(curv > 0)
