"""ctypes binding of the C ABI declared in include/lbfs_abi.h.

Loads lbfs_amd/liblbfs_amd.so (built in-tree by lbfs_amd.build).  There is no
fallback: if the library is missing or a call fails, an exception is raised.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbfs_amd.so")

D3Q15, D3Q19, D2Q9 = 0, 1, 2
E_INVALID_ARG, E_HIP, E_NONFINITE, E_UNSUPPORTED, E_NO_DEVICE, E_STATE = -1, -2, -3, -4, -5, -6
GRAD_HYBRID, GRAD_LS, GRAD_GG = 0, 1, 2
BC_DIRICHLET, BC_NEUMANN, BC_PARABOLIC, BC_SHEAR, BC_EMPTY = 1, 2, 9, 10, 11
TS_LOCAL, TS_CONSTANT, TS_MIN = 1, 2, 3
MESH_OSAY, MESH_MZ = 0, 1
FLAG_UPWIND_BAND, FLAG_NEUMANN_SIGNED, FLAG_GHOST_DIR_GPU, FLAG_LS_DET_COMPAT, FLAG_CLOSED_FACES = 1, 2, 4, 16, 32
FLAG_GHOST_FACE_GRAD = 64
FLAG_HOST_GEOMETRY = 128  # set-up: FP64 geometry and least-squares weights on the host (the same bits)
# the reference GPU solver's face semantics (tie band, ghost direction, ghost gradient at every point)
FLAGS_REFERENCE_GPU = FLAG_UPWIND_BAND | FLAG_GHOST_DIR_GPU | FLAG_GHOST_FACE_GRAD
IBM_LINEAR2, IBM_COS4, IBM_LAGRANGE4 = 0, 1, 2
OPT_GRAPH, OPT_IBM_BINNED, OPT_STEP_SCHEDULE, OPT_FUSED_FLUID, OPT_VISC_SEARCH = 1, 2, 3, 4, 5

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_points", C.c_int), ("points", _dp),
        ("n_cells", C.c_int), ("n_faces", C.c_int), ("n_internal", C.c_int),
        ("face_vertices", _ip), ("owner", _ip), ("neighbour", _ip), ("face_shift", _dp),
        ("bc_rho_type", _ip), ("bc_vel_type", _ip), ("bc_values", _dp),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("origin", C.c_double * 3), ("spacing", C.c_double),
        ("n_blocks", C.c_int), ("block_dims", _ip), ("block_origin", _dp), ("block_spacing", _dp),
        ("block_body_force", _dp),
    ]


class Params(C.Structure):
    _fields_ = [
        ("lattice", C.c_int), ("gradient_type", C.c_int), ("flags", C.c_int),
        ("visc", C.c_double), ("gamma", C.c_double), ("fneq_min", C.c_double), ("dt", C.c_double),
        ("rho0", C.c_double), ("body_force", C.c_double * 3), ("channel_diameter", C.c_double),
        ("time_stepping", C.c_int), ("max_velocity", C.c_double), ("tile", C.c_int * 3), ("block_threads", C.c_int),
    ]


class Diag(C.Structure):
    _fields_ = [
        ("rho_rms", C.c_double), ("u_rms", C.c_double), ("v_rms", C.c_double), ("w_rms", C.c_double),
        ("total_mass", C.c_double), ("max_speed", C.c_double), ("finite", C.c_int),
    ]


class MembraneDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int), ("n_tri", C.c_int), ("nodes", _dp), ("nodes_stress_free", _dp),
        ("nodes_curvature", _dp), ("tri", _ip),
        ("shear_modulus", C.c_double), ("wlc_ratio", C.c_double), ("volume_modulus", C.c_double),
        ("local_area_modulus", C.c_double), ("global_area_modulus", C.c_double),
        ("global_bending_modulus", C.c_double), ("local_bending_modulus", C.c_double),
        ("viscosity", C.c_double), ("curvature_ratio", C.c_double), ("volume_ratio", C.c_double),
        ("sphere_reference", C.c_int), ("mad_compat", C.c_int), ("internal_viscosity_ratio", C.c_double),
        ("n_springs", C.c_int), ("springs", _ip), ("block", C.c_int), ("spring_law", C.c_int),
    ]


class CasePatch(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("rho_type", C.c_int), ("vel_type", C.c_int), ("values", C.c_double * 4)]


class CaseObject(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64), ("type", C.c_int), ("num_nodes", C.c_int), ("radius", C.c_double),
        ("centre", C.c_double * 3), ("mesh_file", C.c_char * 256), ("stress_free_mesh", C.c_char * 256),
        ("curvature_mesh", C.c_char * 256), ("shear_modulus", C.c_double), ("wlc_ratio", C.c_double),
        ("volume_modulus", C.c_double), ("local_area_modulus", C.c_double), ("global_area_modulus", C.c_double),
        ("global_bending_modulus", C.c_double), ("local_bending_modulus", C.c_double), ("viscosity", C.c_double),
        ("curvature_ratio", C.c_double), ("volume_ratio", C.c_double), ("internal_viscosity_ratio", C.c_double),
        ("sphere_reference", C.c_int), ("periodic", C.c_int), ("mz_importer", C.c_int),
    ]


class Case(C.Structure):
    _fields_ = [
        ("import_file", C.c_char * 512), ("import_format", C.c_char * 32), ("mesh_type", C.c_int),
        ("gpu_solver", C.c_int), ("gpu_device", C.c_int), ("output_step", C.c_int), ("time_stepping", C.c_int),
        ("gradient_type", C.c_int), ("max_velocity", C.c_double), ("reynolds_number", C.c_double),
        ("visc", C.c_double), ("gamma", C.c_double), ("fneq_min", C.c_double), ("time_marching_step", C.c_double),
        ("simulation_length", C.c_double), ("scale", C.c_double), ("domain", C.c_double * 3),
        ("spacing", C.c_double * 3), ("origin", C.c_double * 3), ("streaming_dt", C.c_double),
        ("ref_length", C.c_double), ("average_rho", C.c_double), ("initial_velocity", C.c_double * 3),
        ("body_force", C.c_double * 3), ("n_patches", C.c_int), ("patches", CasePatch * 32),
        ("n_objects", C.c_int), ("objects", CaseObject * 64), ("error", C.c_char * 256),
    ]


MEMBRANE_PARAM_NAMES = ("shear_modulus", "wlc_ratio", "volume_modulus", "local_area_modulus", "global_area_modulus",
                        "global_bending_modulus", "local_bending_modulus", "viscosity", "curvature_ratio",
                        "volume_ratio", "sphere_reference", "mad_compat", "internal_viscosity_ratio", "spring_law")
SPRING_WLC, SPRING_WLC_POW = 0, 1

# lattice-unit membrane parameters for a 20-cell RBC in nu = 1/6 flow
RBC_PARAMS = dict(shear_modulus=5e-3, wlc_ratio=2.2, volume_modulus=0.1, local_area_modulus=5e-3,
                  global_area_modulus=0.1, global_bending_modulus=1e-3, local_bending_modulus=0.0, viscosity=0.0,
                  curvature_ratio=1.0, volume_ratio=1.0, sphere_reference=0, mad_compat=0,
                  internal_viscosity_ratio=1.0, spring_law=0)

_lib = None


def load():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblbfs_amd.so not built: run __graft_entry__.build() (expected at %s)" % LIB_PATH)
        lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "lbfs_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(vp)]),
            "lbfs_destroy": (C.c_int, [vp]),
            "lbfs_last_error": (C.c_char_p, [vp]),
            "lbfs_sync": (C.c_int, [vp]),
            "lbfs_set_state": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
            "lbfs_get_state": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
            "lbfs_step": (C.c_int, [vp, C.c_int]),
            "lbfs_stage_gradients": (C.c_int, [vp, _dp]),
            "lbfs_stage_face_flux": (C.c_int, [vp, _dp]),
            "lbfs_stage_residual": (C.c_int, [vp, _dp]),
            "lbfs_reduce_diagnostics": (C.c_int, [vp, C.c_double, C.c_double, C.POINTER(Diag)]),
            "lbfs_step_residual": (C.c_int, [vp, C.c_double, C.c_double, C.POINTER(Diag)]),
            "lbfs_run_to_tolerance": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, _ip, _ip,
                                                C.POINTER(Diag)]),
            "lbfs_get_cell_force": (C.c_int, [vp, _dp]),
            "lbfs_set_option": (C.c_int, [vp, C.c_int, C.c_int]),
            "lbfs_add_membranes": (C.c_int, [vp, C.POINTER(MembraneDesc), C.c_int, C.c_int]),
            "lbfs_get_membrane_nodes": (C.c_int, [vp, _dp, _dp, _dp]),
            "lbfs_get_membrane_globals": (C.c_int, [vp, _dp]),
            "lbfs_get_membrane_energies": (C.c_int, [vp, _dp]),
            "lbfs_get_membrane_energies_noalloc": (C.c_int, [vp, _dp]),
            "lbfs_mesh_icosphere": (C.c_int, [C.c_int, _dp, _ip]),
            "lbfs_mesh_evans_fung": (C.c_int, [C.c_int, _dp, C.c_double, _dp]),
            "lbfs_mesh_springs": (C.c_int, [C.c_int, C.c_int, _ip, _ip, _ip]),
            "lbfs_time_steps": (C.c_int, [vp, C.c_int, _dp]),
            "lbfs_time_kernels": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_char_p), _dp, _ip]),
            "lbfs_time_stage": (C.c_int, [vp, C.c_char_p, C.c_int, _dp]),
            "lbfs_model_bytes": (C.c_int, [vp, C.c_int, _dp, _ip]),
            "lbfs_mesh_hex_counts": (C.c_int, [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip]),
            "lbfs_mesh_hex_generate": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, C.c_double, C.c_uint64,
                                                 _dp, _ip, _ip, _ip, _dp, _ip]),
            "lbfs_mesh_geometry": (C.c_int, [C.POINTER(MeshDesc), C.c_double, C.c_double, C.c_double, C.c_int,
                                             _dp, _dp, _dp, _dp, _dp, _dp]),
            "lbfs_membrane_mesh_counts": (C.c_int, [C.c_char_p, C.c_int, _ip, _ip, _ip]),
            "lbfs_case_read": (C.c_int, [C.c_char_p, C.POINTER(Case)]),
            "lbfs_stage_names": (C.c_int, [vp, C.c_int, C.POINTER(C.c_char_p), _ip]),
            "lbfs_stage_run": (C.c_int, [vp, C.c_char_p]),
            "lbfs_stage_end_step": (C.c_int, [vp]),
            "lbfs_membrane_mesh_read": (C.c_int, [C.c_char_p, C.c_int, _dp, _dp, _ip, _ip]),
            "lbfs_membrane_mesh_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip]),
            "lbfs_set_body_force": (C.c_int, [vp, _dp]),
            "lbfs_set_membrane_moduli": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double]),
            "lbfs_adjoint_forward": (C.c_int, [vp, C.c_int, C.c_int]),
            "lbfs_adjoint_backward": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "lbfs_adjoint_rewind": (C.c_int, [vp]),
            "lbfs_adjoint_run": (C.c_int, [vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
            "lbfs_state_txt_write": (C.c_int, [C.c_char_p, C.c_int, _dp, _dp, _dp, _dp, C.c_int]),
            "lbfs_state_txt_read": (C.c_int, [C.c_char_p, C.c_int, _dp, _dp, _dp, _dp]),
            "lbfs_membrane_dat_write": (C.c_int, [C.c_char_p, C.c_int, _dp, _dp, _dp, C.c_int, _ip, C.c_int, C.c_int,
                                                  _dp, _dp]),
            "lbfs_polymesh_counts": (C.c_int, [C.c_char_p, _ip, _ip, _ip, _ip, _ip]),
            "lbfs_polymesh_read": (C.c_int, [C.c_char_p, _dp, _ip, _ip, _ip, _ip, _dp]),
            "lbfs_polymesh_patch": (C.c_int, [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_int]),
            "lbfs_polymesh_write": (C.c_int, [C.c_char_p, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip,
                                              C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), _ip, _ip, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def ptr(a, kind=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


class LbfsError(RuntimeError):
    """A failed ABI call; `code` is its LBFS_E_* status."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


def check(rc, ctx=None):
    if rc != 0:
        msg = load().lbfs_last_error(ctx)
        raise LbfsError("lbfs call failed (%d): %s" % (rc, msg.decode() if msg else ""), rc)


PATCH_NAMES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


class HexMesh:
    """Seeded jittered hex box in polyMesh form (generated by the native helper),
    or a hexahedral OpenFOAM polyMesh read from disk (read_polymesh)."""

    patch_names = PATCH_NAMES
    patch_types = ("wall",) * 6

    def __init__(self, nx, ny, nz, lengths=None, origin=(0.0, 0.0, 0.0), periodic=(1, 0, 0), jitter=0.1, seed=2):
        lib = load()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.lengths = np.array(lengths if lengths is not None else (nx, ny, nz), dtype=np.float64)
        self.origin = np.array(origin, dtype=np.float64)
        self.periodic = np.array(periodic, dtype=np.int32)
        npnt, nf, ni = C.c_int(), C.c_int(), C.c_int()
        check(lib.lbfs_mesh_hex_counts(nx, ny, nz, ptr(self.periodic, _ip), C.byref(npnt), C.byref(nf), C.byref(ni)))
        self.n_points, self.n_faces, self.n_internal = npnt.value, nf.value, ni.value
        self.n_cells = nx * ny * nz
        self.n_bc = self.n_faces - self.n_internal
        self.points = np.zeros(3 * self.n_points)
        self.face_vertices = np.zeros(4 * self.n_faces, dtype=np.int32)
        self.owner = np.zeros(self.n_faces, dtype=np.int32)
        self.neighbour = np.zeros(max(self.n_internal, 1), dtype=np.int32)
        self.face_shift = np.zeros(3 * self.n_faces)
        self.patch = np.full(self.n_faces, -1, dtype=np.int32)
        check(lib.lbfs_mesh_hex_generate(nx, ny, nz, ptr(self.lengths), ptr(self.origin), ptr(self.periodic, _ip),
                                         float(jitter), int(seed) & (2 ** 64 - 1), ptr(self.points),
                                         ptr(self.face_vertices, _ip), ptr(self.owner, _ip), ptr(self.neighbour, _ip),
                                         ptr(self.face_shift), ptr(self.patch, _ip)))
        self._default_bc()

    def _default_bc(self):
        self.bc_patch = self.patch[self.n_internal:].copy()
        # default: no-slip Dirichlet velocity, Neumann density on every wall patch
        self.bc_rho_type = np.full(max(self.n_bc, 1), BC_NEUMANN, dtype=np.int32)
        self.bc_vel_type = np.full(max(self.n_bc, 1), BC_DIRICHLET, dtype=np.int32)
        self.bc_values = np.zeros(4 * max(self.n_bc, 1))
        self.bc_values[0::4] = 1.0
        self.spacing = float(self.lengths[0] / self.nx) if self.nx else 1.0

    @classmethod
    def read_polymesh(cls, directory, nx=0, ny=0, nz=0, origin=(0.0, 0.0, 0.0), lengths=None):
        """Read constant/polyMesh-style ASCII files (points, faces, owner,
        neighbour, boundary, optional lbfsFaceShift).  nx, ny, nz give the
        logical cell numbering when the mesh has one (needed for tiling and
        membranes); 0 leaves it unknown."""
        lib = load()
        d = os.fsencode(str(directory))
        cnt = [C.c_int() for _ in range(5)]
        check(lib.lbfs_polymesh_counts(d, *[C.byref(c) for c in cnt]))
        m = cls.__new__(cls)
        m.n_points, m.n_cells, m.n_faces, m.n_internal, n_patch = [c.value for c in cnt]
        m.n_bc = m.n_faces - m.n_internal
        m.nx, m.ny, m.nz = nx, ny, nz
        if nx and nx * ny * nz != m.n_cells:
            raise LbfsError("logical size %dx%dx%d does not match %d cells" % (nx, ny, nz, m.n_cells))
        m.origin = np.array(origin, dtype=np.float64)
        m.lengths = np.array(lengths if lengths is not None else (nx, ny, nz), dtype=np.float64)
        m.points = np.zeros(3 * m.n_points)
        m.face_vertices = np.zeros(4 * m.n_faces, dtype=np.int32)
        m.owner = np.zeros(m.n_faces, dtype=np.int32)
        m.neighbour = np.zeros(max(m.n_internal, 1), dtype=np.int32)
        m.patch = np.full(m.n_faces, -1, dtype=np.int32)
        m.face_shift = np.zeros(3 * m.n_faces)
        check(lib.lbfs_polymesh_read(d, ptr(m.points), ptr(m.face_vertices, _ip), ptr(m.owner, _ip),
                                     ptr(m.neighbour, _ip), ptr(m.patch, _ip), ptr(m.face_shift)))
        names, types = [], []
        for k in range(n_patch):
            nb, tb = C.create_string_buffer(256), C.create_string_buffer(256)
            check(lib.lbfs_polymesh_patch(d, k, nb, 256, tb, 256))
            names.append(nb.value.decode())
            types.append(tb.value.decode())
        m.patch_names, m.patch_types = tuple(names), tuple(types)
        m.periodic = np.array([np.any(m.face_shift.reshape(-1, 3)[:, t] != 0.0) for t in range(3)], dtype=np.int32)
        m._default_bc()
        for k, t in enumerate(m.patch_types):
            if t == "empty":
                m.set_patch_bc(k, BC_EMPTY, BC_EMPTY, (1.0, 0.0, 0.0, 0.0))
        return m

    def write_polymesh(self, directory):
        """Write the mesh as ASCII polyMesh files; periodic wrap faces stay
        internal faces and their translations go to lbfsFaceShift."""
        lib = load()
        os.makedirs(directory, exist_ok=True)
        bp = self.patch[self.n_internal:]
        starts, sizes = [], []
        for k in range(len(self.patch_names)):
            idx = np.nonzero(bp == k)[0]
            if len(idx) and not np.array_equal(idx, np.arange(idx[0], idx[0] + len(idx))):
                raise LbfsError("patch %d faces are not contiguous" % k)
            starts.append(self.n_internal + (int(idx[0]) if len(idx) else 0))
            sizes.append(len(idx))
        keep = list(range(len(self.patch_names)))   # empty patches kept so patch ids survive a round trip
        n = len(keep)
        names = (C.c_char_p * max(n, 1))(*[self.patch_names[k].encode() for k in keep])
        types = (C.c_char_p * max(n, 1))(*[self.patch_types[k].encode() for k in keep])
        st = np.array([starts[k] for k in keep] or [0], dtype=np.int32)
        sz = np.array([sizes[k] for k in keep] or [0], dtype=np.int32)
        shift = self.face_shift if np.any(self.face_shift != 0.0) else None
        check(lib.lbfs_polymesh_write(os.fsencode(str(directory)), self.n_points, ptr(self.points), self.n_cells,
                                      self.n_faces, self.n_internal, ptr(self.face_vertices, _ip), ptr(self.owner, _ip),
                                      ptr(self.neighbour, _ip), n, names, types, ptr(st, _ip), ptr(sz, _ip),
                                      ptr(shift)))

    def set_patch_bc(self, patch, rho_type, vel_type, values):
        sel = self.bc_patch == patch
        self.bc_rho_type[: self.n_bc][sel] = rho_type
        self.bc_vel_type[: self.n_bc][sel] = vel_type
        v = self.bc_values[: 4 * self.n_bc].reshape(-1, 4)
        v[sel] = values

    def desc(self):
        d = MeshDesc()
        d.n_points = self.n_points
        d.points = ptr(self.points)
        d.n_cells, d.n_faces, d.n_internal = self.n_cells, self.n_faces, self.n_internal
        d.face_vertices = ptr(self.face_vertices, _ip)
        d.owner = ptr(self.owner, _ip)
        d.neighbour = ptr(self.neighbour, _ip)
        d.face_shift = ptr(self.face_shift)
        d.bc_rho_type = ptr(self.bc_rho_type, _ip)
        d.bc_vel_type = ptr(self.bc_vel_type, _ip)
        d.bc_values = ptr(self.bc_values)
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        for t in range(3):
            d.origin[t] = self.origin[t]
        d.spacing = self.spacing
        return d

    def geometry(self, visc=0.0, gamma=1.0, fneq_min=0.0, flags=0):
        """Product FP64 geometry (host code of the library)."""
        lib = load()
        nf, nc, nb = self.n_faces, self.n_cells, self.n_bc
        out = dict(area=np.zeros(nf), normal=np.zeros(3 * nf), face_centroid=np.zeros(3 * nf), face_dt=np.zeros(nf),
                   centroid=np.zeros(3 * (nc + nb)), volume=np.zeros(nc))
        d = self.desc()
        check(lib.lbfs_mesh_geometry(C.byref(d), visc, gamma, fneq_min, flags, ptr(out["area"]), ptr(out["normal"]),
                                     ptr(out["face_centroid"]), ptr(out["face_dt"]), ptr(out["centroid"]),
                                     ptr(out["volume"])))
        return out


class Ensemble:
    """Independent boxes stepped together as one mesh (cells numbered box
    after box, all internal faces first, then all boundary faces).
    body_forces: one (fx, fy, fz) per box, or None for the solver's
    params.body_force in every box."""

    def __init__(self, members, body_forces=None):
        self.members = list(members)
        self.body_forces = None if body_forces is None else \
            np.ascontiguousarray(np.asarray(body_forces, dtype=np.float64).reshape(len(self.members), 3).reshape(-1))
        pts, fv, own, nb, sh, patch = [], [], [], [], [], []
        bfv, bown, bsh, bpatch, brt, bvt, bval = [], [], [], [], [], [], []
        p0 = c0 = 0
        dims, origin, spacing = [], [], []
        for k, m in enumerate(self.members):
            ni = m.n_internal
            f = m.face_vertices.reshape(-1, 4) + p0
            pts.append(m.points.reshape(-1, 3))
            fv.append(f[:ni]); bfv.append(f[ni:])
            own.append(m.owner[:ni] + c0); bown.append(m.owner[ni:] + c0)
            nb.append(m.neighbour[:ni] + c0)
            s = m.face_shift.reshape(-1, 3)
            sh.append(s[:ni]); bsh.append(s[ni:])
            patch.append(m.patch[:ni]); bpatch.append(m.patch[ni:])
            brt.append(m.bc_rho_type[: m.n_bc]); bvt.append(m.bc_vel_type[: m.n_bc])
            bval.append(m.bc_values[: 4 * m.n_bc])
            dims.append((m.nx, m.ny, m.nz)); origin.append(tuple(m.origin)); spacing.append(m.spacing)
            p0 += m.n_points
            c0 += m.n_cells
        self.points = np.ascontiguousarray(np.concatenate(pts).reshape(-1))
        self.face_vertices = np.ascontiguousarray(np.concatenate(fv + bfv).reshape(-1), dtype=np.int32)
        self.owner = np.ascontiguousarray(np.concatenate(own + bown), dtype=np.int32)
        self.neighbour = np.ascontiguousarray(np.concatenate(nb), dtype=np.int32)
        self.face_shift = np.ascontiguousarray(np.concatenate(sh + bsh).reshape(-1))
        self.patch = np.ascontiguousarray(np.concatenate(patch + bpatch), dtype=np.int32)
        self.n_points = p0
        self.n_cells = c0
        self.n_internal = len(self.neighbour)
        self.n_faces = len(self.owner)
        self.n_bc = self.n_faces - self.n_internal
        self.bc_patch = self.patch[self.n_internal:].copy()
        self.bc_rho_type = np.ascontiguousarray(np.concatenate(brt), dtype=np.int32)
        self.bc_vel_type = np.ascontiguousarray(np.concatenate(bvt), dtype=np.int32)
        self.bc_values = np.ascontiguousarray(np.concatenate(bval))
        self.block_dims = np.array(dims, dtype=np.int32).reshape(-1)
        self.block_origin = np.array(origin, dtype=np.float64).reshape(-1)
        self.block_spacing = np.array(spacing, dtype=np.float64)
        self.cell_start = np.cumsum([0] + [m.n_cells for m in self.members])
        self.nx = self.ny = self.nz = 0
        self.origin = np.zeros(3)
        self.spacing = float(spacing[0])

    def desc(self):
        d = MeshDesc()
        d.n_points = self.n_points
        d.points = ptr(self.points)
        d.n_cells, d.n_faces, d.n_internal = self.n_cells, self.n_faces, self.n_internal
        d.face_vertices = ptr(self.face_vertices, _ip)
        d.owner = ptr(self.owner, _ip)
        d.neighbour = ptr(self.neighbour, _ip)
        d.face_shift = ptr(self.face_shift)
        d.bc_rho_type = ptr(self.bc_rho_type, _ip)
        d.bc_vel_type = ptr(self.bc_vel_type, _ip)
        d.bc_values = ptr(self.bc_values)
        d.n_blocks = len(self.members)
        d.block_dims = ptr(self.block_dims, _ip)
        d.block_origin = ptr(self.block_origin)
        d.block_spacing = ptr(self.block_spacing)
        if self.body_forces is not None:
            d.block_body_force = ptr(self.body_forces)
        return d

    def split(self, arr):
        """Per-member slices of a per-cell array."""
        return [arr[self.cell_start[k]: self.cell_start[k + 1]] for k in range(len(self.members))]


def icosphere(frequency):
    """Frequency-n geodesic unit sphere: (nodes [N,3], triangles [T,3])."""
    n, t = 10 * frequency * frequency + 2, 20 * frequency * frequency
    x = np.zeros(3 * n)
    tri = np.zeros(3 * t, dtype=np.int32)
    check(load().lbfs_mesh_icosphere(int(frequency), ptr(x), ptr(tri, _ip)))
    return x.reshape(n, 3), tri.reshape(t, 3)


def evans_fung(unit_nodes, d0):
    u = np.ascontiguousarray(unit_nodes, dtype=np.float64).reshape(-1)
    out = np.zeros_like(u)
    check(load().lbfs_mesh_evans_fung(len(u) // 3, ptr(u), float(d0), ptr(out)))
    return out.reshape(-1, 3)


def springs(n_nodes, tri):
    t = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1)
    out = np.zeros(4 * (3 * (len(t) // 3) // 2 + 1), dtype=np.int32)
    ns = C.c_int()
    check(load().lbfs_mesh_springs(int(n_nodes), len(t) // 3, ptr(t, _ip), ptr(out, _ip), C.byref(ns)))
    return out[: 4 * ns.value].reshape(-1, 4)


def rotation(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def rbc(frequency=16, d0=20.0, center=(0.0, 0.0, 0.0), rot=None):
    """Evans-Fung red cell on a frequency-n geodesic sphere, rotated and placed."""
    u, tri = icosphere(frequency)
    x = evans_fung(u, d0)
    if rot is not None:
        x = x @ np.asarray(rot).T
    return x + np.asarray(center, dtype=np.float64), tri


def read_case(path):
    """Case file in the reference's XML schema -> Case (ctypes structure)."""
    c = Case()
    rc = load().lbfs_case_read(os.fsencode(str(path)), C.byref(c))
    if rc != 0:
        raise LbfsError("lbfs_case_read failed (%d): %s" % (rc, c.error.decode()))
    return c


def case_params(case, lattice=D3Q15, flags=0):
    """Solver parameters of a case (the reference's GPU lattice by default)."""
    return make_params(lattice=lattice, visc=case.visc, dt=case.time_marching_step, rho0=case.average_rho,
                       gamma=case.gamma, fneq_min=case.fneq_min, flags=flags, gradient_type=case.gradient_type,
                       body_force=tuple(case.body_force), channel_diameter=case.domain[1],
                       time_stepping=case.time_stepping, max_velocity=case.max_velocity)


def apply_case_bcs(mesh, case):
    """Set the mesh's per-face BCs from the case's boundary conditions, matched by patch name."""
    byname = {case.patches[k].name.decode(): case.patches[k] for k in range(case.n_patches)}
    for k, name in enumerate(mesh.patch_names):
        p = byname.get(name)
        if p is not None:
            mesh.set_patch_bc(k, p.rho_type, p.vel_type, tuple(p.values))


def read_membrane_mesh(path, fmt=MESH_OSAY, centre=(0.0, 0.0, 0.0)):
    """Osay or MZ membrane mesh file -> (nodes [N,3], tri [T,3], springs [S,4] or None)."""
    lib = load()
    n, t, sp = C.c_int(), C.c_int(), C.c_int()
    p = os.fsencode(str(path))
    check(lib.lbfs_membrane_mesh_counts(p, fmt, C.byref(n), C.byref(t), C.byref(sp)))
    nodes = np.zeros(3 * n.value)
    tri = np.zeros(3 * t.value, dtype=np.int32)
    springs = np.zeros(4 * max(sp.value, 1), dtype=np.int32)
    c = np.array(centre, dtype=np.float64)
    check(lib.lbfs_membrane_mesh_read(p, fmt, ptr(c), ptr(nodes), ptr(tri, _ip), ptr(springs, _ip)))
    return nodes.reshape(-1, 3), tri.reshape(-1, 3), (springs[: 4 * sp.value].reshape(-1, 4) if sp.value else None)


def write_membrane_mesh(path, nodes, tri, fmt=MESH_OSAY, springs_=None):
    lib = load()
    nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1)
    tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1)
    if fmt == MESH_MZ and springs_ is None:
        springs_ = springs(len(nodes) // 3, tri)
    sp = None if springs_ is None else np.ascontiguousarray(springs_, dtype=np.int32).reshape(-1)
    check(lib.lbfs_membrane_mesh_write(os.fsencode(str(path)), fmt, len(nodes) // 3, ptr(nodes), len(tri) // 3,
                                       ptr(tri, _ip), 0 if sp is None else len(sp) // 4, ptr(sp, _ip)))


def write_state_txt(directory, rho, u, v, w, precision=17):
    """Restart files rho.txt, u.txt, v.txt, w.txt ("i ,value" per cell); precision 0: 6 digits."""
    lib = load()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, w)]
    check(lib.lbfs_state_txt_write(os.fsencode(str(directory)), len(arrs[0]), *[ptr(a) for a in arrs], int(precision)))


def read_state_txt(directory, n_cells):
    lib = load()
    arrs = [np.zeros(n_cells) for _ in range(4)]
    check(lib.lbfs_state_txt_read(os.fsencode(str(directory)), int(n_cells), *[ptr(a) for a in arrs]))
    return tuple(arrs)


def write_membrane_dat(path, X, F, U, tri, time=0, precision=0):
    """Membrane zone file (positions, forces, velocities, nodal area, index, curvature, triangles);
    returns the nodal areas and curvatures written."""
    lib = load()
    X, F, U = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (X, F, U)]
    tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1)
    n = len(X) // 3
    area, curv = np.zeros(n), np.zeros(n)
    check(lib.lbfs_membrane_dat_write(os.fsencode(str(path)), n, ptr(X), ptr(F), ptr(U), len(tri) // 3, ptr(tri, _ip),
                                      int(time), int(precision), ptr(area), ptr(curv)))
    return area, curv


class Membrane:
    """Host-side description of one membrane (keeps its arrays alive)."""

    def __init__(self, nodes, tri, nodes_stress_free=None, nodes_curvature=None, springs=None, block=0, **params):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1)
        self.springs = None if springs is None else np.ascontiguousarray(springs, dtype=np.int32).reshape(-1)
        self.block = int(block)
        self.sf = None if nodes_stress_free is None else np.ascontiguousarray(nodes_stress_free, dtype=np.float64).reshape(-1)
        self.cv = None if nodes_curvature is None else np.ascontiguousarray(nodes_curvature, dtype=np.float64).reshape(-1)
        p = dict(RBC_PARAMS)
        p.update(params)
        self.params = p

    def desc(self):
        d = MembraneDesc()
        d.n_nodes, d.n_tri = len(self.nodes) // 3, len(self.tri) // 3
        d.nodes, d.tri = ptr(self.nodes), ptr(self.tri, _ip)
        d.nodes_stress_free, d.nodes_curvature = ptr(self.sf), ptr(self.cv)
        if self.springs is not None:
            d.n_springs, d.springs = len(self.springs) // 4, ptr(self.springs, _ip)
        d.block = self.block
        for k in MEMBRANE_PARAM_NAMES:
            setattr(d, k, self.params[k])
        return d


def make_params(lattice=D3Q19, visc=1.0 / 6.0, dt=0.5, rho0=1.0, gamma=1.0, fneq_min=0.0, flags=0,
                gradient_type=GRAD_HYBRID, body_force=(0.0, 0.0, 0.0), channel_diameter=0.0, time_stepping=0,
                max_velocity=0.0, tile=(0, 0, 0), block_threads=0):
    p = Params()
    p.lattice, p.gradient_type, p.flags = lattice, gradient_type, flags
    p.visc, p.gamma, p.fneq_min, p.dt, p.rho0 = visc, gamma, fneq_min, dt, rho0
    p.channel_diameter = channel_diameter
    p.time_stepping, p.max_velocity = time_stepping, max_velocity
    p.tile[0], p.tile[1], p.tile[2] = tile
    p.block_threads = block_threads
    for t in range(3):
        p.body_force[t] = body_force[t]
    return p


class Solver:
    """One device context (lbfs_ctx) over a mesh."""

    def __init__(self, mesh, params):
        lib = load()
        self.mesh = mesh
        self.params = params
        self._desc = mesh.desc()
        self.ctx = C.c_void_p()
        check(lib.lbfs_create(C.byref(self._desc), C.byref(params), C.byref(self.ctx)))
        self.n = mesh.n_cells

    def close(self):
        if self.ctx:
            load().lbfs_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, rho, u, v, w):
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, u, v, w)]
        check(load().lbfs_set_state(self.ctx, *[ptr(a) for a in arr]), self.ctx)

    def get_state(self):
        out = [np.zeros(self.n) for _ in range(4)]
        check(load().lbfs_get_state(self.ctx, *[ptr(a) for a in out]), self.ctx)
        return out

    def step(self, n=1):
        check(load().lbfs_step(self.ctx, int(n)), self.ctx)

    def sync(self):
        check(load().lbfs_sync(self.ctx), self.ctx)

    def gradients(self):
        g = np.zeros(12 * self.n)
        check(load().lbfs_stage_gradients(self.ctx, ptr(g)), self.ctx)
        return g

    def face_flux(self):
        f = np.zeros(4 * self.mesh.n_faces)
        check(load().lbfs_stage_face_flux(self.ctx, ptr(f)), self.ctx)
        return f

    def residual(self):
        r = np.zeros(4 * self.n)
        check(load().lbfs_stage_residual(self.ctx, ptr(r)), self.ctx)
        return r

    def set_graph(self, on):
        """Replay steps from a captured hipGraph (default) or launch eagerly."""
        check(load().lbfs_set_option(self.ctx, OPT_GRAPH, 1 if on else 0), self.ctx)

    def add_membranes(self, membranes, ibm_kernel=IBM_COS4):
        self._membranes = list(membranes)
        arr = (MembraneDesc * len(self._membranes))(*[m.desc() for m in self._membranes])
        self._mem_arr = arr
        check(load().lbfs_add_membranes(self.ctx, arr, len(self._membranes), int(ibm_kernel)), self.ctx)
        self.n_mem_nodes = sum(len(m.nodes) // 3 for m in self._membranes)

    def membrane_nodes(self):
        X, U, F = (np.zeros(3 * self.n_mem_nodes) for _ in range(3))
        check(load().lbfs_get_membrane_nodes(self.ctx, ptr(X), ptr(U), ptr(F)), self.ctx)
        return X.reshape(-1, 3), U.reshape(-1, 3), F.reshape(-1, 3)

    def write_restart(self, directory, precision=17):
        """The state as the reference's restart files (rho.txt, u.txt, v.txt, w.txt)."""
        write_state_txt(directory, *self.get_state(), precision=precision)

    def read_restart(self, directory):
        """Set the state from the reference's restart files."""
        self.set_state(*read_state_txt(directory, self.mesh.n_cells))

    def write_membrane_dat(self, path, index=0, time=0, precision=0):
        """Membrane `index` as the reference's zone file; returns nodal areas and curvatures."""
        X, U, F = self.membrane_nodes()
        n0 = sum(len(m.nodes) // 3 for m in self._membranes[:index])
        m = self._membranes[index]
        n = len(m.nodes) // 3
        return write_membrane_dat(path, X[n0:n0 + n], F[n0:n0 + n], U[n0:n0 + n], m.tri, time, precision)

    def membrane_globals(self):
        g = np.zeros(6 * len(self._membranes))
        check(load().lbfs_get_membrane_globals(self.ctx, ptr(g)), self.ctx)
        return g.reshape(-1, 6)

    def set_visc_search(self, on):
        """The internal-viscosity field from binned triangles and nodes (1,
        the default) or from every triangle of each membrane (0); the same
        bits either way."""
        check(load().lbfs_set_option(self.ctx, OPT_VISC_SEARCH, 1 if on else 0), self.ctx)

    def set_fused(self, on):
        """Whole steps with the fused fluid pass (1, the default where the
        tiles fit) or the two passes (0); the same bits either way."""
        check(load().lbfs_set_option(self.ctx, OPT_FUSED_FLUID, 1 if on else 0), self.ctx)

    def set_ibm_binned(self, mode):
        """Spreading path of the next add_membranes: -1 (default) by node density, 0 per node
        into global accumulators, 1 per fluid tile (each tile gathers its nodes, sums in LDS),
        2 per occupied 4x4x4 bin (LDS windows added to the accumulators)."""
        check(load().lbfs_set_option(self.ctx, OPT_IBM_BINNED, int(mode)), self.ctx)

    def set_step_schedule(self, mode):
        """Coupled step schedule: 0 (default) fluid on the main stream from the start, the whole
        coupling chain on a side stream; 1 interpolation alone, then fluid on the main stream and
        membrane forces + spreading on a side stream; 2 fluid on a side stream, forces on a third."""
        check(load().lbfs_set_option(self.ctx, OPT_STEP_SCHEDULE, int(mode)), self.ctx)

    def stage_names(self):
        names = (C.c_char_p * 32)()
        n = C.c_int()
        check(load().lbfs_stage_names(self.ctx, 32, names, C.byref(n)), self.ctx)
        return [names[k].decode() for k in range(n.value)]

    def stage_run(self, name):
        check(load().lbfs_stage_run(self.ctx, name.encode()), self.ctx)

    def stage_end_step(self):
        check(load().lbfs_stage_end_step(self.ctx), self.ctx)

    def membrane_energies(self):
        """Per membrane: spring, local area, local bending, global area,
        volume, ADE, total."""
        out = np.zeros(7 * len(self._membranes))
        check(load().lbfs_get_membrane_energies(self.ctx, ptr(out)), self.ctx)
        return out.reshape(-1, 7)

    def membrane_energies_noalloc(self):
        """The same through lbfs_get_membrane_energies_noalloc (static device
        staging; allocates nothing)."""
        out = np.zeros(7 * len(self._membranes))
        check(load().lbfs_get_membrane_energies_noalloc(self.ctx, ptr(out)), self.ctx)
        return out.reshape(-1, 7)

    def cell_force(self):
        f = np.zeros(3 * self.n)
        check(load().lbfs_get_cell_force(self.ctx, ptr(f)), self.ctx)
        return f

    def time_steps(self, n):
        """Milliseconds for n whole steps (HIP events on the context stream)."""
        ms = C.c_double()
        check(load().lbfs_time_steps(self.ctx, int(n), C.byref(ms)), self.ctx)
        return ms.value

    def time_kernels(self, n, max_count=16):
        """Average ms per launch of each kernel stage over n steps."""
        names = (C.c_char_p * max_count)()
        ms = np.zeros(max_count)
        cnt = C.c_int()
        check(load().lbfs_time_kernels(self.ctx, int(n), max_count, names, ptr(ms), C.byref(cnt)), self.ctx)
        return {names[i].decode(): ms[i] for i in range(cnt.value)}

    def time_stage(self, name, n):
        """Average ms of one fluid pass ("gradients", "flux_update") over n back-to-back launches."""
        ms = C.c_double()
        check(load().lbfs_time_stage(self.ctx, name.encode(), int(n), C.byref(ms)), self.ctx)
        return ms.value

    def model_bytes(self, max_count=16):
        b = np.zeros(max_count)
        cnt = C.c_int()
        check(load().lbfs_model_bytes(self.ctx, max_count, ptr(b), C.byref(cnt)), self.ctx)
        return b[: cnt.value].copy()

    def set_body_force(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        check(load().lbfs_set_body_force(self.ctx, ptr(f)), self.ctx)

    def set_membrane_moduli(self, index, shear, local_bending, global_bending):
        check(load().lbfs_set_membrane_moduli(self.ctx, int(index), float(shear), float(local_bending),
                                              float(global_bending)), self.ctx)

    def adjoint_forward(self, n_steps, checkpoint_every):
        """Run n_steps steps keeping checkpoints every checkpoint_every steps."""
        check(load().lbfs_adjoint_forward(self.ctx, int(n_steps), int(checkpoint_every)), self.ctx)

    def adjoint_rewind(self):
        check(load().lbfs_adjoint_rewind(self.ctx), self.ctx)

    def adjoint_backward(self, lam_nodes=None, lam_state=None):
        """Reverse sweep of the last adjoint_forward for the loss adjoints of
        the final node positions ([n_nodes, 3]) and state ([n_cells, 4]:
        rho, u, v, w).  Returns body_force [3], moduli [n_mem, 3] (shear,
        local bending, global bending), state0 [n_cells, 4], nodes0 [n_nodes, 3]."""
        n_mem = len(getattr(self, "_membranes", []))
        ln = None if lam_nodes is None else np.ascontiguousarray(lam_nodes, dtype=np.float64).reshape(-1)
        ls = None if lam_state is None else np.ascontiguousarray(lam_state, dtype=np.float64).reshape(-1)
        gb = np.zeros(3)
        gm = np.zeros(3 * max(n_mem, 1))
        s0 = np.zeros(4 * self.n)
        x0 = np.zeros(3 * max(getattr(self, "n_mem_nodes", 0), 1))
        check(load().lbfs_adjoint_backward(self.ctx, ptr(ln), ptr(ls), ptr(gb), ptr(gm) if n_mem else None,
                                           ptr(s0), ptr(x0) if n_mem else None), self.ctx)
        return {"body_force": gb, "moduli": gm[: 3 * n_mem].reshape(-1, 3), "state0": s0.reshape(-1, 4),
                "nodes0": x0[: 3 * getattr(self, "n_mem_nodes", 0)].reshape(-1, 3)}

    def adjoint_run(self, n_steps, checkpoint_every, target_nodes):
        """Loss sum |X - target|^2 after n_steps and its gradients."""
        t = np.ascontiguousarray(target_nodes, dtype=np.float64).reshape(-1)
        loss = C.c_double()
        gb = np.zeros(3)
        gm = np.zeros(3 * len(self._membranes))
        check(load().lbfs_adjoint_run(self.ctx, int(n_steps), int(checkpoint_every), ptr(t), C.byref(loss), ptr(gb),
                                      ptr(gm)), self.ctx)
        return loss.value, gb, gm.reshape(-1, 3)

    def diagnostics(self, ref_rho=1.0, ref_u=0.01):
        """Residual RMS of the step that would start from the current state
        (static force only), FP64 mass, largest speed."""
        d = Diag()
        check(load().lbfs_reduce_diagnostics(self.ctx, ref_rho, ref_u, C.byref(d)), self.ctx)
        return {k: getattr(d, k) for k, _ in Diag._fields_}

    def step_residual(self, ref_rho=1.0, ref_u=0.01):
        """One step; the residual RMS of that step (residuals.cpp:114-132)."""
        d = Diag()
        check(load().lbfs_step_residual(self.ctx, ref_rho, ref_u, C.byref(d)), self.ctx)
        return {k: getattr(d, k) for k, _ in Diag._fields_}

    def run_to_tolerance(self, max_steps, tolerance, check_every=1, ref_rho=1.0, ref_u=0.01):
        """Steps until a checked step's max(u, v, w) residual RMS is below
        tolerance (Solver_gpu.cu:1206); returns (steps, converged, diagnostics)."""
        d = Diag()
        done, conv = C.c_int(), C.c_int()
        check(load().lbfs_run_to_tolerance(self.ctx, int(max_steps), float(tolerance), int(check_every), ref_rho,
                                           ref_u, C.byref(done), C.byref(conv), C.byref(d)), self.ctx)
        return done.value, bool(conv.value), {k: getattr(d, k) for k, _ in Diag._fields_}
