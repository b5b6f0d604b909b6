"""Writes the oracle fixtures in tests/golden/oracle_fixtures.npz.

Small inputs whose oracle outputs are kept as data, so that a change of the
oracle's arithmetic shows up as a failing comparison (tests/test_golden.py):
  fluid     4x4x4 jittered fully periodic hex mesh (seed 11), D3Q19: LS
            gradients, face fluxes and residual of a seeded smooth state, and
            the state after 1 and 10 steps (dt 0.25, visc 1/6)
  membrane  12-node icosahedron and 642-node sphere (radius 4), stress-free
            shape the sphere, forces at a seeded perturbation
  ibm       cosine and 2-point weights: velocities interpolated at seeded node
            positions from a seeded cell field, and the spread of seeded node
            forces, on an 8x8x8 periodic box
Run from the repository root: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [REPO, os.path.join(REPO, "lbfs-ibm-ade-sp_amd"), os.path.dirname(HERE)]

import cases  # noqa: E402
from lbfs_amd import abi  # noqa: E402
from oracle import oracle as orc  # noqa: E402

PATH = os.path.join(HERE, "oracle_fixtures.npz")


def fluid_case():
    mesh = abi.HexMesh(4, 4, 4, periodic=(1, 1, 1), jitter=0.1, seed=11)
    st = cases.smooth_state(mesh, seed=11, amp=0.01)
    return mesh, st


def membrane_case(freq):
    x, tri = abi.icosphere(freq)
    x = 4.0 * x.reshape(-1, 3)
    rng = np.random.default_rng(freq)
    xp = x + 0.05 * rng.standard_normal(x.shape)
    params = dict(abi.RBC_PARAMS, local_bending_modulus=1e-3, global_bending_modulus=5e-4, viscosity=0.02)
    return x, xp, tri, params


def ibm_case():
    mesh = abi.HexMesh(8, 8, 8, periodic=(1, 1, 1), jitter=0.05, seed=12)
    rng = np.random.default_rng(12)
    X = rng.uniform(0.5, 7.5, size=(20, 3))
    uc = 0.01 * rng.standard_normal((mesh.n_cells, 3))
    F = rng.standard_normal((20, 3))
    return mesh, X, uc, F


def compute():
    out = {}
    mesh, st = fluid_case()
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=1.0 / 6.0, dt=0.25)
    fl.set_state(*st)
    out["fluid_gradients"] = fl.gradients()
    out["fluid_face_flux"] = fl.face_flux()
    out["fluid_residual"] = fl.residual()
    fl.step(1)
    out["fluid_state_1"] = np.stack(fl.state())
    fl.step(9)
    out["fluid_state_10"] = np.stack(fl.state())
    for freq in (1, 8):
        x, xp, tri, params = membrane_case(freq)
        mo = orc.OracleMembrane(x, tri, params)
        out["membrane_forces_%d" % freq] = np.asarray(mo.forces(X=xp.reshape(-1))).reshape(-1, 3)
    mesh, X, uc, F = ibm_case()
    for kernel in (abi.IBM_COS4, abi.IBM_LINEAR2):
        fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=1.0 / 6.0, dt=0.25)
        fl.set_state(np.ones(mesh.n_cells), uc[:, 0], uc[:, 1], uc[:, 2])
        cp = orc.OracleCoupled(fl, mesh, [], kernel)
        out["ibm_interp_%d" % kernel] = np.asarray(cp.interpolate(X.reshape(-1))).reshape(-1, 3)
        f_cell = np.zeros(3 * mesh.n_cells)
        cp.spread(X.reshape(-1), F.reshape(-1), f_cell)
        out["ibm_spread_%d" % kernel] = f_cell.reshape(-1, 3)
    return out


if __name__ == "__main__":
    np.savez_compressed(PATH, **compute())
    print("wrote", PATH)
