"""Committed fixtures (tests/golden/oracle_fixtures.npz, written by
tests/golden/make_fixtures.py): the oracle must reproduce them (CPU), and the
GPU library must match them within the parity tolerances (GPU)."""
import os
import sys

import numpy as np
import pytest

import cases
from lbfs_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_fixtures  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "oracle_fixtures.npz"))


def test_oracle_reproduces_the_fixtures():
    """Any change of the oracle's arithmetic shows up here."""
    out = make_fixtures.compute()
    assert sorted(out) == sorted(FIX.files)
    for k in FIX.files:
        np.testing.assert_allclose(out[k], FIX[k], rtol=1e-12, atol=1e-15, err_msg=k)


@pytest.mark.gpu
def test_gpu_matches_the_fluid_fixtures():
    mesh, st = make_fixtures.fluid_case()
    s = abi.Solver(mesh, abi.make_params(lattice=abi.D3Q19, visc=1.0 / 6.0, dt=0.25))
    s.set_state(*st)
    g = s.gradients().reshape(-1, 4, 3)
    gf = FIX["fluid_gradients"].reshape(-1, 4, 3)
    for v in range(4):
        assert cases.rel_l2(g[:, v], gf[:, v]) < 2e-5, v
    assert cases.rel_l2(s.face_flux(), FIX["fluid_face_flux"]) < 1e-4
    r = s.residual().reshape(-1, 4)
    rf = FIX["fluid_residual"].reshape(-1, 4)
    for v in range(4):
        assert cases.rel_l2(r[:, v], rf[:, v]) < 1e-4, v
    s.step(1)
    st1 = np.stack(s.get_state())
    s.step(9)
    st10 = np.stack(s.get_state())
    for got, ref in ((st1, FIX["fluid_state_1"]), (st10, FIX["fluid_state_10"])):
        assert cases.rel_l2(got[0] - 1.0, ref[0] - 1.0) < 1e-5
        assert cases.rel_l2(got[1:], ref[1:]) < 1e-5
    s.close()
