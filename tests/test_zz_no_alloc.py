"""lbfs_get_membrane_energies_noalloc (static device staging, no allocation)
against lbfs_get_membrane_energies on two calls in a row, the second with
fewer, smaller membranes, through the C ABI."""
import numpy as np
import pytest

import cases
from lbfs_amd import abi

pytestmark = pytest.mark.gpu

PARAMS = dict(abi.RBC_PARAMS, local_bending_modulus=1e-3, viscosity=0.02, volume_ratio=0.95)


def solver_with(cells, steps):
    mesh = cases.channel_mesh(40, 16, 16, seed=5)
    s = abi.Solver(mesh, abi.make_params(visc=1.0 / 6.0, dt=0.25, body_force=(2e-5, 0.0, 0.0)))
    mems = []
    for freq, d0, c in cells:
        x, tri = abi.rbc(freq, d0, center=c, rot=abi.rotation((1, 1, 0), 0.3))
        mems.append(abi.Membrane(x, tri, **PARAMS))
    s.add_membranes(mems, ibm_kernel=abi.IBM_COS4)
    s.step(steps)
    return s


def test_noalloc_energies_equal_the_allocating_entry():
    big = solver_with([(4, 8.0, (10.0, 8.0, 8.0)), (5, 8.0, (30.0, 8.0, 8.0))], 6)
    a = big.membrane_energies()
    b = big.membrane_energies_noalloc()
    assert a.shape == (2, 7) and np.all(np.isfinite(a)) and np.any(a[:, 6] != 0.0)
    assert np.array_equal(a, b)
    small = solver_with([(3, 7.0, (20.0, 8.0, 8.0))], 4)
    a2 = small.membrane_energies()
    b2 = small.membrane_energies_noalloc()
    assert a2.shape == (1, 7) and np.any(a2[:, 6] != 0.0)
    assert np.array_equal(a2, b2)
    # the first context still reads its own energies after the second call
    assert np.array_equal(big.membrane_energies_noalloc(), a)
    big.close()
    small.close()
