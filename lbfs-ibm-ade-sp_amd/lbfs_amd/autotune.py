"""Tile-shape and workgroup-size autotuner for the flux/update kernel.

The cell tiling (cells per tile along x, y, z) and the workgroup size change
neither the arithmetic nor the summation order of the step, only how faces are
shared between workgroups and how many workgroups fill the chip.  The tuner
builds the solver once per candidate on the caller's mesh, times a few whole
steps on the device, and returns the fastest candidate.
"""
from . import abi

# candidate (tile x, tile y, tile z, workgroup threads)
CANDIDATES = (
    (8, 8, 4, 256),
    (8, 4, 8, 256),
    (8, 8, 8, 256),
    (8, 8, 8, 512),
    (16, 8, 4, 256),
    (16, 8, 4, 512),
    (16, 8, 8, 512),
    (4, 4, 16, 256),
)


def with_tile(params, tile, threads):
    """Copy of `params` with the given tile shape and workgroup size."""
    p = abi.Params()
    C = type(params)
    for name, _ in C._fields_:
        v = getattr(params, name)
        if hasattr(v, "_length_"):
            getattr(p, name)[:] = list(v)
        else:
            setattr(p, name, v)
    p.tile[0], p.tile[1], p.tile[2] = tile
    p.block_threads = threads
    return p


def tune(mesh, params, state=None, steps=20, warmup=5, candidates=CANDIDATES, membranes=None, ibm_kernel=None):
    """Time `steps` whole steps per candidate and return (best_params, table).

    table: list of (tile, threads, ms_per_step).  Candidates whose tiles do not
    fit the mesh or the LDS budget are skipped."""
    table = []
    best = None
    for tx, ty, tz, nt in candidates:
        if mesh.nx and (tx > mesh.nx or ty > mesh.ny or tz > mesh.nz):
            continue
        p = with_tile(params, (tx, ty, tz), nt)
        try:
            s = abi.Solver(mesh, p)
        except abi.LbfsError:
            continue
        if state is not None:
            s.set_state(*state)
        if membranes:
            s.add_membranes(membranes, ibm_kernel=abi.IBM_COS4 if ibm_kernel is None else ibm_kernel)
        s.step(warmup)
        s.sync()
        ms = s.time_steps(steps) / steps
        s.close()
        table.append(((tx, ty, tz), nt, ms))
        if best is None or ms < best[1]:
            best = (p, ms)
    if best is None:
        raise abi.LbfsError("no tile candidate fits this mesh")
    return best[0], table
