"""Tile-shape, workgroup-size and fluid-pass autotuner.

The cell tiling (cells per tile along x, y, z), the workgroup size and the
choice between the fused fluid pass (each tile's gradients in LDS) and the
gradient + face passes change neither the arithmetic nor the summation order
of the step, only how faces and gradients are shared between workgroups and
how many workgroups fill the chip.  The tuner builds the solver once per
candidate on the caller's mesh, times a few whole steps on the device, and
returns the fastest candidate.
"""
from . import abi

# candidate (tile x, tile y, tile z, workgroup threads)
CANDIDATES = (
    (8, 8, 4, 256),
    (8, 4, 8, 256),
    (8, 8, 8, 256),
    (8, 8, 8, 512),
    (8, 8, 8, 576),
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


def tune(mesh, params, state=None, steps=20, warmup=5, candidates=CANDIDATES, membranes=None, ibm_kernel=None,
         margin=0.0):
    """Time `steps` whole steps per candidate (tile, threads) with the fused
    fluid pass where the tiles fit it and with the two passes, and return
    (best_params, best_fused, table).  A later candidate replaces the best so
    far only when it is faster by more than `margin` (relative), so that
    timings within the run-to-run spread keep the earlier candidate.

    table: list of (tile, threads, fused, ms_per_step).  Candidates whose
    tiles do not fit the mesh or the LDS budget are skipped."""
    table = []
    best = None
    for tx, ty, tz, nt in candidates:
        if mesh.nx and (tx > mesh.nx or ty > mesh.ny or tz > mesh.nz):
            continue
        p = with_tile(params, (tx, ty, tz), nt)
        for fused in (True, False):
            try:
                s = abi.Solver(mesh, p)
            except abi.LbfsError:
                break  # the tile does not fit
            try:
                s.set_fused(fused)
            except abi.LbfsError:
                s.close()  # too large for the fused pass
                continue
            if state is not None:
                s.set_state(*state)
            if membranes:
                s.add_membranes(membranes, ibm_kernel=abi.IBM_COS4 if ibm_kernel is None else ibm_kernel)
            s.step(warmup)
            s.sync()
            ms = s.time_steps(steps) / steps
            s.close()
            table.append(((tx, ty, tz), nt, fused, ms))
            if best is None or ms < best[2] * (1.0 - margin):
                best = (p, fused, ms)
    if best is None:
        raise abi.LbfsError("no tile candidate fits this mesh")
    return best[0], best[1], table
