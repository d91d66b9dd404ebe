"""Oracle restatement of the reference step loop.

Test infrastructure only (see ``oracle/__init__.py``).

``src/solver.py:88-151`` advances ``while t <= t_end``: dt from the current
state, one integrator call, then ``t += dt``.  The configurations used for
parity are step counts, so ``advance`` runs a fixed number of those
iterations (same per-step sequence, no snapshot I/O).
"""

from . import grid as _grid


def advance(physics, U, grid, params, space_schemes, integrator, bcs, cfl, n_steps):
    """Run ``n_steps`` iterations of the loop body of ``src/solver.py:137-151``.

    Returns the list of dt values used (one per step).
    """
    dts = []
    for _ in range(n_steps):
        dt = physics.compute_dt(U, params, grid, cfl)
        integrator(physics.compute_rhs, U, dt, params, grid, space_schemes, bcs,
                   _grid.apply_bcs)
        dts.append(float(dt))
    return dts
