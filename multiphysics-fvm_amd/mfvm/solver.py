"""Step loop and ghost fill (drop-in for src/solver.py).

``apply_bcs`` on an mfvm State is one kernel launch for all fields (faces,
edges and corners); ``run`` keeps the reference's ``while t <= t_end`` loop,
and ``advance`` runs a fixed number of steps with dt kept on the device.
"""

import ctypes

from . import _lib
from .grid import geometry
from .grid.boundary_conditions import SIDES
from .physics import _common as c
from .state import State


def apply_bcs(U, bcs):
    """src/solver.py:13-61."""
    if isinstance(U, State) and tuple(bcs) == U.names:
        grid = _grid_of(U)
        _, flat = c.bc_table(U.names, bcs)
        _lib.check(_lib.lib.mfvm_apply_bcs(grid.ptr(), len(U.names), flat, c.dtype_code(U),
                                           c.ptr(U.buf), c.ptr(U.buf), c.stream()))
        return
    for name in bcs:
        f = U[name]
        for side in SIDES:
            bcs[name][side](f, side)
        for a in (0, -1):
            an = 1 if a == 0 else -2
            for b in (0, -1):
                bn = 1 if b == 0 else -2
                f[a, b, 1:-1] = 0.5 * (f[an, b, 1:-1] + f[a, bn, 1:-1])
                f[a, 1:-1, b] = 0.5 * (f[an, 1:-1, b] + f[a, 1:-1, bn])
                f[1:-1, a, b] = 0.5 * (f[1:-1, an, b] + f[1:-1, a, bn])
        for a in (0, -1):
            an = 1 if a == 0 else -2
            for b in (0, -1):
                bn = 1 if b == 0 else -2
                for d in (0, -1):
                    dn = 1 if d == 0 else -2
                    f[a, b, d] = (f[an, b, d] + f[a, bn, d] + f[a, b, dn]) / 3.0


_GRIDS = {}


def _grid_of(U):
    """Any GridDesc of U's size serves the ghost fill (it reads only the sizes)."""
    key = (U.n, str(U.device))
    g = _GRIDS.get(key)
    if g is None:
        cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": U.n[0], "x2_min": 0.0, "x2_max": 1.0,
               "nx2": U.n[1], "x3_min": 0.0, "x3_max": 1.0, "nx3": U.n[2]}
        g = geometry.GridDesc("cartesian", cfg, U.device)
        _GRIDS[key] = g
    return g


def apply_ics(U, params, grid, ics):
    """src/solver.py:64-85."""
    ics["initializer"](U, params, grid, ics["parameters"])


def advance(setup, U, grid, nsteps):
    """Fixed-step form of the loop body of src/solver.py:137-151; returns the dts."""
    physics = setup["physics"]
    return physics.advance(U, grid, setup["physical_parameters"], setup["space_schemes"],
                           setup["boundary_conditions"], setup["time_integrator"]._mfvm_code,
                           setup["time_config"]["CFL"], nsteps)


def run(setup):
    """src/solver.py:88-151 -- the reference loop (dt read back each step)."""
    physics = setup["physics"]
    params = setup["physical_parameters"]
    tc = setup["time_config"]
    grid = geometry.build_grid(setup["coordinate_system"], setup["grid_config"])
    U = physics.allocate_u(setup["grid_config"])
    apply_ics(U, params, grid, setup["initial_conditions"])
    bcs = setup["boundary_conditions"]
    apply_bcs(U, bcs)
    save = setup.get("save_config", {})
    saver, freq = save.get("saver"), save.get("output_frequency", 0)
    t, n = tc["t_start"], 0
    while t <= tc["t_end"]:
        dt = physics.compute_dt(U, params, grid, tc["CFL"])
        setup["time_integrator"](physics.compute_rhs, U, dt, params, grid, setup["space_schemes"],
                                 bcs, apply_bcs)
        if freq and n % freq == 0:
            print(f"n = {n:6d} | t = {t:.2f}/{tc['t_end']:.2f} | dt = {dt:.2e} | {100*t/tc['t_end']:.1f}%")
            if saver is not None:
                saver(U, setup, t, n, save.get("directory"))
        t += dt
        n += 1
    return U
