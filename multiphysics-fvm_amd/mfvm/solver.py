"""Step loop and ghost fill (drop-in for src/solver.py).

``apply_bcs`` on an mfvm State is one kernel launch for all fields (faces,
edges and corners); ``run`` keeps the reference's ``while t <= t_end`` loop,
and ``advance`` runs a fixed number of steps with dt kept on the device.
"""

import ctypes

import numpy as np

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
    """src/solver.py:64-85.

    Initializers are host code written against NumPy arrays (the reference's
    setups index, slice and assign U[var] and read grid coordinates with
    NumPy).  For an mfvm State the initializer therefore runs on a host
    mirror -- U as NumPy arrays of the reference shape, the grid as
    :func:`geometry.build_grid_host` -- and the result is copied to the
    device.  Anything else is passed through unchanged."""
    if not isinstance(U, State):
        ics["initializer"](U, params, grid, ics["parameters"])
        return
    host = {name: U[name].cpu().numpy().copy() for name in U.names}
    g = grid.get("mfvm") if isinstance(grid, dict) else None
    if g is not None:
        hgrid = geometry.build_grid_host(g.coordinate_system, grid["config"])
    else:
        hgrid = grid
    ics["initializer"](host, params, hgrid, ics["parameters"])
    import torch
    for name in U.names:
        U[name].copy_(torch.from_numpy(np.ascontiguousarray(host[name], dtype=np.float64)))


def advance(setup, U, grid, nsteps):
    """Fixed-step form of the loop body of src/solver.py:137-151; returns the dts."""
    physics = setup["physics"]
    return physics.advance(U, grid, setup["physical_parameters"], setup["space_schemes"],
                           setup["boundary_conditions"], setup["time_integrator"]._mfvm_code,
                           setup["time_config"]["CFL"], nsteps)


def _diagnostics(setup, U, grid):
    from .diagnostics import diagnostics
    from .physics import navier_stokes as ns
    gamma = setup["physical_parameters"]["heat_capacity_ratio"] if U.names == ns.VARS else None
    return diagnostics(U, grid, gamma=gamma)


def _report(setup, U, t, n, t_end, grid=None, base=None):
    save = setup.get("save_config", {})
    saver, freq = save.get("saver"), save.get("output_frequency", 0)
    if freq and n % freq == 0:
        print(f"n = {n:6d} | t = {t:.2f}/{t_end:.2f}")
        if base is not None:
            # save_config["diagnostics"] (an extension): the change of every
            # conserved total since the start, and the positivity monitor
            d = _diagnostics(setup, U, grid)
            drift = " ".join(f"{k}={d['total'][k] - base['total'][k]:+.3e}" for k in U.names)
            line = f"           total drift: {drift}"
            if d["nonpositive_cells"] is not None:
                line += f" | min P {d['pressure_min']:.4e}, cells with rho<=0 or P<=0: {d['nonpositive_cells']}"
            print(line)
        if saver is not None:
            saver(U, setup, t, n, save.get("directory"))


def run(setup):
    """src/solver.py:88-151.

    With ``time_config["t_end"]`` only, this is the reference loop: dt from
    ``compute_dt`` (one readback per step), one integrator call, output every
    ``output_frequency`` steps after the step, labelled with the pre-step time
    (Q17: ``while t <= t_end``, no clipping).  Two additions: with
    ``time_config["n_steps"]`` the run takes exactly that many steps with dt
    kept on the device (chunks between outputs, same dts and results as the
    step-by-step loop); with a ``closure`` entry (learned flux correction) it
    advances to exactly ``t_end`` with :func:`mfvm.closure.advance`; with
    ``save_config["diagnostics"]`` each output line is followed by the drift
    of the conserved totals since the start and, for Navier-Stokes, the
    positivity monitor (:mod:`mfvm.diagnostics`).
    Returns U."""
    physics = setup["physics"]
    params = setup["physical_parameters"]
    tc = setup["time_config"]
    grid = geometry.build_grid(setup["coordinate_system"], setup["grid_config"])
    U = physics.allocate_u(setup["grid_config"])
    apply_ics(U, params, grid, setup["initial_conditions"])
    bcs = setup["boundary_conditions"]
    apply_bcs(U, bcs)
    if "closure" in setup:
        return _run_closure(setup, U, grid)
    save = setup.get("save_config", {})
    freq = save.get("output_frequency", 0)
    base = _diagnostics(setup, U, grid) if save.get("diagnostics") else None
    if "n_steps" in tc:
        nsteps = int(tc["n_steps"])
        t, n = tc.get("t_start", 0.0), 0
        while n < nsteps:
            # run up to and including the next output iteration
            chunk = (min(nsteps, (n + freq - 1) // freq * freq + 1) if freq else nsteps) - n
            dts = advance(setup, U, grid, chunk).cpu().numpy()
            last = n + chunk - 1
            if freq and last % freq == 0:
                _report(setup, U, t + float(dts[:chunk - 1].sum()), last, tc.get("t_end", float("inf")),
                        grid, base)
            t += float(dts[:chunk].sum())
            n += chunk
        return U
    t, n = tc["t_start"], 0
    while t <= tc["t_end"]:
        dt = physics.compute_dt(U, params, grid, tc["CFL"])
        setup["time_integrator"](physics.compute_rhs, U, dt, params, grid, setup["space_schemes"],
                                 bcs, apply_bcs)
        _report(setup, U, t, n, tc["t_end"], grid, base)
        t += dt
        n += 1
    return U


def _run_closure(setup, U, grid):
    from . import closure
    tc = setup["time_config"]
    net = closure.FluxCorrectionCNN.load(setup["closure"]["weights"], device=U.device)
    closure._run_to(U, grid, setup["physical_parameters"], setup["space_schemes"], setup["boundary_conditions"],
                    net, tc["CFL"], tc["t_end"] - tc.get("t_start", 0.0))
    save = setup.get("save_config", {})
    if save.get("saver") is not None:
        save["saver"](U, setup, tc["t_end"], 0, save.get("directory"))
    return U


def run_ensemble(setups, nsteps=None, compensated=False, autotune=True):
    """Advance the setups' states together (mfvm.ensemble): one launch per
    stage phase over all members.  Setups must share physics parameters,
    schemes and integrator.  Returns (states, dts)."""
    from .ensemble import Ensemble
    from .physics import navier_stokes as ns
    states, grids = [], []
    for s in setups:
        g = geometry.build_grid(s["coordinate_system"], s["grid_config"])
        U = s["physics"].allocate_u(s["grid_config"])
        apply_ics(U, s["physical_parameters"], g, s["initial_conditions"])
        states.append(ns.to_compensated(U) if compensated else U)
        grids.append(g)
    s0 = setups[0]
    ens = Ensemble(states, grids, s0["physical_parameters"], s0["space_schemes"],
                   [s["boundary_conditions"] for s in setups], s0["time_integrator"], compensated=compensated)
    if autotune:
        ens.autotune(s0["time_config"]["CFL"])
    n = nsteps if nsteps is not None else int(s0["time_config"]["n_steps"])
    dts = ens.advance(s0["time_config"]["CFL"], n)
    return states, dts
