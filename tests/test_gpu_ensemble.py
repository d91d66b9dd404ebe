"""Batched ensemble (mfvm_ens_advance) vs single-instance runs.

Every member of an ensemble must end bit for bit where the single-grid path
(navier_stokes.advance / advance_compensated, itself equal to the oracle)
takes it alone, with the same sequence of time steps -- whatever the launch
shape: tiles, work order, persistent launches, XCD order, graph replay.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPES = [(48, 40, 1), (64, 64, 1), (96, 32, 1), (130, 20, 1), (16, 16, 16), (32, 24, 8)]


def _kh_fields(n, seed):
    rng = np.random.default_rng(seed)
    shear = rng.uniform(0.3, 0.7)
    ratio = rng.uniform(1.5, 3.0)
    phase = rng.uniform(0.0, 2 * np.pi)
    x = (np.arange(n[0]) + 0.5) / n[0]
    y = (np.arange(n[1]) + 0.5) / n[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    X, Y = X[:, :, None], Y[:, :, None]
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, ratio, 1.0) * np.ones(n)
    vx = np.where(inner, shear, -shear) * np.ones(n)
    vy = 0.01 * np.sin(4 * np.pi * X + phase) + rng.uniform(-1e-3, 1e-3, n)
    vz = rng.uniform(-1e-3, 1e-3, n) if n[2] > 1 else np.zeros(n)
    E = 2.5 / 0.4 + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    return [rho, rho * vx, rho * vy, rho * vz, E]


def _setup(shapes, mixed_bcs=True):
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    schemes = {"riemann_solver": riemann_solvers.rusanov, "reconstructor": reconstructors.muscl,
               "slope_limiter": slope_limiters.monotonized_central,
               "diffusive_solver": diffusive_solvers.centered}
    grids, fields, bcs = [], [], []
    for i, n in enumerate(shapes):
        cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": n[0], "x2_min": 0.0, "x2_max": 1.0,
               "nx2": n[1], "x3_min": 0.0, "x3_max": 1.0, "nx3": n[2]}
        grids.append((cfg, build_grid("cartesian", cfg)))
        fields.append(_kh_fields(n, 3 + i))
        if mixed_bcs and i % 2 == 1:
            rule = {k: {"left": bc.periodic, "right": bc.periodic, "bottom": bc.neumann,
                        "top": bc.neumann, "front": bc.periodic, "back": bc.periodic} for k in ns.VARS}
            rule["momentum_x2"]["bottom"] = bc.reflecting
            rule["momentum_x2"]["top"] = bc.reflecting
        else:
            rule = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
        bcs.append(rule)
    return ns, params, schemes, grids, fields, bcs


def _state(ns, cfg, f):
    U = ns.allocate_u(cfg)
    for k, a in zip(ns.VARS, f):
        U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return U


OPTIONS = [
    {},
    {"order": "interleave", "persist": 1, "remap": 0, "graph": 0},
    {"order": "member", "persist": 2, "graph": 4,
     "tiles": {(64, 64, 1): (32, 8, 1), (16, 16, 16): (16, 16, 4), (130, 20, 1): (128, 2, 1)}},
    {"split": 0},
    {"split": 0, "order": "interleave", "persist": 1, "graph": 4,
     "tiles": {(64, 64, 1): (16, 16, 1), (16, 16, 16): (16, 16, 8)}},
]


@pytest.mark.parametrize("opt", range(len(OPTIONS)))
@pytest.mark.parametrize("integ", ["rk3", "rk3_ssp", "rk4"])
def test_ensemble_equals_single_runs(opt, integ):
    from mfvm.ensemble import Ensemble
    from mfvm.numerics import time_integrators as ti
    ns, params, schemes, grids, fields, bcs = _setup(SHAPES)
    integrator = getattr(ti, integ)
    steps = 11
    ens_states = [_state(ns, cfg, f) for (cfg, _), f in zip(grids, fields)]
    ens = Ensemble(ens_states, [g for _, g in grids], params, schemes, bcs, integrator)
    print(f"ensemble {integ} options {OPTIONS[opt]}", flush=True)
    dts = ens.advance(0.4, steps, OPTIONS[opt])
    torch.cuda.synchronize()
    for i, ((cfg, g), f) in enumerate(zip(grids, fields)):
        U = _state(ns, cfg, f)
        ref_dts = ns.advance(U, g, params, schemes, bcs[i], integrator._mfvm_code, 0.4, steps)
        torch.cuda.synchronize()
        assert torch.equal(dts[i], ref_dts), (i, SHAPES[i])
        assert torch.equal(ens_states[i].buf, U.buf), (i, SHAPES[i])


@pytest.mark.parametrize("split", [1, 0])
def test_compensated_ensemble_equals_single_runs(split):
    from mfvm.ensemble import Ensemble
    from mfvm.numerics import time_integrators as ti
    ns, params, schemes, grids, fields, bcs = _setup(SHAPES)
    steps = 9
    ens_states = [ns.to_compensated(_state(ns, cfg, f)) for (cfg, _), f in zip(grids, fields)]
    ens = Ensemble(ens_states, [g for _, g in grids], params, schemes, bcs, ti.rk3, compensated=True)
    print(f"compensated ensemble split={split}", flush=True)
    dts = ens.advance(0.4, steps, {"graph": 4, "persist": 1, "split": split})
    torch.cuda.synchronize()
    for i, ((cfg, g), f) in enumerate(zip(grids, fields)):
        U = ns.to_compensated(_state(ns, cfg, f))
        ref_dts = ns.advance_compensated(U, g, params, schemes, bcs[i], ti.rk3._mfvm_code, 0.4, steps)
        torch.cuda.synchronize()
        assert torch.equal(dts[i], ref_dts), i
        assert torch.equal(ens_states[i].buf, U.buf), i
        shp = tuple(U.buf.shape)
        assert torch.equal(ens_states[i].scratch("lo", shp), U.scratch("lo", shp)), i


def test_autotune_keeps_results():
    from mfvm.ensemble import Ensemble
    from mfvm.numerics import time_integrators as ti
    ns, params, schemes, grids, fields, bcs = _setup(SHAPES[:4] + SHAPES[5:])
    ens_states = [_state(ns, cfg, f) for (cfg, _), f in zip(grids, fields)]
    ens = Ensemble(ens_states, [g for _, g in grids], params, schemes, bcs, ti.rk3)
    before = [U.buf.clone() for U in ens_states]
    best = ens.autotune(0.4, steps=2, verbose=True)
    assert set(best) >= {"tiles", "order", "persist", "graph", "remap", "split"}
    assert len(ens.tuning_log) >= 5
    for U, b in zip(ens_states, before):  # tuning ran on copies
        assert torch.equal(U.buf, b)
    dts = ens.advance(0.4, 5)
    torch.cuda.synchronize()
    for i, ((cfg, g), f) in enumerate(zip(grids, fields)):
        U = _state(ns, cfg, f)
        ref = ns.advance(U, g, params, schemes, bcs[i], ti.rk3._mfvm_code, 0.4, 5)
        torch.cuda.synchronize()
        assert torch.equal(dts[i], ref)
        assert torch.equal(ens_states[i].buf, U.buf)


@pytest.mark.parametrize("kind", ["pow2", "curvilinear"])
@pytest.mark.parametrize("compensated", [False, True])
def test_ensemble_geometry_specialisations_equal_single_runs(kind, compensated):
    """Members whose geometry selects the other kernel forms: all uniform
    power-of-two (the viscous power-of-two specialisation, as in the
    benchmark ensemble), and a mix with cylindrical and spherical members (the
    metric-factor form).  Each member equals its single-grid run bit for bit,
    dt sequence included."""
    from mfvm.ensemble import Ensemble
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import time_integrators as ti
    ns, params, schemes, _, _, _ = _setup([])
    if kind == "pow2":
        members = [("cartesian", (64, 64, 1), (0, 0, 0), (1, 1, 1)),
                   ("cartesian", (32, 128, 1), (0, 0, 0), (1, 1, 1)),
                   ("cartesian", (16, 16, 16), (0, 0, 0), (1, 1, 1))]
    else:
        members = [("cartesian", (48, 40, 1), (0, 0, 0), (1, 1, 1)),
                   ("cylindrical", (40, 24, 8), (0.5, 0.0, 0.0), (1.5, np.pi / 2, 1.0)),
                   ("spherical", (24, 16, 12), (0.5, 0.3, 0.0), (1.5, np.pi - 0.3, np.pi))]
    grids, fields, bcs = [], [], []
    for i, (coord, n, lo, hi) in enumerate(members):
        cfg = {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1],
               "nx2": n[1], "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}
        grids.append((cfg, build_grid(coord, cfg)))
        fields.append(_kh_fields(n, 40 + i))
        if coord == "cartesian":
            bcs.append({k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS})
        else:
            bcs.append({k: {s: bc.neumann for s in bc.SIDES} for k in ns.VARS})

    def make(cfg, f):
        U = _state(ns, cfg, f)
        return ns.to_compensated(U) if compensated else U
    steps = 7
    ens_states = [make(cfg, f) for (cfg, _), f in zip(grids, fields)]
    ens = Ensemble(ens_states, [g for _, g in grids], params, schemes, bcs, ti.rk3, compensated=compensated)
    print(f"ensemble {kind} compensated={compensated}", flush=True)
    dts = ens.advance(0.4, steps, {"graph": 4})
    torch.cuda.synchronize()
    for i, ((cfg, g), f) in enumerate(zip(grids, fields)):
        U = make(cfg, f)
        run = ns.advance_compensated if compensated else ns.advance
        ref_dts = run(U, g, params, schemes, bcs[i], ti.rk3._mfvm_code, 0.4, steps)
        torch.cuda.synchronize()
        assert torch.equal(dts[i], ref_dts), (kind, i)
        assert torch.equal(ens_states[i].buf, U.buf), (kind, i)
        if compensated:
            shp = tuple(U.buf.shape)
            assert torch.equal(ens_states[i].scratch("lo", shp), U.scratch("lo", shp)), (kind, i)
