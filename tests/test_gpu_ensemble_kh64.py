"""The batched-ensemble scheme and workload against the oracle.

The ensemble workload (64 Kelvin-Helmholtz members: 56 2D grids with nx1,
nx2 drawn from {64, 128, 256, 512} and eight 64^3 grids) runs MUSCL + MC
(slope_limiters.py:42-73, with its 1/3 sign-split factor), Rusanov
(riemann_solvers.py:13-48) and the reference rk3 (time_integrators.py:50-103,
the non-Kutta stages).  Bar: bit-exact for float64.

* single grids of that scheme -- power-of-two 64^2 (fused and split 2D
  stages), a non-power-of-two 2D grid, 16^3 and 64^3 -- equal the oracle's
  right-hand side and 3 steps (state and dt sequence);
* the whole 64-member ensemble, advanced together for 3 steps, equals each
  member's single-grid run bitwise, and its smallest 2D member and one 64^3
  member equal the oracle.
"""

import numpy as np
import pytest
import torch

from oracle import grid as og
from oracle import navier_stokes as ons
from oracle import numerics as on
from oracle import solver as osv
from oracle_setup import same

pytestmark = pytest.mark.gpu

PARAMS = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
          "gravitational_acceleration": [0.0, 0.0, 0.0]}
CFL = 0.4
STEPS = 3
_OSCH = {"riemann_solver": on.rusanov, "reconstructor": on.muscl,
         "slope_limiter": on.monotonized_central, "diffusive_solver": on.centered}


def _cfg(n):
    return {"x1_min": 0.0, "x1_max": 1.0, "nx1": n[0], "x2_min": 0.0, "x2_max": 1.0, "nx2": n[1],
            "x3_min": 0.0, "x3_max": 1.0, "nx3": n[2]}


def _member_start(n, seed):
    """Host state (reference shape, ghosts zero) from the ensemble setup's
    initializer on the host grid."""
    from mfvm.grid.geometry import build_grid_host
    from mfvm.setups import ensemble_kh64
    U = ons.allocate_u(_cfg(n))
    ensemble_kh64.initializer(U, PARAMS, build_grid_host("cartesian", _cfg(n)), {"seed": seed})
    return U


def _oracle_run(n, start, steps):
    U = {k: v.copy() for k, v in start.items()}
    g = og.build_grid("cartesian", _cfg(n))
    bcs = {k: {s: og.periodic for s in og.SIDES} for k in ons.VARS}
    rhs = ons.compute_rhs(_filled(U, bcs), g, PARAMS, _OSCH)
    dts = osv.advance(ons, U, g, PARAMS, _OSCH, on.rk3, bcs, CFL, steps)
    return rhs, U, dts


def _filled(U, bcs):
    V = {k: v.copy() for k, v in U.items()}
    og.apply_bcs(V, bcs)
    return V


def _gpu_mods():
    from mfvm import solver
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import navier_stokes as ns
    sch = {"riemann_solver": riemann_solvers.rusanov, "reconstructor": reconstructors.muscl,
           "slope_limiter": slope_limiters.monotonized_central,
           "diffusive_solver": diffusive_solvers.centered}
    bcs = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
    return solver, build_grid, ti, ns, sch, bcs


def _gpu_state(ns, n, start):
    U = ns.allocate_u(_cfg(n))
    for k in ns.VARS:
        U[k] = torch.from_numpy(start[k])
    return U


CASES = [((64, 64, 1), "0"), ((64, 64, 1), "1"), ((48, 40, 1), "0"), ((16, 16, 16), None),
         ((64, 64, 64), None)]


@pytest.mark.parametrize("n,split2d", CASES,
                         ids=[f"{'x'.join(map(str, n))}-split2d{s}" for n, s in CASES])
def test_mc_rusanov_rk3_single_grid_matches_oracle(n, split2d, monkeypatch):
    if split2d is not None:
        monkeypatch.setenv("MFVM_SPLIT_2D", split2d)
    solver, build_grid, ti, ns, sch, bcs = _gpu_mods()
    start = _member_start(n, 11)
    rhs_o, U_o, dts_o = _oracle_run(n, start, STEPS)
    grid = build_grid("cartesian", _cfg(n))
    U = _gpu_state(ns, n, start)
    solver.apply_bcs(U, bcs)
    print(f"mc/rusanov/rk3 {n}: rhs", flush=True)
    rhs = ns.compute_rhs(U, grid, PARAMS, sch)
    for k in ns.VARS:
        assert same(rhs[k].cpu().numpy(), rhs_o[k]), k
    print(f"mc/rusanov/rk3 {n}: {STEPS} rk3 steps", flush=True)
    U = _gpu_state(ns, n, start)
    dts = ns.advance(U, grid, PARAMS, sch, bcs, ti.rk3._mfvm_code, CFL, STEPS)
    torch.cuda.synchronize()
    assert dts.cpu().numpy()[:STEPS].tolist() == dts_o
    for k in ns.VARS:
        assert same(U[k].cpu().numpy(), U_o[k]), k


def test_ensemble_kh64_workload_matches_single_runs_and_oracle():
    """The 64-member workload of the ensemble setup, 3 steps together."""
    from mfvm.ensemble import Ensemble
    from mfvm.setups import ensemble_kh64
    solver, build_grid, ti, ns, sch, bcs = _gpu_mods()
    setups = ensemble_kh64.get_setups()
    shapes = ensemble_kh64.member_shapes()
    assert len(setups) == 64 and sum(s[2] == 64 for s in shapes) == 8
    starts, states, grids = [], [], []
    for i, s in enumerate(setups):
        n = tuple(s["grid_config"][f"nx{a}"] for a in (1, 2, 3))
        start = _member_start(n, s["initial_conditions"]["parameters"]["seed"])
        starts.append(start)
        states.append(_gpu_state(ns, n, start))
        grids.append(build_grid("cartesian", s["grid_config"]))
    s0 = setups[0]
    ens = Ensemble(states, grids, s0["physical_parameters"], s0["space_schemes"],
                   [s["boundary_conditions"] for s in setups], s0["time_integrator"])
    print(f"ensemble kh64: 64 members, {STEPS} steps", flush=True)
    dts = ens.advance(s0["time_config"]["CFL"], STEPS)
    torch.cuda.synchronize()
    for i, (n, start) in enumerate(zip(shapes, starts)):
        U = _gpu_state(ns, n, start)
        ref = ns.advance(U, grids[i], PARAMS, sch, bcs, ti.rk3._mfvm_code, CFL, STEPS)
        torch.cuda.synchronize()
        assert torch.equal(dts[i], ref), (i, n)
        assert torch.equal(states[i].buf, U.buf), (i, n)
    small = min(range(56), key=lambda i: (shapes[i][0] * shapes[i][1], i))
    big = shapes.index((64, 64, 64))
    for i in (small, big):
        print(f"ensemble kh64: member {i} {shapes[i]} vs the oracle", flush=True)
        _, U_o, dts_o = _oracle_run(shapes[i], starts[i], STEPS)
        assert dts[i].cpu().numpy()[:STEPS].tolist() == dts_o, i
        for k in ns.VARS:
            assert same(states[i][k].cpu().numpy(), U_o[k]), (i, k)
