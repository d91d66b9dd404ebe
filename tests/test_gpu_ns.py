"""Navier-Stokes HIP path vs the reference golden vectors and the oracle.

Bar: bit-exact (every value ==, NaN == NaN) for the float64 path; the
float64 tolerance of the workload (1e-12 relative L-inf after one step,
1e-10 after 200) is asserted as well where a long run is compared.
"""

import os

import numpy as np
import pytest
import torch

from oracle import grid as og
from oracle import navier_stokes as ons
from oracle import numerics as on
from oracle_setup import group, golden_files, load, make_bcs, ns_setup, same

pytestmark = pytest.mark.gpu

NS_FILES = golden_files("ns_*.npz")


def _mods():
    from mfvm import solver
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import (diffusive_solvers, reconstructors, riemann_solvers,
                               slope_limiters, time_integrators)
    from mfvm.physics import navier_stokes as ns
    return solver, bc, build_grid, diffusive_solvers, reconstructors, riemann_solvers, \
        slope_limiters, time_integrators, ns


def _gpu_setup(z):
    solver, bc, build_grid, ds, rec, rs, sl, ti, ns = _mods()
    s = ns_setup(z)
    schemes = {"riemann_solver": {"hll": rs.hll, "rusanov": rs.rusanov}[s["rs"]],
               "reconstructor": {"muscl": rec.muscl, "pc": rec.piecewise_constant}[s["recon"]],
               "slope_limiter": {"minmod": sl.minmod, "mc": sl.monotonized_central,
                                 "vanleer": sl.van_leer}[s["lim"]],
               "diffusive_solver": ds.centered}
    bmap = {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}
    bcs = {v: {side: bmap[int(z["bcs"][i, j])] for j, side in enumerate(bc.SIDES)}
           for i, v in enumerate(ns.VARS)}
    integ = {"euler": ti.euler_explicit, "rk3": ti.rk3, "ssprk3": ti.rk3_ssp, "rk4": ti.rk4}[s["integ"]]
    grid = build_grid(s["coord"], s["gc"])
    U = ns.allocate_u(s["gc"])
    for k, v in group(z, "U_in/").items():
        U[k] = torch.from_numpy(v)
    return s, schemes, bcs, integ, grid, U


@pytest.mark.parametrize("path", NS_FILES, ids=[os.path.basename(p)[:-4] for p in NS_FILES])
def test_rhs_dt_and_steps_match_reference(path):
    solver, *_, ns = _mods()
    z = load(path)
    s, schemes, bcs, integ, grid, U = _gpu_setup(z)
    print(f"case {os.path.basename(path)} n={s['gc']['nx1']}x{s['gc']['nx2']}x{s['gc']['nx3']}",
          flush=True)
    rhs = ns.compute_rhs(U, grid, s["params"], schemes)
    ref = group(z, "rhs/")
    for k in ns.VARS:
        assert same(rhs[k].cpu().numpy(), ref[k]), k
    assert ns.compute_dt(U, s["params"], grid, 0.4) == float(z["dt"])
    for dt in z["dts"]:
        integ(ns.compute_rhs, U, float(dt), s["params"], grid, schemes, bcs, solver.apply_bcs)
    out = group(z, "U_out/")
    for k in ns.VARS:
        assert same(U[k].cpu().numpy(), out[k]), k


@pytest.mark.parametrize("path", NS_FILES, ids=[os.path.basename(p)[:-4] for p in NS_FILES])
def test_advance_with_device_dt_matches_reference(path):
    solver, *_, ns = _mods()
    z = load(path)
    s, schemes, bcs, integ, grid, U = _gpu_setup(z)
    dts = ns.advance(U, grid, s["params"], schemes, bcs, integ._mfvm_code, 0.4, len(z["dts"]))
    assert np.array_equal(dts.cpu().numpy()[:-1], z["dts"])
    out = group(z, "U_out/")
    for k in ns.VARS:
        assert same(U[k].cpu().numpy(), out[k]), k


def test_apply_bcs_matches_reference():
    solver, bc, build_grid, *_ = _mods()
    from mfvm.state import State
    z = load(os.path.join(os.path.dirname(NS_FILES[0]), "bcs.npz"))
    for code, name in ((0, "neumann"), (1, "periodic"), (2, "reflecting")):
        f = z[f"{name}/in"]
        n = tuple(x - 2 for x in f.shape)
        cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": n[0], "x2_min": 0.0, "x2_max": 1.0,
               "nx2": n[1], "x3_min": 0.0, "x3_max": 1.0, "nx3": n[2]}
        U = State(("q",), cfg)
        U["q"] = torch.from_numpy(f)
        rule = {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}[code]
        solver.apply_bcs(U, {"q": {s: rule for s in bc.SIDES}})
        assert same(U["q"].cpu().numpy(), z[f"{name}/out"]), name


def _kh_state(n, seed=0):
    """Kelvin-Helmholtz initial state of the 128x128x1 configuration."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n[0]) + 0.5) / n[0]
    y = (np.arange(n[1]) + 0.5) / n[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    X, Y = X[:, :, None], Y[:, :, None]
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    vy = 0.01 * np.sin(4 * np.pi * X) + rng.uniform(-1e-3, 1e-3, n)
    p = 2.5
    E = p / 0.4 + 0.5 * rho * (vx * vx + vy * vy)
    return [rho, rho * vx, rho * vy, np.zeros(n), E]


def test_kh128_200_steps_match_oracle():
    """The 2D Kelvin-Helmholtz configuration: 200 SSPRK3 steps, device dt."""
    solver, bc, build_grid, ds, rec, rs, sl, ti, ns = _mods()
    n = (128, 128, 1)
    cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": 128, "x2_min": 0.0, "x2_max": 1.0,
           "nx2": 128, "x3_min": 0.0, "x3_max": 1.0, "nx3": 1}
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    fields = _kh_state(n)
    Uo = ons.allocate_u(cfg)
    for k, v in zip(ons.VARS, fields):
        Uo[k][1:-1, 1:-1, 1:-1] = v
    bo = {k: {s: og.periodic for s in og.SIDES} for k in ons.VARS}
    og.apply_bcs(Uo, bo)
    so = {"riemann_solver": on.hll, "reconstructor": on.muscl, "slope_limiter": on.minmod,
          "diffusive_solver": on.centered}
    go = og.build_grid("cartesian", cfg)

    U = ns.allocate_u(cfg)
    for k in ns.VARS:
        U[k] = torch.from_numpy(Uo[k])
    bg = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
    sg = {"riemann_solver": rs.hll, "reconstructor": rec.muscl, "slope_limiter": sl.minmod,
          "diffusive_solver": ds.centered}
    grid = build_grid("cartesian", cfg)

    # one step: 1e-12 relative L-inf (here: bitwise)
    from oracle import solver as osv
    dts_o = osv.advance(ons, Uo, go, params, so, on.rk3_ssp, bo, 0.4, 1)
    dts = ns.advance(U, grid, params, sg, bg, ti.rk3_ssp._mfvm_code, 0.4, 1)
    assert dts.cpu().numpy()[0] == dts_o[0]
    for k in ns.VARS:
        g, o = U[k].cpu().numpy(), Uo[k]
        assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o)), k
        assert same(g, o), k
    dts_o += osv.advance(ons, Uo, go, params, so, on.rk3_ssp, bo, 0.4, 199)
    dts = ns.advance(U, grid, params, sg, bg, ti.rk3_ssp._mfvm_code, 0.4, 199)
    for k in ns.VARS:
        g, o = U[k].cpu().numpy(), Uo[k]
        assert np.max(np.abs(g - o)) <= 1e-10 * np.max(np.abs(o)), k
        assert same(g, o), k


@pytest.mark.parametrize("coord,lim,rs", [("cylindrical", "minmod", "hll"), ("spherical", "mc", "rusanov"),
                                          ("spherical", "vanleer", "hll")])
def test_curvilinear_tiles_and_chunks_match_oracle(coord, lim, rs):
    """Curvilinear Navier-Stokes on a grid of several x1/x2 tiles and x3
    chunks, partial at every edge (70 x 9 x 37), viscous, with gravity and a
    reflecting wall: the right-hand side (fused stage) and two SSPRK3 steps
    with device dt (split stages) equal the oracle bit for bit."""
    solver, bc, build_grid, ds, rec, rsm, sl, ti, ns = _mods()
    from oracle import solver as osv
    n = (70, 9, 37)
    if coord == "cylindrical":
        lo, hi = (0.1, 0.0, 0.0), (1.0, 2 * np.pi, 1.0)
        sides = {"left": 0, "right": 0, "bottom": 1, "top": 1, "front": 2, "back": 2}
        wall = "momentum_x3"
    else:
        lo, hi = (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi)
        sides = {"left": 0, "right": 0, "bottom": 2, "top": 2, "front": 1, "back": 1}
        wall = "momentum_x2"
    cfg = {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1],
           "nx2": n[1], "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-3,
              "gravitational_acceleration": [-0.1, 0.0, 0.0]}
    so = {"riemann_solver": {"hll": on.hll, "rusanov": on.rusanov}[rs], "reconstructor": on.muscl,
          "slope_limiter": {"minmod": on.minmod, "mc": on.monotonized_central, "vanleer": on.van_leer}[lim],
          "diffusive_solver": on.centered}
    sg = {"riemann_solver": {"hll": rsm.hll, "rusanov": rsm.rusanov}[rs], "reconstructor": rec.muscl,
          "slope_limiter": {"minmod": sl.minmod, "mc": sl.monotonized_central, "vanleer": sl.van_leer}[lim],
          "diffusive_solver": ds.centered}

    def rule(v, code, mods):
        # reflecting only on the wall-normal momentum, neumann for the rest
        if code == 2:
            return mods[2] if v == wall else mods[0]
        return mods[code]

    bo = {v: {s: rule(v, c, {0: og.neumann, 1: og.periodic, 2: og.reflecting}) for s, c in sides.items()}
          for v in ons.VARS}
    bg = {v: {s: rule(v, c, {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}) for s, c in sides.items()}
          for v in ns.VARS}

    def fresh():
        rng = np.random.default_rng(7093)
        Uo = ons.allocate_u(cfg)
        shp = Uo["mass_density"].shape
        rho = 1.0 + rng.random(shp)
        v = [0.3 * (2 * rng.random(shp) - 1) for _ in range(3)]
        p = 1.0 + rng.random(shp)
        Uo["mass_density"][...] = rho
        for a in range(3):
            Uo[f"momentum_x{a + 1}"][...] = rho * v[a]
        Uo["energy_density"][...] = p / 0.4 + 0.5 * rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
        og.apply_bcs(Uo, bo)
        return Uo

    go = og.build_grid(coord, cfg)
    grid = build_grid(coord, cfg)
    Uo = fresh()
    ro = ons.compute_rhs(Uo, go, params, so)
    U = ns.allocate_u(cfg)
    for k in ns.VARS:
        U[k] = torch.from_numpy(Uo[k])
    print(f"ns {coord} {n} {lim} {rs}: rhs", flush=True)
    r = ns.compute_rhs(U, grid, params, sg)
    for k in ns.VARS:
        assert same(r[k].cpu().numpy(), ro[k]), k
    dts_o = osv.advance(ons, Uo, go, params, so, on.rk3_ssp, bo, 0.4, 2)
    U = ns.allocate_u(cfg)
    Uf = fresh()
    for k in ns.VARS:
        U[k] = torch.from_numpy(Uf[k])
    print(f"ns {coord} {n} {lim} {rs}: 2 ssprk3 steps", flush=True)
    dts = ns.advance(U, grid, params, sg, bg, ti.rk3_ssp._mfvm_code, 0.4, 2)
    assert dts.cpu().numpy()[:2].tolist() == list(dts_o)
    for k in ns.VARS:
        assert same(U[k].cpu().numpy(), Uo[k]), k
