"""Per-configuration measurements beside bench.py (one JSON line per config).

  kh128    2D Kelvin-Helmholtz 128x128x1, MUSCL+minmod, HLL, SSPRK3, CFL 0.4,
           200 steps, FP64 (with the NumPy oracle's wall time for the same run
           unless --no-cpu).
  maxwell  Maxwell pulse on a spherical grid 128x128x256 (r in [0.1,1],
           theta in [0.05, pi-0.05], phi in [0, 2pi)), MUSCL+van Leer, Rusanov,
           RK4, 300 steps, FP64 and FP32 (error of FP32 vs FP64 reported).
  heat     heat diffusion on a cylindrical grid 256x128x128, kappa=1e-2, hot
           disc T=2 in T=1, piecewise-constant + centred flux, Euler, 300 steps,
           FP64 and FP32.
  ensemble 64 Kelvin-Helmholtz instances: 56 2D grids with nx1, nx2 drawn from
           {64,128,256,512} (seed 2) and eight 64^3 grids; per instance (seed
           3+i) shear U[0.3,0.7], density ratio U[1.5,3], phase U[0,2pi);
           MUSCL+MC, Rusanov, RK3, mu=1e-4, periodic, 500 steps; autotuned
           launch shape, graph replay; FP64 and FP32-compensated, against
           advancing the members one by one.
  diagnostics  mfvm_diagnostics on the Rayleigh-Taylor 256^3 state (FP64
           alone, FP64 against a reference state, FP32-compensated against it):
           kernel time and HBM rate.

Usage: python bench_configs.py [kh128|maxwell|heat|ensemble|diagnostics|all] [--steps N] [--no-cpu]
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "multiphysics-fvm_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402


def _cfg(n, lo, hi):
    return {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1],
            "nx2": n[1], "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}


def _time(fn, torch):
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def kh128(args, torch):
    from mfvm import _lib
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    n = (128, 128, 1)
    cfg = _cfg(n, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    rng = np.random.default_rng(0)
    x = (np.arange(128) + 0.5) / 128
    X, Y = np.meshgrid(x, x, indexing="ij")
    X, Y = X[:, :, None], Y[:, :, None]
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    vy = 0.01 * np.sin(4 * np.pi * X) + rng.uniform(-1e-3, 1e-3, n)
    fields = [rho, rho * vx, rho * vy, np.zeros(n), 2.5 / 0.4 + 0.5 * rho * (vx * vx + vy * vy)]
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    grid = build_grid("cartesian", cfg)
    U = ns.allocate_u(cfg)
    for k, f in zip(ns.VARS, fields):
        U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(f).cuda()
    bcs = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
    schemes = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
               "slope_limiter": slope_limiters.minmod, "diffusive_solver": diffusive_solvers.centered}
    steps = args.steps or 200
    ns.advance(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 2)  # warm-up on a copy below
    for k, f in zip(ns.VARS, fields):
        U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(f).cuda()
    ms = _time(lambda: ns.advance(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, steps), torch)
    out = {"config": "kh128", "steps": steps, "ms": round(ms, 3),
           "Mcell_updates_per_s": round(128 * 128 * steps / (ms / 1e3) / 1e6, 2), "dtype": "f64"}
    # The same run replayed from captured hipGraphs (50 steps per graph): the
    # plain scheme through mfvm_closure_advance without a network, which takes
    # the launch overhead of this small grid out of the step.
    from mfvm import closure as cl
    Ug = ns.allocate_u(cfg)
    for k, f in zip(ns.VARS, fields):
        Ug[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(f).cuda()
    Uw = ns.allocate_u(cfg)
    Uw.buf.copy_(Ug.buf)
    cl.advance(Uw, grid, params, schemes, bcs, None, 0.4, 50, graph=50)  # warm-up on a copy
    msg = _time(lambda: cl.advance(Ug, grid, params, schemes, bcs, None, 0.4, steps, graph=50), torch)
    out["graph_replay"] = {"ms": round(msg, 3),
                           "Mcell_updates_per_s": round(128 * 128 * steps / (msg / 1e3) / 1e6, 2),
                           "bitwise_equal_to_direct": bool(torch.equal(Ug.buf, U.buf))}
    if not args.no_cpu:
        from oracle import grid as og
        from oracle import navier_stokes as ons
        from oracle import numerics as on
        from oracle import solver as osv
        Uo = ons.allocate_u(cfg)
        for k, f in zip(ons.VARS, fields):
            Uo[k][1:-1, 1:-1, 1:-1] = f
        bo = {k: {s: og.periodic for s in og.SIDES} for k in ons.VARS}
        og.apply_bcs(Uo, bo)
        so = {"riemann_solver": on.hll, "reconstructor": on.muscl, "slope_limiter": on.minmod,
              "diffusive_solver": on.centered}
        t0 = time.perf_counter()
        osv.advance(ons, Uo, og.build_grid("cartesian", cfg), params, so, on.rk3_ssp, bo, 0.4, steps)
        wall = time.perf_counter() - t0
        same = all(np.array_equal(U[k].cpu().numpy(), Uo[k]) for k in ns.VARS)
        out.update({"cpu_oracle_s": round(wall, 2), "cpu_cores": 1, "bitwise_equal_to_oracle": same,
                    "speedup_vs_cpu": round(wall * 1e3 / ms, 1)})
    return out


def maxwell(args, torch):
    from mfvm import _lib
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import maxwell as mx
    n = (128, 128, 256)
    cfg = _cfg(n, (0.1, 0.05, 0.0), (1.0, math.pi - 0.05, 2 * math.pi))
    grid = build_grid("spherical", cfg)
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
              "electrical_conductivity": 0.0, "charge_density": 0.0,
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 1e-3}
    r = 0.1 + (np.arange(n[0]) + 0.5) * (0.9 / n[0])
    th = 0.05 + (np.arange(n[1]) + 0.5) * ((math.pi - 0.1) / n[1])
    R, TH = np.meshgrid(r, th, indexing="ij")
    Ez = np.exp(-((R - 0.5) / 0.05) ** 2)
    er = (Ez * np.cos(TH))[:, :, None] * np.ones(n)
    eth = (-Ez * np.sin(TH))[:, :, None] * np.ones(n)
    schemes = {"riemann_solver": riemann_solvers.rusanov, "reconstructor": reconstructors.muscl,
               "slope_limiter": slope_limiters.van_leer, "diffusive_solver": diffusive_solvers.centered}
    bcs = {k: {"left": bc.neumann, "right": bc.neumann, "bottom": bc.neumann, "top": bc.neumann,
               "front": bc.periodic, "back": bc.periodic} for k in mx.VARS}
    steps = args.steps or 300
    res = {}
    for dt_name, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        U = mx.allocate_u(cfg, dtype=dtype)
        U["electric_d_field_x1"][1:-1, 1:-1, 1:-1] = torch.from_numpy(er).to(dtype).cuda()
        U["electric_d_field_x2"][1:-1, 1:-1, 1:-1] = torch.from_numpy(eth).to(dtype).cuda()
        ms = _time(lambda: mx.advance(U, grid, params, schemes, bcs, _lib.INT_RK4, 0.4, steps), torch)
        res[dt_name] = (ms, U.buf.to(torch.float64))
    d = (res["f32"][1] - res["f64"][1])[:, 1:-1, 1:-1, 1:-1].abs()
    cells = n[0] * n[1] * n[2]
    return {"config": "maxwell_spherical_128x128x256_rk4", "steps": steps,
            "f64_ms": round(res["f64"][0], 2), "f32_ms": round(res["f32"][0], 2),
            "f64_Mcell_updates_per_s": round(cells * steps / res["f64"][0] * 1e-3, 1),
            "f32_Mcell_updates_per_s": round(cells * steps / res["f32"][0] * 1e-3, 1),
            "f32_vs_f64_L1": float(d.mean()), "f32_vs_f64_Linf": float(d.max())}


def heat(args, torch):
    from mfvm import _lib
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import heat as ht
    n = (256, 128, 128)
    cfg = _cfg(n, (0.0, 0.0, 0.0), (1.0, 2 * math.pi, 1.0))
    grid = build_grid("cylindrical", cfg)
    params = {"mass_density": 1.0, "specific_heat_capacity": 1.0, "thermal_conductivity": 1e-2,
              "volumetric_heat_source": 0.0}
    r = (np.arange(n[0]) + 0.5) / n[0]
    T0 = np.where(r < 0.25, 2.0, 1.0)[:, None, None] * np.ones(n)
    schemes = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.piecewise_constant,
               "slope_limiter": slope_limiters.minmod, "diffusive_solver": diffusive_solvers.centered}
    bcs = {"internal_energy_density": {"left": bc.neumann, "right": bc.neumann, "bottom": bc.periodic,
                                       "top": bc.periodic, "front": bc.neumann, "back": bc.neumann}}
    steps = args.steps or 300
    res = {}
    for dt_name, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        U = ht.allocate_u(cfg, dtype=dtype)
        U["internal_energy_density"][1:-1, 1:-1, 1:-1] = torch.from_numpy(T0).to(dtype).cuda()
        ms = _time(lambda: ht.advance(U, grid, params, schemes, bcs, _lib.INT_EULER, 0.4, steps), torch)
        res[dt_name] = (ms, U.buf.to(torch.float64))
    d = (res["f32"][1] - res["f64"][1])[:, 1:-1, 1:-1, 1:-1].abs()
    cells = n[0] * n[1] * n[2]
    return {"config": "heat_cylindrical_256x128x128_euler", "steps": steps,
            "f64_ms": round(res["f64"][0], 2), "f32_ms": round(res["f32"][0], 2),
            "f64_Mcell_updates_per_s": round(cells * steps / res["f64"][0] * 1e-3, 1),
            "f32_Mcell_updates_per_s": round(cells * steps / res["f32"][0] * 1e-3, 1),
            "f64_hbm_frac_16B": round(cells * steps * 16 / (res["f64"][0] * 1e-3) / 8e12, 4),
            "f32_vs_f64_L1": float(d.mean()), "f32_vs_f64_Linf": float(d.max())}


def ensemble_members(seed=2):
    """Sizes of the 64-member Kelvin-Helmholtz ensemble."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 128, 256, 512], size=(56, 2))
    return [(int(a), int(b), 1) for a, b in sizes] + [(64, 64, 64)] * 8


def kh_instance(n, seed):
    """Kelvin-Helmholtz instance: shear U[0.3,0.7], density ratio U[1.5,3],
    perturbation phase U[0,2pi), 1e-3 velocity noise (vz only in 3D)."""
    rng = np.random.default_rng(seed)
    shear = rng.uniform(0.3, 0.7)
    ratio = rng.uniform(1.5, 3.0)
    phase = rng.uniform(0.0, 2 * math.pi)
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


def ensemble(args, torch):
    from mfvm.ensemble import Ensemble
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import navier_stokes as ns
    shapes = ensemble_members()
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    schemes = {"riemann_solver": riemann_solvers.rusanov, "reconstructor": reconstructors.muscl,
               "slope_limiter": slope_limiters.monotonized_central,
               "diffusive_solver": diffusive_solvers.centered}
    bcs = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
    steps = args.steps or 500
    grids, fields = [], []
    for i, n in enumerate(shapes):
        grids.append(build_grid("cartesian", _cfg(n, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))))
        fields.append(kh_instance(n, 3 + i))

    def states():
        out = []
        for n, f in zip(shapes, fields):
            U = ns.allocate_u({"nx1": n[0], "nx2": n[1], "nx3": n[2]})
            for k, a in zip(ns.VARS, f):
                U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(np.ascontiguousarray(a)).cuda()
            out.append(U)
        return out

    cells = sum(a * b * c for a, b, c in shapes)
    out = {"config": "ensemble_kh64_rk3", "members": len(shapes), "cells": cells, "steps": steps}
    for leg in ("f64", "f32_compensated"):
        comp = leg != "f64"
        S = states()
        if comp:
            S = [ns.to_compensated(U) for U in S]
        ens = Ensemble(S, grids, params, schemes, bcs, ti.rk3, compensated=comp)
        t0 = time.perf_counter()
        best = ens.autotune(0.4, steps=3)
        tune_s = time.perf_counter() - t0
        ms = _time(lambda: ens.advance(0.4, steps), torch)
        # the autotuner's timed trials of the persistent launch (k workgroups
        # per CU over the whole work list) against one workgroup per tile,
        # with the other options as chosen
        pers = {}
        for o, t in ens.tuning_log:
            if all(o.get(k) == best[k] for k in ("split", "order")) and o["tiles"] == best["tiles"]:
                pers[f"persist={o['persist']},remap={o['remap']}"] = round(t, 3)
        out[leg] = {"ms": round(ms, 2), "Mcell_updates_per_s": round(cells * steps / ms * 1e-3, 1),
                    "autotune_s": round(tune_s, 2), "trials": len(ens.tuning_log),
                    "autotune_ms_per_step": pers,
                    "options": {"split": best["split"], "order": best["order"], "persist": best["persist"],
                                "remap": best["remap"], "graph": best["graph"],
                                "tiles": {"x".join(map(str, k)): list(v) for k, v in best["tiles"].items()}}}
        del ens, S
    # the same members advanced one at a time (single-grid launches), on a
    # bounded number of steps
    seq_steps = min(steps, 20)
    S = states()
    rk3 = ti.rk3._mfvm_code
    ms = _time(lambda: [ns.advance(U, g, params, schemes, bcs, rk3, 0.4, seq_steps)
                        for U, g in zip(S, grids)], torch)
    seq = cells * seq_steps / ms * 1e-3
    out["one_by_one"] = {"steps": seq_steps, "ms": round(ms, 2), "Mcell_updates_per_s": round(seq, 1)}
    out["batching_speedup_f64"] = round(out["f64"]["Mcell_updates_per_s"] / seq, 2)
    return out


def diagnostics(args, torch):
    """State diagnostics (mfvm_diagnostics) on the Rayleigh-Taylor 256^3 state:
    kernel time per call and HBM rate over the compulsory bytes (each
    interior value read once)."""
    import bench
    from mfvm import _lib
    from mfvm.diagnostics import raw
    from mfvm.grid.geometry import build_grid
    from mfvm.physics import navier_stokes as ns
    n = 256
    cfg = bench.rt_config(n)
    grid = build_grid("cartesian", cfg)
    U = ns.allocate_u(cfg)
    for k, f in zip(ns.VARS, bench.rt_fields(n)):
        U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(f).cuda()
    ref = ns.allocate_u(cfg)
    ref.buf.copy_(U.buf)
    Uc = ns.to_compensated(U)
    lo = Uc.scratch("lo", tuple(Uc.buf.shape))
    cells = n ** 3
    calls = args.steps or 50
    cases = [("f64_pressure", U, dict(gamma=1.4), 40),
             ("f64_vs_ref", U, dict(ref=ref, gamma=1.4), 80),
             ("f32_compensated_vs_ref", Uc, dict(ref=ref, gamma=1.4, lo=lo), 80)]
    out = {"config": "diagnostics_rt256", "cells": cells, "calls": calls}
    for name, X, kw, bpc in cases:
        raw(X, grid, **kw)
        torch.cuda.synchronize()
        _lib.timing(1)
        ms = _time(lambda: [raw(X, grid, **kw) for _ in range(calls)], torch)
        kms, kcnt = _lib.timing(2)
        _lib.timing(0)
        k_ms = kms[_lib.TK_OTHER] / max(kcnt[_lib.TK_OTHER], 1)
        out[name] = {"bytes_per_cell": bpc, "kernel_ms": round(k_ms, 4),
                     "call_ms": round(ms / calls, 4),
                     "GB_per_s": round(bpc * cells / (k_ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(bpc * cells / (k_ms * 1e-3) / 8e12, 3)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", nargs="?", default="all")
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    import torch
    todo = ["kh128", "maxwell", "heat", "ensemble", "diagnostics"] if args.which == "all" else [args.which]
    for name in todo:
        print(json.dumps(globals()[name](args, torch)), flush=True)


if __name__ == "__main__":
    main()
