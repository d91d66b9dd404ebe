"""Benchmark: 3D Rayleigh-Taylor, compressible Navier-Stokes, 256^3, FP64.

One "step" is one SSPRK3 time step (three stages plus the CFL step)
over the 256^3 grid: MUSCL + van Leer, HLL, centred viscous flux
(mu = 1e-4), gravity g = -0.1 in x3, periodic x1/x2 and reflecting x3 walls
(momentum_x3 reflecting, the other fields Neumann).  The state is resident
in HBM before the timed region; dt stays on the device.

Prints one JSON line (metric Mcell-updates/s) with the roofline of the
dominant kernel (the x1/x2 pass of each stage, ns_xyrow_kernel, timed with
HIP events around each launch inside the timed region, against the
compulsory 320/3 B per cell and stage) and the CPU baseline (the NumPy oracle
on a bounded sample of the same scheme, timed on this host).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "multiphysics-fvm_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Compulsory HBM traffic of one SSPRK3 step per cell (float64, 5 fields):
# stage 1 reads U^n and writes U1 (80 B); stages 2 and 3 read U^s and U^n and
# write (120 B each).  The roofline of a stage kernel is taken against a third
# of it, 320/3 B per cell and stage.
BYTES_PER_CELL_STEP = 320.0
STAGES = 3
# Bytes per cell that the x1/x2 kernel itself must move (read U 5 x 8 B and
# the closure P, T11, T12, T13, T22, T23 6 x 8 B, write the in-plane
# divergence 5 x 8 B): reported beside the roofline, not used for it.
XY_OWN_BYTES_PER_CELL = 128.0
# HBM bytes per launch of the dominant kernel at 256^3 (default tile), from the
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs summarised in the named
# profile (KiB per dispatch, scaled by the 1 GiB copy calibration: read =
# FETCH_SIZE * 1024 * 2.0, written = WRITE_SIZE * 1024 * 1.0).  The profiled
# kernel build is named with it; None where no such profile exists.
PROFILED_TRAFFIC = {
    ("ns_xyrow_kernel", 256): ((1024564 * 2.0 + 655360 * 1.0) * 1024,
                               "profiles/ns_xyrow_kernel_rt256.md: prof_rt4 FETCH_SIZE / WRITE_SIZE runs"),
}


def _shape(n):
    return (n, n, n) if isinstance(n, int) else tuple(int(x) for x in n)


def rt_config(n):
    n1, n2, n3 = _shape(n)
    return {"x1_min": 0.0, "x1_max": 1.0, "nx1": n1, "x2_min": 0.0, "x2_max": 1.0, "nx2": n2,
            "x3_min": 0.0, "x3_max": 1.0, "nx3": n3}


def rt_fields(n, seed=1):
    """Heavy fluid (rho=2) over light (rho=1), hydrostatic p with p=2.5 at z=0.5,
    vz = 0.01 (1+cos 2pi x)(1+cos 2pi y)/4 + U(-1e-3, 1e-3), on the unit cube
    with n^3 cells (n an int) or n1 x n2 x n3 cells (n a triple)."""
    shp = _shape(n)
    rng = np.random.default_rng(seed)
    c = [(np.arange(m) + 0.5) / m for m in shp]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    heavy = Z > 0.5
    rho = np.where(heavy, 2.0, 1.0)
    g = 0.1
    p = np.where(heavy, 2.5 - g * 2.0 * (Z - 0.5), 2.5 + g * 1.0 * (0.5 - Z))
    vz = 0.01 * (1 + np.cos(2 * np.pi * X)) * (1 + np.cos(2 * np.pi * Y)) / 4 \
        + rng.uniform(-1e-3, 1e-3, shp)
    E = p / 0.4 + 0.5 * rho * vz * vz
    z = np.zeros_like(rho)
    return [rho, z, z, rho * vz, E]


PARAMS = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
          "gravitational_acceleration": [0.0, 0.0, -0.1]}
CFL = 0.4


def bc_rows():
    # [left,right,bottom,top,front,back]: periodic x1/x2; x3 walls
    out = []
    for v in range(5):
        out.append([1, 1, 1, 1, 2, 2] if v == 3 else [1, 1, 1, 1, 0, 0])
    return out


def cpu_baseline(target=256, n=120, steps=2):
    """The NumPy oracle on a bounded sample of the same scheme (n^3 cells,
    `steps` steps, about 10 s of one core).  Its cell-update rate is the
    baseline figure for the target^3 workload: the oracle's whole-array
    passes cost the same per cell at either size (a 256^3 step would take
    minutes and ~35 GB), so the rate is carried over per cell-update."""
    from oracle import grid as og
    from oracle import navier_stokes as ons
    from oracle import numerics as on
    from oracle import solver as osv
    cfg = rt_config(n)
    U = ons.allocate_u(cfg)
    for k, f in zip(ons.VARS, rt_fields(n)):
        U[k][1:-1, 1:-1, 1:-1] = f
    rule = {0: og.neumann, 1: og.periodic, 2: og.reflecting}
    bcs = {k: {s: rule[c] for s, c in zip(og.SIDES, row)} for k, row in zip(ons.VARS, bc_rows())}
    og.apply_bcs(U, bcs)
    g = og.build_grid("cartesian", cfg)
    schemes = {"riemann_solver": on.hll, "reconstructor": on.muscl, "slope_limiter": on.van_leer,
               "diffusive_solver": on.centered}
    t0 = time.perf_counter()
    osv.advance(ons, U, g, PARAMS, schemes, on.rk3_ssp, bcs, CFL, steps)
    wall = time.perf_counter() - t0
    return {"value": n ** 3 * steps / wall / 1e6, "unit": "Mcell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"NumPy oracle, RT {n}^3, {steps} SSPRK3 steps (same scheme), "
                      f"{wall:.1f} s wall on 1 host core",
            "extrapolation": f"rate per cell-update measured at {n}^3 and applied to the "
                             f"{target}^3 workload (not timed at {target}^3)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=str, default="")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-compensated leg")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the 1 GiB copies that calibrate the HBM byte counters")
    args = ap.parse_args()
    if args.calibrate:
        calibrate()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    import torch

    from mfvm import _lib
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns

    n = args.n
    cfg = rt_config(n)
    grid = build_grid("cartesian", cfg)
    U = ns.allocate_u(cfg)
    for k, f in zip(ns.VARS, rt_fields(n)):
        U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(f).cuda()
    rule = {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}
    bcs = {k: {s: rule[c] for s, c in zip(bc.SIDES, row)} for k, row in zip(ns.VARS, bc_rows())}
    schemes = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
               "slope_limiter": slope_limiters.van_leer,
               "diffusive_solver": diffusive_solvers.centered}
    tile = tuple(int(x) for x in args.tile.split(",")) if args.tile else None

    def run(k):
        return ns.advance(U, grid, PARAMS, schemes, bcs, _lib.INT_SSPRK3, CFL, k, tile=tile)

    run(args.warmup)
    torch.cuda.synchronize()
    _lib.timing(1)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    dts = run(args.steps)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    kms, kcnt = _lib.timing(2)
    _lib.timing(0)
    assert bool(torch.isfinite(dts).all()), "non-finite dt in the timed run"

    cells = n ** 3
    value = cells * args.steps / (ms / 1e3) / 1e6
    split = ns.split_stages()
    names = ("stage_xy" if split else "stage", "closure", "bc", "stage_z" if split else "other")
    stage_ms = kms[_lib.TK_STAGE] / max(kcnt[_lib.TK_STAGE], 1)
    z_ms = kms[_lib.TK_OTHER] / max(kcnt[_lib.TK_OTHER], 1) if split else 0.0
    kernel = "ns_xyrow_kernel" if split else "ns_stage_kernel"
    # roofline of the dominant kernel against the compulsory 320/3 B per cell
    # and stage (one launch advances every cell through one stage's share)
    stage_bytes = BYTES_PER_CELL_STEP / STAGES * cells
    achieved = stage_bytes / (stage_ms / 1e3) / 1e9
    prof = PROFILED_TRAFFIC.get((kernel, n)) if tile is None else None
    step_frac = value * 1e6 * BYTES_PER_CELL_STEP / (HBM_PEAK_GBS * 1e9)
    out = {
        "metric": "Mcell-updates/s", "value": round(value, 2), "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D Rayleigh-Taylor compressible NS {n}^3, MUSCL+van Leer, HLL, "
                               "SSPRK3, mu=1e-4, g=-0.1 x3, periodic x1/x2, reflecting x3",
                   "cells": cells, "integrator": "SSPRK3", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": prof[0] if prof else None,
                     "traffic_profile": prof[1] if prof else None,
                     "kernel": kernel, "avg_launch_ms": round(stage_ms, 4),
                     "bytes_per_cell": round(BYTES_PER_CELL_STEP / STAGES, 3),
                     "bytes_basis": "compulsory 320 B per SSPRK3 cell-update / 3 stages",
                     "step_frac": round(step_frac, 4),
                     "own_bytes_per_cell": XY_OWN_BYTES_PER_CELL if split else None,
                     "own_bytes_frac": round(XY_OWN_BYTES_PER_CELL * cells / (stage_ms / 1e3) / 1e9
                                             / HBM_PEAK_GBS, 4) if split else None},
        # the whole stage (both kernels when split) against the same 320/3 B
        "stage_roofline": {"ms": round(stage_ms + z_ms, 4),
                           "achieved": round(stage_bytes / ((stage_ms + z_ms) / 1e3) / 1e9, 1),
                           "unit": "GB/s"},
        "step_hbm_frac": round(step_frac, 4),
        "kernel_ms_per_step": {nm: round(kms[i] / args.steps, 4) for i, nm in enumerate(names)},
    }
    if not args.no_fp32:
        # FP32 reconstruction/fluxes with the FP64 compensated accumulation,
        # same workload, same timing protocol, same number of steps from the
        # same initial state as the FP64 leg (warm-up included), so that the
        # two final states are comparable field by field.
        U64 = ns.allocate_u(cfg)
        for k, f in zip(ns.VARS, rt_fields(n)):
            U64[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(f).cuda()
        Uc = ns.to_compensated(U64)
        del U64

        def runc(k):
            return ns.advance_compensated(Uc, grid, PARAMS, schemes, bcs, _lib.INT_SSPRK3, CFL, k,
                                          tile=tile)

        runc(args.warmup)
        torch.cuda.synchronize()
        _lib.timing(1)
        ev0.record()
        runc(args.steps)
        ev1.record()
        torch.cuda.synchronize()
        msc = ev0.elapsed_time(ev1)
        kmc, _ = _lib.timing(2)
        _lib.timing(0)
        vc = cells * args.steps / (msc / 1e3) / 1e6
        err = {}
        val = ns.compensated_value(Uc)
        for v, name in enumerate(ns.VARS):
            d = (val[v, 1:-1, 1:-1, 1:-1] - U.buf[v, 1:-1, 1:-1, 1:-1]).abs()
            err[name] = {"L1": float(d.mean()), "Linf": float(d.max())}
        del val
        out["fp32_compensated"] = {"value": round(vc, 2), "unit": "Mcell-updates/s",
                                   "ms_per_step": round(msc / args.steps, 4),
                                   "speedup_vs_f64": round(vc / value, 3),
                                   "error_vs_f64": err,
                                   "error_after_steps": args.warmup + args.steps,
                                   # state traffic per cell-step (FP32 words, 5 fields):
                                   # stage 1 reads hi, writes d and the next state (60 B);
                                   # stage 2 reads the state, hi and d, writes d and the
                                   # next state (100 B); the last stage reads the state,
                                   # hi, lo and d and writes hi and lo (120 B)
                                   "bytes_per_cell_step": 280,
                                   "kernel_ms_per_step": {nm: round(kmc[i] / args.steps, 4)
                                                          for i, nm in enumerate(names)}}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(out), flush=True)


def calibrate():
    """Plain device copies of a known size (1 GiB each, 3 of each kind): the
    reference kernels for scaling the FETCH_SIZE / WRITE_SIZE counters into
    bytes (run under rocprofv3 --pmc)."""
    import torch
    m = 1 << 27  # 128 Mi float64 = 1 GiB
    a = torch.ones(m, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)  # __amd_rocclr_copyBuffer: reads 1 GiB, writes 1 GiB
    for _ in range(3):
        torch.add(a, a, out=b)  # elementwise: reads a (1 GiB), writes 1 GiB
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "1 GiB device copies", "bytes": a.numel() * 8, "repeats": 3}),
          flush=True)

if __name__ == "__main__":
    main()
