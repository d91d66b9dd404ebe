"""Write golden vectors for the FV right-hand-side / time-step path.

Runs the reference implementation (multiphysics-fvm, read-only at
/root/reference) on small seeded inputs and stores inputs and outputs as
``.npz`` fixtures next to this script.  Only this script touches the
reference; the fixtures are plain arrays (no reference code, no pickles).

Usage (in the build container only):  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.dont_write_bytecode = True
sys.path.insert(0, REF)

from src.grid import boundary_conditions as bc  # noqa: E402
from src.grid import geometry  # noqa: E402
from src.numerics import (diffusive_solvers, reconstructors, riemann_solvers,  # noqa: E402
                          slope_limiters, time_integrators)
from src.physics import heat, maxwell, navier_stokes  # noqa: E402
from src import solver  # noqa: E402

LIMITERS = {"minmod": slope_limiters.minmod, "mc": slope_limiters.monotonized_central,
            "vanleer": slope_limiters.van_leer}
RIEMANN = {"hll": riemann_solvers.hll, "rusanov": riemann_solvers.rusanov}
INTEG = {"euler": time_integrators.euler_explicit, "rk3": time_integrators.rk3,
         "ssprk3": time_integrators.rk3_ssp, "rk4": time_integrators.rk4}
BC = {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}
SIDES = ("left", "right", "bottom", "top", "front", "back")
NS_VARS = ("mass_density", "momentum_x1", "momentum_x2", "momentum_x3", "energy_density")
MX_VARS = ("magnetic_field_x1", "magnetic_field_x2", "magnetic_field_x3",
           "electric_d_field_x1", "electric_d_field_x2", "electric_d_field_x3")


def grid_config(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    return {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0],
            "x2_min": lo[1], "x2_max": hi[1], "nx2": n[1],
            "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}


def make_bcs(names, codes):
    """codes: array [nvar, 6] of 0 neumann / 1 periodic / 2 reflecting."""
    return {v: {s: BC[int(codes[i, j])] for j, s in enumerate(SIDES)} for i, v in enumerate(names)}


def ns_state(n, seed, kind="noisy"):
    rng = np.random.default_rng(seed)
    U = navier_stokes.allocate_u(grid_config(n))
    ing = (slice(1, -1),) * 3
    shp = tuple(n)
    if kind == "kh":
        g = geometry.build_grid("cartesian", grid_config(n))
        x = g["cell"]["coordinates"]["x1"]
        y = g["cell"]["coordinates"]["x2"]
        inner = np.abs(y - 0.5) < 0.25
        rho = np.where(inner, 2.0, 1.0)
        vx = np.where(inner, 0.5, -0.5)
        vy = 0.01 * np.sin(4 * np.pi * x) + rng.uniform(-1e-3, 1e-3, shp)
        vz = np.zeros(shp)
        p = np.full(shp, 2.5)
    else:
        rho = 1.0 + 0.5 * rng.random(shp)
        vx = 0.4 * (rng.random(shp) - 0.5)
        vy = 0.4 * (rng.random(shp) - 0.5)
        vz = 0.4 * (rng.random(shp) - 0.5) if n[2] > 1 else np.zeros(shp)
        p = 1.0 + 0.5 * rng.random(shp)
    U["mass_density"][ing] = rho
    U["momentum_x1"][ing] = rho * vx
    U["momentum_x2"][ing] = rho * vy
    U["momentum_x3"][ing] = rho * vz
    U["energy_density"][ing] = p / 0.4 + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    return U


def pack(prefix, d, out):
    for k, v in d.items():
        out[f"{prefix}{k}"] = np.array(v, copy=True)


def ns_cases():
    cases = []
    # 2D periodic KH-like, all limiter x Riemann combinations, viscous.
    for lim in LIMITERS:
        for rs in RIEMANN:
            cases.append(dict(name=f"ns_cart2d_{lim}_{rs}", coord="cartesian", n=(12, 10, 1),
                              kind="kh", seed=11, lim=lim, rs=rs, recon="muscl", mu=1e-4,
                              g=(0.0, 0.0, 0.0), bcs=np.ones((5, 6), int), integ="ssprk3", steps=3))
    # 3D noisy, mixed BCs, gravity, viscous; each integrator.
    rt_bcs = np.array([[1, 1, 1, 1, 0, 0], [1, 1, 1, 1, 0, 0], [1, 1, 1, 1, 0, 0],
                       [1, 1, 1, 1, 2, 2], [1, 1, 1, 1, 0, 0]])
    for integ in INTEG:
        cases.append(dict(name=f"ns_cart3d_vanleer_hll_{integ}", coord="cartesian", n=(8, 6, 5),
                          kind="noisy", seed=21, lim="vanleer", rs="hll", recon="muscl", mu=2e-3,
                          g=(0.0, 0.0, -0.1), bcs=rt_bcs, integ=integ, steps=2))
    cases.append(dict(name="ns_cart3d_mc_rusanov_pc", coord="cartesian", n=(7, 6, 5), kind="noisy",
                      seed=22, lim="mc", rs="rusanov", recon="pc", mu=0.0, g=(0.05, 0.0, 0.0),
                      bcs=np.zeros((5, 6), int), integ="rk3", steps=2))
    cases.append(dict(name="ns_cart3d_minmod_hll_nonpow2", coord="cartesian", n=(6, 7, 5),
                      kind="noisy", seed=23, lim="minmod", rs="hll", recon="muscl", mu=1e-3,
                      g=(0.0, -0.2, 0.0), bcs=rt_bcs, integ="ssprk3", steps=2,
                      lo=(0.0, -0.3, 0.1), hi=(0.7, 1.1, 0.6)))
    cyl_bcs = np.array([[0, 0, 1, 1, 0, 0]] * 5)
    cases.append(dict(name="ns_cyl_vanleer_hll", coord="cylindrical", n=(8, 6, 4), kind="noisy",
                      seed=31, lim="vanleer", rs="hll", recon="muscl", mu=1e-3, g=(0.0, 0.0, 0.0),
                      bcs=cyl_bcs, integ="ssprk3", steps=2, lo=(0.2, 0.0, 0.0), hi=(1.0, 2 * np.pi, 1.0)))
    sph_bcs = np.array([[0, 0, 0, 0, 1, 1]] * 5)
    cases.append(dict(name="ns_sph_mc_rusanov", coord="spherical", n=(8, 6, 4), kind="noisy",
                      seed=32, lim="mc", rs="rusanov", recon="muscl", mu=1e-3, g=(0.0, 0.0, 0.0),
                      bcs=sph_bcs, integ="rk4", steps=1, lo=(0.3, 0.2, 0.0), hi=(1.0, np.pi - 0.2, 2 * np.pi)))
    cases.append(dict(name="ns_kh32_ssprk3_steps", coord="cartesian", n=(32, 32, 1), kind="kh", seed=0,
                      lim="minmod", rs="hll", recon="muscl", mu=1e-4, g=(0.0, 0.0, 0.0),
                      bcs=np.ones((5, 6), int), integ="ssprk3", steps=10))
    return cases


def run_ns(c):
    gc = grid_config(c["n"], c.get("lo", (0.0, 0.0, 0.0)), c.get("hi", (1.0, 1.0, 1.0)))
    grid = geometry.build_grid(c["coord"], gc)
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": c["mu"],
              "gravitational_acceleration": list(c["g"])}
    schemes = {"riemann_solver": RIEMANN[c["rs"]],
               "reconstructor": reconstructors.muscl if c["recon"] == "muscl" else reconstructors.piecewise_constant,
               "slope_limiter": LIMITERS[c["lim"]], "diffusive_solver": diffusive_solvers.centered}
    bcs = make_bcs(NS_VARS, c["bcs"])
    U = ns_state(c["n"], c["seed"], c["kind"])
    solver.apply_bcs(U, bcs)
    out = {"bcs": c["bcs"], "n": np.array(c["n"]), "mu": c["mu"], "g": np.array(c["g"]),
           "lo": np.array([gc["x1_min"], gc["x2_min"], gc["x3_min"]]),
           "hi": np.array([gc["x1_max"], gc["x2_max"], gc["x3_max"]]),
           "meta": np.array([c["coord"], c["lim"], c["rs"], c["recon"], c["integ"], str(c["steps"])])}
    pack("U_in/", U, out)
    rhs = navier_stokes.compute_rhs(U, grid, params, schemes)
    pack("rhs/", rhs, out)
    out["dt"] = navier_stokes.compute_dt(U, params, grid, 0.4)
    dts = []
    for _ in range(c["steps"]):
        dt = navier_stokes.compute_dt(U, params, grid, 0.4)
        INTEG[c["integ"]](navier_stokes.compute_rhs, U, dt, params, grid, schemes, bcs, solver.apply_bcs)
        dts.append(dt)
    out["dts"] = np.array(dts)
    pack("U_out/", U, out)
    return out


def run_heat():
    gc = grid_config((8, 6, 4), (0.1, 0.0, 0.0), (1.0, 2 * np.pi, 1.0))
    grid = geometry.build_grid("cylindrical", gc)
    params = {"mass_density": 1.0, "specific_heat_capacity": 1.0, "thermal_conductivity": 1e-2,
              "volumetric_heat_source": 0.05}
    rng = np.random.default_rng(41)
    U = heat.allocate_u(gc)
    U["internal_energy_density"][1:-1, 1:-1, 1:-1] = 1.0 + rng.random((8, 6, 4))
    codes = np.array([[0, 0, 1, 1, 0, 0]])
    bcs = make_bcs(("internal_energy_density",), codes)
    solver.apply_bcs(U, bcs)
    out = {"bcs": codes}
    pack("U_in/", U, out)
    out_rhs = {}
    for lim in ("pc", "vanleer"):
        schemes = {"riemann_solver": None, "slope_limiter": LIMITERS["vanleer"],
                   "reconstructor": reconstructors.piecewise_constant if lim == "pc" else reconstructors.muscl,
                   "diffusive_solver": diffusive_solvers.centered}
        out_rhs[lim] = heat.compute_rhs(U, grid, params, schemes)
        pack(f"rhs_{lim}/", out_rhs[lim], out)
    out["dt"] = heat.compute_dt(U, params, grid, 0.4)
    schemes = {"riemann_solver": None, "slope_limiter": LIMITERS["vanleer"],
               "reconstructor": reconstructors.piecewise_constant,
               "diffusive_solver": diffusive_solvers.centered}
    for _ in range(3):
        dt = heat.compute_dt(U, params, grid, 0.4)
        time_integrators.euler_explicit(heat.compute_rhs, U, dt, params, grid, schemes, bcs, solver.apply_bcs)
    pack("U_out/", U, out)
    return out


def run_maxwell():
    gc = grid_config((8, 8, 8), (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi))
    grid = geometry.build_grid("spherical", gc)
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
              "electrical_conductivity": 0.1, "charge_density": 0.0,
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 2e-3}
    rng = np.random.default_rng(51)
    U = maxwell.allocate_u(gc)
    for v in MX_VARS:
        U[v][1:-1, 1:-1, 1:-1] = rng.standard_normal((8, 8, 8)) * 0.1
    codes = np.array([[0, 0, 0, 0, 1, 1]] * 6)
    bcs = make_bcs(MX_VARS, codes)
    solver.apply_bcs(U, bcs)
    out = {"bcs": codes}
    pack("U_in/", U, out)
    for rs in ("rusanov", "hll"):
        schemes = {"riemann_solver": RIEMANN[rs], "reconstructor": reconstructors.muscl,
                   "slope_limiter": LIMITERS["vanleer"], "diffusive_solver": None}
        pack(f"rhs_{rs}/", maxwell.compute_rhs(U, grid, params, schemes), out)
    out["dt"] = maxwell.compute_dt(U, params, grid, 0.4)
    schemes = {"riemann_solver": RIEMANN["rusanov"], "reconstructor": reconstructors.muscl,
               "slope_limiter": LIMITERS["vanleer"], "diffusive_solver": None}
    dt = maxwell.compute_dt(U, params, grid, 0.4)
    time_integrators.rk4(maxwell.compute_rhs, U, dt, params, grid, schemes, bcs, solver.apply_bcs)
    pack("U_out/", U, out)
    return out


def run_grids():
    out = {}
    cfgs = {"cyl": ("cylindrical", grid_config((5, 4, 3), (0.1, 0.0, -1.0), (2.0, 2 * np.pi, 1.0))),
            "sph": ("spherical", grid_config((5, 4, 3), (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi))),
            "cart": ("cartesian", grid_config((5, 4, 3), (0.0, -0.3, 0.1), (0.7, 1.1, 0.6)))}
    for tag, (cs, gc) in cfgs.items():
        g = geometry.build_grid(cs, gc)
        out[f"{tag}/V"] = g["cell"]["geometry"]["volume"]
        for d in ("x1", "x2", "x3"):
            out[f"{tag}/S{d}"] = g["cell"]["geometry"]["surface"][d]
            out[f"{tag}/L{d}"] = g["cell"]["geometry"]["length"][d]
            out[f"{tag}/h{d}"] = g["lame"]["coefficients"][d]
            for k, v in g["lame"]["derivatives"][d].items():
                out[f"{tag}/d{d}_{k}"] = v
    return out


def run_bcs():
    rng = np.random.default_rng(61)
    out = {}
    for code, name in ((0, "neumann"), (1, "periodic"), (2, "reflecting")):
        f = rng.standard_normal((5, 4, 3))
        out[f"{name}/in"] = f.copy()  # apply_bcs mutates f
        solver.apply_bcs({"q": f}, {"q": {s: BC[code] for s in SIDES}})
        out[f"{name}/out"] = f
    return out


def main():
    for c in ns_cases():
        np.savez_compressed(os.path.join(HERE, f"{c['name']}.npz"), **run_ns(c))
    np.savez_compressed(os.path.join(HERE, "heat_cyl.npz"), **run_heat())
    np.savez_compressed(os.path.join(HERE, "maxwell_sph.npz"), **run_maxwell())
    np.savez_compressed(os.path.join(HERE, "grids.npz"), **run_grids())
    np.savez_compressed(os.path.join(HERE, "bcs.npz"), **run_bcs())
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
