"""The CPU oracle reproduces the reference's own outputs bit for bit.

The fixtures in tests/golden/ were written by the reference implementation
(tests/golden/make_golden.py); this file checks the oracle against them.
"""

import os

import numpy as np
import pytest

from oracle import grid as og
from oracle import heat as oh
from oracle import maxwell as om
from oracle import navier_stokes as ons
from oracle import numerics as on
from oracle import solver as osv
from oracle_setup import (GOLDEN, group, golden_files, grid_config, load, make_bcs, ns_setup,
                          same)

NS_FILES = golden_files("ns_*.npz")


@pytest.mark.parametrize("path", NS_FILES, ids=[os.path.basename(p)[:-4] for p in NS_FILES])
def test_ns_rhs_dt_and_steps(path):
    z = load(path)
    s = ns_setup(z)
    grid = og.build_grid(s["coord"], s["gc"])
    U = {k: v.copy() for k, v in group(z, "U_in/").items()}
    rhs = ons.compute_rhs(U, grid, s["params"], s["schemes"])
    ref = group(z, "rhs/")
    for k in ons.VARS:
        assert same(rhs[k], ref[k]), k
    assert ons.compute_dt(U, s["params"], grid, 0.4) == z["dt"]
    bcs = make_bcs(ons.VARS, z["bcs"])
    dts = osv.advance(ons, U, grid, s["params"], s["schemes"],
                      {"euler": on.euler_explicit, "rk3": on.rk3, "ssprk3": on.rk3_ssp,
                       "rk4": on.rk4}[s["integ"]], bcs, 0.4, s["steps"])
    assert np.array_equal(np.array(dts), z["dts"])
    out = group(z, "U_out/")
    for k in ons.VARS:
        assert same(U[k], out[k]), k


def test_heat_cylindrical():
    z = load(os.path.join(GOLDEN, "heat_cyl.npz"))
    gc = grid_config((8, 6, 4), (0.1, 0.0, 0.0), (1.0, 2 * np.pi, 1.0))
    grid = og.build_grid("cylindrical", gc)
    params = {"mass_density": 1.0, "specific_heat_capacity": 1.0, "thermal_conductivity": 1e-2,
              "volumetric_heat_source": 0.05}
    U = {k: v.copy() for k, v in group(z, "U_in/").items()}
    for tag, rec in (("pc", on.piecewise_constant), ("vanleer", on.muscl)):
        schemes = {"riemann_solver": None, "slope_limiter": on.van_leer, "reconstructor": rec,
                   "diffusive_solver": on.centered}
        r = oh.compute_rhs(U, grid, params, schemes)
        assert same(r["internal_energy_density"], z[f"rhs_{tag}/internal_energy_density"])
    assert oh.compute_dt(U, params, grid, 0.4) == z["dt"]
    schemes = {"riemann_solver": None, "slope_limiter": on.van_leer,
               "reconstructor": on.piecewise_constant, "diffusive_solver": on.centered}
    bcs = make_bcs(oh.VARS, z["bcs"])
    osv.advance(oh, U, grid, params, schemes, on.euler_explicit, bcs, 0.4, 3)
    assert same(U["internal_energy_density"], z["U_out/internal_energy_density"])


def test_maxwell_spherical():
    z = load(os.path.join(GOLDEN, "maxwell_sph.npz"))
    gc = grid_config((8, 8, 8), (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi))
    grid = og.build_grid("spherical", gc)
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
              "electrical_conductivity": 0.1, "charge_density": 0.0,
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 2e-3}
    U = {k: v.copy() for k, v in group(z, "U_in/").items()}
    for rs, fn in (("rusanov", on.rusanov), ("hll", on.hll)):
        schemes = {"riemann_solver": fn, "reconstructor": on.muscl, "slope_limiter": on.van_leer,
                   "diffusive_solver": None}
        r = om.compute_rhs(U, grid, params, schemes)
        for k in om.VARS:
            assert same(r[k], z[f"rhs_{rs}/{k}"]), (rs, k)
    assert om.compute_dt(U, params, grid, 0.4) == z["dt"]
    schemes = {"riemann_solver": on.rusanov, "reconstructor": on.muscl,
               "slope_limiter": on.van_leer, "diffusive_solver": None}
    bcs = make_bcs(om.VARS, z["bcs"])
    osv.advance(om, U, grid, params, schemes, on.rk4, bcs, 0.4, 1)
    for k in om.VARS:
        assert same(U[k], z[f"U_out/{k}"]), k


def test_grids():
    z = load(os.path.join(GOLDEN, "grids.npz"))
    cfgs = {"cyl": ("cylindrical", grid_config((5, 4, 3), (0.1, 0.0, -1.0), (2.0, 2 * np.pi, 1.0))),
            "sph": ("spherical", grid_config((5, 4, 3), (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi))),
            "cart": ("cartesian", grid_config((5, 4, 3), (0.0, -0.3, 0.1), (0.7, 1.1, 0.6)))}
    for tag, (cs, gc) in cfgs.items():
        g = og.build_grid(cs, gc)
        assert same(g["cell"]["geometry"]["volume"], z[f"{tag}/V"])
        for d in ("x1", "x2", "x3"):
            assert same(g["cell"]["geometry"]["surface"][d], z[f"{tag}/S{d}"]), (tag, d)
            assert same(g["cell"]["geometry"]["length"][d], z[f"{tag}/L{d}"]), (tag, d)
            assert same(g["lame"]["coefficients"][d], z[f"{tag}/h{d}"]), (tag, d)
            for k, v in g["lame"]["derivatives"][d].items():
                assert same(v, z[f"{tag}/d{d}_{k}"]), (tag, d, k)


def test_bcs():
    z = load(os.path.join(GOLDEN, "bcs.npz"))
    for code, name in ((0, "neumann"), (1, "periodic"), (2, "reflecting")):
        f = z[f"{name}/in"].copy()
        og.apply_bcs({"q": f}, {"q": {s: {0: og.neumann, 1: og.periodic, 2: og.reflecting}[code]
                                      for s in og.SIDES}})
        assert same(f, z[f"{name}/out"]), name
