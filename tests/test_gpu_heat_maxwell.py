"""Heat and Maxwell HIP paths vs the reference golden vectors (bit-exact)."""

import os

import numpy as np
import pytest
import torch

from oracle_setup import GOLDEN, group, grid_config, load, same

pytestmark = pytest.mark.gpu


def _schemes(rec, lim, rs=None):
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    return {"reconstructor": getattr(reconstructors, rec), "slope_limiter": getattr(slope_limiters, lim),
            "riemann_solver": getattr(riemann_solvers, rs) if rs else None,
            "diffusive_solver": diffusive_solvers.centered}


def _bcs(names, codes):
    from mfvm.grid import boundary_conditions as bc
    m = {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}
    return {v: {s: m[int(codes[i, j])] for j, s in enumerate(bc.SIDES)} for i, v in enumerate(names)}


def test_heat_cylindrical_matches_reference():
    from mfvm import solver
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import heat
    z = load(os.path.join(GOLDEN, "heat_cyl.npz"))
    gc = grid_config((8, 6, 4), (0.1, 0.0, 0.0), (1.0, 2 * np.pi, 1.0))
    grid = build_grid("cylindrical", gc)
    params = {"mass_density": 1.0, "specific_heat_capacity": 1.0, "thermal_conductivity": 1e-2,
              "volumetric_heat_source": 0.05}
    U = heat.allocate_u(gc)
    for k, v in group(z, "U_in/").items():
        U[k] = torch.from_numpy(v)
    for tag, rec in (("pc", "piecewise_constant"), ("vanleer", "muscl")):
        r = heat.compute_rhs(U, grid, params, _schemes(rec, "van_leer"))
        assert same(r["internal_energy_density"].cpu().numpy(),
                    z[f"rhs_{tag}/internal_energy_density"]), tag
    assert heat.compute_dt(U, params, grid, 0.4) == float(z["dt"])
    bcs = _bcs(heat.VARS, z["bcs"])
    schemes = _schemes("piecewise_constant", "van_leer")
    for _ in range(3):
        dt = heat.compute_dt(U, params, grid, 0.4)
        ti.euler_explicit(heat.compute_rhs, U, dt, params, grid, schemes, bcs, solver.apply_bcs)
    assert same(U["internal_energy_density"].cpu().numpy(), z["U_out/internal_energy_density"])
    # the same three steps through the device-dt driver
    U2 = heat.allocate_u(gc)
    for k, v in group(z, "U_in/").items():
        U2[k] = torch.from_numpy(v)
    heat.advance(U2, grid, params, schemes, bcs, ti.euler_explicit._mfvm_code, 0.4, 3)
    assert same(U2["internal_energy_density"].cpu().numpy(), z["U_out/internal_energy_density"])


def test_maxwell_spherical_matches_reference():
    from mfvm import solver
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import maxwell
    z = load(os.path.join(GOLDEN, "maxwell_sph.npz"))
    gc = grid_config((8, 8, 8), (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi))
    grid = build_grid("spherical", gc)
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
              "electrical_conductivity": 0.1, "charge_density": 0.0,
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 2e-3}
    U = maxwell.allocate_u(gc)
    for k, v in group(z, "U_in/").items():
        U[k] = torch.from_numpy(v)
    for rs in ("rusanov", "hll"):
        r = maxwell.compute_rhs(U, grid, params, _schemes("muscl", "van_leer", rs))
        for k in maxwell.VARS:
            assert same(r[k].cpu().numpy(), z[f"rhs_{rs}/{k}"]), (rs, k)
    assert maxwell.compute_dt(U, params, grid, 0.4) == float(z["dt"])
    bcs = _bcs(maxwell.VARS, z["bcs"])
    schemes = _schemes("muscl", "van_leer", "rusanov")
    dt = maxwell.compute_dt(U, params, grid, 0.4)
    ti.rk4(maxwell.compute_rhs, U, dt, params, grid, schemes, bcs, solver.apply_bcs)
    for k in maxwell.VARS:
        assert same(U[k].cpu().numpy(), z[f"U_out/{k}"]), k
