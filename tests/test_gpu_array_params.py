"""Array-valued physical parameters against the oracle, bit for bit.

The reference broadcasts arrays where its parameters enter interior-shaped
expressions: the heat source q (heat.py:294-297, added to the right-hand
side), Maxwell's conductivity sigma (maxwell.py:71, J = -sigma E) and charge
density (maxwell.py:475, rD = div D - rho, then edge-padded).  The drop-ins
take numbers or arrays broadcastable to (nx1, nx2, nx3); here full-shape and
broadcast arrays on curvilinear grids of several tiles, through the
right-hand side and whole steps with the device dt.
"""

import numpy as np
import pytest
import torch

from oracle import grid as og
from oracle import heat as oht
from oracle import maxwell as omx
from oracle import numerics as on
from oracle import solver as osv
from oracle_setup import grid_config, same

pytestmark = pytest.mark.gpu


def _mods():
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.numerics import time_integrators as ti
    return bc, build_grid, diffusive_solvers, reconstructors, riemann_solvers, slope_limiters, ti


@pytest.mark.parametrize("qshape", ["full", "radial"])
def test_heat_array_source_matches_oracle(qshape):
    bc, build_grid, ds, rec, rs, sl, ti = _mods()
    from mfvm.physics import heat
    n = (70, 9, 6)
    gc = grid_config(n, (0.1, 0.0, 0.0), (1.0, 2 * np.pi, 1.0))
    rng = np.random.default_rng(2911)
    q = rng.uniform(-0.5, 1.0, n if qshape == "full" else (n[0], 1, 1))
    params = {"mass_density": 1.0, "specific_heat_capacity": 1.0, "thermal_conductivity": 1e-2,
              "volumetric_heat_source": q}
    Uo = oht.allocate_u(gc)
    Uo["internal_energy_density"][...] = 1.0 + rng.random(Uo["internal_energy_density"].shape)
    sides = {"left": 0, "right": 0, "bottom": 1, "top": 1, "front": 0, "back": 0}
    bo = {"internal_energy_density": {s: {0: og.neumann, 1: og.periodic}[c] for s, c in sides.items()}}
    bg = {"internal_energy_density": {s: {0: bc.neumann, 1: bc.periodic}[c] for s, c in sides.items()}}
    og.apply_bcs(Uo, bo)
    go = og.build_grid("cylindrical", gc)
    so = {"reconstructor": on.muscl, "slope_limiter": on.van_leer, "riemann_solver": None,
          "diffusive_solver": on.centered}
    sg = {"reconstructor": rec.muscl, "slope_limiter": sl.van_leer, "riemann_solver": None,
          "diffusive_solver": ds.centered}
    grid = build_grid("cylindrical", gc)
    U = heat.allocate_u(gc)
    U["internal_energy_density"] = torch.from_numpy(Uo["internal_energy_density"])
    print(f"heat array source {qshape}: rhs", flush=True)
    r = heat.compute_rhs(U, grid, params, sg)
    ro = oht.compute_rhs(Uo, go, params, so)
    assert same(r["internal_energy_density"].cpu().numpy(), ro["internal_energy_density"])
    print(f"heat array source {qshape}: 3 euler steps", flush=True)
    dts_o = osv.advance(oht, Uo, go, params, so, on.euler_explicit, bo, 0.4, 3)
    dts = heat.advance(U, grid, params, sg, bg, ti.euler_explicit._mfvm_code, 0.4, 3)
    assert dts.cpu().numpy()[:3].tolist() == dts_o
    assert same(U["internal_energy_density"].cpu().numpy(), Uo["internal_energy_density"])


@pytest.mark.parametrize("shapes", [("full", "full"), ("theta", "radial")])
def test_maxwell_array_sigma_and_charge_match_oracle(shapes):
    bc, build_grid, ds, rec, rs, sl, ti = _mods()
    from mfvm.physics import maxwell
    n = (70, 9, 12)
    gc = grid_config(n, (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi))
    rng = np.random.default_rng(3117)
    form = {"full": n, "radial": (n[0], 1, 1), "theta": (1, n[1], 1)}
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
              "electrical_conductivity": rng.uniform(0.0, 0.3, form[shapes[0]]),
              "charge_density": rng.uniform(-0.2, 0.2, form[shapes[1]]),
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 2e-3}
    Uo = omx.allocate_u(gc)
    for k in omx.VARS:
        Uo[k][...] = rng.standard_normal(Uo[k].shape)
    sides = {"left": 0, "right": 0, "bottom": 0, "top": 0, "front": 1, "back": 1}
    bo = {k: {s: {0: og.neumann, 1: og.periodic}[c] for s, c in sides.items()} for k in omx.VARS}
    bg = {k: {s: {0: bc.neumann, 1: bc.periodic}[c] for s, c in sides.items()} for k in maxwell.VARS}
    og.apply_bcs(Uo, bo)
    go = og.build_grid("spherical", gc)
    so = {"reconstructor": on.muscl, "slope_limiter": on.van_leer, "riemann_solver": on.rusanov,
          "diffusive_solver": on.centered}
    sg = {"reconstructor": rec.muscl, "slope_limiter": sl.van_leer, "riemann_solver": rs.rusanov,
          "diffusive_solver": ds.centered}
    grid = build_grid("spherical", gc)
    U = maxwell.allocate_u(gc)
    for k in maxwell.VARS:
        U[k] = torch.from_numpy(Uo[k])
    print(f"maxwell array sigma/rho {shapes}: rhs", flush=True)
    r = maxwell.compute_rhs(U, grid, params, sg)
    ro = omx.compute_rhs(Uo, go, params, so)
    for k in maxwell.VARS:
        assert same(r[k].cpu().numpy(), ro[k]), k
    print(f"maxwell array sigma/rho {shapes}: 2 rk4 steps", flush=True)
    dts_o = osv.advance(omx, Uo, go, params, so, on.rk4, bo, 0.4, 2)
    dts = maxwell.advance(U, grid, params, sg, bg, ti.rk4._mfvm_code, 0.4, 2)
    assert dts.cpu().numpy()[:2].tolist() == dts_o
    for k in maxwell.VARS:
        assert same(U[k].cpu().numpy(), Uo[k]), k
