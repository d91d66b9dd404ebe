"""Oracle restatement of the heat-equation module.

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``src/physics/heat.py`` of multiphysics-fvm: Fourier heat flux from
centred temperature differences (interior only, zero ghosts), reconstruction
of the heat flux itself, the centred diffusive face flux and the diffusive CFL
step.  There is no convective flux and no geometric source (heat.py:436-457).
"""

import numpy as np

from .navier_stokes import _CM, _CP, divergence

VARS = ("internal_energy_density",)


def allocate_u(grid_config):
    """heat.py:13-37."""
    shape = tuple(grid_config[f"nx{a}"] + 2 for a in (1, 2, 3))
    return {"internal_energy_density": np.zeros(shape)}


def compute_closure(U, params, grid):
    """heat.py:40-102 -- phi = -k dT/dx, centred, zero-padded ghosts."""
    cp = params["specific_heat_capacity"]
    k = params["thermal_conductivity"]
    rho = params["mass_density"]
    L = grid["cell"]["geometry"]["length"]
    T = U["internal_energy_density"] / (rho * cp)
    Ls = (L["x1"], L["x2"], L["x3"])
    phi = [np.pad(-k * ((T[_CP[d]] - T[_CM[d]]) / (2 * Ls[d])), 1) for d in range(3)]
    return {"phix1": phi[0], "phix2": phi[1], "phix3": phi[2]}


def compute_diffusive_flux(U, C, n):
    """heat.py:139-174 -- n . phi on the reconstructed closure."""
    return {"internal_energy_density": n[0] * C["phix1"] + n[1] * C["phix2"] + n[2] * C["phix3"]}


def compute_dt(U, params, grid, cfl):
    """heat.py:305-351 -- dt = CFL / max(2 nu sum_d 1/L_d^2), nu = k/(rho cp)."""
    L = grid["cell"]["geometry"]["length"]
    nu = params["thermal_conductivity"] / (params["mass_density"] * params["specific_heat_capacity"])
    rate_d = 2 * nu * (1 / L["x1"] ** 2 + 1 / L["x2"] ** 2 + 1 / L["x3"] ** 2)
    return cfl / (0.0 + np.max(rate_d))


def compute_rhs(U, grid, params, space_schemes):
    """heat.py:354-473 -- -sum(F.S)/V + q with F = 0 + centred(Fd_L, Fd_R)."""
    V = grid["cell"]["geometry"]["volume"]
    S = grid["cell"]["geometry"]["surface"]
    L = grid["cell"]["geometry"]["length"]
    n = grid["cell"]["normal"]
    C = compute_closure(U, params, grid)
    C_L, C_R = space_schemes["reconstructor"](C, L, space_schemes["slope_limiter"])
    F = {}
    for d in ("x1", "x2", "x3"):
        fdl = compute_diffusive_flux(None, C_L[d], n[d])
        fdr = compute_diffusive_flux(None, C_R[d], n[d])
        fd = space_schemes["diffusive_solver"](fdl, fdr)
        F[d] = {"internal_energy_density": 0.0 + fd["internal_energy_density"]}
    src = {"internal_energy_density": params["volumetric_heat_source"]}
    return divergence(F, S, V, src, VARS)
