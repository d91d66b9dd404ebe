"""Heat diffusion on a cylindrical grid (r, phi, z) 256 x 128 x 128,
r in [0, 1], phi in [0, 2 pi), z in [0, 1]: rho = cp = 1, k = 1e-2, no
source; T = 2 for r < 0.25, T = 1 elsewhere; piecewise-constant
reconstruction, centred diffusive flux, Euler, CFL 0.4; phi periodic, r and
z Neumann; 300 steps."""

import numpy as np

from ..grid import boundary_conditions as bc
from ..numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters, time_integrators
from ..physics import heat
from ._common import grid_config, save_config


def initializer(U, params, grid, ics_params):
    r = grid["cell"]["coordinates"]["x1"]
    T = np.where(r < ics_params["radius"], ics_params["hot"], ics_params["cold"])
    U["internal_energy_density"][(slice(1, -1),) * 3] = (
        params["mass_density"] * params["specific_heat_capacity"] * T)


def get_setup(n=(256, 128, 128)):
    bcs = {"internal_energy_density": {"left": bc.neumann, "right": bc.neumann, "bottom": bc.periodic,
                                       "top": bc.periodic, "front": bc.neumann, "back": bc.neumann}}
    return {"physics": heat,
            "physical_parameters": {"mass_density": 1.0, "specific_heat_capacity": 1.0,
                                    "thermal_conductivity": 1e-2, "volumetric_heat_source": 0.0},
            "coordinate_system": "cylindrical",
            "grid_config": grid_config(n, (0.0, 0.0, 0.0), (1.0, 2 * np.pi, 1.0)),
            "time_config": {"t_start": 0.0, "t_end": 10.0, "CFL": 0.4, "n_steps": 300},
            "time_integrator": time_integrators.euler_explicit,
            "space_schemes": {"riemann_solver": riemann_solvers.hll,
                              "reconstructor": reconstructors.piecewise_constant,
                              "slope_limiter": slope_limiters.minmod,
                              "diffusive_solver": diffusive_solvers.centered},
            "boundary_conditions": bcs,
            "initial_conditions": {"initializer": initializer,
                                   "parameters": {"radius": 0.25, "hot": 2.0, "cold": 1.0}},
            "save_config": save_config("./outputs/heated_disc_cylindrical")}
