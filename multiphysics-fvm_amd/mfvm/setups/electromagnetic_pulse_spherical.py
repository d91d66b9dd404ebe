"""Maxwell pulse on a spherical grid, r in [0.1, 1], theta in
[0.05, pi - 0.05], phi in [0, 2 pi), 128 x 128 x 256: eps = mu = 1,
sigma = 0, cleaning coefficients 1e-3; D from a Gaussian Ez of width 0.05 at
r = 0.5 projected on (r, theta): D_r = Ez cos(theta), D_theta = -Ez
sin(theta); MUSCL + van Leer, Rusanov, RK4, CFL 0.4; phi periodic, r and
theta Neumann; 300 steps."""

import numpy as np

from ..grid import boundary_conditions as bc
from ..numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters, time_integrators
from ..physics import maxwell
from ._common import grid_config, save_config


def initializer(U, params, grid, ics_params):
    r = grid["cell"]["coordinates"]["x1"]
    th = grid["cell"]["coordinates"]["x2"]
    ez = np.exp(-((r - ics_params["radius"]) / ics_params["width"]) ** 2)
    ing = (slice(1, -1),) * 3
    U["electric_d_field_x1"][ing] = ez * np.cos(th)
    U["electric_d_field_x2"][ing] = -ez * np.sin(th)


def get_setup(n=(128, 128, 256)):
    bcs = {k: {"left": bc.neumann, "right": bc.neumann, "bottom": bc.neumann, "top": bc.neumann,
               "front": bc.periodic, "back": bc.periodic} for k in maxwell.VARS}
    return {"physics": maxwell,
            "physical_parameters": {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
                                    "electrical_conductivity": 0.0, "charge_density": 0.0,
                                    "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 1e-3},
            "coordinate_system": "spherical",
            "grid_config": grid_config(n, (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi)),
            "time_config": {"t_start": 0.0, "t_end": 10.0, "CFL": 0.4, "n_steps": 300},
            "time_integrator": time_integrators.rk4,
            "space_schemes": {"riemann_solver": riemann_solvers.rusanov, "reconstructor": reconstructors.muscl,
                              "slope_limiter": slope_limiters.van_leer,
                              "diffusive_solver": diffusive_solvers.centered},
            "boundary_conditions": bcs,
            "initial_conditions": {"initializer": initializer, "parameters": {"radius": 0.5, "width": 0.05}},
            "save_config": save_config("./outputs/electromagnetic_pulse_spherical")}
