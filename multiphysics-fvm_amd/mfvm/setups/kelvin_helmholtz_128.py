"""2D Kelvin-Helmholtz, 128 x 128 x 1, FP64: rho = 2, vx = +0.5 for
|y - 0.5| < 0.25, else rho = 1, vx = -0.5; p = 2.5; vy = 0.01 sin(4 pi x) +
U(-1e-3, 1e-3) (seed 0); MUSCL + minmod, HLL, SSPRK3, CFL 0.4, mu = 1e-4,
periodic; 200 steps (t_end set to the time of 200 steps' worth: the run stops
on the step count when ``n_steps`` is given)."""

import numpy as np

from ..grid import boundary_conditions as bc
from ..numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters, time_integrators
from ..physics import navier_stokes
from ._common import bcs_all, grid_config, save_config


def initializer(U, params, grid, ics_params):
    x = grid["cell"]["coordinates"]["x1"]
    y = grid["cell"]["coordinates"]["x2"]
    rng = np.random.default_rng(ics_params["seed"])
    inner = np.abs(y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    vy = ics_params["amplitude"] * np.sin(4 * np.pi * x) + rng.uniform(-1e-3, 1e-3, x.shape)
    gamma = params["heat_capacity_ratio"]
    ing = (slice(1, -1),) * 3
    U["mass_density"][ing] = rho
    U["momentum_x1"][ing] = rho * vx
    U["momentum_x2"][ing] = rho * vy
    U["momentum_x3"][ing] = 0.0
    U["energy_density"][ing] = 2.5 / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy)


def get_setup(n=128):
    return {"physics": navier_stokes,
            "physical_parameters": {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
                                    "gravitational_acceleration": [0.0, 0.0, 0.0]},
            "coordinate_system": "cartesian",
            "grid_config": grid_config((n, n, 1)),
            "time_config": {"t_start": 0.0, "t_end": 1.0, "CFL": 0.4, "n_steps": 200},
            "time_integrator": time_integrators.rk3_ssp,
            "space_schemes": {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
                              "slope_limiter": slope_limiters.minmod,
                              "diffusive_solver": diffusive_solvers.centered},
            "boundary_conditions": bcs_all(navier_stokes.VARS, bc.periodic),
            "initial_conditions": {"initializer": initializer, "parameters": {"seed": 0, "amplitude": 0.01}},
            "save_config": save_config("./outputs/kelvin_helmholtz_128")}
