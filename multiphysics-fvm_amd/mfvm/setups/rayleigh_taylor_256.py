"""3D Rayleigh-Taylor, 256^3, FP64: heavy rho = 2 over light rho = 1 at
z = 0.5, g = -0.1 in x3, hydrostatic p with p = 2.5 at the interface,
vz = 0.01 (1 + cos 2 pi x)(1 + cos 2 pi y) / 4 + U(-1e-3, 1e-3) (seed 1);
mu = 1e-4, MUSCL + van Leer, HLL, SSPRK3, CFL 0.4; periodic x1/x2, x3 walls
(momentum_x3 reflecting, the rest Neumann); 100 steps."""

import numpy as np

from ..grid import boundary_conditions as bc
from ..numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters, time_integrators
from ..physics import navier_stokes
from ._common import grid_config, save_config


def initializer(U, params, grid, ics_params):
    x = grid["cell"]["coordinates"]["x1"]
    y = grid["cell"]["coordinates"]["x2"]
    z = grid["cell"]["coordinates"]["x3"]
    rng = np.random.default_rng(ics_params["seed"])
    g = -params["gravitational_acceleration"][2]
    heavy = z > 0.5
    rho = np.where(heavy, 2.0, 1.0)
    p = np.where(heavy, 2.5 - g * 2.0 * (z - 0.5), 2.5 + g * 1.0 * (0.5 - z))
    vz = 0.01 * (1 + np.cos(2 * np.pi * x)) * (1 + np.cos(2 * np.pi * y)) / 4 + rng.uniform(-1e-3, 1e-3, x.shape)
    gamma = params["heat_capacity_ratio"]
    ing = (slice(1, -1),) * 3
    U["mass_density"][ing] = rho
    U["momentum_x1"][ing] = 0.0
    U["momentum_x2"][ing] = 0.0
    U["momentum_x3"][ing] = rho * vz
    U["energy_density"][ing] = p / (gamma - 1.0) + 0.5 * rho * vz * vz


def get_setup(n=256):
    bcs = {}
    for k in navier_stokes.VARS:
        wall = bc.reflecting if k == "momentum_x3" else bc.neumann
        bcs[k] = {"left": bc.periodic, "right": bc.periodic, "bottom": bc.periodic, "top": bc.periodic,
                  "front": wall, "back": wall}
    return {"physics": navier_stokes,
            "physical_parameters": {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
                                    "gravitational_acceleration": [0.0, 0.0, -0.1]},
            "coordinate_system": "cartesian",
            "grid_config": grid_config((n, n, n)),
            "time_config": {"t_start": 0.0, "t_end": 10.0, "CFL": 0.4, "n_steps": 100},
            "time_integrator": time_integrators.rk3_ssp,
            "space_schemes": {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
                              "slope_limiter": slope_limiters.van_leer,
                              "diffusive_solver": diffusive_solvers.centered},
            "boundary_conditions": bcs,
            "initial_conditions": {"initializer": initializer, "parameters": {"seed": 1}},
            "save_config": save_config("./outputs/rayleigh_taylor_256")}
