"""Batched ensemble: 64 Kelvin-Helmholtz instances, 56 2D grids with nx1,
nx2 drawn uniformly from {64, 128, 256, 512} (seed 2) and eight 64^3 grids;
per instance (seed 3 + i) shear U[0.3, 0.7], density ratio U[1.5, 3],
perturbation phase U[0, 2 pi), 1e-3 velocity noise; MUSCL + MC, Rusanov,
RK3, CFL 0.4, mu = 1e-4, periodic; 500 steps.  ``get_setups()`` returns one
setup dict per member; :func:`mfvm.solver.run_ensemble` advances them
together."""

import numpy as np

from ..grid import boundary_conditions as bc
from ..numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters, time_integrators
from ..physics import navier_stokes
from ._common import bcs_all, grid_config, save_config


def member_shapes(seed=2):
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 128, 256, 512], size=(56, 2))
    return [(int(a), int(b), 1) for a, b in sizes] + [(64, 64, 64)] * 8


def initializer(U, params, grid, ics_params):
    x = grid["cell"]["coordinates"]["x1"]
    y = grid["cell"]["coordinates"]["x2"]
    rng = np.random.default_rng(ics_params["seed"])
    shear = rng.uniform(0.3, 0.7)
    ratio = rng.uniform(1.5, 3.0)
    phase = rng.uniform(0.0, 2 * np.pi)
    n = x.shape
    inner = np.abs(y - 0.5) < 0.25
    rho = np.where(inner, ratio, 1.0)
    vx = np.where(inner, shear, -shear)
    vy = 0.01 * np.sin(4 * np.pi * x + phase) + rng.uniform(-1e-3, 1e-3, n)
    vz = rng.uniform(-1e-3, 1e-3, n) if n[2] > 1 else np.zeros(n)
    gamma = params["heat_capacity_ratio"]
    ing = (slice(1, -1),) * 3
    U["mass_density"][ing] = rho
    U["momentum_x1"][ing] = rho * vx
    U["momentum_x2"][ing] = rho * vy
    U["momentum_x3"][ing] = rho * vz
    U["energy_density"][ing] = 2.5 / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)


def get_setups():
    out = []
    for i, n in enumerate(member_shapes()):
        out.append({"physics": navier_stokes,
                    "physical_parameters": {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
                                            "gravitational_acceleration": [0.0, 0.0, 0.0]},
                    "coordinate_system": "cartesian",
                    "grid_config": grid_config(n),
                    "time_config": {"t_start": 0.0, "t_end": 10.0, "CFL": 0.4, "n_steps": 500},
                    "time_integrator": time_integrators.rk3,
                    "space_schemes": {"riemann_solver": riemann_solvers.rusanov,
                                      "reconstructor": reconstructors.muscl,
                                      "slope_limiter": slope_limiters.monotonized_central,
                                      "diffusive_solver": diffusive_solvers.centered},
                    "boundary_conditions": bcs_all(navier_stokes.VARS, bc.periodic),
                    "initial_conditions": {"initializer": initializer, "parameters": {"seed": 3 + i}},
                    "save_config": save_config(f"./outputs/ensemble_kh64/member_{i:02d}")})
    return out


def get_setup():
    return get_setups()[0]
