"""Kelvin-Helmholtz 128^2 (as kelvin_helmholtz_128, seeded perturbation
phase) advanced to t = 2 with the learned flux-correction closure (weights
``mfvm/data/closure_kh128.npz``) and the plain-scheme positivity fallback."""

import os

from . import kelvin_helmholtz_128 as kh

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data", "closure_kh128.npz")


def initializer(U, params, grid, ics_params):
    import numpy as np
    from ..closure import kh_initial
    n = grid["config"]["nx1"]
    for k, a in zip(("mass_density", "momentum_x1", "momentum_x2", "momentum_x3", "energy_density"),
                    kh_initial(n, ics_params["seed"])):
        U[k][(slice(1, -1),) * 3] = np.asarray(a)


def get_setup(n=128, seed=5000):
    s = kh.get_setup(n)
    s["time_config"] = {"t_start": 0.0, "t_end": 2.0, "CFL": 0.4}
    s["initial_conditions"] = {"initializer": initializer, "parameters": {"seed": seed}}
    s["closure"] = {"weights": WEIGHTS}
    s["save_config"]["directory"] = "./outputs/kelvin_helmholtz_closure"
    return s
