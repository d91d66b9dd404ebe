"""Helpers that rebuild oracle setups from the golden fixtures (test-only)."""

import glob
import os

import numpy as np

from oracle import grid as og
from oracle import numerics as on

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIDES = og.SIDES
LIMITERS = {"minmod": on.minmod, "mc": on.monotonized_central, "vanleer": on.van_leer}
RIEMANN = {"hll": on.hll, "rusanov": on.rusanov}
RECON = {"muscl": on.muscl, "pc": on.piecewise_constant}
INTEG = {"euler": on.euler_explicit, "rk3": on.rk3, "ssprk3": on.rk3_ssp, "rk4": on.rk4}
BC = {0: og.neumann, 1: og.periodic, 2: og.reflecting}


def golden_files(pattern):
    return sorted(glob.glob(os.path.join(GOLDEN, pattern)))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def group(z, prefix):
    return {k[len(prefix):]: v for k, v in z.items() if k.startswith(prefix)}


def grid_config(n, lo, hi):
    return {"x1_min": float(lo[0]), "x1_max": float(hi[0]), "nx1": int(n[0]),
            "x2_min": float(lo[1]), "x2_max": float(hi[1]), "nx2": int(n[1]),
            "x3_min": float(lo[2]), "x3_max": float(hi[2]), "nx3": int(n[2])}


def make_bcs(names, codes):
    return {v: {s: BC[int(codes[i, j])] for j, s in enumerate(SIDES)} for i, v in enumerate(names)}


def ns_setup(z):
    coord, lim, rs, recon, integ, steps = [str(x) for x in z["meta"]]
    gc = grid_config(z["n"], z["lo"], z["hi"])
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": float(z["mu"]),
              "gravitational_acceleration": [float(x) for x in z["g"]]}
    schemes = {"riemann_solver": RIEMANN[rs], "reconstructor": RECON[recon],
               "slope_limiter": LIMITERS[lim], "diffusive_solver": on.centered}
    return dict(coord=coord, gc=gc, params=params, schemes=schemes, integ=integ,
                steps=int(steps), lim=lim, rs=rs, recon=recon)


def same(a, b):
    """Bitwise-equal values (+0 == -0, NaN == NaN)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))
