"""Shared pieces of the setups."""

from ..grid import boundary_conditions as bc


def grid_config(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    return {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1], "nx2": n[1],
            "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}


def bcs_all(names, rule):
    return {k: {s: rule for s in bc.SIDES} for k in names}


def save_config(directory, freq=0):
    from .. import io
    return {"saver": io.save_snapshot if freq else None, "output_frequency": freq, "directory": directory}
