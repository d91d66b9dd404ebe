"""Snapshots (drop-in for src/io.py).

The reference writes HDF5 (``src/io.py:14-63``): group ``U`` with one dataset
per conserved variable *including ghost cells* in the reference shape,
group ``setup`` with the coordinate system, the grid configuration and the
physical parameters as attributes, and root attributes ``time`` and
``iteration``, in ``{save_dir}/snapshot_{iteration:06d}.h5``.  h5py is not
part of this build's environment, so the same content is written to
``snapshot_{iteration:06d}.npz`` under the keys ``U/<var>``,
``setup/<key>``, ``time`` and ``iteration``; :func:`load_snapshot` returns
the same tuple as the reference's (``src/io.py:66-96``), and :func:`restore`
copies a loaded U into a device state to resume a run.
"""

import os

import numpy as np

from .state import State


def _host(U):
    if isinstance(U, State):
        return {name: U[name].cpu().numpy() for name in U.names}
    return {k: np.asarray(v) for k, v in U.items()}


def save_snapshot(U, setup, time, iteration, save_dir):
    """src/io.py:14-63 (npz container, same keys)."""
    os.makedirs(save_dir, exist_ok=True)
    data = {}
    for name, arr in _host(U).items():
        data[f"U/{name}"] = np.ascontiguousarray(arr, dtype=np.float64)
    data["setup/coordinate_system"] = np.asarray(setup["coordinate_system"])
    for key, val in setup["grid_config"].items():
        data[f"setup/{key}"] = np.asarray(val)
    for key, val in setup["physical_parameters"].items():
        data[f"setup/{key}"] = np.asarray(val)
    data["time"] = np.asarray(float(time))
    data["iteration"] = np.asarray(int(iteration))
    path = os.path.join(save_dir, f"snapshot_{int(iteration):06d}.npz")
    np.savez(path, **data)
    return path


def load_snapshot(filepath):
    """src/io.py:66-96 -> (U, setup, time, iteration) with NumPy arrays."""
    U, setup = {}, {}
    with np.load(filepath, allow_pickle=False) as f:
        for key in f.files:
            v = f[key]
            if key.startswith("U/"):
                U[key[2:]] = v
            elif key.startswith("setup/"):
                setup[key[6:]] = v.item() if v.ndim == 0 else v
        time = float(f["time"])
        iteration = int(f["iteration"])
    return U, setup, time, iteration


def restore(state, U):
    """Copy loaded arrays (reference shape, with ghosts) into a device State."""
    import torch
    for name in state.names:
        state[name].copy_(torch.from_numpy(np.ascontiguousarray(U[name])))
    return state
