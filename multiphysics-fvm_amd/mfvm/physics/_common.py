"""Scheme resolution shared by the physics modules."""

import ctypes
import numbers

import numpy as np
import torch

from .. import _lib
from ..grid.boundary_conditions import bc_codes
from ..state import State


def require_state(U, names):
    if not isinstance(U, State):
        raise TypeError("U must be an mfvm State (use allocate_u); got " + type(U).__name__)
    if U.names != tuple(names):
        raise ValueError(f"U holds {U.names}, expected {tuple(names)}")
    if U.device.type != "cuda":
        raise RuntimeError("mfvm kernels need U on a HIP device")
    return U


def require_grid(grid, U):
    g = grid.get("mfvm") if isinstance(grid, dict) else None
    if g is None:
        raise TypeError("grid must come from mfvm.grid.geometry.build_grid")
    if tuple(g.n) != tuple(U.n):
        raise ValueError(f"grid is {g.n}, state is {U.n}")
    return g


def limiter_code(space_schemes):
    recon = space_schemes["reconstructor"]
    kind = getattr(recon, "_mfvm_recon", None)
    if kind is None:
        raise NotImplementedError(f"reconstructor {recon!r} has no fused kernel")
    if kind == "pc":
        return _lib.LIM_NONE
    lim = space_schemes["slope_limiter"]
    code = getattr(lim, "_mfvm_code", None)
    if code is None:
        raise NotImplementedError(f"slope limiter {lim!r} has no fused kernel")
    return code


def riemann_code(space_schemes):
    rs = space_schemes["riemann_solver"]
    code = getattr(rs, "_mfvm_code", None)
    if code is None:
        raise NotImplementedError(f"Riemann solver {rs!r} has no fused kernel")
    return code


def dtype_code(U):
    if U.dtype == torch.float64:
        return _lib.F64
    if U.dtype == torch.float32:
        return _lib.F32
    raise TypeError("state dtype must be float64 or float32")


def bc_table(names, bcs):
    rows = bc_codes(names, bcs)
    flat = (ctypes.c_int32 * (6 * len(names)))(*[c for r in rows for c in r])
    return rows, flat


def cell_param(U, value):
    """A physical parameter that the reference broadcasts against interior
    arrays: a real number -> (float, None); an array (NumPy, torch or nested
    sequence) broadcastable to (nx1, nx2, nx3) -> (0.0, device tensor of the
    value at every interior cell, [nx3][nx2][nx1] in U's dtype).  A 0-d array
    counts as a number."""
    if isinstance(value, numbers.Real):
        return float(value), None
    t = value if isinstance(value, torch.Tensor) else torch.as_tensor(np.asarray(value, dtype=np.float64))
    if t.dim() == 0:
        return float(t), None
    n1, n2, n3 = U.n
    try:
        full = torch.broadcast_to(t, (n1, n2, n3))
    except RuntimeError as e:
        raise ValueError(f"parameter of shape {tuple(t.shape)} does not broadcast to the "
                         f"interior shape {(n1, n2, n3)}") from e
    field = full.permute(2, 1, 0).to(device=U.device, dtype=U.dtype).contiguous()
    return 0.0, field


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return _lib.stream_ptr(torch)
