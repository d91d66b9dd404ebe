"""State diagnostics on the device (``mfvm_diagnostics``).

The reference framework has no run-time monitor: checking a run there means
NumPy over the whole state on the host.  Here one call reduces a state on the
GPU, in float64 over the interior cells, and reads back a few numbers:

* ``total``: the conserved totals, sum of U * V per field;
* ``min`` / ``max`` per field (NaN propagates);
* ``l1`` / ``linf`` against a float64 reference state: the volume-weighted
  mean and the maximum of |U - ref| per field;
* ``volume``: the sum of the cell volumes;
* Navier-Stokes states with ``gamma``: ``pressure_min``, the minimum of the
  ideal-gas pressure of the reference closure (navier_stokes.py:94-100), and
  ``nonpositive_cells``, the number of cells with rho <= 0 or P <= 0 (NaN
  counts), the positivity monitor.

The reduction order is fixed, so equal states give equal results.
"""

import torch

from . import _lib
from .physics import _common as c
from .state import State


def _checked(U, grid, ref, gamma, lo):
    """Validate the arguments the kernel cannot check (it takes plain
    pointers); returns (grid descriptor, nvar, partials size)."""
    if not isinstance(U, State):
        raise TypeError("U must be an mfvm State; got " + type(U).__name__)
    if U.device.type != "cuda":
        raise RuntimeError("mfvm kernels need U on a HIP device")
    g = c.require_grid(grid, U)
    if ref is not None:
        if not isinstance(ref, State) or ref.names != U.names or tuple(ref.n) != tuple(U.n):
            raise ValueError("ref must be a State with the fields and shape of U")
        if ref.dtype != torch.float64 or ref.device != U.device:
            raise ValueError("ref must be a float64 state on U's device")
    if lo is not None:
        if (U.dtype != torch.float32 or lo.dtype != torch.float32 or tuple(lo.shape) != tuple(U.buf.shape)
                or lo.device != U.device or not lo.is_contiguous()):
            raise ValueError("lo must be a contiguous float32 array shaped like the float32 state U")
    if gamma is not None and not float(gamma) > 0.0:
        raise ValueError("gamma must be > 0")
    nv = len(U.names)
    n = int(_lib.lib.mfvm_diag_partials_size(g.ptr(), nv))
    if n < 0:
        _lib.check(_lib.E_ARG)
    return g, nv, n


def _call(U, g, nv, ref, gamma, lo, partials, out):
    _lib.check(_lib.lib.mfvm_diagnostics(
        g.ptr(), nv, c.dtype_code(U), c.ptr(U.buf), None if lo is None else c.ptr(lo),
        None if ref is None else c.ptr(ref.buf), 0.0 if gamma is None else float(gamma),
        c.ptr(partials), c.ptr(out), c.stream()))


def diagnostics(U, grid, ref=None, gamma=None, compensated=False):
    """Diagnostics of the state ``U`` (any physics, 1..6 fields).

    ref         -- None or a float64 :class:`State` of the same fields and shape.
    gamma       -- heat_capacity_ratio; enables the positivity monitor
                   (Navier-Stokes states only).
    compensated -- U is the float32 state of
                   ``navier_stokes.to_compensated``: each value is hi + lo.
    """
    lo = None
    if compensated:
        if not isinstance(U, State) or U.dtype != torch.float32 or "lo" not in U._work:
            raise ValueError("compensated diagnostics need the float32 state of to_compensated")
        lo = U.scratch("lo", tuple(U.buf.shape))
    g, nv, n = _checked(U, grid, ref, gamma, lo)
    nout = 5 * nv + 3
    out = U.scratch64("diag_out", nout)
    _call(U, g, nv, ref, gamma, lo, U.scratch64("diag_partials", n), out)
    h = out[:nout].cpu().tolist()

    def per(base):
        return {name: h[base + q] for q, name in enumerate(U.names)}

    return {
        "total": per(0),
        "min": per(nv),
        "max": per(2 * nv),
        "l1": per(3 * nv) if ref is not None else None,
        "linf": per(4 * nv) if ref is not None else None,
        "volume": h[5 * nv],
        "pressure_min": h[5 * nv + 1] if gamma is not None else None,
        "nonpositive_cells": int(h[5 * nv + 2]) if gamma is not None else None,
    }


def raw(U, grid, ref=None, gamma=None, lo=None):
    """The C-ABI call with no host read-back (for timed regions):
    returns the device array of 5*nvar+3 float64 results (the per-state
    scratch, overwritten by the next call on U)."""
    g, nv, n = _checked(U, grid, ref, gamma, lo)
    out = U.scratch64("diag_out", 5 * nv + 3)
    _call(U, g, nv, ref, gamma, lo, U.scratch64("diag_partials", n), out)
    return out[:5 * nv + 3]
