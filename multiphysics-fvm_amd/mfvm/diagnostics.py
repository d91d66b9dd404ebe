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

import ctypes

import torch

from . import _lib
from .physics import _common as c
from .state import State


def diagnostics(U, grid, ref=None, gamma=None, compensated=False):
    """Diagnostics of the state ``U`` (any physics, 1..6 fields).

    ref         -- None or a float64 :class:`State` of the same fields and shape.
    gamma       -- heat_capacity_ratio; enables the positivity monitor
                   (Navier-Stokes states only).
    compensated -- U is the float32 state of
                   ``navier_stokes.to_compensated``: each value is hi + lo.
    """
    if not isinstance(U, State):
        raise TypeError("U must be an mfvm State; got " + type(U).__name__)
    if U.device.type != "cuda":
        raise RuntimeError("mfvm kernels need U on a HIP device")
    g = c.require_grid(grid, U)
    nv = len(U.names)
    dtype = c.dtype_code(U)
    ref_ptr = None
    if ref is not None:
        if not isinstance(ref, State) or ref.names != U.names or tuple(ref.n) != tuple(U.n):
            raise ValueError("ref must be a State with the fields and shape of U")
        if ref.dtype != torch.float64 or ref.device != U.device:
            raise ValueError("ref must be a float64 state on U's device")
        ref_ptr = c.ptr(ref.buf)
    lo_ptr = None
    if compensated:
        if U.dtype != torch.float32 or "lo" not in U._work:
            raise ValueError("compensated diagnostics need the float32 state of to_compensated")
        lo_ptr = c.ptr(U.scratch("lo", tuple(U.buf.shape)))
    g_val = 0.0
    if gamma is not None:
        g_val = float(gamma)
        if not g_val > 0.0:
            raise ValueError("gamma must be > 0")
    n = int(_lib.lib.mfvm_diag_partials_size(g.ptr(), nv))
    if n < 0:
        _lib.check(_lib.E_ARG)
    nout = 5 * nv + 3
    partials = U.scratch64("diag_partials", n)
    out = U.scratch64("diag_out", nout)
    _lib.check(_lib.lib.mfvm_diagnostics(g.ptr(), nv, dtype, c.ptr(U.buf), lo_ptr, ref_ptr, g_val,
                                         c.ptr(partials), c.ptr(out), c.stream()))
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


def raw(U, grid, ref=None, gamma=None, lo=None, out=None, partials=None):
    """The C-ABI call on explicit device buffers, no host read-back: returns
    ``out`` (float64, 5*nvar+3 entries) for use inside a timed or captured
    region."""
    g = c.require_grid(grid, U)
    nv = len(U.names)
    nout = 5 * nv + 3
    n = int(_lib.lib.mfvm_diag_partials_size(g.ptr(), nv))
    if n < 0:
        _lib.check(_lib.E_ARG)
    if partials is None:
        partials = U.scratch64("diag_partials", n)
    if out is None:
        out = U.scratch64("diag_out", nout)
    _lib.check(_lib.lib.mfvm_diagnostics(
        g.ptr(), nv, c.dtype_code(U), c.ptr(U.buf), None if lo is None else c.ptr(lo),
        None if ref is None else c.ptr(ref.buf), 0.0 if gamma is None else float(gamma),
        c.ptr(partials), c.ptr(out), c.stream()))
    return out[:nout]
