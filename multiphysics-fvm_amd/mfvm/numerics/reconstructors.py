"""Interface reconstructors (drop-in for src/numerics/reconstructors.py).

Used through ``compute_rhs`` they select the fused kernel variant
(piecewise constant: no slopes; MUSCL: limited slopes).  Called directly they
return the reference's face-state dicts computed with tensor ops.
"""

import torch

DIRS = ("x1", "x2", "x3")


def _sl(axis, along):
    idx = [slice(1, -1)] * 3
    idx[axis] = along
    return tuple(idx)


_LO = [_sl(a, slice(None, -1)) for a in range(3)]
_HI = [_sl(a, slice(1, None)) for a in range(3)]
_M = [_sl(a, slice(None, -2)) for a in range(3)]
_P = [_sl(a, slice(2, None)) for a in range(3)]
_IN = (slice(1, -1),) * 3


def piecewise_constant(fields, L, slope_limiter):
    """reconstructors.py:13-64."""
    out_l = {d: {k: f[_LO[a]] for k, f in fields.items()} for a, d in enumerate(DIRS)}
    out_r = {d: {k: f[_HI[a]] for k, f in fields.items()} for a, d in enumerate(DIRS)}
    return out_l, out_r


def muscl(fields, L, slope_limiter):
    """reconstructors.py:67-158 (ghost cells carry zero slope)."""
    out_l = {d: {} for d in DIRS}
    out_r = {d: {} for d in DIRS}
    for k, f in fields.items():
        c = f[_IN]
        for a, d in enumerate(DIRS):
            s = torch.zeros_like(f)
            s[_IN] = slope_limiter((c - f[_M[a]]) / L[d], (f[_P[a]] - c) / L[d]) * L[d]
            out_l[d][k] = f[_LO[a]] + s[_LO[a]] / 2.0
            out_r[d][k] = f[_HI[a]] - s[_HI[a]] / 2.0
    return out_l, out_r


piecewise_constant._mfvm_recon = "pc"
muscl._mfvm_recon = "muscl"
