"""Slope limiters (drop-in for src/numerics/slope_limiters.py).

Inside ``compute_rhs`` the limiter runs fused in the stage kernel
(``limit<>`` in csrc/common.h), selected by the ``_mfvm_code`` tag.  Called
directly, each function evaluates the same formula elementwise on tensors.
"""

import torch

from .. import _lib


def minmod(bwd_slope, fwd_slope):
    """slope_limiters.py:13-39."""
    agree = (torch.sign(bwd_slope) + torch.sign(fwd_slope)) / 2.0
    return agree * torch.minimum(torch.abs(bwd_slope), torch.abs(fwd_slope))


def monotonized_central(bwd_slope, fwd_slope):
    """slope_limiters.py:42-73 (sign sum over three candidates divided by 3)."""
    c = 0.5 * (bwd_slope + fwd_slope)
    b = 2.0 * bwd_slope
    f = 2.0 * fwd_slope
    agree = (torch.sign(c) + torch.sign(b) + torch.sign(f)) / 3.0
    return agree * torch.minimum(torch.abs(c), torch.minimum(torch.abs(b), torch.abs(f)))


def van_leer(bwd_slope, fwd_slope):
    """slope_limiters.py:76-105 (harmonic mean where |b|+|f| > 0)."""
    agree = (torch.sign(bwd_slope) + torch.sign(fwd_slope)) / 2.0
    denom = torch.abs(bwd_slope) + torch.abs(fwd_slope)
    ok = denom > 0.0
    harm = torch.where(ok, 2.0 * bwd_slope * fwd_slope / torch.where(ok, denom, torch.ones_like(denom)),
                       torch.zeros_like(denom))
    return agree * harm


minmod._mfvm_code = _lib.LIM_MINMOD
monotonized_central._mfvm_code = _lib.LIM_MC
van_leer._mfvm_code = _lib.LIM_VANLEER
