"""Riemann solvers (drop-in for src/numerics/riemann_solvers.py).

Fused into the stage kernel (``face_flux`` in csrc/ns.hip) when used through
``compute_rhs``; callable directly on dicts of tensors.
"""

import torch

from .. import _lib


def rusanov(Fc_L, Fc_R, U_L, U_R, lambda_L, lambda_R):
    """riemann_solvers.py:13-48."""
    al = torch.stack([torch.abs(x) for x in lambda_L.values()]).amax(0)
    ar = torch.stack([torch.abs(x) for x in lambda_R.values()]).amax(0)
    vs = torch.maximum(al, ar)
    return {k: (Fc_L[k] + Fc_R[k]) / 2.0 - vs * (U_R[k] - U_L[k]) / 2.0 for k in Fc_L}


def hll(Fc_L, Fc_R, U_L, U_R, lambda_L, lambda_R):
    """riemann_solvers.py:51-89."""
    smin = torch.minimum(torch.stack(list(lambda_L.values())).amin(0),
                         torch.stack(list(lambda_R.values())).amin(0))
    smax = torch.maximum(torch.stack(list(lambda_L.values())).amax(0),
                         torch.stack(list(lambda_R.values())).amax(0))
    out = {}
    for k in Fc_L:
        mid = (smax * Fc_L[k] - smin * Fc_R[k] + smin * smax * (U_R[k] - U_L[k])) / (smax - smin)
        out[k] = torch.where(smin >= 0.0, Fc_L[k], torch.where(smax <= 0.0, Fc_R[k], mid))
    return out


rusanov._mfvm_code = _lib.RS_RUSANOV
hll._mfvm_code = _lib.RS_HLL
