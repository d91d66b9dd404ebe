"""Diffusive solver (drop-in for src/numerics/diffusive_solvers.py)."""


def centered(Fd_l, Fd_r):
    """diffusive_solvers.py:11-30 -- mean of the left and right face fluxes."""
    return {k: (Fd_l[k] + Fd_r[k]) / 2.0 for k in Fd_l}


centered._mfvm_code = 0
