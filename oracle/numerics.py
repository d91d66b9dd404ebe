"""Oracle restatement of the reference numerical building blocks.

Test infrastructure only (see ``oracle/__init__.py``).

Covers ``src/numerics/slope_limiters.py``, ``reconstructors.py``,
``riemann_solvers.py``, ``diffusive_solvers.py`` and ``time_integrators.py``
of multiphysics-fvm.  Operation order follows the reference expression by
expression; comments give the reference line each block restates.
"""

import numpy as np

DIRS = ("x1", "x2", "x3")


# ----------------------------------------------------------------------------
# Slope limiters (src/numerics/slope_limiters.py)
# ----------------------------------------------------------------------------

def minmod(bwd_slope, fwd_slope):
    """slope_limiters.py:13-39 -- sign agreement halved, times min |slope|."""
    agree = (np.sign(bwd_slope) + np.sign(fwd_slope)) / 2.0
    least = np.minimum(np.abs(bwd_slope), np.abs(fwd_slope))
    return agree * least


def monotonized_central(bwd_slope, fwd_slope):
    """slope_limiters.py:42-73 -- three candidates; sign sum divided by 3.

    The sign factor is (s_c + s_b + s_f)/3, so a 2-to-1 sign split yields
    +-1/3 of the smallest candidate instead of zero (reference behaviour).
    """
    cand_c = 0.5 * (bwd_slope + fwd_slope)
    cand_b = 2.0 * bwd_slope
    cand_f = 2.0 * fwd_slope
    agree = (np.sign(cand_c) + np.sign(cand_b) + np.sign(cand_f)) / 3.0
    least = np.minimum(np.abs(cand_c), np.minimum(np.abs(cand_b), np.abs(cand_f)))
    return agree * least


def van_leer(bwd_slope, fwd_slope):
    """slope_limiters.py:76-105 -- masked harmonic mean 2ab/(|a|+|b|)."""
    agree = (np.sign(bwd_slope) + np.sign(fwd_slope)) / 2.0
    denom = np.abs(bwd_slope) + np.abs(fwd_slope)
    harm = np.zeros_like(bwd_slope)
    ok = denom > 0.0
    harm[ok] = 2.0 * bwd_slope[ok] * fwd_slope[ok] / denom[ok]
    return agree * harm


# ----------------------------------------------------------------------------
# Reconstructors (src/numerics/reconstructors.py)
# ----------------------------------------------------------------------------

def _sl(axis, along, rest=slice(1, -1)):
    """Index tuple: ``along`` on ``axis``, ``rest`` on the other two axes."""
    idx = [rest, rest, rest]
    idx[axis] = along
    return tuple(idx)


# Face-adjacent cells: left cell of each face (":-1") and right cell ("1:")
# along the face normal, transverse ghosts excluded (reconstructors.py:37-42,
# 108-113).
_FACE_LO = [_sl(a, slice(None, -1)) for a in range(3)]
_FACE_HI = [_sl(a, slice(1, None)) for a in range(3)]
_CELL_M = [_sl(a, slice(None, -2)) for a in range(3)]
_CELL_P = [_sl(a, slice(2, None)) for a in range(3)]
_INTERIOR = (slice(1, -1), slice(1, -1), slice(1, -1))


def piecewise_constant(fields, L, slope_limiter):
    """reconstructors.py:13-64 -- face states are the adjacent cell values."""
    out_l = {d: {} for d in DIRS}
    out_r = {d: {} for d in DIRS}
    for ax, d in enumerate(DIRS):
        for name, f in fields.items():
            out_l[d][name] = f[_FACE_LO[ax]]
            out_r[d][name] = f[_FACE_HI[ax]]
    return out_l, out_r


def muscl(fields, L, slope_limiter):
    """reconstructors.py:67-158 -- limited linear reconstruction.

    back/front differences are divided by the centre cell length, the limited
    slope is multiplied back by it, and ghost cells carry a zero slope
    (``np.zeros_like`` filled on the interior only, :127-132), so the ghost
    side of each boundary face is first order.
    """
    out_l = {d: {} for d in DIRS}
    out_r = {d: {} for d in DIRS}
    for name, f in fields.items():
        centre = f[_INTERIOR]
        for ax, d in enumerate(DIRS):
            length = L[d]
            back = (centre - f[_CELL_M[ax]]) / length
            front = (f[_CELL_P[ax]] - centre) / length
            slope = np.zeros_like(f)
            slope[_INTERIOR] = slope_limiter(back, front) * length
            out_l[d][name] = f[_FACE_LO[ax]] + slope[_FACE_LO[ax]] / 2.0
            out_r[d][name] = f[_FACE_HI[ax]] - slope[_FACE_HI[ax]] / 2.0
    return out_l, out_r


# ----------------------------------------------------------------------------
# Riemann solvers (src/numerics/riemann_solvers.py)
# ----------------------------------------------------------------------------

def rusanov(Fc_L, Fc_R, U_L, U_R, lambda_L, lambda_R):
    """riemann_solvers.py:13-48 -- local Lax-Friedrichs flux."""
    amax_l = np.maximum.reduce([np.abs(w) for w in lambda_L.values()])
    amax_r = np.maximum.reduce([np.abs(w) for w in lambda_R.values()])
    speed = np.maximum(amax_l, amax_r)
    return {k: (Fc_L[k] + Fc_R[k]) / 2.0 - speed * (U_R[k] - U_L[k]) / 2.0
            for k in Fc_L}


def hll(Fc_L, Fc_R, U_L, U_R, lambda_L, lambda_R):
    """riemann_solvers.py:51-89 -- two-wave HLL flux with upwind branches."""
    smin = np.minimum(np.minimum.reduce(list(lambda_L.values())),
                      np.minimum.reduce(list(lambda_R.values())))
    smax = np.maximum(np.maximum.reduce(list(lambda_L.values())),
                      np.maximum.reduce(list(lambda_R.values())))
    out = {}
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in Fc_L:
            mid = (smax * Fc_L[k] - smin * Fc_R[k]
                   + smin * smax * (U_R[k] - U_L[k])) / (smax - smin)
            out[k] = np.where(smin >= 0.0, Fc_L[k],
                              np.where(smax <= 0.0, Fc_R[k], mid))
    return out


# ----------------------------------------------------------------------------
# Diffusive solver (src/numerics/diffusive_solvers.py)
# ----------------------------------------------------------------------------

def centered(Fd_l, Fd_r):
    """diffusive_solvers.py:11-30 -- arithmetic mean of the two face fluxes."""
    return {k: (Fd_l[k] + Fd_r[k]) / 2.0 for k in Fd_l}


# ----------------------------------------------------------------------------
# Time integrators (src/numerics/time_integrators.py)
# ----------------------------------------------------------------------------

def _interior_views(U):
    return {k: U[k][_INTERIOR] for k in U}


def euler_explicit(rhs, U, dt, params, grid, space_schemes, bcs, apply_bcs):
    """time_integrators.py:10-47 -- U += dt*k1, in place on the interior."""
    core = _interior_views(U)
    apply_bcs(U, bcs)
    k1 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] += dt * k1[k]


def rk3(rhs, U, dt, params, grid, space_schemes, bcs, apply_bcs):
    """time_integrators.py:50-103 -- the reference's three-stage scheme.

    Stages: U0 + dt/2 k1, U0 + dt k2, U0 + dt/6 (k1 + 4 k2 + k3).  This is
    not Kutta's third-order method (its linear amplification factor is
    1 + z + z^2/2 + z^3/12).
    """
    core = _interior_views(U)
    base = {k: core[k].copy() for k in core}
    apply_bcs(U, bcs)
    k1 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = base[k] + 0.5 * dt * k1[k]
    apply_bcs(U, bcs)
    k2 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = base[k] + dt * k2[k]
    apply_bcs(U, bcs)
    k3 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = base[k] + (dt / 6.0) * (k1[k] + 4.0 * k2[k] + k3[k])


def rk3_ssp(rhs, U, dt, params, grid, space_schemes, bcs, apply_bcs):
    """time_integrators.py:106-159 -- Shu-Osher SSPRK3."""
    core = _interior_views(U)
    base = {k: core[k].copy() for k in core}
    apply_bcs(U, bcs)
    k1 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = base[k] + dt * k1[k]
    apply_bcs(U, bcs)
    k2 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = 0.75 * base[k] + 0.25 * (core[k] + dt * k2[k])
    apply_bcs(U, bcs)
    k3 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = (1.0 / 3.0) * base[k] + (2.0 / 3.0) * (core[k] + dt * k3[k])


def rk4(rhs, U, dt, params, grid, space_schemes, bcs, apply_bcs):
    """time_integrators.py:162-221 -- classical fourth-order Runge-Kutta."""
    core = _interior_views(U)
    base = {k: core[k].copy() for k in core}
    apply_bcs(U, bcs)
    k1 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = base[k] + 0.5 * dt * k1[k]
    apply_bcs(U, bcs)
    k2 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = base[k] + 0.5 * dt * k2[k]
    apply_bcs(U, bcs)
    k3 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = base[k] + dt * k3[k]
    apply_bcs(U, bcs)
    k4 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] = base[k] + (dt / 6.0) * (k1[k] + 2.0 * k2[k] + 2.0 * k3[k] + k4[k])
