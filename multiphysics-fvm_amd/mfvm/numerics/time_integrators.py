"""Time integrators (drop-in for src/numerics/time_integrators.py).

When ``rhs`` is an mfvm physics ``compute_rhs``, ``U`` an mfvm State and
``apply_bcs`` mfvm's, one call runs the whole step in libmfvm.so
(``mfvm_ns_step``: per stage a ghost-fill kernel, a closure kernel and one
fused flux/divergence/update kernel).  Any other ``rhs``/``apply_bcs``
callables are driven exactly as the reference does, with tensor updates of
the interior views.
"""

from .. import _lib

_IN = (slice(1, -1),) * 3


def _fused(rhs, U, apply_bcs):
    from .. import solver
    from ..state import State
    return (getattr(rhs, "_mfvm_physics", None) is not None and isinstance(U, State)
            and apply_bcs is solver.apply_bcs)


def _run_fused(code, rhs, U, dt, params, grid, space_schemes, bcs):
    import importlib
    mod = importlib.import_module(f"mfvm.physics.{rhs._mfvm_physics}")
    mod.fused_step(U, grid, params, space_schemes, bcs, code, dt)


def euler_explicit(rhs, U, dt, params, grid, space_schemes, bcs, apply_bcs):
    """time_integrators.py:10-47."""
    if _fused(rhs, U, apply_bcs):
        return _run_fused(_lib.INT_EULER, rhs, U, dt, params, grid, space_schemes, bcs)
    core = {k: U[k][_IN] for k in U}
    apply_bcs(U, bcs)
    k1 = rhs(U, grid, params, space_schemes)
    for k in core:
        core[k][:] += dt * k1[k]


def rk3(rhs, U, dt, params, grid, space_schemes, bcs, apply_bcs):
    """time_integrators.py:50-103 (stages U0+dt/2 k1, U0+dt k2, U0+dt/6(k1+4k2+k3))."""
    if _fused(rhs, U, apply_bcs):
        return _run_fused(_lib.INT_RK3, rhs, U, dt, params, grid, space_schemes, bcs)
    core = {k: U[k][_IN] for k in U}
    base = {k: core[k].clone() for k in core}
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
    """time_integrators.py:106-159 (Shu-Osher SSPRK3)."""
    if _fused(rhs, U, apply_bcs):
        return _run_fused(_lib.INT_SSPRK3, rhs, U, dt, params, grid, space_schemes, bcs)
    core = {k: U[k][_IN] for k in U}
    base = {k: core[k].clone() for k in core}
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
    """time_integrators.py:162-221 (classical RK4)."""
    if _fused(rhs, U, apply_bcs):
        return _run_fused(_lib.INT_RK4, rhs, U, dt, params, grid, space_schemes, bcs)
    core = {k: U[k][_IN] for k in U}
    base = {k: core[k].clone() for k in core}
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


euler_explicit._mfvm_code = _lib.INT_EULER
rk3._mfvm_code = _lib.INT_RK3
rk3_ssp._mfvm_code = _lib.INT_SSPRK3
rk4._mfvm_code = _lib.INT_RK4
