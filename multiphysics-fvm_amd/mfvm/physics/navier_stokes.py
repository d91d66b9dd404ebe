"""Compressible Navier-Stokes (drop-in for src/physics/navier_stokes.py).

``compute_rhs`` and ``compute_dt`` run entirely in libmfvm.so (closure
kernel, fused stage kernel in right-hand-side mode, CFL reduction); the
integrators in :mod:`mfvm.numerics.time_integrators` recognise this module's
``compute_rhs`` and run whole fused stages instead.
"""

import ctypes
import os

import torch

from .. import _lib
from ..state import State
from . import _common as c

VARS = ("mass_density", "momentum_x1", "momentum_x2", "momentum_x3", "energy_density")
NCLOSURE = 7


def allocate_u(grid_config, dtype=torch.float64, device="cuda"):
    """navier_stokes.py:13-41 -- five zero fields with one ghost layer."""
    return State(VARS, grid_config, dtype=dtype, device=device)


def make_params(U, params, space_schemes, bcs=None, tile=None):
    """C descriptor from the reference's params / space_schemes / bcs dicts."""
    p = _lib.NSParams()
    p.dtype = c.dtype_code(U)
    p.limiter = c.limiter_code(space_schemes)
    p.riemann = c.riemann_code(space_schemes)
    ds = space_schemes["diffusive_solver"]
    if getattr(ds, "_mfvm_code", None) != 0:
        raise NotImplementedError(f"diffusive solver {ds!r} has no fused kernel")
    p.gamma = float(params["heat_capacity_ratio"])
    p.mu = float(params["dynamic_viscosity"])
    g = params["gravitational_acceleration"]
    p.g[:] = [float(g[0]), float(g[1]), float(g[2])]
    if bcs is not None:
        rows, _ = c.bc_table(VARS, bcs)
        for v in range(5):
            p.bc[v][:] = rows[v]
    if tile is not None:
        p.tile[:] = list(tile)
    return p


def _closure_buf(U):
    n1, n2, n3 = U.n
    return U.scratch("C", (NCLOSURE, n3 + 2, n2 + 2, n1 + 2), zero=True)


def compute_rhs(U, grid, params, space_schemes):
    """navier_stokes.py:588-711 -- returns fresh (nx1, nx2, nx3) tensors per field."""
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes)
    n1, n2, n3 = U.n
    rhs = torch.empty((5, n3, n2, n1), dtype=U.dtype, device=U.device)
    _lib.check(_lib.lib.mfvm_ns_rhs(g.ptr(), ctypes.byref(p), c.ptr(U.buf), c.ptr(_closure_buf(U)),
                                    c.ptr(rhs), c.stream()))
    return U.interior_views(rhs)


compute_rhs._mfvm_physics = "navier_stokes"


def compute_dt(U, params, grid, cfl):
    """navier_stokes.py:512-585 -- CFL step as a Python float (one device sync)."""
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = _lib.NSParams()
    p.dtype = c.dtype_code(U)
    p.gamma = float(params["heat_capacity_ratio"])
    p.mu = float(params["dynamic_viscosity"])
    n = _lib.lib.mfvm_ns_partials_size(g.ptr(), ctypes.byref(p))
    partials = U.scratch64("partials", max(int(n), 1))
    dt = U.scratch64("dt", 1)
    _lib.check(_lib.lib.mfvm_ns_dt(g.ptr(), ctypes.byref(p), c.ptr(U.buf), float(cfl),
                                   c.ptr(partials), c.ptr(dt), c.stream()))
    return float(dt[0].item())


# 2D grids split from this many cells on: below it the fused stage is faster
# (Kelvin-Helmholtz runs to t = 2 measured 324 ms fused against 379 split at
# 256^2, 1918 against 1467 at 512^2); MFVM_SPLIT_2D=0/1 forces either way.
SPLIT_2D_MIN_CELLS = 512 * 512


def split_2d(U):
    """Whether a 2D (nx3 == 1) state's stages split as well."""
    env = os.environ.get("MFVM_SPLIT_2D")
    if env is not None:
        return env != "0"
    return U.n[0] * U.n[1] >= SPLIT_2D_MIN_CELLS


def split_stages():
    """Whether 3D float64 stages run split in two kernels (x1/x2 faces, then
    x3 faces and the update; identical results).  MFVM_SPLIT=0/1 overrides."""
    return os.environ.get("MFVM_SPLIT", "1") != "0"


def work(U, nsteps=0):
    """Scratch buffers of whole fused steps (mfvm_ns_work)."""
    n1, n2, n3 = U.n
    w = _lib.NSWork()
    w.U = U.buf.data_ptr()
    w.S1 = U.scratch("S1", tuple(U.buf.shape)).data_ptr()
    w.S2 = U.scratch("S2", tuple(U.buf.shape)).data_ptr()
    w.C = _closure_buf(U).data_ptr()
    w.kacc = U.scratch("kacc", (5, n3, n2, n1)).data_ptr()
    if U.dtype == torch.float64 and (n3 > 1 or split_2d(U)) and split_stages():
        w.xy = _xy_buf(U).data_ptr()
    return w


def _xy_buf(U):
    """Interior scratch [5, nx3, nx2, nx1] of split stages (state dtype)."""
    n1, n2, n3 = U.n
    return U.scratch("xy", (5, n3, n2, n1))


def fused_step(U, grid, params, space_schemes, bcs, integrator, dt):
    """One integrator call through mfvm_ns_step."""
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes, bcs)
    w = work(U)
    _lib.check(_lib.lib.mfvm_ns_step(g.ptr(), ctypes.byref(p), ctypes.byref(w), integrator,
                                     float(dt), c.stream()))


def advance(U, grid, params, space_schemes, bcs, integrator, cfl, nsteps, tile=None):
    """``nsteps`` iterations of the reference loop body (dt on the device).

    Returns the device tensor of the nsteps+1 time steps (dts[s] used by step s).
    """
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes, bcs, tile)
    w = work(U)
    n = _lib.lib.mfvm_ns_partials_size(g.ptr(), ctypes.byref(p))
    partials = U.scratch64("partials", max(int(n), 1))
    dts = torch.zeros(nsteps + 1, dtype=torch.float64, device=U.device)
    w.partials = partials.data_ptr()
    w.dts = dts.data_ptr()
    _lib.check(_lib.lib.mfvm_ns_advance(g.ptr(), ctypes.byref(p), ctypes.byref(w), integrator,
                                        float(cfl), int(nsteps), c.stream()))
    return dts


# ----------------------------------------------------------------------------
# Compensated FP32 mode (no reference counterpart: the reduced-precision form)
# ----------------------------------------------------------------------------

def to_compensated(U64):
    """FP32 state (hi) with its compensation words (lo) from an FP64 state:
    hi = fl32(u), lo = fl32(u - hi), so hi + lo carries ~48 significant bits."""
    c.require_state(U64, VARS)
    if U64.dtype != torch.float64:
        raise TypeError("to_compensated expects a float64 state")
    cfg = {"nx1": U64.n[0], "nx2": U64.n[1], "nx3": U64.n[2]}
    U = State(VARS, cfg, dtype=torch.float32, device=U64.device)
    U.buf.copy_(U64.buf.to(torch.float32))
    lo = U.scratch("lo", tuple(U.buf.shape))
    lo.copy_((U64.buf - U.buf.to(torch.float64)).to(torch.float32))
    return U


def compensated_value(U):
    """hi + lo in FP64 (the state the compensated path represents)."""
    lo = U.scratch("lo", tuple(U.buf.shape))
    return U.buf.to(torch.float64) + lo.to(torch.float64)


def advance_compensated(U, grid, params, space_schemes, bcs, integrator, cfl, nsteps, tile=None):
    """``nsteps`` steps with FP32 reconstruction and fluxes and the FP64
    compensated accumulation (mfvm_ns_comp_advance).  U from to_compensated."""
    c.require_state(U, VARS)
    if U.dtype != torch.float32:
        raise TypeError("advance_compensated expects the float32 state of to_compensated")
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes, bcs, tile)
    n1, n2, n3 = U.n
    w = _lib.NSCompWork()
    w.hi = U.buf.data_ptr()
    w.lo = U.scratch("lo", tuple(U.buf.shape)).data_ptr()
    w.Sa = U.scratch("S1", tuple(U.buf.shape)).data_ptr()
    w.Sb = U.scratch("S2", tuple(U.buf.shape)).data_ptr()
    w.C = _closure_buf(U).data_ptr()
    w.dinc = U.scratch("dinc", (5, n3, n2, n1)).data_ptr()
    w.kacc = U.scratch("kacc", (5, n3, n2, n1)).data_ptr()
    if n3 > 1 and os.environ.get("MFVM_SPLIT_COMP", "1") != "0":
        w.xy = _xy_buf(U).data_ptr()
    n = _lib.lib.mfvm_ns_partials_size(g.ptr(), ctypes.byref(p))
    partials = U.scratch64("partials", max(int(n), 1))
    dts = torch.zeros(nsteps + 1, dtype=torch.float64, device=U.device)
    w.partials = partials.data_ptr()
    w.dts = dts.data_ptr()
    _lib.check(_lib.lib.mfvm_ns_comp_advance(g.ptr(), ctypes.byref(p), ctypes.byref(w), integrator,
                                             float(cfl), int(nsteps), c.stream()))
    return dts
