"""Heat equation (drop-in for src/physics/heat.py), computed in libmfvm.so."""

import ctypes

import torch

from .. import _lib
from ..state import State
from . import _common as c

VARS = ("internal_energy_density",)


def allocate_u(grid_config, dtype=torch.float64, device="cuda"):
    """heat.py:13-37."""
    return State(VARS, grid_config, dtype=dtype, device=device)


def make_params(U, params, space_schemes, bcs=None):
    p = _lib.HeatParams()
    p.dtype = c.dtype_code(U)
    p.limiter = c.limiter_code(space_schemes)
    ds = space_schemes["diffusive_solver"]
    if getattr(ds, "_mfvm_code", None) != 0:
        raise NotImplementedError(f"diffusive solver {ds!r} has no fused kernel")
    p.rho_cp = float(params["mass_density"]) * float(params["specific_heat_capacity"])
    p.conductivity = float(params["thermal_conductivity"])
    # heat.py:294-297: q is a number or an array broadcast over the interior
    p.source, field = c.cell_param(U, params["volumetric_heat_source"])
    if field is not None:
        p.source_field = field.data_ptr()
        p._keep = field  # alive as long as the descriptor
    if bcs is not None:
        rows, _ = c.bc_table(VARS, bcs)
        p.bc[0][:] = rows[0]
    return p


def compute_rhs(U, grid, params, space_schemes):
    """heat.py:354-473 -- fresh (nx1, nx2, nx3) tensor."""
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes)
    n1, n2, n3 = U.n
    rhs = torch.empty((1, n3, n2, n1), dtype=U.dtype, device=U.device)
    _lib.check(_lib.lib.mfvm_heat_rhs(g.ptr(), ctypes.byref(p), c.ptr(U.buf), c.ptr(rhs), c.stream()))
    return U.interior_views(rhs)


compute_rhs._mfvm_physics = "heat"


def _dt_params(U, params):
    p = _lib.HeatParams()
    p.dtype = c.dtype_code(U)
    p.rho_cp = float(params["mass_density"]) * float(params["specific_heat_capacity"])
    p.conductivity = float(params["thermal_conductivity"])
    return p


def compute_dt(U, params, grid, cfl):
    """heat.py:305-351."""
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = _dt_params(U, params)
    dt = U.scratch64("dt", 1)
    _lib.check(_lib.lib.mfvm_heat_dt(g.ptr(), ctypes.byref(p), float(cfl), c.ptr(dt), c.stream()))
    return float(dt[0].item())


def work(U):
    n1, n2, n3 = U.n
    w = _lib.NSWork()
    w.U = U.buf.data_ptr()
    w.S1 = U.scratch("S1", tuple(U.buf.shape)).data_ptr()
    w.S2 = U.scratch("S2", tuple(U.buf.shape)).data_ptr()
    w.C = None
    w.kacc = U.scratch("kacc", (1, n3, n2, n1)).data_ptr()
    return w


def fused_step(U, grid, params, space_schemes, bcs, integrator, dt):
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes, bcs)
    w = work(U)
    _lib.check(_lib.lib.mfvm_heat_step(g.ptr(), ctypes.byref(p), ctypes.byref(w), integrator,
                                       float(dt), c.stream()))


def advance(U, grid, params, space_schemes, bcs, integrator, cfl, nsteps):
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes, bcs)
    w = work(U)
    dts = torch.zeros(nsteps + 1, dtype=torch.float64, device=U.device)
    w.dts = dts.data_ptr()
    _lib.check(_lib.lib.mfvm_heat_advance(g.ptr(), ctypes.byref(p), ctypes.byref(w), integrator,
                                          float(cfl), int(nsteps), c.stream()))
    return dts
