"""Maxwell in B/D form with cleaning (drop-in for src/physics/maxwell.py)."""

import ctypes
import os

import torch

from .. import _lib
from ..state import State
from . import _common as c

VARS = ("magnetic_field_x1", "magnetic_field_x2", "magnetic_field_x3",
        "electric_d_field_x1", "electric_d_field_x2", "electric_d_field_x3")


def allocate_u(grid_config, dtype=torch.float64, device="cuda"):
    """maxwell.py:13-42."""
    return State(VARS, grid_config, dtype=dtype, device=device)


def make_params(U, params, space_schemes=None, bcs=None):
    """C descriptor from the reference dicts.  Keys are read as the reference
    reads them (a missing one raises KeyError): the closure's permittivity
    and conductivity always (maxwell.py:70-71, also on the dt path), the
    charge density where the right-hand side needs it (:475).  Conductivity
    and charge density may be numbers or arrays broadcast over the interior."""
    p = _lib.MaxwellParams()
    p.dtype = c.dtype_code(U)
    if space_schemes is not None:
        p.limiter = c.limiter_code(space_schemes)
        p.riemann = c.riemann_code(space_schemes)
    p.mu = float(params["magnetic_permeability"])
    p.ep = float(params["electric_permittivity"])
    keep = []
    p.sigma, field = c.cell_param(U, params["electrical_conductivity"])
    if field is not None:
        p.sigma_field = field.data_ptr()
        keep.append(field)
    if space_schemes is not None:
        p.rho, field = c.cell_param(U, params["charge_density"])
        if field is not None:
            p.rho_field = field.data_ptr()
            keep.append(field)
    p._keep = keep  # alive as long as the descriptor
    p.chi_b = float(params["magnetic_cleaning_coeff"])
    p.chi_d = float(params["electric_cleaning_coeff"])
    if bcs is not None:
        rows, _ = c.bc_table(VARS, bcs)
        for v in range(6):
            p.bc[v][:] = rows[v]
    return p


def _res_buf(U):
    n1, n2, n3 = U.n
    return U.scratch("R", (2, n3 + 2, n2 + 2, n1 + 2), zero=True)


def compute_rhs(U, grid, params, space_schemes):
    """maxwell.py:603-729 -- fresh (nx1, nx2, nx3) tensors per field."""
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes)
    n1, n2, n3 = U.n
    rhs = torch.empty((6, n3, n2, n1), dtype=U.dtype, device=U.device)
    _lib.check(_lib.lib.mfvm_maxwell_rhs(g.ptr(), ctypes.byref(p), c.ptr(U.buf), c.ptr(_res_buf(U)),
                                         c.ptr(rhs), c.stream()))
    return U.interior_views(rhs)


compute_rhs._mfvm_physics = "maxwell"


def compute_dt(U, params, grid, cfl):
    """maxwell.py:533-600."""
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = make_params(U, params)
    dt = U.scratch64("dt", 1)
    _lib.check(_lib.lib.mfvm_maxwell_dt(g.ptr(), ctypes.byref(p), float(cfl), c.ptr(dt), c.stream()))
    return float(dt[0].item())


def work(U):
    n1, n2, n3 = U.n
    w = _lib.NSWork()
    w.U = U.buf.data_ptr()
    w.S1 = U.scratch("S1", tuple(U.buf.shape)).data_ptr()
    w.S2 = U.scratch("S2", tuple(U.buf.shape)).data_ptr()
    w.C = _res_buf(U).data_ptr()
    w.kacc = U.scratch("kacc", (6, n3, n2, n1)).data_ptr()
    # split stages (x1/x2 flux differences, then the x3 pass; identical
    # results); MFVM_SPLIT=0 selects the fused stage kernel
    if os.environ.get("MFVM_SPLIT", "1") != "0":
        w.xy = U.scratch("xy", (6, n3, n2, n1)).data_ptr()
    return w


def fused_step(U, grid, params, space_schemes, bcs, integrator, dt):
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes, bcs)
    w = work(U)
    _lib.check(_lib.lib.mfvm_maxwell_step(g.ptr(), ctypes.byref(p), ctypes.byref(w), integrator,
                                          float(dt), c.stream()))


def advance(U, grid, params, space_schemes, bcs, integrator, cfl, nsteps):
    c.require_state(U, VARS)
    g = c.require_grid(grid, U)
    p = make_params(U, params, space_schemes, bcs)
    w = work(U)
    dts = torch.zeros(nsteps + 1, dtype=torch.float64, device=U.device)
    w.dts = dts.data_ptr()
    _lib.check(_lib.lib.mfvm_maxwell_advance(g.ptr(), ctypes.byref(p), ctypes.byref(w), integrator,
                                             float(cfl), int(nsteps), c.stream()))
    return dts
