"""Oracle restatement of the Maxwell module (B/D form with cleaning).

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``src/physics/maxwell.py`` of multiphysics-fvm: curl fluxes of
H = B/mu and E = D/eps, light-speed wave set, Ohm current source (zero-padded
closure), curvilinear geometric source, divergence-cleaning source with an
edge-padded residual, and the light-speed plus cleaning CFL step.
"""

import numpy as np

from .navier_stokes import _CM, _CP, _ING, divergence

VARS = ("magnetic_field_x1", "magnetic_field_x2", "magnetic_field_x3",
        "electric_d_field_x1", "electric_d_field_x2", "electric_d_field_x3")


def allocate_u(grid_config):
    """maxwell.py:13-42."""
    shape = tuple(grid_config[f"nx{a}"] + 2 for a in (1, 2, 3))
    return {k: np.zeros(shape) for k in VARS}


def compute_closure(U, params, grid):
    """maxwell.py:45-97 -- J = -sigma E on the interior, zero-padded."""
    ep = params["electric_permittivity"]
    sigma = params["electrical_conductivity"]
    J = [np.pad(-sigma * (U[f"electric_d_field_x{a}"] / ep)[_ING], 1) for a in (1, 2, 3)]
    return {"Jx1": J[0], "Jx2": J[1], "Jx3": J[2]}


def _flux_tensors(U, params):
    """Curl flux tensors FB[a][b], FD[a][b] (maxwell.py:137-165); the diagonal
    entries are the Python float 0.0, as in the reference."""
    mu = params["magnetic_permeability"]
    ep = params["electric_permittivity"]
    H = [U[f"magnetic_field_x{a}"] / mu for a in (1, 2, 3)]
    E = [U[f"electric_d_field_x{a}"] / ep for a in (1, 2, 3)]
    fb12, fb13, fb23 = E[2], -E[1], E[0]
    fd12, fd13, fd23 = -H[2], H[1], -H[0]
    FB = [[0.0, fb12, fb13], [-fb12, 0.0, fb23], [-fb13, -fb23, 0.0]]
    FD = [[0.0, fd12, fd13], [-fd12, 0.0, fd23], [-fd13, -fd23, 0.0]]
    return FB, FD


def compute_convective_flux(U, C, params, n):
    """maxwell.py:100-183 -- n-weighted sum of the curl flux components."""
    FB, FD = _flux_tensors(U, params)
    out = {}
    for a in range(3):
        out[VARS[a]] = n[0] * FB[a][0] + n[1] * FB[a][1] + n[2] * FB[a][2]
    for a in range(3):
        out[VARS[3 + a]] = n[0] * FD[a][0] + n[1] * FD[a][1] + n[2] * FD[a][2]
    return out


def compute_characteristic_velocity(U, C, params):
    """maxwell.py:229-262 -- +-1/sqrt(mu eps) everywhere."""
    c = 1 / np.sqrt(params["magnetic_permeability"] * params["electric_permittivity"])
    shape = U["magnetic_field_x1"].shape
    return {"light_minus": np.full(shape, -c), "light_plus": np.full(shape, +c)}


def compute_geometric_source(U, C, params, grid):
    """maxwell.py:265-376 -- curvilinear source of the curl fluxes."""
    lame = grid["lame"]
    h1, h2, h3 = (lame["coefficients"][d] for d in ("x1", "x2", "x3"))
    dv = lame["derivatives"]
    d1h2, d1h3 = dv["x1"]["ln_hx2"], dv["x1"]["ln_hx3"]
    d2h1, d2h3 = dv["x2"]["ln_hx1"], dv["x2"]["ln_hx3"]
    d3h1, d3h2 = dv["x3"]["ln_hx1"], dv["x3"]["ln_hx2"]
    out = {}
    for base, T in zip((0, 3), _flux_tensors(U, params)):
        t = [[x[_ING] if isinstance(x, np.ndarray) else x for x in row] for row in T]
        out[VARS[base + 0]] = (t[0][1] * d2h1 / h2 + t[0][2] * d3h1 / h3
                               - t[1][1] * d1h2 / h1 - t[2][2] * d1h3 / h1)
        out[VARS[base + 1]] = (t[1][0] * d1h2 / h1 + t[1][2] * d3h2 / h3
                               - t[0][0] * d2h1 / h2 - t[2][2] * d2h3 / h2)
        out[VARS[base + 2]] = (t[2][0] * d1h3 / h1 + t[2][1] * d2h3 / h2
                               - t[0][0] * d3h1 / h3 - t[1][1] * d3h2 / h3)
    return out


def compute_physical_source(U, C, params):
    """maxwell.py:379-425 -- J on the D equations."""
    return {VARS[0]: 0.0, VARS[1]: 0.0, VARS[2]: 0.0,
            VARS[3]: C["Jx1"][_ING], VARS[4]: C["Jx2"][_ING], VARS[5]: C["Jx3"][_ING]}


def compute_constraint_cleaning_source(U, params, grid):
    """maxwell.py:428-530 -- -chi grad(div B) and -chi grad(div D - rho),
    the residual gradient taken on an edge-padded residual."""
    L = grid["cell"]["geometry"]["length"]
    Ls = (L["x1"], L["x2"], L["x3"])
    lame = grid["lame"]
    h1, h2, h3 = (lame["coefficients"][d] for d in ("x1", "x2", "x3"))
    dv = lame["derivatives"]
    d1h2, d1h3 = dv["x1"]["ln_hx2"], dv["x1"]["ln_hx3"]
    d2h1, d2h3 = dv["x2"]["ln_hx1"], dv["x2"]["ln_hx3"]
    d3h1, d3h2 = dv["x3"]["ln_hx1"], dv["x3"]["ln_hx2"]
    chi = (params["magnetic_cleaning_coeff"], params["electric_cleaning_coeff"])
    out = {}
    for base, name, rho in ((0, "magnetic_field_x", None),
                            (3, "electric_d_field_x", params["charge_density"])):
        W = [U[f"{name}{a}"] for a in (1, 2, 3)]
        grad = [(W[a][_CP[a]] - W[a][_CM[a]]) / (2.0 * Ls[a]) for a in range(3)]
        c1, c2, c3 = (w[_ING] for w in W)
        div = (grad[0] + grad[1] + grad[2] + c1 * (d1h2 + d1h3) / h1
               + c2 * (d2h1 + d2h3) / h2 + c3 * (d3h1 + d3h2) / h3)
        res = div if rho is None else div - rho
        rg = np.pad(res, 1, mode="edge")
        c = chi[0] if rho is None else chi[1]
        for a in range(3):
            out[VARS[base + a]] = -c * ((rg[_CP[a]] - rg[_CM[a]]) / (2.0 * Ls[a]))
    return out


def compute_dt(U, params, grid, cfl):
    """maxwell.py:533-600 -- light speed over L plus 2 max(chi) sum 1/L^2."""
    L = grid["cell"]["geometry"]["length"]
    kappa = compute_characteristic_velocity(U, None, params)
    amax = np.maximum.reduce([np.abs(x[_ING]) for x in kappa.values()])
    rate_c = np.maximum.reduce([amax / L["x1"], amax / L["x2"], amax / L["x3"]])
    chi_max = max(params["magnetic_cleaning_coeff"], params["electric_cleaning_coeff"])
    rate_d = 2.0 * chi_max * (1.0 / L["x1"] ** 2 + 1.0 / L["x2"] ** 2 + 1.0 / L["x3"] ** 2)
    return cfl / (np.max(rate_c) + np.max(rate_d))


def compute_rhs(U, grid, params, space_schemes):
    """maxwell.py:603-729 -- -sum(F.S)/V + geometric + current + cleaning."""
    V = grid["cell"]["geometry"]["volume"]
    S = grid["cell"]["geometry"]["surface"]
    L = grid["cell"]["geometry"]["length"]
    n = grid["cell"]["normal"]
    riemann = space_schemes["riemann_solver"]
    C = compute_closure(U, params, grid)
    U_L, U_R = space_schemes["reconstructor"](U, L, space_schemes["slope_limiter"])
    F = {}
    for d in ("x1", "x2", "x3"):
        fcl = compute_convective_flux(U_L[d], None, params, n[d])
        fcr = compute_convective_flux(U_R[d], None, params, n[d])
        laml = compute_characteristic_velocity(U_L[d], None, params)
        lamr = compute_characteristic_velocity(U_R[d], None, params)
        fc = riemann(fcl, fcr, U_L[d], U_R[d], laml, lamr)
        F[d] = {k: fc[k] + 0.0 for k in fc}
    sg = compute_geometric_source(U, C, params, grid)
    sp = compute_physical_source(U, C, params)
    sc = compute_constraint_cleaning_source(U, params, grid)
    src = {k: sg[k] + sp[k] + sc[k] for k in sg}
    return divergence(F, S, V, src, VARS)
