"""Oracle restatement of the compressible Navier-Stokes module.

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``src/physics/navier_stokes.py`` of multiphysics-fvm: closure
(pressure, Stokes viscous tensor with Lame terms), convective and viscous face
fluxes, wave speeds, geometric and gravity sources, the CFL time step and the
finite-volume right-hand side, all on whole arrays in the reference order.
"""

import numpy as np

VARS = ("mass_density", "momentum_x1", "momentum_x2", "momentum_x3", "energy_density")
_ING = (slice(1, -1), slice(1, -1), slice(1, -1))
_CM = [(slice(None, -2), slice(1, -1), slice(1, -1)),
       (slice(1, -1), slice(None, -2), slice(1, -1)),
       (slice(1, -1), slice(1, -1), slice(None, -2))]
_CP = [(slice(2, None), slice(1, -1), slice(1, -1)),
       (slice(1, -1), slice(2, None), slice(1, -1)),
       (slice(1, -1), slice(1, -1), slice(2, None))]
_T_KEYS = ("Tx1x1", "Tx2x2", "Tx3x3", "Tx1x2", "Tx1x3", "Tx2x3", "Tx2x1", "Tx3x1", "Tx3x2")


def allocate_u(grid_config):
    """navier_stokes.py:13-41 -- five zero arrays with one ghost layer."""
    shape = tuple(grid_config[f"nx{a}"] + 2 for a in (1, 2, 3))
    return {k: np.zeros(shape) for k in VARS}


def _lame(grid):
    lame = grid["lame"]
    h = [lame["coefficients"][d] for d in ("x1", "x2", "x3")]
    dv = lame["derivatives"]
    return h, dv


def compute_closure(U, params, grid):
    """navier_stokes.py:44-156 -- P on all cells; viscous T on the interior,
    zero-padded to the ghost layer (so ghost T is exactly 0)."""
    rho = U["mass_density"]
    gamma = params["heat_capacity_ratio"]
    mu = params["dynamic_viscosity"]
    L = grid["cell"]["geometry"]["length"]
    (h1, h2, h3), dv = _lame(grid)
    d1h2, d1h3 = dv["x1"]["ln_hx2"], dv["x1"]["ln_hx3"]
    d2h1, d2h3 = dv["x2"]["ln_hx1"], dv["x2"]["ln_hx3"]
    d3h1, d3h2 = dv["x3"]["ln_hx1"], dv["x3"]["ln_hx2"]

    v = [U["momentum_x1"] / rho, U["momentum_x2"] / rho, U["momentum_x3"] / rho]
    E = U["energy_density"] / rho
    P = (gamma - 1.0) * rho * (E - (v[0] ** 2 + v[1] ** 2 + v[2] ** 2) / 2.0)   # :100

    # g[a][b] = d v_a / d x_b, centred, over 2*L_b of the centre cell (:111-119)
    Ls = [L["x1"], L["x2"], L["x3"]]
    g = [[(v[a][_CP[b]] - v[a][_CM[b]]) / (2 * Ls[b]) for b in range(3)] for a in range(3)]
    v1, v2, v3 = (vv[_ING] for vv in v)

    divv = (g[0][0] + g[1][1] + g[2][2] + v1 * (d1h2 + d1h3) / h1
            + v2 * (d2h1 + d2h3) / h2 + v3 * (d3h1 + d3h2) / h3)                    # :128
    t11 = mu * (2 * g[0][0] - (2.0 / 3.0) * divv)
    t22 = mu * (2 * (g[1][1] + v1 * d1h2 / h1) - (2.0 / 3.0) * divv)
    t33 = mu * (2 * (g[2][2] + v1 * d1h3 / h1 + v2 * d2h3 / h2) - (2.0 / 3.0) * divv)
    t12 = mu * (g[0][1] + g[1][0] - v2 * d1h2 / h1 + v1 * d2h1 / h2)
    t13 = mu * (g[0][2] + g[2][0] - v3 * d1h3 / h1 + v1 * d3h1 / h3)
    t23 = mu * (g[1][2] + g[2][1] - v3 * d2h3 / h2 + v2 * d3h2 / h3)              # :131-136
    t11, t22, t33, t12, t13, t23 = (np.pad(t, 1) for t in (t11, t22, t33, t12, t13, t23))
    # Symmetric entries are the same arrays (:137-139).
    return {"P": P, "Tx1x1": t11, "Tx2x2": t22, "Tx3x3": t33, "Tx1x2": t12,
            "Tx1x3": t13, "Tx2x3": t23, "Tx2x1": t12, "Tx3x1": t13, "Tx3x2": t23}


def compute_convective_flux(U, C, params, n):
    """navier_stokes.py:159-214 -- Euler flux with the reconstructed P,
    projected on the face normal n (n-weighted sum of the three components)."""
    rho = U["mass_density"]
    P = C["P"]
    v = [U["momentum_x1"] / rho, U["momentum_x2"] / rho, U["momentum_x3"] / rho]
    E = U["energy_density"] / rho
    F = {"mass_density": [rho * v[d] for d in range(3)]}
    for a in range(3):
        F[VARS[1 + a]] = [rho * v[a] * v[d] + P if a == d else rho * v[a] * v[d]
                          for d in range(3)]
    F["energy_density"] = [(rho * E + P) * v[d] for d in range(3)]
    return {k: n[0] * F[k][0] + n[1] * F[k][1] + n[2] * F[k][2] for k in VARS}


def compute_diffusive_flux(U, C, n):
    """navier_stokes.py:217-277 -- viscous momentum and energy flux, row d of
    T for the d-direction, velocities from the reconstructed U."""
    rho = U["mass_density"]
    v = [U["momentum_x1"] / rho, U["momentum_x2"] / rho, U["momentum_x3"] / rho]
    T = [[C["Tx1x1"], C["Tx1x2"], C["Tx1x3"]],
         [C["Tx2x1"], C["Tx2x2"], C["Tx2x3"]],
         [C["Tx3x1"], C["Tx3x2"], C["Tx3x3"]]]
    F = {"mass_density": [0.0, 0.0, 0.0]}
    for a in range(3):
        F[VARS[1 + a]] = [-T[a][d] for d in range(3)]
    F["energy_density"] = [-T[d][0] * v[0] - T[d][1] * v[1] - T[d][2] * v[2]
                           for d in range(3)]
    return {k: n[0] * F[k][0] + n[1] * F[k][1] + n[2] * F[k][2] for k in VARS}


def compute_characteristic_velocity(U, C, params):
    """navier_stokes.py:280-316 -- {0, -cs, +cs} with cs = sqrt(gamma P/rho)."""
    rho = U["mass_density"]
    cs = np.sqrt(params["heat_capacity_ratio"] * C["P"] / rho)
    return {"contact": np.full(rho.shape, 0), "acoustic_minus": -cs, "acoustic_plus": cs}


def compute_normal_velocity(U, n):
    """navier_stokes.py:319-355."""
    rho = U["mass_density"]
    return (n[0] * (U["momentum_x1"] / rho) + n[1] * (U["momentum_x2"] / rho)
            + n[2] * (U["momentum_x3"] / rho))


def compute_geometric_source(U, C, params, grid):
    """navier_stokes.py:358-454 -- convective and viscous curvilinear sources
    (exactly +-0 on Cartesian grids)."""
    (h1, h2, h3), dv = _lame(grid)
    d1h, d1h2, d1h3 = dv["x1"]["ln_hx1hx2hx3"], dv["x1"]["ln_hx2"], dv["x1"]["ln_hx3"]
    d2h, d2h1, d2h3 = dv["x2"]["ln_hx1hx2hx3"], dv["x2"]["ln_hx1"], dv["x2"]["ln_hx3"]
    d3h, d3h1, d3h2 = dv["x3"]["ln_hx1hx2hx3"], dv["x3"]["ln_hx1"], dv["x3"]["ln_hx2"]
    rho_g = U["mass_density"]
    v1 = (U["momentum_x1"] / rho_g)[_ING]
    v2 = (U["momentum_x2"] / rho_g)[_ING]
    v3 = (U["momentum_x3"] / rho_g)[_ING]
    rho = rho_g[_ING]
    P = C["P"][_ING]
    T11, T22, T33 = C["Tx1x1"][_ING], C["Tx2x2"][_ING], C["Tx3x3"][_ING]
    T12, T13, T23 = C["Tx1x2"][_ING], C["Tx1x3"][_ING], C["Tx2x3"][_ING]
    T21, T31, T32 = C["Tx2x1"][_ING], C["Tx3x1"][_ING], C["Tx3x2"][_ING]
    conv = {
        "mass_density": 0.0,
        "momentum_x1": (P * d1h / h1 + rho * v2 * v2 * d1h2 / h1 + rho * v3 * v3 * d1h3 / h1
                        - rho * v1 * v2 * d2h1 / h2 - rho * v1 * v3 * d3h1 / h3),
        "momentum_x2": (P * d2h / h2 + rho * v1 * v1 * d2h1 / h2 + rho * v3 * v3 * d2h3 / h2
                        - rho * v2 * v1 * d1h2 / h1 - rho * v2 * v3 * d3h2 / h3),
        "momentum_x3": (P * d3h / h3 + rho * v1 * v1 * d3h1 / h3 + rho * v2 * v2 * d3h2 / h3
                        - rho * v3 * v1 * d1h3 / h1 - rho * v3 * v2 * d2h3 / h2),
        "energy_density": 0.0}
    visc = {
        "mass_density": 0.0,
        "momentum_x1": T12 * d2h1 / h2 + T13 * d3h1 / h3 - T22 * d1h2 / h1 - T33 * d1h3 / h1,
        "momentum_x2": T21 * d1h2 / h1 + T23 * d3h2 / h3 - T11 * d2h1 / h2 - T33 * d2h3 / h2,
        "momentum_x3": T31 * d1h3 / h1 + T32 * d2h3 / h2 - T11 * d3h1 / h3 - T22 * d3h2 / h3,
        "energy_density": 0.0}
    return {k: conv[k] + visc[k] for k in conv}


def compute_physical_source(U, C, params):
    """navier_stokes.py:457-509 -- gravity: rho g and rho (g . v)."""
    g1, g2, g3 = params["gravitational_acceleration"][:3]
    rho_g = U["mass_density"]
    v1 = (U["momentum_x1"] / rho_g)[_ING]
    v2 = (U["momentum_x2"] / rho_g)[_ING]
    v3 = (U["momentum_x3"] / rho_g)[_ING]
    rho = rho_g[_ING]
    return {"mass_density": 0.0, "momentum_x1": rho * g1, "momentum_x2": rho * g2,
            "momentum_x3": rho * g3, "energy_density": rho * (g1 * v1 + g2 * v2 + g3 * v3)}


def compute_dt(U, params, grid, cfl):
    """navier_stokes.py:512-585 -- CFL step from interior cells: max over cells
    of max_d(max|lambda_d|/L_d) plus 2 max(mu/rho) sum_d 1/L_d^2."""
    rho = U["mass_density"]
    L = grid["cell"]["geometry"]["length"]
    mu = params["dynamic_viscosity"]
    v = {"x1": U["momentum_x1"] / rho, "x2": U["momentum_x2"] / rho,
         "x3": U["momentum_x3"] / rho}
    C = compute_closure(U, params, grid)
    kappa = compute_characteristic_velocity(U, C, params)
    lam = {d: {w: v[d][_ING] + kappa[w][_ING] for w in kappa} for d in v}
    amax = {d: np.maximum.reduce([np.abs(x) for x in lam[d].values()]) for d in lam}
    nu = np.max(mu / rho[_ING])
    rate_c = np.maximum.reduce([amax["x1"] / L["x1"], amax["x2"] / L["x2"],
                                amax["x3"] / L["x3"]])
    rate_d = 2 * nu * (1 / L["x1"] ** 2 + 1 / L["x2"] ** 2 + 1 / L["x3"] ** 2)
    return cfl / (np.max(rate_c) + np.max(rate_d))


_FLO = [(slice(1, None), slice(None), slice(None)),
        (slice(None), slice(1, None), slice(None)),
        (slice(None), slice(None), slice(1, None))]
_FHI = [(slice(None, -1), slice(None), slice(None)),
        (slice(None), slice(None, -1), slice(None)),
        (slice(None), slice(None), slice(None, -1))]


def divergence(F, S, V, src, keys):
    """navier_stokes.py:693-709 (also heat.py:459-471, maxwell.py:710-727):
    -(((((FS1[+]-FS1[-]) + FS2[+]) - FS2[-]) + FS3[+]) - FS3[-]) / V + src."""
    out = {}
    for k in keys:
        fs = [F[d][k] * S[d] for d in ("x1", "x2", "x3")]
        acc = fs[0][_FLO[0]] - fs[0][_FHI[0]] + fs[1][_FLO[1]] - fs[1][_FHI[1]] \
            + fs[2][_FLO[2]] - fs[2][_FHI[2]]
        out[k] = -acc / V + src[k]
    return out


def compute_rhs(U, grid, params, space_schemes):
    """navier_stokes.py:588-711 -- the finite-volume right-hand side."""
    V = grid["cell"]["geometry"]["volume"]
    S = grid["cell"]["geometry"]["surface"]
    L = grid["cell"]["geometry"]["length"]
    n = grid["cell"]["normal"]
    limiter = space_schemes["slope_limiter"]
    riemann = space_schemes["riemann_solver"]
    diffusive = space_schemes["diffusive_solver"]
    recon = space_schemes["reconstructor"]
    dirs = ("x1", "x2", "x3")

    C = compute_closure(U, params, grid)
    U_L, U_R = recon(U, L, limiter)
    C_L, C_R = recon(C, L, limiter)
    F = {}
    for d in dirs:
        fcl = compute_convective_flux(U_L[d], C_L[d], params, n[d])
        fcr = compute_convective_flux(U_R[d], C_R[d], params, n[d])
        fdl = compute_diffusive_flux(U_L[d], C_L[d], n[d])
        fdr = compute_diffusive_flux(U_R[d], C_R[d], n[d])
        kl = compute_characteristic_velocity(U_L[d], C_L[d], params)
        kr = compute_characteristic_velocity(U_R[d], C_R[d], params)
        vnl = compute_normal_velocity(U_L[d], n[d])
        vnr = compute_normal_velocity(U_R[d], n[d])
        laml = {w: vnl + kl[w] for w in kl}
        lamr = {w: vnr + kr[w] for w in kr}
        fc = riemann(fcl, fcr, U_L[d], U_R[d], laml, lamr)
        fd = diffusive(fdl, fdr)
        F[d] = {k: fc[k] + fd[k] for k in fc}
    sg = compute_geometric_source(U, C, params, grid)
    sp = compute_physical_source(U, C, params)
    src = {k: sg[k] + sp[k] for k in sg}
    return divergence(F, S, V, src, VARS)
