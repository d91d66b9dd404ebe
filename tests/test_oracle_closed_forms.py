"""Known answers and closed forms that pin the oracle independently of the
golden vectors (each check would fail for a dropped term, a sign slip or a
swapped operand)."""

import numpy as np
import pytest

from oracle import grid as og
from oracle import heat as oh
from oracle import maxwell as om
from oracle import navier_stokes as ons
from oracle import numerics as on

A = np.array


# --- limiters: hand-computed values --------------------------------------------------------

def test_minmod_known_answers():
    b = A([1.0, 1.0, -2.0, 0.0, 3.0, -1.0])
    f = A([2.0, -1.0, -1.0, 5.0, 3.0, -4.0])
    assert np.array_equal(on.minmod(b, f), A([1.0, 0.0, -1.0, 0.0, 3.0, -1.0]))


def test_mc_known_answers_including_one_third_split():
    b = A([1.0, 1.0, 0.0, -1.0, 2.0])
    f = A([3.0, -0.5, 7.0, -3.0, 2.0])
    out = on.monotonized_central(b, f)
    # (1,3): min(2, 2, 6) = 2; (1,-0.5): signs (+,+,-) -> 1/3 * min(0.25, 2, 1)
    assert out[0] == 2.0
    assert out[1] == (1.0 / 3.0) * 0.25
    assert out[2] == 0.0 and out[3] == -2.0 and out[4] == 2.0


def test_van_leer_known_answers():
    b = A([1.0, -1.0, 1.0, 0.0, 2.0])
    f = A([3.0, -3.0, -1.0, 0.0, 2.0])
    assert np.array_equal(on.van_leer(b, f), A([1.5, -1.5, 0.0, 0.0, 2.0]))


def test_limiters_propagate_nan():
    for lim in (on.minmod, on.monotonized_central, on.van_leer):
        assert np.isnan(lim(A([np.nan]), A([1.0]))[0])


# --- Riemann solvers --------------------------------------------------------------------------

def _waves(vn, cs):
    return {"c": vn + 0.0, "m": vn - cs, "p": vn + cs}


def test_hll_supersonic_branches_pick_upwind_flux():
    FL, FR = {"q": A([1.0, 1.0])}, {"q": A([5.0, 5.0])}
    UL, UR = {"q": A([0.0, 0.0])}, {"q": A([1.0, 1.0])}
    lam_l = _waves(A([3.0, -3.0]), A([1.0, 1.0]))
    lam_r = _waves(A([3.0, -3.0]), A([1.0, 1.0]))
    out = on.hll(FL, FR, UL, UR, lam_l, lam_r)["q"]
    assert out[0] == 1.0 and out[1] == 5.0


def test_hll_subsonic_closed_form():
    # sL = -1, sR = 2: F = (2*FL + FR - 2*(UR-UL)) / 3
    FL, FR = {"q": A([1.0])}, {"q": A([4.0])}
    UL, UR = {"q": A([0.5])}, {"q": A([2.0])}
    lam = _waves(A([0.5]), A([1.5]))
    out = on.hll(FL, FR, UL, UR, lam, lam)["q"][0]
    assert out == pytest.approx((2.0 * 1.0 + 4.0 - 2.0 * 1.5) / 3.0, rel=1e-15)


def test_rusanov_closed_form():
    FL, FR = {"q": A([1.0])}, {"q": A([3.0])}
    UL, UR = {"q": A([1.0])}, {"q": A([2.0])}
    lam_l = _waves(A([-0.5]), A([1.0]))   # max |.| = 1.5
    lam_r = _waves(A([0.25]), A([1.0]))   # max |.| = 1.25
    out = on.rusanov(FL, FR, UL, UR, lam_l, lam_r)["q"][0]
    assert out == 2.0 - 1.5 * 1.0 / 2.0


def test_centered_is_mean():
    assert on.centered({"a": A([1.0])}, {"a": A([4.0])})["a"][0] == 2.5


# --- integrators: amplification factors with a mock linear RHS -------------------------------

@pytest.mark.parametrize("name,factor", [
    ("euler_explicit", lambda z: 1 + z),
    ("rk3", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 12),
    ("rk3_ssp", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6),
    ("rk4", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24),
])
def test_integrator_amplification(name, factor):
    lam, dt = -0.7, 0.3
    U = {"q": np.ones((3, 3, 3))}

    def rhs(U, grid, params, schemes):
        return {"q": lam * U["q"][1:-1, 1:-1, 1:-1]}

    getattr(on, name)(rhs, U, dt, None, None, None, None, lambda U, b: None)
    assert U["q"][1, 1, 1] == pytest.approx(factor(lam * dt), rel=1e-14)
    assert U["q"][0, 0, 0] == 1.0  # ghosts untouched by the update


# --- grid closed forms ------------------------------------------------------------------------

def _gc(n, lo, hi):
    return {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1],
            "nx2": n[1], "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}


def test_grid_volumes_sum_to_domain():
    g = og.build_grid("cylindrical", _gc((6, 5, 4), (0.5, 0.0, 0.0), (2.0, np.pi, 3.0)))
    assert g["cell"]["geometry"]["volume"].sum() == pytest.approx((4.0 - 0.25) / 2 * np.pi * 3.0)
    g = og.build_grid("spherical", _gc((6, 5, 4), (0.5, 0.1, 0.0), (2.0, 3.0, 2 * np.pi)))
    exact = (8.0 - 0.125) / 3 * (np.cos(0.1) - np.cos(3.0)) * 2 * np.pi
    assert g["cell"]["geometry"]["volume"].sum() == pytest.approx(exact)
    assert np.allclose(g["lame"]["coefficients"]["x3"],
                       g["cell"]["coordinates"]["x1"] * np.sin(g["cell"]["coordinates"]["x2"]))


def test_unknown_coordinate_system_raises():
    with pytest.raises(ValueError, match="Unknown coordinate system"):
        og.build_grid("toroidal", _gc((2, 2, 2), (0, 0, 0), (1, 1, 1)))


def test_unknown_side_raises():
    with pytest.raises(ValueError, match="Unknown boundary side"):
        og.neumann(np.zeros((3, 3, 3)), "up")


def test_bc_brute_force_small_array():
    rng = np.random.default_rng(5)
    f = rng.standard_normal((5, 4, 3))
    g = f.copy()
    og.apply_bcs({"q": g}, {"q": {s: og.periodic for s in og.SIDES}})
    assert np.array_equal(g[0, 1:-1, 1:-1], f[-2, 1:-1, 1:-1])
    assert np.array_equal(g[1:-1, -1, 1:-1], f[1:-1, 1, 1:-1])
    assert g[0, 0, 1] == 0.5 * (g[1, 0, 1] + g[0, 1, 1])
    assert g[-1, -1, -1] == (g[-2, -1, -1] + g[-1, -2, -1] + g[-1, -1, -2]) / 3.0
    h = f.copy()
    og.apply_bcs({"q": h}, {"q": {s: og.reflecting for s in og.SIDES}})
    assert np.array_equal(h[1:-1, 1:-1, 0], -f[1:-1, 1:-1, 1])


# --- physics invariants -------------------------------------------------------------------------

def _uniform_ns(n, rho, v, p, gamma=1.4):
    U = ons.allocate_u({"nx1": n[0], "nx2": n[1], "nx3": n[2]})
    U["mass_density"][:] = rho
    for a in range(3):
        U[f"momentum_x{a + 1}"][:] = rho * v[a]
    U["energy_density"][:] = p / (gamma - 1) + 0.5 * rho * sum(x * x for x in v)
    return U


def test_ns_uniform_state_has_zero_rhs_except_gravity():
    n = (6, 5, 4)
    g = og.build_grid("cartesian", _gc(n, (0, 0, 0), (1, 1, 1)))
    U = _uniform_ns(n, 1.3, (0.2, -0.1, 0.05), 2.0)
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-3,
              "gravitational_acceleration": [0.0, 0.0, -0.5]}
    for rs in (on.hll, on.rusanov):
        schemes = {"riemann_solver": rs, "reconstructor": on.muscl, "slope_limiter": on.van_leer,
                   "diffusive_solver": on.centered}
        r = ons.compute_rhs(U, g, params, schemes)
        assert np.allclose(r["mass_density"], 0.0, atol=1e-12)
        assert np.allclose(r["momentum_x3"], 1.3 * -0.5, atol=1e-12)
        assert np.allclose(r["energy_density"], 1.3 * (-0.5 * 0.05), atol=1e-12)


def test_ns_dt_uniform_closed_form():
    n = (4, 4, 4)
    g = og.build_grid("cartesian", _gc(n, (0, 0, 0), (1, 2, 4)))
    rho, v, p, mu = 1.5, (0.3, -0.2, 0.1), 2.0, 1e-2
    U = _uniform_ns(n, rho, v, p)
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": mu,
              "gravitational_acceleration": [0, 0, 0]}
    c = np.sqrt(1.4 * p / rho)
    L = (0.25, 0.5, 1.0)
    rate = max((abs(v[d]) + c) / L[d] for d in range(3)) + 2 * (mu / rho) * sum(1 / x ** 2 for x in L)
    assert ons.compute_dt(U, params, g, 0.4) == pytest.approx(0.4 / rate, rel=1e-14)


def test_ns_density_advection_is_upwinded():
    # Pure x1 advection of a density ramp at constant p and v: the first-order
    # (piecewise-constant) HLL update equals the closed-form HLL mass flux.
    n = (8, 1, 1)
    g = og.build_grid("cartesian", _gc(n, (0, 0, 0), (1, 1, 1)))
    U = _uniform_ns(n, 1.0, (0.5, 0.0, 0.0), 1.0)
    rho = 1.0 + 0.1 * np.arange(10.0)
    U["mass_density"][:] = rho[:, None, None]
    U["momentum_x1"][:] = (rho * 0.5)[:, None, None]
    U["energy_density"][:] = (1.0 / 0.4 + 0.5 * rho * 0.25)[:, None, None]
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 0.0,
              "gravitational_acceleration": [0, 0, 0]}
    schemes = {"riemann_solver": on.hll, "reconstructor": on.piecewise_constant,
               "slope_limiter": on.minmod, "diffusive_solver": on.centered}
    r = ons.compute_rhs(U, g, params, schemes)["mass_density"][:, 0, 0]
    c = np.sqrt(1.4 * 1.0 / rho)
    sl = np.minimum(0.5 - c[:-1], 0.5 - c[1:])
    sr = np.maximum(0.5 + c[:-1], 0.5 + c[1:])
    F = (sr * rho[:-1] * 0.5 - sl * rho[1:] * 0.5 + sl * sr * (rho[1:] - rho[:-1])) / (sr - sl)
    assert np.allclose(r, -(F[1:] - F[:-1]) / 0.125, rtol=1e-12)


def test_heat_uniform_temperature_gives_source_only():
    n = (5, 4, 3)
    g = og.build_grid("cylindrical", _gc(n, (0.5, 0.0, 0.0), (1.5, 2 * np.pi, 1.0)))
    U = oh.allocate_u({"nx1": 5, "nx2": 4, "nx3": 3})
    U["internal_energy_density"][:] = 3.0
    params = {"mass_density": 1.0, "specific_heat_capacity": 2.0, "thermal_conductivity": 0.1,
              "volumetric_heat_source": 0.25}
    schemes = {"riemann_solver": None, "slope_limiter": on.minmod,
               "reconstructor": on.piecewise_constant, "diffusive_solver": on.centered}
    # Interior cells away from the zero-padded closure ghosts see zero flux.
    r = oh.compute_rhs(U, g, params, schemes)["internal_energy_density"]
    assert np.allclose(r[1:-1, 1:-1, 1:-1], 0.25, atol=1e-14)


def test_maxwell_dt_closed_form():
    n = (4, 4, 4)
    g = og.build_grid("cartesian", _gc(n, (0, 0, 0), (1, 1, 2)))
    U = om.allocate_u({"nx1": 4, "nx2": 4, "nx3": 4})
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 4.0,
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 3e-3}
    c = 1 / np.sqrt(4.0)
    rate = c / 0.25 + 2 * 3e-3 * (16 + 16 + 4)
    assert om.compute_dt(U, params, g, 0.5) == pytest.approx(0.5 / rate, rel=1e-14)
