"""Device diagnostics (mfvm_diagnostics) vs the oracle.

Bar: min, max, Linf, the pressure minimum and the violation count are equal
to the oracle's (same float64 values, same operand order); totals, volumes
and L1 are float64 sums in another order, held to 1e-13 of the sum of
absolute terms.  At the benchmark's full size (256^3) the checks are
size-independent: totals against torch's float64 sums, and conservation of
mass, momentum and energy over steps of a periodic flow.
"""

import math

import numpy as np
import pytest
import torch

from oracle import diagnostics as od
from oracle import grid as og

pytestmark = pytest.mark.gpu

NS = ("mass_density", "momentum_x1", "momentum_x2", "momentum_x3", "energy_density")
MX = ("B_x1", "B_x2", "B_x3", "D_x1", "D_x2", "D_x3")


def _cfg(n, lo, hi):
    return {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1],
            "nx2": n[1], "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}


def _state(names, n, dtype, seed):
    from mfvm.state import State
    rng = np.random.default_rng(seed)
    U = State(names, {"nx1": n[0], "nx2": n[1], "nx3": n[2]}, dtype=dtype)
    # ghosts hold junk far outside the interior range: they must not count
    junk = rng.uniform(-1e6, 1e6, tuple(U.buf.shape))
    U.buf.copy_(torch.from_numpy(junk))
    for name in names:
        lo, hi = (0.5, 2.0) if name == "mass_density" else ((4.0, 6.0) if name == "energy_density"
                                                            else (-1.0, 1.0))
        U[name][1:-1, 1:-1, 1:-1] = torch.from_numpy(rng.uniform(lo, hi, n)).to(U.device)
    return U


def _host(U, value=None):
    buf = (U.buf if value is None else value).double().cpu().numpy()
    return {name: buf[v].transpose(2, 1, 0) for v, name in enumerate(U.names)}


def _check(got, want, Uh, names, V):
    for name in names:
        scale = math.fsum((np.abs(Uh[name][1:-1, 1:-1, 1:-1]) * V).ravel())
        assert abs(got["total"][name] - want["total"][name]) <= 1e-13 * scale + 1e-300, name
        assert got["min"][name] == want["min"][name], name
        assert got["max"][name] == want["max"][name], name
        if want["l1"] is not None:
            assert got["linf"][name] == want["linf"][name], name
            assert abs(got["l1"][name] - want["l1"][name]) <= 1e-13 * max(want["linf"][name], 1e-300), name
    assert got["volume"] == pytest.approx(want["volume"], rel=1e-13)
    assert got["pressure_min"] == want["pressure_min"]
    assert got["nonpositive_cells"] == want["nonpositive_cells"]


CASES = [
    ("cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (24, 20, 12), NS, torch.float64, True, True),
    ("cartesian", (0.0, -0.3, 0.1), (0.7, 1.1, 0.6), (33, 7, 5), NS, torch.float64, False, True),
    ("cylindrical", (0.1, 0.0, -1.0), (2.0, 2 * np.pi, 1.0), (16, 12, 8), NS, torch.float64, True, False),
    ("spherical", (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi), (16, 12, 10), MX, torch.float64, True,
     False),
    ("cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (64, 48, 1), ("T",), torch.float32, True, False),
    ("spherical", (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi), (20, 9, 7), MX[:3], torch.float32,
     False, False),
]


@pytest.mark.parametrize("coord,lo,hi,n,names,dtype,with_ref,with_p", CASES,
                         ids=[f"{c[0]}-{c[3][0]}x{c[3][1]}x{c[3][2]}-{len(c[4])}f-{str(c[5])[6:]}" for c in CASES])
def test_diagnostics_match_oracle(coord, lo, hi, n, names, dtype, with_ref, with_p):
    from mfvm.diagnostics import diagnostics
    from mfvm.grid.geometry import build_grid
    print(f"case {coord} {n} {len(names)} fields {dtype}", flush=True)
    cfg = _cfg(n, lo, hi)
    grid = build_grid(coord, cfg)
    U = _state(names, n, dtype, 11)
    ref = None
    if with_ref:
        ref = _state(names, n, torch.float64, 11)
        ref.buf.add_(1e-3 * torch.rand_like(ref.buf))
    gamma = 1.4 if with_p else None
    got = diagnostics(U, grid, ref=ref, gamma=gamma)
    g = og.build_grid(coord, cfg)
    Uh = _host(U)
    want = od.diagnostics(Uh, names, g, ref=None if ref is None else _host(ref), gamma=gamma)
    _check(got, want, Uh, names, g["cell"]["geometry"]["volume"])
    again = diagnostics(U, grid, ref=ref, gamma=gamma)
    assert again == got  # fixed reduction order


def test_compensated_state_uses_hi_plus_lo():
    from mfvm.diagnostics import diagnostics
    from mfvm.grid.geometry import build_grid
    from mfvm.physics import navier_stokes as ns
    n = (40, 24, 16)
    cfg = _cfg(n, (0, 0, 0), (1, 1, 1))
    grid = build_grid("cartesian", cfg)
    U64 = _state(NS, n, torch.float64, 12)
    U64.buf.mul_(1.0 + 1e-9 * torch.rand_like(U64.buf))   # values with more than 24 bits
    Uc = ns.to_compensated(U64)
    got = diagnostics(Uc, grid, ref=U64, gamma=1.4, compensated=True)
    val = ns.compensated_value(Uc)
    Uh = _host(Uc, val)
    g = og.build_grid("cartesian", cfg)
    want = od.diagnostics(Uh, NS, g, ref=_host(U64), gamma=1.4)
    _check(got, want, Uh, NS, g["cell"]["geometry"]["volume"])
    # hi + lo carries the float64 interior values to ~48 bits (|u| <= 6 here)
    assert max(got["linf"].values()) <= 1e-13
    plain = diagnostics(Uc, grid, ref=U64)     # hi alone: float32 rounding
    assert max(plain["linf"].values()) > 1e-9


def test_nan_and_violations_in_the_interior_only():
    from mfvm.diagnostics import diagnostics
    from mfvm.grid.geometry import build_grid
    n = (16, 8, 4)
    cfg = _cfg(n, (0, 0, 0), (1, 1, 1))
    grid = build_grid("cartesian", cfg)
    U = _state(NS, n, torch.float64, 13)
    U["energy_density"][2, 2, 2] = 0.0
    U["mass_density"][3, 1, 1] = -1.0
    U["momentum_x1"][1, 3, 2] = float("nan")
    U["mass_density"][0, 2, 2] = float("nan")    # ghost
    got = diagnostics(U, grid, gamma=1.4)
    assert got["nonpositive_cells"] == 3
    assert math.isnan(got["pressure_min"]) and math.isnan(got["min"]["momentum_x1"])
    assert math.isnan(got["total"]["momentum_x1"])
    assert not math.isnan(got["total"]["mass_density"])
    want = od.diagnostics(_host(U), NS, og.build_grid("cartesian", cfg), gamma=1.4)
    assert want["nonpositive_cells"] == 3 and got["min"]["mass_density"] == want["min"]["mass_density"]


def test_argument_errors():
    from mfvm.diagnostics import diagnostics
    from mfvm.grid.geometry import build_grid
    n = (8, 8, 8)
    grid = build_grid("cartesian", _cfg(n, (0, 0, 0), (1, 1, 1)))
    U = _state(MX, n, torch.float64, 14)
    with pytest.raises(ValueError):
        diagnostics(U, grid, gamma=1.4)           # pressure needs the NS fields
    with pytest.raises(ValueError):
        diagnostics(U, grid, compensated=True)    # float64 state
    with pytest.raises(ValueError):
        diagnostics(U, grid, ref=_state(MX, n, torch.float32, 14))
    # the no-read-back form checks what the kernel cannot (plain pointers)
    from mfvm.diagnostics import raw
    with pytest.raises(ValueError):
        raw(U, grid, ref=_state(MX, n, torch.float32, 14))     # half the bytes of float64
    V = _state(MX, n, torch.float32, 15)
    with pytest.raises(ValueError):
        raw(V, grid, lo=torch.zeros(3, device=V.device))        # too short
    with pytest.raises(ValueError):
        raw(U, grid, lo=torch.zeros_like(V.buf))                # a float64 state has no lo
    out = raw(V, grid, lo=torch.zeros_like(V.buf))
    assert out.shape == (5 * 6 + 3,) and out.dtype == torch.float64


def test_rt256_totals_and_conservation_at_full_size():
    """256^3 (the benchmark grid): totals equal torch's float64 sums to 1e-12,
    extrema equal torch's.  Over 5 SSPRK3 steps of a fully periodic inviscid
    flow without gravity and with piecewise-constant reconstruction, the
    mass, momentum and energy totals are conserved to round-off (FP64) and
    to 1e-9 of their scale (FP32-compensated).  (With MUSCL, or with
    viscosity, the reference scheme itself is not conservative across a
    periodic seam: ghost cells are unsloped and closure ghosts are zero.)"""
    import bench
    from mfvm import _lib
    from mfvm.diagnostics import diagnostics
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    n = 256
    cfg = bench.rt_config(n)
    grid = build_grid("cartesian", cfg)
    U = ns.allocate_u(cfg)
    for k, f in zip(ns.VARS, bench.rt_fields(n)):
        U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(f).cuda()
    bcs = {k: {s: bc.periodic for s in bc.SIDES} for k in ns.VARS}
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 0.0,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    schemes = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.piecewise_constant,
               "slope_limiter": slope_limiters.van_leer, "diffusive_solver": diffusive_solvers.centered}
    print("rt256 diagnostics: totals vs torch", flush=True)
    d0 = diagnostics(U, grid, gamma=1.4)
    inner = U.buf[:, 1:-1, 1:-1, 1:-1]
    vol = 1.0 / n ** 3
    sums = (inner * vol).sum(dim=(1, 2, 3)).cpu().tolist()
    abss = (inner.abs() * vol).sum(dim=(1, 2, 3)).cpu().tolist()
    mins = inner.amin(dim=(1, 2, 3)).cpu().tolist()
    maxs = inner.amax(dim=(1, 2, 3)).cpu().tolist()
    for q, name in enumerate(ns.VARS):
        assert abs(d0["total"][name] - sums[q]) <= 1e-12 * abss[q] + 1e-300, name
        assert d0["min"][name] == mins[q] and d0["max"][name] == maxs[q], name
    assert d0["volume"] == 1.0 and d0["nonpositive_cells"] == 0 and d0["pressure_min"] > 0
    Uc = ns.to_compensated(U)
    print("rt256 diagnostics: 5 periodic FP64 steps", flush=True)
    ns.advance(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 5)
    d1 = diagnostics(U, grid, gamma=1.4)
    # the scale of a field's round-off: its largest magnitude before or after
    # the steps; for a momentum component, the largest over the three
    # components (x1/x2 start at exactly 0 and stay near it, while their
    # fluxes round at the scale of the whole momentum vector)
    scale = {name: max(abs(d[k][name]) for d in (d0, d1) for k in ("min", "max")) for name in ns.VARS}
    mom = max(scale[name] for name in ns.VARS[1:4])
    scale.update({name: mom for name in ns.VARS[1:4]})
    for name in ns.VARS:
        drift = abs(d1["total"][name] - d0["total"][name])
        print(f"fp64 {name}: drift {drift:.3e} of scale {scale[name]:.3e}", flush=True)
        assert drift <= 1e-12 * scale[name], name
    assert d1["nonpositive_cells"] == 0
    print("rt256 diagnostics: 5 periodic FP32-compensated steps", flush=True)
    ns.advance_compensated(Uc, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 5)
    d2 = diagnostics(Uc, grid, ref=U, gamma=1.4, compensated=True)
    for name in ns.VARS:
        drift = abs(d2["total"][name] - d0["total"][name])
        print(f"comp {name}: drift {drift:.3e}, L1 vs fp64 {d2['l1'][name]:.3e}", flush=True)
        assert drift <= 1e-9 * scale[name], name
        assert d2["l1"][name] <= 1e-6, name
    assert d2["nonpositive_cells"] == 0
