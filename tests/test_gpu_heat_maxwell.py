"""Heat and Maxwell HIP paths vs the reference golden vectors (bit-exact)."""

import os

import numpy as np
import pytest
import torch

from oracle_setup import GOLDEN, group, grid_config, load, same

pytestmark = pytest.mark.gpu


def _schemes(rec, lim, rs=None):
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    return {"reconstructor": getattr(reconstructors, rec), "slope_limiter": getattr(slope_limiters, lim),
            "riemann_solver": getattr(riemann_solvers, rs) if rs else None,
            "diffusive_solver": diffusive_solvers.centered}


def _bcs(names, codes):
    from mfvm.grid import boundary_conditions as bc
    m = {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}
    return {v: {s: m[int(codes[i, j])] for j, s in enumerate(bc.SIDES)} for i, v in enumerate(names)}


def test_heat_cylindrical_matches_reference():
    from mfvm import solver
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import heat
    z = load(os.path.join(GOLDEN, "heat_cyl.npz"))
    gc = grid_config((8, 6, 4), (0.1, 0.0, 0.0), (1.0, 2 * np.pi, 1.0))
    grid = build_grid("cylindrical", gc)
    params = {"mass_density": 1.0, "specific_heat_capacity": 1.0, "thermal_conductivity": 1e-2,
              "volumetric_heat_source": 0.05}
    U = heat.allocate_u(gc)
    for k, v in group(z, "U_in/").items():
        U[k] = torch.from_numpy(v)
    for tag, rec in (("pc", "piecewise_constant"), ("vanleer", "muscl")):
        r = heat.compute_rhs(U, grid, params, _schemes(rec, "van_leer"))
        assert same(r["internal_energy_density"].cpu().numpy(),
                    z[f"rhs_{tag}/internal_energy_density"]), tag
    assert heat.compute_dt(U, params, grid, 0.4) == float(z["dt"])
    bcs = _bcs(heat.VARS, z["bcs"])
    schemes = _schemes("piecewise_constant", "van_leer")
    for _ in range(3):
        dt = heat.compute_dt(U, params, grid, 0.4)
        ti.euler_explicit(heat.compute_rhs, U, dt, params, grid, schemes, bcs, solver.apply_bcs)
    assert same(U["internal_energy_density"].cpu().numpy(), z["U_out/internal_energy_density"])
    # the same three steps through the device-dt driver
    U2 = heat.allocate_u(gc)
    for k, v in group(z, "U_in/").items():
        U2[k] = torch.from_numpy(v)
    heat.advance(U2, grid, params, schemes, bcs, ti.euler_explicit._mfvm_code, 0.4, 3)
    assert same(U2["internal_energy_density"].cpu().numpy(), z["U_out/internal_energy_density"])


def test_maxwell_spherical_matches_reference():
    from mfvm import solver
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import maxwell
    z = load(os.path.join(GOLDEN, "maxwell_sph.npz"))
    gc = grid_config((8, 8, 8), (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi))
    grid = build_grid("spherical", gc)
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
              "electrical_conductivity": 0.1, "charge_density": 0.0,
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 2e-3}
    U = maxwell.allocate_u(gc)
    for k, v in group(z, "U_in/").items():
        U[k] = torch.from_numpy(v)
    for rs in ("rusanov", "hll"):
        r = maxwell.compute_rhs(U, grid, params, _schemes("muscl", "van_leer", rs))
        for k in maxwell.VARS:
            assert same(r[k].cpu().numpy(), z[f"rhs_{rs}/{k}"]), (rs, k)
    assert maxwell.compute_dt(U, params, grid, 0.4) == float(z["dt"])
    bcs = _bcs(maxwell.VARS, z["bcs"])
    schemes = _schemes("muscl", "van_leer", "rusanov")
    dt = maxwell.compute_dt(U, params, grid, 0.4)
    ti.rk4(maxwell.compute_rhs, U, dt, params, grid, schemes, bcs, solver.apply_bcs)
    for k in maxwell.VARS:
        assert same(U[k].cpu().numpy(), z[f"U_out/{k}"]), k


@pytest.mark.parametrize("lim,rs", [("piecewise_constant", "rusanov"), ("minmod", "hll"),
                                    ("monotonized_central", "rusanov"), ("van_leer", "hll")])
def test_maxwell_tiles_and_chunks_match_oracle(lim, rs):
    """A spherical grid of several x1/x2 tiles and x3 chunks, partial at every
    edge (70 x 9 x 37): the right-hand side equals the oracle's bit for bit,
    and so do two RK4 steps with device dt."""
    from oracle import grid as og
    from oracle import maxwell as omx
    from oracle import numerics as on
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import maxwell
    n = (70, 9, 37)
    gc = grid_config(n, (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi))
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
              "electrical_conductivity": 0.1, "charge_density": 0.0,
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 2e-3}
    rng = np.random.default_rng(70937)
    Uo = omx.allocate_u(gc)
    for k in omx.VARS:
        Uo[k][...] = rng.standard_normal(Uo[k].shape)
    go = og.build_grid("spherical", gc)
    rec = "piecewise_constant" if lim == "piecewise_constant" else "muscl"
    so = {"reconstructor": getattr(on, rec), "slope_limiter": getattr(on, "minmod" if rec != "muscl" else lim),
          "riemann_solver": getattr(on, rs), "diffusive_solver": on.centered}
    ro = omx.compute_rhs(Uo, go, params, so)

    grid = build_grid("spherical", gc)
    U = maxwell.allocate_u(gc)
    for k in maxwell.VARS:
        U[k] = torch.from_numpy(Uo[k])
    sg = _schemes(rec, "minmod" if rec != "muscl" else lim, rs)
    print(f"maxwell {n} {lim} {rs}: rhs", flush=True)
    r = maxwell.compute_rhs(U, grid, params, sg)
    for k in maxwell.VARS:
        assert same(r[k].cpu().numpy(), ro[k]), k

    # two RK4 steps: phi periodic, r and theta Neumann
    sides = {"left": og.neumann, "right": og.neumann, "bottom": og.neumann, "top": og.neumann,
             "front": og.periodic, "back": og.periodic}
    bo = {k: dict(sides) for k in omx.VARS}
    from oracle import solver as osv
    og.apply_bcs(Uo, bo)
    dts_o = osv.advance(omx, Uo, go, params, so, on.rk4, bo, 0.4, 2)
    gsides = {"left": bc.neumann, "right": bc.neumann, "bottom": bc.neumann, "top": bc.neumann,
              "front": bc.periodic, "back": bc.periodic}
    bg = {k: dict(gsides) for k in maxwell.VARS}
    U2 = maxwell.allocate_u(gc)
    rng = np.random.default_rng(70937)
    Ur = omx.allocate_u(gc)
    for k in omx.VARS:
        Ur[k][...] = rng.standard_normal(Ur[k].shape)
    og.apply_bcs(Ur, bo)
    for k in maxwell.VARS:
        U2[k] = torch.from_numpy(Ur[k])
    print(f"maxwell {n} {lim} {rs}: rk4 steps", flush=True)
    dts = maxwell.advance(U2, grid, params, sg, bg, ti.rk4._mfvm_code, 0.4, 2)
    assert dts.cpu().numpy()[:2].tolist() == list(dts_o)
    for k in maxwell.VARS:
        assert same(U2[k].cpu().numpy(), Uo[k]), k


def test_maxwell_fp32_from_zero_fields_stays_finite():
    """A pulse in D with B exactly zero: B's slopes pass through the float32
    denormal range in the first steps, where the van Leer harmonic mean must
    not turn into NaN.  The float32 run stays finite and close to float64
    (error bar of the reduced-precision mode: max |diff| <= 1e-4, mean <=
    1e-6, for fields of order 1)."""
    import math
    from mfvm import _lib
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.physics import maxwell as mx
    n = (32, 32, 16)
    gc = grid_config(n, (0.1, 0.05, 0.0), (1.0, math.pi - 0.05, 2 * math.pi))
    grid = build_grid("spherical", gc)
    params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
              "electrical_conductivity": 0.0, "charge_density": 0.0,
              "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 1e-3}
    r = 0.1 + (np.arange(n[0]) + 0.5) * (0.9 / n[0])
    th = 0.05 + (np.arange(n[1]) + 0.5) * ((math.pi - 0.1) / n[1])
    R, TH = np.meshgrid(r, th, indexing="ij")
    ez = np.exp(-((R - 0.5) / 0.1) ** 2)
    sides = {"left": bc.neumann, "right": bc.neumann, "bottom": bc.neumann, "top": bc.neumann,
             "front": bc.periodic, "back": bc.periodic}
    bcs = {k: dict(sides) for k in mx.VARS}
    out = {}
    for dtype in (torch.float64, torch.float32):
        U = mx.allocate_u(gc, dtype=dtype)
        U["electric_d_field_x1"][1:-1, 1:-1, 1:-1] = torch.from_numpy((ez * np.cos(TH))[:, :, None] * np.ones(n)).to(dtype)
        U["electric_d_field_x2"][1:-1, 1:-1, 1:-1] = torch.from_numpy((-ez * np.sin(TH))[:, :, None] * np.ones(n)).to(dtype)
        mx.advance(U, grid, params, _schemes("muscl", "van_leer", "rusanov"), bcs, _lib.INT_RK4, 0.4, 10)
        out[dtype] = U.buf[:, 1:-1, 1:-1, 1:-1].to(torch.float64).cpu()
    assert bool(torch.isfinite(out[torch.float32]).all())
    d = (out[torch.float32] - out[torch.float64]).abs()
    assert float(d.max()) <= 1e-4 and float(d.mean()) <= 1e-6, (float(d.max()), float(d.mean()))


@pytest.mark.parametrize("rec,lim", [("piecewise_constant", "minmod"), ("muscl", "minmod"),
                                     ("muscl", "monotonized_central"), ("muscl", "van_leer")])
def test_heat_tiles_match_oracle(rec, lim):
    """A cylindrical grid of several x1/x2 tiles, partial at every edge (70 x
    9 x 6): the right-hand side equals the oracle's bit for bit, and so do 3
    and 4 Euler steps (the device driver alternates its two buffers, so odd
    and even step counts end differently)."""
    from oracle import grid as og
    from oracle import heat as oht
    from oracle import numerics as on
    from oracle import solver as osv
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import heat
    n = (70, 9, 6)
    gc = grid_config(n, (0.1, 0.0, 0.0), (1.0, 2 * np.pi, 1.0))
    params = {"mass_density": 1.0, "specific_heat_capacity": 1.5, "thermal_conductivity": 1e-2,
              "volumetric_heat_source": 0.05}
    so = {"reconstructor": getattr(on, rec), "slope_limiter": getattr(on, lim),
          "riemann_solver": None, "diffusive_solver": on.centered}
    sg = _schemes(rec, lim)
    go = og.build_grid("cylindrical", gc)
    grid = build_grid("cylindrical", gc)

    def fresh():
        rng = np.random.default_rng(7096)
        Uo = oht.allocate_u(gc)
        for k in oht.VARS:
            Uo[k][...] = 1.0 + rng.random(Uo[k].shape)
        return Uo

    Uo = fresh()
    ro = oht.compute_rhs(Uo, go, params, so)
    U = heat.allocate_u(gc)
    for k in heat.VARS:
        U[k] = torch.from_numpy(Uo[k])
    print(f"heat {n} {rec} {lim}: rhs", flush=True)
    r = heat.compute_rhs(U, grid, params, sg)
    for k in heat.VARS:
        assert same(r[k].cpu().numpy(), ro[k]), k

    sides = {"left": og.neumann, "right": og.neumann, "bottom": og.periodic, "top": og.periodic,
             "front": og.neumann, "back": og.neumann}
    gsides = {"left": bc.neumann, "right": bc.neumann, "bottom": bc.periodic, "top": bc.periodic,
              "front": bc.neumann, "back": bc.neumann}
    bo = {k: dict(sides) for k in oht.VARS}
    bg = {k: dict(gsides) for k in heat.VARS}
    for steps in (3, 4):
        Uo = fresh()
        dts_o = osv.advance(oht, Uo, go, params, so, on.euler_explicit, bo, 0.4, steps)
        Ur = fresh()
        U2 = heat.allocate_u(gc)
        for k in heat.VARS:
            U2[k] = torch.from_numpy(Ur[k])
        print(f"heat {n} {rec} {lim}: {steps} euler steps", flush=True)
        dts = heat.advance(U2, grid, params, sg, bg, ti.euler_explicit._mfvm_code, 0.4, steps)
        assert dts.cpu().numpy()[:steps].tolist() == list(dts_o)
        for k in heat.VARS:
            assert same(U2[k].cpu().numpy(), Uo[k]), (steps, k)


@pytest.mark.parametrize("rec,lim", [("piecewise_constant", "minmod"), ("muscl", "van_leer")])
@pytest.mark.parametrize("rho_cp,z_max", [(1.0, 0.5625), (1.5, 1.0)])
def test_heat_march_chunks_match_oracle(rec, lim, rho_cp, z_max):
    """A cylindrical grid whose workgroups march several x3 planes (256 x 64
    x 72: four planes per workgroup, 18 chunks): the right-hand side and 3
    Euler steps equal the oracle's bit for bit.  rho*cp = 1 with dz = 2^-7
    takes the specialised power-of-two Euler form with a power-of-two x1 and
    x3 length (as the cylindrical heat benchmark does); rho*cp = 1.5 the
    general form."""
    _heat_march_case(rec, lim, rho_cp, z_max, 72)


@pytest.mark.parametrize("rho_cp", [1.0, 1.5])
def test_heat_march_odd_and_partial_chunks_match_oracle(rho_cp):
    """256 x 64 x 50: three planes per workgroup (an odd count, so the march's
    two phi and T buffers end on the other parity) and a last chunk of two:
    the right-hand side and 3 Euler steps equal the oracle's bit for bit."""
    _heat_march_case("piecewise_constant", "minmod", rho_cp, 50 / 128, 50)


def _heat_march_case(rec, lim, rho_cp, z_max, n3):
    from oracle import grid as og
    from oracle import heat as oht
    from oracle import numerics as on
    from oracle import solver as osv
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import time_integrators as ti
    from mfvm.physics import heat
    n = (256, 64, n3)
    gc = grid_config(n, (0.0, 0.0, 0.0), (1.0, 2 * np.pi, z_max))
    params = {"mass_density": 1.0, "specific_heat_capacity": rho_cp, "thermal_conductivity": 1e-2,
              "volumetric_heat_source": 0.0}
    so = {"reconstructor": getattr(on, rec), "slope_limiter": getattr(on, lim),
          "riemann_solver": None, "diffusive_solver": on.centered}
    sg = _schemes(rec, lim)
    go = og.build_grid("cylindrical", gc)
    grid = build_grid("cylindrical", gc)

    def fresh():
        rng = np.random.default_rng(4411)
        Uo = oht.allocate_u(gc)
        for k in oht.VARS:
            Uo[k][...] = 1.0 + rng.random(Uo[k].shape)
        return Uo

    Uo = fresh()
    ro = oht.compute_rhs(Uo, go, params, so)
    U = heat.allocate_u(gc)
    for k in heat.VARS:
        U[k] = torch.from_numpy(Uo[k])
    print(f"heat march {n} {rec} rho_cp={rho_cp}: rhs", flush=True)
    r = heat.compute_rhs(U, grid, params, sg)
    for k in heat.VARS:
        assert same(r[k].cpu().numpy(), ro[k]), k
    sides = {"left": og.neumann, "right": og.neumann, "bottom": og.periodic, "top": og.periodic,
             "front": og.neumann, "back": og.neumann}
    gsides = {"left": bc.neumann, "right": bc.neumann, "bottom": bc.periodic, "top": bc.periodic,
              "front": bc.neumann, "back": bc.neumann}
    bo = {k: dict(sides) for k in oht.VARS}
    bg = {k: dict(gsides) for k in heat.VARS}
    steps = 3
    dts_o = osv.advance(oht, Uo, go, params, so, on.euler_explicit, bo, 0.4, steps)
    Ur = fresh()
    U2 = heat.allocate_u(gc)
    for k in heat.VARS:
        U2[k] = torch.from_numpy(Ur[k])
    print(f"heat march {n} {rec} rho_cp={rho_cp}: {steps} euler steps", flush=True)
    dts = heat.advance(U2, grid, params, sg, bg, ti.euler_explicit._mfvm_code, 0.4, steps)
    assert dts.cpu().numpy()[:steps].tolist() == list(dts_o)
    for k in heat.VARS:
        assert same(U2[k].cpu().numpy(), Uo[k]), k


COORDS = {
    "cartesian": ((0.0, -0.5, 0.25), (1.0, 0.75, 1.5)),
    "cylindrical": ((0.1, 0.0, -1.0), (1.0, 2 * np.pi, 1.0)),
    "spherical": ((0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi)),
}


@pytest.mark.parametrize("physics", ["heat", "maxwell"])
@pytest.mark.parametrize("coord", sorted(COORDS))
def test_rhs_on_every_coordinate_system(physics, coord):
    """Heat and Maxwell right-hand sides on Cartesian, cylindrical and
    spherical grids of several x1/x2 tiles and x3 chunks, partial at every
    edge (70 x 9 x 37), MUSCL + van Leer (HLL for Maxwell): equal to the
    oracle's bit for bit."""
    from oracle import grid as og
    from oracle import heat as oht
    from oracle import maxwell as omx
    from oracle import numerics as on
    from mfvm.grid.geometry import build_grid
    from mfvm.physics import heat, maxwell
    n = (70, 9, 37)
    lo, hi = COORDS[coord]
    gc = grid_config(n, lo, hi)
    if physics == "heat":
        om, gm, rs = oht, heat, None
        params = {"mass_density": 1.0, "specific_heat_capacity": 1.5, "thermal_conductivity": 1e-2,
                  "volumetric_heat_source": 0.05}
    else:
        om, gm, rs = omx, maxwell, "hll"
        params = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0,
                  "electrical_conductivity": 0.1, "charge_density": 0.0,
                  "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 2e-3}
    rng = np.random.default_rng(37970)
    Uo = om.allocate_u(gc)
    for k in om.VARS:
        Uo[k][...] = 1.0 + rng.random(Uo[k].shape)
    so = {"reconstructor": on.muscl, "slope_limiter": on.van_leer,
          "riemann_solver": getattr(on, rs) if rs else None, "diffusive_solver": on.centered}
    ro = om.compute_rhs(Uo, og.build_grid(coord, gc), params, so)
    U = gm.allocate_u(gc)
    for k in gm.VARS:
        U[k] = torch.from_numpy(Uo[k])
    print(f"{physics} {coord} {n}: rhs", flush=True)
    r = gm.compute_rhs(U, build_grid(coord, gc), params, _schemes("muscl", "van_leer", rs))
    for k in gm.VARS:
        assert same(r[k].cpu().numpy(), ro[k]), k


def _heated_disc(n, dtype):
    """The heated-disc cylindrical state (r in [0, 1], T = 2 for r < 0.25)."""
    from mfvm.grid.geometry import build_grid
    from mfvm.physics import heat
    from mfvm.setups import heated_disc_cylindrical as hd
    s = hd.get_setup(n)
    grid = build_grid("cylindrical", s["grid_config"])
    U = heat.allocate_u(s["grid_config"], dtype=dtype)
    r = (np.arange(n[0]) + 0.5) / n[0]
    T0 = np.where(r < 0.25, 2.0, 1.0)[:, None, None] * np.ones(n)
    U["internal_energy_density"][1:-1, 1:-1, 1:-1] = torch.from_numpy(T0).to(dtype).cuda()
    return s, grid, U, T0


def test_heat_fp32_cylindrical_full_size_near_fp64():
    """The FP32 leg of the cylindrical heat configuration (256 x 128 x 128,
    300 Euler steps, heat.py:354-473): against the FP64 run of the same
    steps, L1 <= 1e-7 and L-inf <= 5e-6 for a field of order 1 (FP32 carries
    about 7 digits and every step rounds the update once; measured on MI355X:
    L1 1.4e-8, L-inf 7.1e-7, while the field itself moves by 9e-3)."""
    from mfvm import _lib
    from mfvm.physics import heat
    n = (256, 128, 128)
    out = {}
    for dtype in (torch.float64, torch.float32):
        s, grid, U, T0 = _heated_disc(n, dtype)
        print(f"heat fp32 leg: {dtype} 300 steps", flush=True)
        dts = heat.advance(U, grid, s["physical_parameters"], s["space_schemes"], s["boundary_conditions"],
                           _lib.INT_EULER, 0.4, 300)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(dts).all())
        out[dtype] = U.buf[:, 1:-1, 1:-1, 1:-1].to(torch.float64)
    moved = (out[torch.float64] - out[torch.float64].new_tensor(T0.transpose(2, 1, 0)[None])).abs().max()
    d = (out[torch.float32] - out[torch.float64]).abs()
    l1, linf = float(d.mean()), float(d.max())
    print(f"heat fp32 vs fp64 (300 steps): L1 {l1:.3e} Linf {linf:.3e}; "
          f"largest change of the field {float(moved):.3e}", flush=True)
    assert float(moved) > 1e-3  # the run did diffuse the disc
    assert l1 <= 1e-7 and linf <= 5e-6, (l1, linf)


def test_heat_fp32_reduced_size_near_oracle():
    """FP32 heat against the oracle's FP64 run on a reduced grid (64 x 32 x
    16, 50 Euler steps): L1 <= 2e-7, L-inf <= 5e-6 (measured on MI355X: L1
    4.1e-8, L-inf 6.8e-7); the FP64 run of the same steps equals the oracle
    bitwise."""
    from mfvm import _lib
    from mfvm.physics import heat
    from oracle import heat as oht
    from oracle import numerics as on
    from oracle import solver as osv
    from oracle import grid as og
    n = (64, 32, 16)
    s, _, _, T0 = _heated_disc(n, torch.float64)
    Uo = oht.allocate_u(s["grid_config"])
    Uo["internal_energy_density"][1:-1, 1:-1, 1:-1] = T0
    sides = {"left": og.neumann, "right": og.neumann, "bottom": og.periodic, "top": og.periodic,
             "front": og.neumann, "back": og.neumann}
    bo = {"internal_energy_density": sides}
    go = og.build_grid("cylindrical", s["grid_config"])
    so = {"reconstructor": on.piecewise_constant, "slope_limiter": on.minmod, "riemann_solver": on.hll,
          "diffusive_solver": on.centered}
    dts_o = osv.advance(oht, Uo, go, s["physical_parameters"], so, on.euler_explicit, bo, 0.4, 50)
    ref = Uo["internal_energy_density"][1:-1, 1:-1, 1:-1].transpose(2, 1, 0)
    for dtype in (torch.float64, torch.float32):
        s, grid, U, _ = _heated_disc(n, dtype)
        print(f"heat reduced size: {dtype} 50 steps", flush=True)
        dts = heat.advance(U, grid, s["physical_parameters"], s["space_schemes"], s["boundary_conditions"],
                           _lib.INT_EULER, 0.4, 50)
        torch.cuda.synchronize()
        got = U.buf[0, 1:-1, 1:-1, 1:-1].to(torch.float64).cpu().numpy()
        if dtype == torch.float64:
            assert dts.cpu().numpy()[:50].tolist() == dts_o
            assert same(got, ref)
        else:
            d = np.abs(got - ref)
            print(f"heat fp32 vs oracle (50 steps): L1 {d.mean():.3e} Linf {d.max():.3e}", flush=True)
            assert d.mean() <= 2e-7 and d.max() <= 5e-6, (d.mean(), d.max())
