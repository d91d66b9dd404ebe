"""Callers on the GPU: a reference-format setup runs through mfvm.solver.run
(initializer on the host mirror, device stepping) and ends bit for bit where
the oracle's loop ends; snapshots resume a run exactly."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _small_kh(n=32, steps=5, freq=0, directory=None):
    from mfvm import io
    from mfvm.setups import kelvin_helmholtz_128 as kh
    s = kh.get_setup(n)
    s["time_config"]["n_steps"] = steps
    s["save_config"] = {"saver": io.save_snapshot if freq else None, "output_frequency": freq,
                        "directory": directory}
    return s


def test_setup_run_matches_oracle():
    from mfvm import solver
    from mfvm.setups import kelvin_helmholtz_128 as kh
    from oracle import grid as og
    from oracle import navier_stokes as ons
    from oracle import numerics as on
    from oracle import solver as osv
    s = _small_kh()
    U = solver.run(s)
    torch.cuda.synchronize()
    cfg = s["grid_config"]
    Uo = ons.allocate_u(cfg)
    g = og.build_grid("cartesian", cfg)
    kh.initializer(Uo, s["physical_parameters"], g, s["initial_conditions"]["parameters"])
    bo = {k: {side: og.periodic for side in og.SIDES} for k in ons.VARS}
    og.apply_bcs(Uo, bo)
    so = {"riemann_solver": on.hll, "reconstructor": on.muscl, "slope_limiter": on.minmod,
          "diffusive_solver": on.centered}
    osv.advance(ons, Uo, g, s["physical_parameters"], so, on.rk3_ssp, bo, 0.4, 5)
    for k in ons.VARS:
        assert np.array_equal(U[k].cpu().numpy(), Uo[k]), k


def test_setup_run_reports_diagnostics(capsys):
    """save_config["diagnostics"]: every output line is followed by the drift
    of the conserved totals and the positivity monitor; the run itself is
    unchanged (bitwise equal to one without diagnostics)."""
    from mfvm import solver
    plain = solver.run(_small_kh(steps=5))
    s = _small_kh(steps=5, freq=2)
    s["save_config"]["saver"] = None
    s["save_config"]["diagnostics"] = True
    capsys.readouterr()
    U = solver.run(s)
    out = capsys.readouterr().out.splitlines()
    torch.cuda.synchronize()
    assert torch.equal(U.buf, plain.buf)
    lines = [l for l in out if "total drift:" in l]
    assert len(lines) == len([l for l in out if l.startswith("n = ")]) >= 2
    for l in lines:
        assert "mass_density=" in l and "energy_density=" in l
        assert l.rstrip().endswith("cells with rho<=0 or P<=0: 0")
        mass = float(l.split("mass_density=")[1].split()[0])
        assert abs(mass) < 1e-6


def test_snapshot_resume_is_exact(tmp_path):
    from mfvm import io, solver
    from mfvm.grid.geometry import build_grid
    from mfvm.physics import navier_stokes as ns
    full = solver.run(_small_kh(steps=6))
    solver.run(_small_kh(steps=4, freq=3, directory=str(tmp_path)))
    U, setup, t, it = io.load_snapshot(str(tmp_path / "snapshot_000003.npz"))
    assert it == 3 and setup["nx1"] == 32
    s = _small_kh(steps=2)
    V = io.restore(ns.allocate_u(s["grid_config"]), U)
    solver.advance(s, V, build_grid("cartesian", s["grid_config"]), 2)
    torch.cuda.synchronize()
    assert torch.equal(V.buf[:, 1:-1, 1:-1, 1:-1], full.buf[:, 1:-1, 1:-1, 1:-1])


def _hot_spot(U, params, grid, ics):
    # a reference-format initializer (host NumPy on the host mirror of U)
    e = U["internal_energy_density"][1:-1, 1:-1, 1:-1]
    x = grid["cell"]["coordinates"]["x1"]
    y = grid["cell"]["coordinates"]["x2"]
    T = ics["T0"] + ics["A"] * np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / (2.0 * ics["w"] ** 2))
    e[:] = params["mass_density"] * params["specific_heat_capacity"] * T


def test_reference_format_heat_setup_t_end_loop():
    """A setup in the reference's own format (heat, Cartesian 2D, SSPRK3,
    MUSCL + van Leer, Neumann walls, host initializer) through the reference
    loop of mfvm.solver.run: dt read back every step, ``while t <= t_end``
    with no clipping.  It ends bit for bit where the same loop over the
    oracle ends."""
    from mfvm import solver
    from mfvm.grid import boundary_conditions as bc
    from mfvm.numerics import diffusive_solvers, reconstructors, slope_limiters, time_integrators
    from mfvm.physics import heat
    from oracle import grid as og
    from oracle import heat as oht
    from oracle import numerics as on
    cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": 40, "x2_min": 0.0, "x2_max": 1.0, "nx2": 24,
           "x3_min": 0.0, "x3_max": 1.0, "nx3": 1}
    params = {"mass_density": 10.0, "specific_heat_capacity": 2.0, "thermal_conductivity": 8.0,
              "volumetric_heat_source": 0.0}
    ics = {"T0": 300.0, "A": 1000.0, "w": 0.08}
    t_end = 3.5e-4
    setup = {"physics": heat, "physical_parameters": params, "coordinate_system": "cartesian",
             "grid_config": cfg, "time_config": {"t_start": 0.0, "t_end": t_end, "CFL": 0.2},
             "time_integrator": time_integrators.rk3_ssp,
             "space_schemes": {"reconstructor": reconstructors.muscl, "slope_limiter": slope_limiters.van_leer,
                               "diffusive_solver": diffusive_solvers.centered},
             "boundary_conditions": {"internal_energy_density": {s: bc.neumann for s in bc.SIDES}},
             "initial_conditions": {"initializer": _hot_spot, "parameters": ics},
             "save_config": {"saver": None, "output_frequency": 0, "directory": None}}
    U = solver.run(setup)
    torch.cuda.synchronize()

    g = og.build_grid("cartesian", cfg)
    Uo = oht.allocate_u(cfg)
    _hot_spot(Uo, params, g, ics)
    bo = {"internal_energy_density": {s: og.neumann for s in og.SIDES}}
    og.apply_bcs(Uo, bo)
    so = {"reconstructor": on.muscl, "slope_limiter": on.van_leer, "diffusive_solver": on.centered,
          "riemann_solver": None}
    t, n = 0.0, 0
    while t <= t_end:
        dt = oht.compute_dt(Uo, params, g, 0.2)
        on.rk3_ssp(oht.compute_rhs, Uo, dt, params, g, so, bo, og.apply_bcs)
        t += dt
        n += 1
    assert n >= 3
    assert np.array_equal(U["internal_energy_density"].cpu().numpy(), Uo["internal_energy_density"])


@pytest.mark.parametrize("which", ["electromagnetic_pulse_spherical", "heated_disc_cylindrical"])
def test_curvilinear_setups_run_match_oracle(which):
    """The curvilinear benchmark setups at a reduced size, through the
    fixed-step driver of mfvm.solver.run, against the oracle loop."""
    import importlib
    from mfvm import solver
    from oracle import grid as og
    from oracle import heat as oht
    from oracle import maxwell as omx
    from oracle import numerics as on
    from oracle import solver as osv
    mod = importlib.import_module(f"mfvm.setups.{which}")
    n = (24, 16, 20) if which.startswith("electro") else (24, 20, 8)
    s = mod.get_setup(n)
    steps = 3
    s["time_config"]["n_steps"] = steps
    s["save_config"] = {"saver": None, "output_frequency": 0, "directory": None}
    U = solver.run(s)
    torch.cuda.synchronize()

    ophys = omx if which.startswith("electro") else oht
    cfg = s["grid_config"]
    g = og.build_grid(s["coordinate_system"], cfg)
    Uo = ophys.allocate_u(cfg)
    s["initial_conditions"]["initializer"](Uo, s["physical_parameters"], g, s["initial_conditions"]["parameters"])
    names = {"neumann": og.neumann, "periodic": og.periodic, "reflecting": og.reflecting}
    bo = {k: {side: names[f.__name__] for side, f in sides.items()} for k, sides in s["boundary_conditions"].items()}
    og.apply_bcs(Uo, bo)
    sc = s["space_schemes"]
    so = {"reconstructor": getattr(on, sc["reconstructor"].__name__),
          "slope_limiter": getattr(on, sc["slope_limiter"].__name__),
          "riemann_solver": getattr(on, sc["riemann_solver"].__name__) if sc.get("riemann_solver") else None,
          "diffusive_solver": on.centered}
    integ = getattr(on, s["time_integrator"].__name__)
    osv.advance(ophys, Uo, g, s["physical_parameters"], so, integ, bo, s["time_config"]["CFL"], steps)
    for k in ophys.VARS:
        assert np.array_equal(U[k].cpu().numpy(), Uo[k]), k
