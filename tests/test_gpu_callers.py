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
