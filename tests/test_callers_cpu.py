"""Callers around the hot path, CPU side: the host grid dict, snapshot files
and the setup dicts."""

import numpy as np
import pytest


@pytest.mark.parametrize("coord,lo,hi,n", [
    ("cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (8, 6, 4)),
    ("cylindrical", (0.0, 0.0, 0.0), (1.0, 2 * np.pi, 1.0), (8, 6, 4)),
    ("spherical", (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi), (8, 6, 4)),
])
def test_host_grid_equals_oracle_grid(coord, lo, hi, n):
    from mfvm.grid.geometry import build_grid_host
    from oracle import grid as og
    cfg = {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1], "nx2": n[1],
           "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}
    got = build_grid_host(coord, cfg)
    ref = og.build_grid(coord, cfg)

    def walk(a, b, path):
        if isinstance(b, dict):
            for k in b:
                walk(a[k], b[k], path + (k,))
        else:
            assert np.array_equal(np.asarray(a), np.asarray(b)), path
    walk(got["cell"], ref["cell"], ("cell",))
    walk(got["lame"], ref["lame"], ("lame",))


def test_snapshot_roundtrip(tmp_path):
    from mfvm import io
    rng = np.random.default_rng(0)
    U = {"mass_density": rng.random((6, 5, 3)), "energy_density": rng.random((6, 5, 3))}
    setup = {"coordinate_system": "cartesian",
             "grid_config": {"x1_min": 0.0, "x1_max": 1.0, "nx1": 4, "x2_min": 0.0, "x2_max": 1.0, "nx2": 3,
                             "x3_min": 0.0, "x3_max": 1.0, "nx3": 1},
             "physical_parameters": {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4}}
    path = io.save_snapshot(U, setup, 0.125, 42, str(tmp_path))
    assert path.endswith("snapshot_000042.npz")
    V, s, t, it = io.load_snapshot(path)
    assert t == 0.125 and it == 42
    assert s["coordinate_system"] == "cartesian" and s["nx1"] == 4 and s["dynamic_viscosity"] == 1e-4
    for k in U:
        assert np.array_equal(U[k], V[k])


def test_setups_have_the_reference_keys():
    import importlib
    from mfvm import setups
    keys = {"physics", "physical_parameters", "coordinate_system", "grid_config", "time_config",
            "time_integrator", "space_schemes", "boundary_conditions", "initial_conditions", "save_config"}
    for name in setups.NAMES:
        mod = importlib.import_module(f"mfvm.setups.{name}")
        s = mod.get_setup()
        assert keys <= set(s), name
        assert set(s["boundary_conditions"]) == set(s["physics"].VARS), name
    from mfvm.setups import ensemble_kh64
    shapes = ensemble_kh64.member_shapes()
    assert len(shapes) == 64 and shapes.count((64, 64, 64)) == 8
    assert all(a in (64, 128, 256, 512) and b in (64, 128, 256, 512) for a, b, c in shapes[:56])
