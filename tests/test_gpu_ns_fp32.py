"""FP32-compensated Navier-Stokes path vs the FP64 path / oracle.

Tolerance: L1 (mean absolute) error <= 1e-6 per field against the float64
result, on the 2D Kelvin-Helmholtz configuration (200 SSPRK3 steps) and a
3D Rayleigh-Taylor case.  L-inf is printed for every field.  Plain FP32
(no compensation) is run beside it for comparison and only reported.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _setup(n, kind):
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": n[0], "x2_min": 0.0, "x2_max": 1.0, "nx2": n[1],
           "x3_min": 0.0, "x3_max": 1.0, "nx3": n[2]}
    grid = build_grid("cartesian", cfg)
    U = ns.allocate_u(cfg)
    rng = np.random.default_rng(0)
    if kind == "kh":
        x = (np.arange(n[0]) + 0.5) / n[0]
        y = (np.arange(n[1]) + 0.5) / n[1]
        X, Y = np.meshgrid(x, y, indexing="ij")
        X, Y = X[:, :, None], Y[:, :, None]
        inner = np.abs(Y - 0.5) < 0.25
        rho = np.where(inner, 2.0, 1.0)
        vx = np.where(inner, 0.5, -0.5)
        vy = 0.01 * np.sin(4 * np.pi * X) + rng.uniform(-1e-3, 1e-3, n)
        fields = [rho, rho * vx, rho * vy, np.zeros(n), 2.5 / 0.4 + 0.5 * rho * (vx * vx + vy * vy)]
        params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-4,
                  "gravitational_acceleration": [0.0, 0.0, 0.0]}
        rows = [[1] * 6] * 5
        lim = slope_limiters.minmod
    else:
        import bench
        fields = bench.rt_fields(n[0])
        params = bench.PARAMS
        rows = bench.bc_rows()
        lim = slope_limiters.van_leer
    for k, f in zip(ns.VARS, fields):
        U[k][1:-1, 1:-1, 1:-1] = torch.from_numpy(np.asarray(f, dtype=np.float64)).cuda()
    rule = {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}
    bcs = {k: {s: rule[c] for s, c in zip(bc.SIDES, row)} for k, row in zip(ns.VARS, rows)}
    schemes = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
               "slope_limiter": lim, "diffusive_solver": diffusive_solvers.centered}
    return ns, grid, U, params, schemes, bcs


def _errors(ref, got, names):
    out = {}
    for v, name in enumerate(names):
        r = ref[v, 1:-1, 1:-1, 1:-1]
        g = got[v, 1:-1, 1:-1, 1:-1]
        d = (g - r).abs()
        out[name] = (float(d.mean()), float(d.max()))
    return out


@pytest.mark.parametrize("case,n,steps", [("kh", (128, 128, 1), 200), ("rt", (32, 32, 32), 40)])
def test_compensated_fp32_matches_fp64(case, n, steps):
    from mfvm import _lib
    ns, grid, U, params, schemes, bcs = _setup(n, case)
    Uc = ns.to_compensated(U)
    Up = ns.allocate_u({"nx1": n[0], "nx2": n[1], "nx3": n[2]}, dtype=torch.float32)
    Up.buf.copy_(U.buf.to(torch.float32))
    ns.advance(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, steps)
    ns.advance_compensated(Uc, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, steps)
    ns.advance(Up, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, steps)
    torch.cuda.synchronize()
    comp = _errors(U.buf, ns.compensated_value(Uc), ns.VARS)
    plain = _errors(U.buf, Up.buf.to(torch.float64), ns.VARS)
    for name in ns.VARS:
        print(f"{case} {name}: comp L1 {comp[name][0]:.3e} Linf {comp[name][1]:.3e} | "
              f"plain fp32 L1 {plain[name][0]:.3e} Linf {plain[name][1]:.3e}", flush=True)
    for name in ns.VARS:
        assert comp[name][0] <= 1e-6, (name, comp[name])


@pytest.mark.parametrize("comp", [False, True])
def test_split_stages_equal_fused_stages(comp, monkeypatch):
    """Split stages (x1/x2 kernel + x3 kernel) give the fused stage's results
    bit for bit, FP64 and FP32-compensated (3D, curvilinear-free RT case with
    gravity, viscosity and walls)."""
    from mfvm import _lib
    ns, grid, U, params, schemes, bcs = _setup((24, 24, 24), "rt")
    V = ns.allocate_u({"nx1": 24, "nx2": 24, "nx3": 24})
    V.buf.copy_(U.buf)
    if comp:
        U, V = ns.to_compensated(U), ns.to_compensated(V)
        run = ns.advance_compensated
    else:
        run = ns.advance
    monkeypatch.setenv("MFVM_SPLIT", "1")
    monkeypatch.setenv("MFVM_SPLIT_COMP", "1")
    d1 = run(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 6)
    monkeypatch.setenv("MFVM_SPLIT", "0")
    monkeypatch.setenv("MFVM_SPLIT_COMP", "0")
    d2 = run(V, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 6)
    torch.cuda.synchronize()
    assert torch.equal(d1, d2)
    assert torch.equal(U.buf, V.buf)
    if comp:
        shp = tuple(U.buf.shape)
        assert torch.equal(U.scratch("lo", shp), V.scratch("lo", shp))


def test_split_stages_equal_fused_stages_2d(monkeypatch):
    """The same for a 2D grid (split forced below its size threshold)."""
    from mfvm import _lib
    ns, grid, U, params, schemes, bcs = _setup((96, 64, 1), "kh")
    V = ns.allocate_u({"nx1": 96, "nx2": 64, "nx3": 1})
    V.buf.copy_(U.buf)
    monkeypatch.setenv("MFVM_SPLIT_2D", "1")
    d1 = ns.advance(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 8)
    monkeypatch.setenv("MFVM_SPLIT_2D", "0")
    d2 = ns.advance(V, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 8)
    torch.cuda.synchronize()
    assert torch.equal(d1, d2)
    assert torch.equal(U.buf, V.buf)


def test_rt256_full_size_split_fused_and_compensated(monkeypatch):
    """The benchmark's workload at its full size (Rayleigh-Taylor 256^3),
    where the oracle would take hours: properties that do not depend on the
    size.  Split stages equal the fused stage bit for bit over 3 SSPRK3 steps
    (dt sequence included), and 20 FP32-compensated steps stay within the
    L1 <= 1e-6 bar of the FP64 result."""
    from mfvm import _lib
    n = (256, 256, 256)
    ns, grid, U, params, schemes, bcs = _setup(n, "rt")
    cfg = {"nx1": n[0], "nx2": n[1], "nx3": n[2]}
    V = ns.allocate_u(cfg)
    V.buf.copy_(U.buf)
    U0 = U.buf.clone()
    print("rt256: split vs fused, 3 steps", flush=True)
    monkeypatch.setenv("MFVM_SPLIT", "1")
    d1 = ns.advance(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 3)
    monkeypatch.setenv("MFVM_SPLIT", "0")
    d2 = ns.advance(V, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 3)
    torch.cuda.synchronize()
    assert torch.equal(d1, d2)
    assert torch.equal(U.buf, V.buf)
    del V
    print("rt256: compensated vs fp64, 20 steps", flush=True)
    monkeypatch.setenv("MFVM_SPLIT", "1")
    U.buf.copy_(U0)
    Uc = ns.to_compensated(U)
    ns.advance(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 20)
    ns.advance_compensated(Uc, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 20)
    torch.cuda.synchronize()
    err = _errors(U.buf, ns.compensated_value(Uc), ns.VARS)
    for name in ns.VARS:
        print(f"rt256 {name}: comp L1 {err[name][0]:.3e} Linf {err[name][1]:.3e}", flush=True)
        assert err[name][0] <= 1e-6, (name, err[name])


@pytest.mark.parametrize("comp", [False, True])
def test_split_stages_with_runtime_tile_equal_fused(comp, monkeypatch):
    """The x1/x2 kernel's runtime-tile form (any tile but 64 x 4, and arrays
    above 4 GiB) gives the fused stage's results bit for bit as well."""
    from mfvm import _lib
    ns, grid, U, params, schemes, bcs = _setup((40, 40, 40), "rt")
    V = ns.allocate_u({"nx1": 40, "nx2": 40, "nx3": 40})
    V.buf.copy_(U.buf)
    if comp:
        U, V = ns.to_compensated(U), ns.to_compensated(V)
        run = ns.advance_compensated
    else:
        run = ns.advance
    tile = (32, 8, 8)
    monkeypatch.setenv("MFVM_SPLIT", "1")
    monkeypatch.setenv("MFVM_SPLIT_COMP", "1")
    d1 = run(U, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 4, tile=tile)
    monkeypatch.setenv("MFVM_SPLIT", "0")
    monkeypatch.setenv("MFVM_SPLIT_COMP", "0")
    d2 = run(V, grid, params, schemes, bcs, _lib.INT_SSPRK3, 0.4, 4, tile=tile)
    torch.cuda.synchronize()
    assert torch.equal(d1, d2)
    assert torch.equal(U.buf, V.buf)
