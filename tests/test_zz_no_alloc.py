"""The allocation-free CFL step, mfvm_ns_dt_noalloc, against mfvm_ns_dt.

Both compute navier_stokes.py:512-585's dt of the same state; the
allocation-free form keeps its CFL partials in the library's fixed device
array instead of a caller-supplied buffer.  Called through the C-ABI twice
in a row, the second time on a smaller grid (the array keeps the first
call's contents), in float64 and float32; the dt must be bitwise equal.  A
grid beyond the array's fixed size is rejected with an error before any
launch.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(n, dtype, seed):
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": n[0], "x2_min": 0.0, "x2_max": 2.0, "nx2": n[1],
           "x3_min": 0.0, "x3_max": 1.0, "nx3": n[2]}
    grid = build_grid("cartesian", cfg)
    U = ns.allocate_u(cfg, dtype=dtype)
    rng = np.random.default_rng(seed)
    shp = tuple(U.buf.shape[1:])
    rho = 1.0 + rng.random(shp)
    v = [0.5 * (2 * rng.random(shp) - 1) for _ in range(3)]
    p = 1.0 + rng.random(shp)
    vals = [rho, rho * v[0], rho * v[1], rho * v[2], p / 0.4 + 0.5 * rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)]
    U.buf.copy_(torch.from_numpy(np.stack(vals)).to(dtype))
    params = {"heat_capacity_ratio": 1.4, "dynamic_viscosity": 1e-3,
              "gravitational_acceleration": [0.0, 0.0, 0.0]}
    sch = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
           "slope_limiter": slope_limiters.van_leer, "diffusive_solver": diffusive_solvers.centered}
    return ns, grid, U, ns.make_params(U, params, sch)


def _both(n, dtype, seed):
    from mfvm import _lib
    from mfvm.physics import _common as c
    ns, grid, U, p = _case(n, dtype, seed)
    g = grid["mfvm"]
    nslot = _lib.lib.mfvm_ns_partials_size(g.ptr(), ctypes.byref(p))
    partials = torch.empty(int(nslot), dtype=torch.float64, device="cuda")
    dts = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    print(f"ns_dt vs ns_dt_noalloc: {n} {dtype}", flush=True)
    _lib.check(_lib.lib.mfvm_ns_dt(g.ptr(), ctypes.byref(p), c.ptr(U.buf), ctypes.c_double(0.4),
                                   c.ptr(partials), ctypes.c_void_p(dts.data_ptr()), c.stream()))
    _lib.check(_lib.lib.mfvm_ns_dt_noalloc(g.ptr(), ctypes.byref(p), c.ptr(U.buf), ctypes.c_double(0.4),
                                           ctypes.c_void_p(dts.data_ptr() + 8), c.stream()))
    torch.cuda.synchronize()
    a, b = dts.cpu().tolist()
    return a, b


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_dt_equals_dt_on_two_calls(dtype):
    for n, seed in (((96, 64, 40), 1), ((20, 12, 7), 2)):
        a, b = _both(n, dtype, seed)
        assert np.isfinite(a) and a > 0, (n, a)
        assert a == b, (n, dtype, a, b)


def test_noalloc_dt_rejects_grids_beyond_its_scratch():
    from mfvm import _lib
    from mfvm.physics import _common as c
    ns, grid, U, p = _case((8, 8, 8), torch.float64, 3)
    from mfvm.grid.geometry import build_grid
    big = build_grid("cartesian", {"x1_min": 0.0, "x1_max": 1.0, "nx1": 2048, "x2_min": 0.0, "x2_max": 1.0,
                                   "nx2": 2048, "x3_min": 0.0, "x3_max": 1.0, "nx3": 64})
    p.tile[:] = [64, 4, 1]
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = _lib.lib.mfvm_ns_dt_noalloc(big["mfvm"].ptr(), ctypes.byref(p), c.ptr(U.buf), ctypes.c_double(0.4),
                                     c.ptr(dt), c.stream())
    assert rc == _lib.E_ARG
    assert b"fixed maximum" in _lib.lib.mfvm_last_error()
