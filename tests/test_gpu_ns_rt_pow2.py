"""The benchmarked Rayleigh-Taylor kernels against the oracle, bit for bit.

The benchmark workload (3D Rayleigh-Taylor: MUSCL + van Leer, HLL, centred
viscous flux with mu = 1e-4, gravity -0.1 along x3, periodic x1/x2 and
reflecting x3 walls) runs on a uniform power-of-two grid, which selects the
specialised stage kernels: exact reciprocal lengths and volume, the viscous
terms compiled in, the fixed 64 x 4 tile with buffer-resource reads.  Here the
same workload runs on power-of-two unit-cube grids small enough for the oracle
(32^3, 64^3 and the non-cubic 64 x 32 x 16), through the split stages (the
default) and the fused stage:

* ``compute_rhs`` equals the oracle's right-hand side bitwise
  (navier_stokes.py:588-711);
* 1 and 5 SSPRK3 steps with the device dt (time_integrators.py:106-159,
  boundary_conditions.py:67-91 walls, navier_stokes.py:500-504 gravity)
  equal the oracle's state and dt sequence bitwise;
* the FP32-compensated path stays within L1 <= 1e-6 per field of the
  oracle's state after 5 steps (L-inf printed).
"""

import numpy as np
import pytest
import torch

import bench
from oracle import grid as og
from oracle import navier_stokes as ons
from oracle import numerics as on
from oracle import solver as osv
from oracle_setup import same

pytestmark = pytest.mark.gpu

SHAPES = [(32, 32, 32), (64, 64, 64), (64, 32, 16)]
_ORACLE = {}


def _oracle_bcs():
    rule = {0: og.neumann, 1: og.periodic, 2: og.reflecting}
    return {k: {s: rule[c] for s, c in zip(og.SIDES, row)} for k, row in zip(ons.VARS, bench.bc_rows())}


def _oracle(shape):
    """Oracle right-hand side, and state and dts after 1 and 5 steps."""
    if shape in _ORACLE:
        return _ORACLE[shape]
    cfg = bench.rt_config(shape)
    U = ons.allocate_u(cfg)
    for k, f in zip(ons.VARS, bench.rt_fields(shape)):
        U[k][1:-1, 1:-1, 1:-1] = f
    bcs = _oracle_bcs()
    og.apply_bcs(U, bcs)
    g = og.build_grid("cartesian", cfg)
    sch = {"riemann_solver": on.hll, "reconstructor": on.muscl, "slope_limiter": on.van_leer,
           "diffusive_solver": on.centered}
    start = {k: v.copy() for k, v in U.items()}
    rhs = ons.compute_rhs(U, g, bench.PARAMS, sch)
    dts = osv.advance(ons, U, g, bench.PARAMS, sch, on.rk3_ssp, bcs, bench.CFL, 1)
    one = {k: v.copy() for k, v in U.items()}
    dts += osv.advance(ons, U, g, bench.PARAMS, sch, on.rk3_ssp, bcs, bench.CFL, 4)
    _ORACLE[shape] = (start, rhs, one, dts, U)
    return _ORACLE[shape]


def _gpu(shape, start):
    from mfvm.grid import boundary_conditions as bc
    from mfvm.grid.geometry import build_grid
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    from mfvm.physics import navier_stokes as ns
    cfg = bench.rt_config(shape)
    U = ns.allocate_u(cfg)
    for k in ns.VARS:
        U[k] = torch.from_numpy(start[k])
    rule = {0: bc.neumann, 1: bc.periodic, 2: bc.reflecting}
    bcs = {k: {s: rule[c] for s, c in zip(bc.SIDES, row)} for k, row in zip(ns.VARS, bench.bc_rows())}
    sch = {"riemann_solver": riemann_solvers.hll, "reconstructor": reconstructors.muscl,
           "slope_limiter": slope_limiters.van_leer, "diffusive_solver": diffusive_solvers.centered}
    return ns, U, build_grid("cartesian", cfg), sch, bcs


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def test_rt_pow2_rhs_matches_oracle(shape):
    start, rhs_o, *_ = _oracle(shape)
    ns, U, grid, sch, _ = _gpu(shape, start)
    print(f"rt pow2 {shape}: compute_rhs", flush=True)
    rhs = ns.compute_rhs(U, grid, bench.PARAMS, sch)
    for k in ns.VARS:
        assert same(rhs[k].cpu().numpy(), rhs_o[k]), k


@pytest.mark.parametrize("split", ["1", "0"], ids=["split", "fused"])
@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def test_rt_pow2_steps_match_oracle(shape, split, monkeypatch):
    from mfvm import _lib
    start, _, one, dts_o, five = _oracle(shape)
    monkeypatch.setenv("MFVM_SPLIT", split)
    ns, U, grid, sch, bcs = _gpu(shape, start)
    print(f"rt pow2 {shape} split={split}: 1 step", flush=True)
    dts = ns.advance(U, grid, bench.PARAMS, sch, bcs, _lib.INT_SSPRK3, bench.CFL, 1)
    torch.cuda.synchronize()
    assert dts.cpu().numpy()[0] == dts_o[0]
    for k in ns.VARS:
        g = U[k].cpu().numpy()
        assert np.max(np.abs(g - one[k])) <= 1e-12 * np.max(np.abs(one[k])), k
        assert same(g, one[k]), k
    print(f"rt pow2 {shape} split={split}: 4 more steps", flush=True)
    dts4 = ns.advance(U, grid, bench.PARAMS, sch, bcs, _lib.INT_SSPRK3, bench.CFL, 4)
    torch.cuda.synchronize()
    assert dts4.cpu().numpy()[:4].tolist() == dts_o[1:5]
    for k in ns.VARS:
        assert same(U[k].cpu().numpy(), five[k]), k


@pytest.mark.parametrize("split", ["1", "0"], ids=["split", "fused"])
@pytest.mark.parametrize("lim", ["vanleer", "minmod", "mc"])
def test_rt_pow2_limiter_edge_slopes_match_oracle(lim, split, monkeypatch):
    """Slopes at the edges of the limiters (slope_limiters.py) through the
    power-of-two viscous stage kernels, planted in x1, x2 and x3 rows of an
    RT 32^3 state: opposite-signed slopes with a subnormal product, one-signed
    slopes with a subnormal numerator (2 b) f, one-signed slopes whose
    numerator underflows to zero, and zero slopes beside nonzero ones.
    compute_rhs equals the oracle's bitwise.

    (States whose flux components overflow are outside the bar: the reference
    projects every flux on the face normal as n1 F1 + n2 F2 + n3 F3, and
    0 * inf in a tangential component makes the face flux NaN where these
    kernels, which form the normal component only, keep it finite.)"""
    from mfvm.numerics import slope_limiters
    shape = (32, 32, 32)
    cfg = bench.rt_config(shape)
    params = bench.PARAMS
    fields = [f.copy() for f in bench.rt_fields(shape)]
    fields[2][10:14, 5, 7] = [0.0, 0.0, 3e-162, 0.0]
    fields[3][20, 14:17, 9] = [0.0, 3e-162, 0.0]
    fields[1][3, 25, 20:23] = [0.0, 1e-162, 2e-162]
    fields[1][7:10, 9, 11] = [0.0, 1e-172, 2e-172]
    fields[3][15, 3, 27:30] = [-1e-172, -2e-172, -2e-172]
    U = ons.allocate_u(cfg)
    for k, f in zip(ons.VARS, fields):
        U[k][1:-1, 1:-1, 1:-1] = f
    og.apply_bcs(U, _oracle_bcs())
    start = {k: v.copy() for k, v in U.items()}
    limiter = {"vanleer": on.van_leer, "minmod": on.minmod, "mc": on.monotonized_central}[lim]
    sch = {"riemann_solver": on.hll, "reconstructor": on.muscl, "slope_limiter": limiter,
           "diffusive_solver": on.centered}
    with np.errstate(under="ignore"):
        rhs_o = ons.compute_rhs(U, og.build_grid("cartesian", cfg), params, sch)
    monkeypatch.setenv("MFVM_SPLIT", split)
    ns, Ug, grid, gsch, _ = _gpu(shape, start)
    gsch["slope_limiter"] = {"vanleer": slope_limiters.van_leer, "minmod": slope_limiters.minmod,
                             "mc": slope_limiters.monotonized_central}[lim]
    print(f"rt pow2 limiter edges {lim} split={split}: compute_rhs", flush=True)
    rhs = ns.compute_rhs(Ug, grid, params, gsch)
    for k in ns.VARS:
        assert same(rhs[k].cpu().numpy(), rhs_o[k]), k


@pytest.mark.parametrize("shape", SHAPES[:1] + SHAPES[2:], ids=["32x32x32", "64x32x16"])
def test_rt_pow2_compensated_near_oracle(shape):
    """FP32-compensated steps against the oracle's FP64 state: L1 <= 1e-6
    per field after 5 steps."""
    from mfvm import _lib
    start, _, _, _, five = _oracle(shape)
    ns, U, grid, sch, bcs = _gpu(shape, start)
    Uc = ns.to_compensated(U)
    print(f"rt pow2 {shape}: compensated, 5 steps", flush=True)
    ns.advance_compensated(Uc, grid, bench.PARAMS, sch, bcs, _lib.INT_SSPRK3, bench.CFL, 5)
    torch.cuda.synchronize()
    val = ns.compensated_value(Uc).cpu().numpy()
    for v, k in enumerate(ns.VARS):
        d = np.abs(val[v, 1:-1, 1:-1, 1:-1] - five[k][1:-1, 1:-1, 1:-1].transpose(2, 1, 0))
        print(f"  {k}: L1 {d.mean():.3e} Linf {d.max():.3e}", flush=True)
        assert d.mean() <= 1e-6, (k, d.mean())
