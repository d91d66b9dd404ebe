"""CPU-side checks of the C-ABI library and the host-side grid factors.

No kernel runs here: argument validation returns before any HIP call, and
the factor tests evaluate the 1D products on the host.
"""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mfvm.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mfvm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from mfvm import _lib
    names = _declared()
    assert len(names) >= 10
    for name in names:
        assert hasattr(_lib.lib, name), name
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert bound == set(names)


def test_abi_version_and_error_paths():
    from mfvm import _lib
    assert _lib.lib.mfvm_abi_version() == 1
    g = _lib.Grid()  # zero sizes -> rejected before any HIP call
    p = _lib.NSParams()
    rc = _lib.lib.mfvm_ns_closure(ctypes.byref(g), ctypes.byref(p), None, None, None)
    assert rc == _lib.E_ARG
    assert "grid sizes" in _lib.lib.mfvm_last_error().decode()
    with pytest.raises(ValueError):
        _lib.check(rc)
    g.n[:] = [4, 4, 4]
    g.coord = 7
    g.fac = 1
    g.fstride = 5
    rc = _lib.lib.mfvm_ns_closure(ctypes.byref(g), ctypes.byref(p), None, None, None)
    assert rc == _lib.E_ARG and "coordinate system" in _lib.lib.mfvm_last_error().decode()
    g.coord = 0
    p.limiter = 9
    rc = _lib.lib.mfvm_ns_closure(ctypes.byref(g), ctypes.byref(p), None, None, None)
    assert rc == _lib.E_UNSUP
    p.limiter = 1
    p.bc[0][0] = 5
    assert _lib.lib.mfvm_ns_closure(ctypes.byref(g), ctypes.byref(p), None, None, None) == _lib.E_ARG


@pytest.mark.parametrize("coord,lo,hi,n", [
    ("cartesian", (0.0, -0.3, 0.1), (0.7, 1.1, 0.6), (6, 7, 5)),
    ("cylindrical", (0.1, 0.0, -1.0), (2.0, 2 * np.pi, 1.0), (16, 12, 8)),
    ("spherical", (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi), (16, 12, 10)),
    ("spherical", (0.1, 0.05, 0.0), (1.0, np.pi - 0.05, 2 * np.pi), (32, 32, 64)),
])
def test_metric_factors_reproduce_reference_grid(coord, lo, hi, n):
    """(a[i]*b[j])*c[k] over the 1D factors == the oracle's 3D metrics, bitwise
    (the oracle's grids equal the reference's golden grids)."""
    from mfvm import _lib
    from mfvm.grid.geometry import GridDesc
    from oracle import grid as og
    cfg = {"x1_min": lo[0], "x1_max": hi[0], "nx1": n[0], "x2_min": lo[1], "x2_max": hi[1],
           "nx2": n[1], "x3_min": lo[2], "x3_max": hi[2], "nx3": n[2]}
    gd = GridDesc(coord, cfg, "cpu")
    ref = og.build_grid(coord, cfg)
    geo = ref["cell"]["geometry"]
    dv = ref["lame"]["derivatives"]
    want = {_lib.MET_V: geo["volume"],
            _lib.MET_S1: geo["surface"]["x1"], _lib.MET_S2: geo["surface"]["x2"],
            _lib.MET_S3: geo["surface"]["x3"],
            _lib.MET_L1: geo["length"]["x1"], _lib.MET_L2: geo["length"]["x2"],
            _lib.MET_L3: geo["length"]["x3"],
            _lib.MET_H1: ref["lame"]["coefficients"]["x1"],
            _lib.MET_H2: ref["lame"]["coefficients"]["x2"],
            _lib.MET_H3: ref["lame"]["coefficients"]["x3"],
            _lib.MET_D1H: dv["x1"]["ln_hx1hx2hx3"], _lib.MET_D1H2: dv["x1"]["ln_hx2"],
            _lib.MET_D1H3: dv["x1"]["ln_hx3"], _lib.MET_D2H: dv["x2"]["ln_hx1hx2hx3"],
            _lib.MET_D2H1: dv["x2"]["ln_hx1"], _lib.MET_D2H3: dv["x2"]["ln_hx3"],
            _lib.MET_D3H: dv["x3"]["ln_hx1hx2hx3"], _lib.MET_D3H1: dv["x3"]["ln_hx1"],
            _lib.MET_D3H2: dv["x3"]["ln_hx2"]}
    for m, ref_arr in want.items():
        a, b, c = gd.factors_host[m]
        got = (a[:, None, None] * b[None, :, None]) * c[None, None, :]
        assert got.shape == ref_arr.shape, m
        assert np.array_equal(got, ref_arr), m
    L = [geo["length"][d] for d in ("x1", "x2", "x3")]
    smax = np.max(1 / L[0] ** 2 + 1 / L[1] ** 2 + 1 / L[2] ** 2)
    assert gd.desc.smax == smax


def test_uniform_power_of_two_detection():
    from mfvm.grid.geometry import GridDesc
    cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": 256, "x2_min": 0.0, "x2_max": 1.0,
           "nx2": 128, "x3_min": 0.0, "x3_max": 1.0, "nx3": 1}
    d = GridDesc("cartesian", cfg, "cpu").desc
    assert list(d.lpow2) == [1, 1, 1] and d.vpow2 == 1
    assert list(d.lval) == [1 / 256, 1 / 128, 1.0] and d.vval == 1 / 256 / 128
    cfg["x1_max"] = 0.7
    d = GridDesc("cartesian", cfg, "cpu").desc
    assert d.lpow2[0] == 0 and d.vpow2 == 0


def test_ensemble_argument_checks():
    """mfvm_ens_*: sizes, tiles, integrators and buffers are validated on the
    host (no HIP call before the checks pass; nsteps=0 returns early)."""
    from mfvm import _lib
    from mfvm.grid.geometry import GridDesc
    cfg = {"x1_min": 0.0, "x1_max": 1.0, "nx1": 100, "x2_min": 0.0, "x2_max": 1.0, "nx2": 30,
           "x3_min": 0.0, "x3_max": 1.0, "nx3": 1}
    gd = GridDesc("cartesian", cfg, "cpu")
    m = (_lib.EnsMember * 2)()
    for i in range(2):
        m[i].grid = gd.desc
        for v in range(5):
            m[i].bc[v][:] = [1] * 6
    assert _lib.lib.mfvm_ens_member_tiles(m) == 2 * 8          # 64 x 4 tiles
    m[0].tile[:] = [32, 8, 0]
    assert _lib.lib.mfvm_ens_member_tiles(m) == 4 * 4
    m[0].tile[:] = [32, 4, 0]                                     # 128 threads: rejected
    assert _lib.lib.mfvm_ens_member_tiles(m) == -1
    assert "256" in _lib.lib.mfvm_last_error().decode()
    m[0].tile[:] = [0, 0, 0]
    assert _lib.lib.mfvm_ens_table_bytes(m, 2, _lib.INT_EULER) == -1
    nb = _lib.lib.mfvm_ens_table_bytes(m, 2, _lib.INT_RK3)
    assert nb > 0 and nb % 256 == 0
    p = _lib.NSParams()
    p.dtype = _lib.F64
    p.limiter = _lib.LIM_MC
    p.riemann = _lib.RS_RUSANOV
    p.gamma = 1.4
    o = _lib.EnsOpts()
    args = lambda integ, comp, table, nbytes, steps: (m, 2, ctypes.byref(p), ctypes.byref(o), integ, comp,
                                                      0.4, steps, table, nbytes, None)
    # NULL member buffers
    assert _lib.lib.mfvm_ens_advance(*args(_lib.INT_RK3, 0, 1, nb, 0)) == _lib.E_ARG
    for i in range(2):
        for f in ("U", "S1", "S2", "C", "kacc", "partials", "dts"):
            setattr(m[i], f, 4096)
    assert _lib.lib.mfvm_ens_advance(*args(_lib.INT_RK3, 0, 1, nb, 0)) == _lib.OK
    assert _lib.lib.mfvm_ens_advance(*args(_lib.INT_EULER, 0, 1, nb, 0)) == _lib.E_UNSUP
    assert _lib.lib.mfvm_ens_advance(*args(_lib.INT_RK3, 0, 1, nb - 1, 0)) == _lib.E_ARG
    assert "table" in _lib.lib.mfvm_last_error().decode()
    assert _lib.lib.mfvm_ens_advance(*args(_lib.INT_RK3, 1, 1, nb, 0)) == _lib.E_ARG   # comp needs f32
    p.dtype = _lib.F32
    assert _lib.lib.mfvm_ens_advance(*args(_lib.INT_RK3, 1, 1, nb, 0)) == _lib.E_ARG   # lo / dinc
    o.order = 5
    p.dtype = _lib.F64
    assert _lib.lib.mfvm_ens_advance(*args(_lib.INT_RK3, 0, 1, nb, 0)) == _lib.E_ARG


def test_diagnostics_argument_checks():
    from mfvm import _lib
    g = _lib.Grid()
    g.n[:] = [64, 32, 16]
    g.coord = 0
    g.fac = 1          # never dereferenced: every call below is rejected first
    g.fstride = 65
    P = ctypes.c_void_p
    assert _lib.lib.mfvm_diag_partials_size(ctypes.byref(g), 0) == -1
    assert _lib.lib.mfvm_diag_partials_size(ctypes.byref(g), 7) == -1
    # 32768 cells at 1024 per block -> 32 blocks of 5*5+3 slots
    assert _lib.lib.mfvm_diag_partials_size(ctypes.byref(g), 5) == 32 * 28
    big = _lib.Grid()
    big.n[:] = [2048, 1024, 1024]
    big.fac, big.fstride = 1, 2049
    assert _lib.lib.mfvm_diag_partials_size(ctypes.byref(big), 1) == -1
    assert "2^31" in _lib.lib.mfvm_last_error().decode()

    def call(nvar, dtype, U, lo, ref, gamma, part, out):
        return _lib.lib.mfvm_diagnostics(ctypes.byref(g), nvar, dtype, U, lo, ref, gamma, part, out, None)
    assert call(5, _lib.F64, None, None, None, 0.0, P(8), P(8)) == _lib.E_ARG
    assert "NULL" in _lib.lib.mfvm_last_error().decode()
    assert call(5, 9, P(8), None, None, 0.0, P(8), P(8)) == _lib.E_ARG
    assert call(5, _lib.F64, P(8), P(8), None, 0.0, P(8), P(8)) == _lib.E_ARG
    assert "compensation" in _lib.lib.mfvm_last_error().decode()
    assert call(6, _lib.F64, P(8), None, None, 1.4, P(8), P(8)) == _lib.E_ARG
    assert "pressure" in _lib.lib.mfvm_last_error().decode()
