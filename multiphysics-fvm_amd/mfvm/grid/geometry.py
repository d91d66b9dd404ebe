"""Grid geometry as 1D metric factors on the device.

Drop-in for ``src/grid/geometry.py:14-194`` (``build_grid``).  Every metric
of the reference (volume, face areas, cell lengths, Lame coefficients and
their log-derivatives) is a broadcast product of 1D factors; here each metric
is stored as three 1D factor arrays a(x1), b(x2), c(x3) and evaluated as
``(a[i] * b[j]) * c[k]`` -- the reference's own left-to-right product order,
so values are identical.  The kernels read only the factors (a few KB); the
returned dict additionally holds the full 3D tensors on the device, in the
reference layout, for callers that use the dict directly.

Trigonometric factors are evaluated with NumPy on arrays of the same shape as
in the reference, so that they carry NumPy's last-ulp behaviour; the GPU never
evaluates trigonometric functions.
"""

import ctypes
import math

import numpy as np
import torch

from .. import _lib


def _axis(grid_config, a):
    lo = grid_config[f"x{a}_min"]
    hi = grid_config[f"x{a}_max"]
    n = grid_config[f"nx{a}"]
    h = (hi - lo) / n                                   # geometry.py:37-39
    centre = lo + (np.arange(n) + 0.5) * h              # :42-44
    face = lo + np.arange(n + 1) * h                    # :47-49
    return n, centre, face, face[1:] - face[:-1]


def _pow2(x):
    m, _ = math.frexp(x)
    return x > 0 and m == 0.5


def _factors(coordinate_system, grid_config):
    """Return {metric slot: (a, b, c)} following geometry.py:80-171."""
    n1, c1, f1, d1 = _axis(grid_config, 1)
    n2, c2, f2, d2 = _axis(grid_config, 2)
    n3, c3, f3, d3 = _axis(grid_config, 3)
    one = [np.ones(n1), np.ones(n2), np.ones(n3)]
    zero1 = np.zeros(n1)
    F = {}
    if coordinate_system == "cartesian":
        F[_lib.MET_V] = (d1, d2, d3)
        F[_lib.MET_S1] = (np.ones(n1 + 1), d2, d3)
        F[_lib.MET_S2] = (d1, np.ones(n2 + 1), d3)
        F[_lib.MET_S3] = (d1, d2, np.ones(n3 + 1))
        F[_lib.MET_L1] = (d1, one[1], one[2])
        F[_lib.MET_L2] = (one[0], d2, one[2])
        F[_lib.MET_L3] = (one[0], one[1], d3)
        F[_lib.MET_H1] = (one[0], one[1], one[2])
        F[_lib.MET_H2] = (one[0], one[1], one[2])
        F[_lib.MET_H3] = (one[0], one[1], one[2])
        for m in range(_lib.MET_D1H, _lib.NMET):
            F[m] = (zero1, one[1], one[2])
    elif coordinate_system == "cylindrical":
        a2 = (f1[1:] ** 2 - f1[:-1] ** 2) / 2.0
        F[_lib.MET_V] = (a2, d2, d3)                        # :114
        F[_lib.MET_S1] = (f1, d2, d3)                       # :117
        F[_lib.MET_S2] = (d1, np.ones(n2 + 1), d3)          # :118
        F[_lib.MET_S3] = (a2, d2, np.ones(n3 + 1))          # :119
        F[_lib.MET_L1] = (d1, one[1], one[2])
        F[_lib.MET_L2] = (c1, d2, one[2])                   # :123
        F[_lib.MET_L3] = (one[0], one[1], d3)
        F[_lib.MET_H1] = (one[0], one[1], one[2])
        F[_lib.MET_H2] = (c1, one[1], one[2])               # :128
        F[_lib.MET_H3] = (one[0], one[1], one[2])
        for m in range(_lib.MET_D1H, _lib.NMET):
            F[m] = (zero1, one[1], one[2])
        F[_lib.MET_D1H2] = (1.0 / c1, one[1], one[2])        # :134
        F[_lib.MET_D1H] = (1.0 / c1, one[1], one[2])         # :138
    elif coordinate_system == "spherical":
        # Trig on arrays shaped as in geometry.py:145-170, then sliced to 1D.
        x2f3 = f2[None, :, None] * np.ones((n1, n2 + 1, n3))
        x2c3 = np.meshgrid(c1, c2, c3, indexing="ij")[1]
        cosd3 = np.cos(x2f3[:, :-1]) - np.cos(x2f3[:, 1:])
        cosd1 = np.cos(f2[None, :, None][:, :-1]) - np.cos(f2[None, :, None][:, 1:])
        sinf = np.sin(f2[None, :, None])[0, :, 0]
        sinc = np.sin(x2c3)[0, :, 0]
        cotc = (1.0 / np.tan(x2c3))[0, :, 0]
        a2 = (f1[1:] ** 2 - f1[:-1] ** 2) / 2.0
        a3 = (f1[1:] ** 3 - f1[:-1] ** 3) / 3.0
        F[_lib.MET_V] = (a3, cosd3[0, :, 0], d3)            # :145
        F[_lib.MET_S1] = (f1 ** 2, cosd1[0, :, 0], d3)       # :148
        F[_lib.MET_S2] = (a2, sinf, d3)                      # :149
        F[_lib.MET_S3] = (a2, d2, np.ones(n3 + 1))           # :150
        F[_lib.MET_L1] = (d1, one[1], one[2])
        F[_lib.MET_L2] = (c1, d2, one[2])                    # :154
        F[_lib.MET_L3] = (c1, sinc, d3)                      # :155
        F[_lib.MET_H1] = (one[0], one[1], one[2])
        F[_lib.MET_H2] = (c1, one[1], one[2])
        F[_lib.MET_H3] = (c1, sinc, one[2])                  # :160
        for m in range(_lib.MET_D1H, _lib.NMET):
            F[m] = (zero1, one[1], one[2])
        F[_lib.MET_D1H2] = (1.0 / c1, one[1], one[2])        # :165
        F[_lib.MET_D1H3] = (1.0 / c1, one[1], one[2])        # :167
        F[_lib.MET_D2H3] = (one[0], cotc, one[2])            # :168
        F[_lib.MET_D1H] = (2.0 / c1, one[1], one[2])         # :169
        F[_lib.MET_D2H] = (one[0], cotc, one[2])             # :170
    else:
        raise ValueError(f"Unknown coordinate system: '{coordinate_system}'. Expected one of: "
                         "cartesian, cylindrical, spherical.")
    return F, (n1, n2, n3), (c1, c2, c3)


def _uniform_pow2(fa, fb, fc):
    """(flag, value) when a*b*c is the same power of two in every cell."""
    vals = []
    for f in (fa, fb, fc):
        if not np.all(f == f[0]):
            return 0, 0.0
        vals.append(float(f[0]))
    v = (vals[0] * vals[1]) * vals[2]
    if all(_pow2(x) for x in vals) and _pow2(v):
        return 1, v
    return 0, 0.0


class GridDesc:
    """Device-resident 1D factors plus the C descriptor for libmfvm."""

    def __init__(self, coordinate_system, grid_config, device):
        F, n, centres = _factors(coordinate_system, grid_config)
        self.coordinate_system = coordinate_system
        self.n = n
        self.device = torch.device(device)
        stride = max(n) + 1
        table = np.zeros((_lib.NMET * 3, stride), dtype=np.float64)
        for m in range(_lib.NMET):
            for a, arr in enumerate(F[m]):
                table[3 * m + a, :len(arr)] = arr
        self.factors_host = F
        self.fac = torch.from_numpy(table.reshape(-1)).to(self.device)
        d = _lib.Grid()
        d.n[:] = list(n)
        d.coord = {"cartesian": _lib.CART, "cylindrical": _lib.CYL, "spherical": _lib.SPH}[coordinate_system]
        d.fac = self.fac.data_ptr()
        d.fstride = stride
        for a in range(3):
            flag, val = _uniform_pow2(*F[_lib.MET_L1 + a])
            d.lpow2[a] = flag
            d.lval[a] = val
        d.vpow2, d.vval = _uniform_pow2(*F[_lib.MET_V])
        uni = coordinate_system == "cartesian" and all(
            np.all(x == x[0]) for m in range(_lib.MET_V, _lib.MET_L3 + 1) for x in F[m])
        d.uniform = int(uni)
        if uni:
            def val(m):
                a, b, c = F[m]
                return (float(a[0]) * float(b[0])) * float(c[0])
            d.lval[:] = [val(_lib.MET_L1), val(_lib.MET_L2), val(_lib.MET_L3)]
            d.vval = val(_lib.MET_V)
            d.sval[:] = [val(_lib.MET_S1), val(_lib.MET_S2), val(_lib.MET_S3)]
        d.smax = self._smax(F)
        for a in range(3):
            fa, fb, fc = F[_lib.MET_L1 + a]
            d.lmin[a] = float(np.min((fa[:, None, None] * fb[None, :, None]) * fc[None, None, :]))
        self.desc = d

    @staticmethod
    def _smax(F):
        """max over cells of 1/L1**2 + 1/L2**2 + 1/L3**2 (navier_stokes.py:575)."""
        L = [F[_lib.MET_L1 + a] for a in range(3)]
        # Evaluate plane by plane in x3 to bound host memory.
        best = -np.inf
        n3 = len(L[0][2])
        for k in range(n3):
            s = None
            for a in range(3):
                fa, fb, fc = L[a]
                Lk = (fa[:, None] * fb[None, :]) * fc[k]
                t = 1 / Lk ** 2
                s = t if s is None else s + t
            best = max(best, float(np.max(s)))
        return best

    def ptr(self):
        return ctypes.byref(self.desc)

    def metric(self, m):
        """Full 3D metric tensor (a[i]*b[j])*c[k] on the device."""
        fa, fb, fc = (torch.from_numpy(np.asarray(x, dtype=np.float64)).to(self.device)
                      for x in self.factors_host[m])
        return (fa[:, None, None] * fb[None, :, None]) * fc[None, None, :]


class _Lazy(dict):
    """dict whose tensor values are produced on first access."""

    def __init__(self, makers):
        super().__init__((k, None) for k in makers)
        self._makers = makers

    def __getitem__(self, key):
        v = dict.__getitem__(self, key)
        if v is None:
            v = self._makers[key]()
            dict.__setitem__(self, key, v)
        return v

    def values(self):
        return [self[k] for k in self]

    def items(self):
        return [(k, self[k]) for k in self]


def build_grid(coordinate_system, grid_config, device="cuda"):
    """Drop-in for geometry.build_grid (geometry.py:14-194).

    Returns the reference dict structure with device tensors, plus
    ``"coordinate_system"`` and ``"mfvm"`` (the :class:`GridDesc` the kernels
    use).  Raises ``ValueError`` for an unknown coordinate system, with the
    reference message.
    """
    g = GridDesc(coordinate_system, grid_config, device)
    dev = g.device
    _, _, centres = _factors(coordinate_system, grid_config)

    def X(a):
        return lambda: torch.from_numpy(np.meshgrid(*centres, indexing="ij")[a]).to(dev)

    def M(m):
        return lambda: g.metric(m)

    def L(d):
        return _Lazy(d)
    normal = {"x1": torch.tensor([1.0, 0.0, 0.0], dtype=torch.float64),
              "x2": torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64),
              "x3": torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64)}
    return {
        "config": grid_config,
        "coordinate_system": coordinate_system,
        "mfvm": g,
        "cell": {"coordinates": _Lazy({"x1": X(0), "x2": X(1), "x3": X(2)}),
                 "normal": normal,
                 "geometry": L({"volume": M(_lib.MET_V),
                                "surface": lambda: L({"x1": M(_lib.MET_S1), "x2": M(_lib.MET_S2),
                                                      "x3": M(_lib.MET_S3)}),
                                "length": lambda: L({"x1": M(_lib.MET_L1), "x2": M(_lib.MET_L2),
                                                     "x3": M(_lib.MET_L3)})})},
        "lame": {"coefficients": L({"x1": M(_lib.MET_H1), "x2": M(_lib.MET_H2),
                                    "x3": M(_lib.MET_H3)}),
                 "derivatives": {
                     "x1": L({"ln_hx1hx2hx3": M(_lib.MET_D1H), "ln_hx2": M(_lib.MET_D1H2),
                              "ln_hx3": M(_lib.MET_D1H3)}),
                     "x2": L({"ln_hx1hx2hx3": M(_lib.MET_D2H), "ln_hx1": M(_lib.MET_D2H1),
                              "ln_hx3": M(_lib.MET_D2H3)}),
                     "x3": L({"ln_hx1hx2hx3": M(_lib.MET_D3H), "ln_hx1": M(_lib.MET_D3H1),
                              "ln_hx2": M(_lib.MET_D3H2)})}},
    }


def build_grid_host(coordinate_system, grid_config):
    """The reference's grid dict (geometry.py:181-192) with NumPy arrays on
    the host, for host-side code written against the reference (initial
    conditions, analysis).  Metrics are (a[i]*b[j])*c[k] over the 1D factors,
    equal to the reference's arrays value for value."""
    F, _, centres = _factors(coordinate_system, grid_config)

    def M(m):
        a, b, c = (np.asarray(x, dtype=np.float64) for x in F[m])
        return (a[:, None, None] * b[None, :, None]) * c[None, None, :]

    X = np.meshgrid(*centres, indexing="ij")
    return {
        "config": grid_config,
        "cell": {"coordinates": {"x1": X[0], "x2": X[1], "x3": X[2]},
                 "normal": {"x1": np.array([1.0, 0.0, 0.0]), "x2": np.array([0.0, 1.0, 0.0]),
                            "x3": np.array([0.0, 0.0, 1.0])},
                 "geometry": {"volume": M(_lib.MET_V),
                              "surface": {"x1": M(_lib.MET_S1), "x2": M(_lib.MET_S2), "x3": M(_lib.MET_S3)},
                              "length": {"x1": M(_lib.MET_L1), "x2": M(_lib.MET_L2), "x3": M(_lib.MET_L3)}}},
        "lame": {"coefficients": {"x1": M(_lib.MET_H1), "x2": M(_lib.MET_H2), "x3": M(_lib.MET_H3)},
                 "derivatives": {
                     "x1": {"ln_hx1hx2hx3": M(_lib.MET_D1H), "ln_hx2": M(_lib.MET_D1H2),
                            "ln_hx3": M(_lib.MET_D1H3)},
                     "x2": {"ln_hx1hx2hx3": M(_lib.MET_D2H), "ln_hx1": M(_lib.MET_D2H1),
                            "ln_hx3": M(_lib.MET_D2H3)},
                     "x3": {"ln_hx1hx2hx3": M(_lib.MET_D3H), "ln_hx1": M(_lib.MET_D3H1),
                            "ln_hx2": M(_lib.MET_D3H2)}}},
    }
