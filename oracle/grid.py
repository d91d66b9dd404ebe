"""Oracle restatement of the reference grid and boundary conditions.

Test infrastructure only (see ``oracle/__init__.py``).

Covers ``src/grid/geometry.py`` (``build_grid``), ``src/grid/
boundary_conditions.py`` (``neumann``, ``periodic``, ``reflecting``) and
``src/solver.py:13-61`` (``apply_bcs`` with ghost edge/corner averaging).
"""

import numpy as np

SIDES = ("left", "right", "bottom", "top", "front", "back")
_SIDE_ERR = ("Unknown boundary side: '{}'. Expected one of: left, right, bottom, "
             "top, front, back.")


# ----------------------------------------------------------------------------
# Geometry (src/grid/geometry.py:14-194)
# ----------------------------------------------------------------------------

def build_grid(coordinate_system, grid_config):
    """geometry.py:14-194 -- cell volumes, face areas, lengths, Lame terms.

    Every quantity is a broadcast product of 1D factors; the products are
    formed in the reference order so that values match to the bit.
    """
    lo = [grid_config[f"x{a}_min"] for a in (1, 2, 3)]
    hi = [grid_config[f"x{a}_max"] for a in (1, 2, 3)]
    n = [grid_config[f"nx{a}"] for a in (1, 2, 3)]
    h = [(hi[a] - lo[a]) / n[a] for a in range(3)]                      # :37-39
    centre = [lo[a] + (np.arange(n[a]) + 0.5) * h[a] for a in range(3)]  # :42-44
    face = [lo[a] + np.arange(n[a] + 1) * h[a] for a in range(3)]        # :47-49
    x1, x2, x3 = np.meshgrid(*centre, indexing="ij")                     # :52

    def b(arr, axis):  # 1D -> broadcastable 3D (geometry.py:63-65)
        shape = [1, 1, 1]
        shape[axis] = arr.shape[0]
        return arr.reshape(shape)

    shp = tuple(n)
    fshape = [tuple(n[a] + (1 if a == ax else 0) for a in range(3)) for ax in range(3)]
    # 3D face coordinate arrays (geometry.py:76-78); multiplying by ones is exact.
    f3 = [b(face[a], a) * np.ones(fshape[a]) for a in range(3)]
    dface3 = [np.diff(f3[a], axis=a) for a in range(3)]  # x_f[p] - x_f[m], :83
    dface1 = [b(face[a][1:] - face[a][:-1], a) for a in range(3)]
    zero = np.full(shp, 0.0)
    one = np.full(shp, 1.0)

    if coordinate_system == "cartesian":                                # :80-109
        V = dface3[0] * dface3[1] * dface3[2]
        S1 = np.full(fshape[0], 1.0) * dface1[1] * dface1[2]
        S2 = dface1[0] * np.full(fshape[1], 1.0) * dface1[2]
        S3 = dface1[0] * dface1[1] * np.full(fshape[2], 1.0)
        L1, L2, L3 = dface3[0], dface3[1], dface3[2]
        hx = [one, np.full(shp, 1.0), np.full(shp, 1.0)]
        d = dict(d2h1=zero, d3h1=zero.copy(), d1h2=zero.copy(), d3h2=zero.copy(),
                 d1h3=zero.copy(), d2h3=zero.copy(), d1h=zero.copy(),
                 d2h=zero.copy(), d3h=zero.copy())
    elif coordinate_system == "cylindrical":                            # :111-140
        V = ((f3[0][1:] ** 2 - f3[0][:-1] ** 2) / 2.0) * dface3[1] * dface3[2]
        S1 = b(face[0], 0) * dface1[1] * dface1[2]
        S2 = dface1[0] * np.full(fshape[1], 1.0) * dface1[2]
        S3 = ((b(face[0], 0)[1:] ** 2 - b(face[0], 0)[:-1] ** 2) / 2.0) * dface1[1] \
            * np.full(fshape[2], 1.0)
        L1 = dface3[0]
        L2 = x1 * dface3[1]
        L3 = dface3[2]
        hx = [np.full(shp, 1.0), x1, np.full(shp, 1.0)]
        d = dict(d2h1=zero, d3h1=zero.copy(), d1h2=1.0 / x1, d3h2=zero.copy(),
                 d1h3=zero.copy(), d2h3=zero.copy(), d1h=1.0 / x1,
                 d2h=zero.copy(), d3h=zero.copy())
    elif coordinate_system == "spherical":                              # :142-171
        V = ((f3[0][1:] ** 3 - f3[0][:-1] ** 3) / 3.0) \
            * (np.cos(f3[1][:, :-1]) - np.cos(f3[1][:, 1:])) * dface3[2]
        S1 = (b(face[0], 0) ** 2) * (np.cos(b(face[1], 1)[:, :-1])
                                     - np.cos(b(face[1], 1)[:, 1:])) * dface1[2]
        S2 = ((b(face[0], 0)[1:] ** 2 - b(face[0], 0)[:-1] ** 2) / 2.0) \
            * np.sin(b(face[1], 1)) * dface1[2]
        S3 = ((b(face[0], 0)[1:] ** 2 - b(face[0], 0)[:-1] ** 2) / 2.0) * dface1[1] \
            * np.full(fshape[2], 1.0)
        L1 = dface3[0]
        L2 = x1 * dface3[1]
        L3 = x1 * np.sin(x2) * dface3[2]
        hx = [np.full(shp, 1.0), x1, x1 * np.sin(x2)]
        d = dict(d2h1=zero, d3h1=zero.copy(), d1h2=1.0 / x1, d3h2=zero.copy(),
                 d1h3=1.0 / x1, d2h3=1.0 / np.tan(x2), d1h=2.0 / x1,
                 d2h=1.0 / np.tan(x2), d3h=zero.copy())
    else:
        raise ValueError(f"Unknown coordinate system: '{coordinate_system}'. Expected "
                         "one of: cartesian, cylindrical, spherical.")

    normal = {"x1": np.array([1.0, 0.0, 0.0]), "x2": np.array([0.0, 1.0, 0.0]),
              "x3": np.array([0.0, 0.0, 1.0])}
    return {
        "config": grid_config,
        "cell": {"coordinates": {"x1": x1, "x2": x2, "x3": x3},
                 "normal": normal,
                 "geometry": {"volume": V,
                              "surface": {"x1": S1, "x2": S2, "x3": S3},
                              "length": {"x1": L1, "x2": L2, "x3": L3}}},
        "lame": {"coefficients": {"x1": hx[0], "x2": hx[1], "x3": hx[2]},
                 "derivatives": {
                     "x1": {"ln_hx1hx2hx3": d["d1h"], "ln_hx2": d["d1h2"], "ln_hx3": d["d1h3"]},
                     "x2": {"ln_hx1hx2hx3": d["d2h"], "ln_hx1": d["d2h1"], "ln_hx3": d["d2h3"]},
                     "x3": {"ln_hx1hx2hx3": d["d3h"], "ln_hx1": d["d3h1"], "ln_hx2": d["d3h2"]}}},
    }


# ----------------------------------------------------------------------------
# Boundary conditions (src/grid/boundary_conditions.py)
# ----------------------------------------------------------------------------

# side -> (axis, ghost index, adjacent interior index, opposite interior index)
_SIDE_GEOM = {"left": (0, 0, 1, -2), "right": (0, -1, -2, 1),
              "bottom": (1, 0, 1, -2), "top": (1, -1, -2, 1),
              "front": (2, 0, 1, -2), "back": (2, -1, -2, 1)}


def _face(axis, index):
    idx = [slice(1, -1)] * 3
    idx[axis] = index
    return tuple(idx)


def _side(side):
    if side not in _SIDE_GEOM:
        raise ValueError(_SIDE_ERR.format(side))
    return _SIDE_GEOM[side]


def neumann(field, side):
    """boundary_conditions.py:13-37 -- ghost = adjacent interior cell."""
    ax, g, adj, _ = _side(side)
    field[_face(ax, g)] = field[_face(ax, adj)]


def periodic(field, side):
    """boundary_conditions.py:40-64 -- ghost = opposite interior cell."""
    ax, g, _, opp = _side(side)
    field[_face(ax, g)] = field[_face(ax, opp)]


def reflecting(field, side):
    """boundary_conditions.py:67-91 -- ghost = minus the adjacent interior cell."""
    ax, g, adj, _ = _side(side)
    field[_face(ax, g)] = -field[_face(ax, adj)]


def apply_bcs(U, bcs):
    """src/solver.py:13-61 -- six faces, then 12 edge and 8 corner averages."""
    for name in bcs:
        f = U[name]
        for side in SIDES:
            bcs[name][side](f, side)
        # Edges: mean of the two face ghosts that meet there (solver.py:40-51).
        for a in (0, -1):
            an = 1 if a == 0 else -2
            for c in (0, -1):
                cn = 1 if c == 0 else -2
                f[a, c, 1:-1] = 0.5 * (f[an, c, 1:-1] + f[a, cn, 1:-1])
        for a in (0, -1):
            an = 1 if a == 0 else -2
            for c in (0, -1):
                cn = 1 if c == 0 else -2
                f[a, 1:-1, c] = 0.5 * (f[an, 1:-1, c] + f[a, 1:-1, cn])
        for a in (0, -1):
            an = 1 if a == 0 else -2
            for c in (0, -1):
                cn = 1 if c == 0 else -2
                f[1:-1, a, c] = 0.5 * (f[1:-1, an, c] + f[1:-1, a, cn])
        # Corners: mean of the three adjacent edge ghosts (solver.py:54-61).
        for a in (0, -1):
            an = 1 if a == 0 else -2
            for b_ in (0, -1):
                bn = 1 if b_ == 0 else -2
                for c in (0, -1):
                    cn = 1 if c == 0 else -2
                    f[a, b_, c] = (f[an, b_, c] + f[a, bn, c] + f[a, b_, cn]) / 3.0
