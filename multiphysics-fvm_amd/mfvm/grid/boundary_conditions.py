"""Boundary conditions (drop-in for src/grid/boundary_conditions.py:13-91).

Each function applies its rule to one face of one ghosted field and is
tagged with the code the ghost-fill kernel uses; ``mfvm.solver.apply_bcs``
fills all fields of a :class:`mfvm.state.State` in one kernel launch.
"""

from .. import _lib

_SIDES = {"left": (0, 0, 1, -2), "right": (0, -1, -2, 1), "bottom": (1, 0, 1, -2),
          "top": (1, -1, -2, 1), "front": (2, 0, 1, -2), "back": (2, -1, -2, 1)}
SIDES = ("left", "right", "bottom", "top", "front", "back")


def _face(axis, index):
    idx = [slice(1, -1)] * 3
    idx[axis] = index
    return tuple(idx)


def _geom(side):
    if side not in _SIDES:
        raise ValueError(f"Unknown boundary side: '{side}'. Expected one of: left, right, bottom, "
                         "top, front, back.")
    return _SIDES[side]


def neumann(field, side):
    """Zero normal gradient: ghost = adjacent interior cell (boundary_conditions.py:13-37)."""
    ax, g, adj, _ = _geom(side)
    field[_face(ax, g)] = field[_face(ax, adj)]


def periodic(field, side):
    """Ghost = opposite interior cell (boundary_conditions.py:40-64)."""
    ax, g, _, opp = _geom(side)
    field[_face(ax, g)] = field[_face(ax, opp)]


def reflecting(field, side):
    """Ghost = minus the adjacent interior cell (boundary_conditions.py:67-91)."""
    ax, g, adj, _ = _geom(side)
    field[_face(ax, g)] = -field[_face(ax, adj)]


neumann._mfvm_bc = _lib.BC_NEUMANN
periodic._mfvm_bc = _lib.BC_PERIODIC
reflecting._mfvm_bc = _lib.BC_REFLECTING


def bc_codes(names, bcs):
    """[nvar][6] table of kernel codes for bcs[var][side]; NotImplementedError for
    a callable without a kernel rule."""
    out = []
    for name in names:
        row = []
        for side in SIDES:
            fn = bcs[name][side]
            code = getattr(fn, "_mfvm_bc", None)
            if code is None:
                raise NotImplementedError(f"boundary condition {fn!r} has no kernel rule")
            row.append(code)
        out.append(row)
    return out
