"""mfvm -- MI355X-native finite-volume RHS and time stepping.

Drop-in replacement for the hot path of multiphysics-fvm (``src/physics``,
``src/numerics``, ``src/grid``, ``src/solver.py``): the same module and
function names, operating on device tensors, with every right-hand side,
stage update, ghost fill and time step computed by hand-written HIP kernels
in ``libmfvm.so`` (``include/mfvm.h``).
"""

from . import _lib  # noqa: F401  (fails loudly when libmfvm.so is missing)

__all__ = ["grid", "numerics", "physics", "solver", "state"]
