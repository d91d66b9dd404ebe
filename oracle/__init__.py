"""CPU oracle for the finite-volume right-hand-side and time-step path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports, calls or
links this code; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and there only as the checker
(or as the timed CPU baseline).

It is a NumPy restatement of the reference algorithm (multiphysics-fvm,
``src/``): every function keeps the reference's whole-array, unfused
structure and -- what matters for parity -- the exact floating-point
operation order, so that an IEEE-faithful GPU kernel can agree with it to the
last bit.  Each function cites the reference file:line it restates.

How each part is pinned (beyond retyping the formulas):

* limiters, Riemann solvers, diffusive solver -- hand-computed known answers
  and closed forms; boundary conditions and ghost edge/corner averaging --
  brute force on tiny arrays; grid metrics -- volume sums and the
  unknown-system error; integrators -- amplification factors with a mock
  linear right-hand side; uniform-state, advection and closed-form dt checks
  of the physics modules (``tests/test_oracle_closed_forms.py``);
* the composed right-hand sides, time steps, grids, ghost fills and full
  steps of the three physics modules -- golden vectors written by the
  reference itself (``tests/golden/``, checked in
  ``tests/test_oracle_golden.py``);
* the learned flux-correction network (``cnn.py``, no reference
  counterpart) -- torch CPU autograd in float64 against central finite
  differences, and the periodic convolution against an explicit loop
  (``tests/test_oracle_cnn.py``).

Parity unpinned: nothing in this package is left without one of the checks
above; the spherical/cylindrical metric factors use ``np.sin``/``np.cos``/
``np.tan`` whose last-ulp results depend on the host libm, so the GPU path
consumes these factors as computed on the host and never evaluates trig.
"""
