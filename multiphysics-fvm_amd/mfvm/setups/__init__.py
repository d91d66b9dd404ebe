"""Setups of the benchmark configurations, in the reference's setup-dict
format (``setups/kelvin_helmholtz.py:90-163``): ``get_setup()`` returns the
dict that :func:`mfvm.solver.run` consumes.  Initializers are host NumPy code
(run by :func:`mfvm.solver.apply_ics` on a host mirror of U)."""

NAMES = ("kelvin_helmholtz_128", "rayleigh_taylor_256", "ensemble_kh64",
         "electromagnetic_pulse_spherical", "heated_disc_cylindrical", "kelvin_helmholtz_closure")
