"""Parameter handling of the drop-ins (host only, no kernel launch).

The reference reads its parameters with ``params[key]`` (a missing key is a
KeyError) and broadcasts array-valued ones against interior arrays
(heat.py:294-297, maxwell.py:71 and :475).  The descriptors must do the same.
"""

import numpy as np
import pytest
import torch


def _state(mod, n=(6, 5, 4)):
    cfg = {"nx1": n[0], "nx2": n[1], "nx3": n[2]}
    return mod.allocate_u(cfg, device="cpu")


def _schemes(rs=True):
    from mfvm.numerics import diffusive_solvers, reconstructors, riemann_solvers, slope_limiters
    return {"reconstructor": reconstructors.muscl, "slope_limiter": slope_limiters.van_leer,
            "riemann_solver": riemann_solvers.rusanov if rs else None,
            "diffusive_solver": diffusive_solvers.centered}


MX = {"magnetic_permeability": 1.0, "electric_permittivity": 1.0, "electrical_conductivity": 0.1,
      "charge_density": 0.0, "magnetic_cleaning_coeff": 1e-3, "electric_cleaning_coeff": 2e-3}


@pytest.mark.parametrize("missing", ["electrical_conductivity", "charge_density", "electric_permittivity"])
def test_maxwell_missing_parameter_raises_keyerror(missing):
    from mfvm.physics import maxwell
    U = _state(maxwell)
    params = {k: v for k, v in MX.items() if k != missing}
    with pytest.raises(KeyError):
        maxwell.make_params(U, params, _schemes())


def test_maxwell_dt_descriptor_needs_no_charge_density():
    """The reference's compute_dt reads the closure's parameters but not the
    charge density (maxwell.py:566-574)."""
    from mfvm.physics import maxwell
    U = _state(maxwell)
    params = {k: v for k, v in MX.items() if k != "charge_density"}
    p = maxwell.make_params(U, params)
    assert p.sigma == 0.1 and not p.rho_field
    with pytest.raises(KeyError):
        maxwell.make_params(U, {k: v for k, v in params.items() if k != "electrical_conductivity"})


def test_heat_missing_source_raises_keyerror():
    from mfvm.physics import heat
    U = _state(heat)
    params = {"mass_density": 1.0, "specific_heat_capacity": 1.0, "thermal_conductivity": 1e-2}
    with pytest.raises(KeyError):
        heat.make_params(U, params, _schemes(False))


@pytest.mark.parametrize("shape", [(6, 5, 4), (6, 1, 1), (1, 5, 1), (5, 4), ()])
def test_array_parameters_broadcast_over_the_interior(shape):
    """A value broadcastable to (nx1, nx2, nx3) becomes the per-cell field in
    the interior SoA layout [nx3][nx2][nx1]; a 0-d value stays a number."""
    from mfvm.physics import _common as c
    from mfvm.physics import heat
    U = _state(heat)
    rng = np.random.default_rng(5)
    q = rng.random(shape) if shape else np.float64(0.25)
    val, field = c.cell_param(U, q)
    if shape == ():
        assert val == 0.25 and field is None
        return
    assert field.shape == (4, 5, 6) and field.dtype == torch.float64 and field.is_contiguous()
    full = np.broadcast_to(q, (6, 5, 4))
    assert np.array_equal(field.numpy(), full.transpose(2, 1, 0))
    p = heat.make_params(U, {"mass_density": 1.0, "specific_heat_capacity": 1.0,
                             "thermal_conductivity": 1e-2, "volumetric_heat_source": q},
                         _schemes(False))
    assert p.source_field


def test_array_parameter_of_wrong_shape_is_rejected():
    from mfvm.physics import _common as c
    from mfvm.physics import maxwell
    U = _state(maxwell)
    with pytest.raises(ValueError):
        c.cell_param(U, np.ones((3, 5, 4)))
