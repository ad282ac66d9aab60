"""Host side of the differentiable energy (no GPU): the per-type water
parameterisation expands to exactly the per-particle arrays of the force
(traceless quadrupole and octopole from their independent components), the
finite-difference fixture's parameterisation is the same map, and autograd
sums per-particle gradients into the per-type ones as the chain rule says."""
import numpy as np
import torch

from mpidopenmmplugin_amd.autograd import (expand_water_parameters, octopole_from_free, quadrupole_from_free,
                                           water_type_parameters)
from mpidopenmmplugin_amd.waterbox import make_waterbox
from tests.golden.make_adjoint_fixture import apply_types, type_vector


def test_expansion_reproduces_the_force():
    force, pos, box = make_waterbox("water216x")
    flat = force.flatten()
    p = water_type_parameters(force, dtype=torch.float64)
    arrays = expand_water_parameters(p, flat["n"] // 3)
    for name, t in zip(("charges", "dipoles", "quadrupoles", "octopoles", "polarity", "tholes"), arrays):
        np.testing.assert_allclose(t.detach().numpy().reshape(-1), flat[name], rtol=0, atol=1e-15, err_msg=name)
    assert float(p["default_thole"].detach()) == flat["default_thole"]


def test_fixture_parameterisation_matches():
    force, pos, box = make_waterbox("water216x")
    flat = force.flatten()
    v = type_vector(flat)
    assert v.size == 23
    p = water_type_parameters(force, dtype=torch.float64)
    order = ["q_O", "q_H", "d_O", "Q_O", "O_O", "alpha_O", "thole_O", "thole_H", "default_thole"]
    v2 = np.concatenate([np.atleast_1d(p[k].detach().numpy()) for k in order])
    np.testing.assert_array_equal(v, v2)
    rng = np.random.default_rng(3)
    w = v * (1.0 + 0.1 * rng.normal(size=v.size))
    f2 = apply_types(flat, w)
    p2 = {k: torch.tensor(np.asarray(val)) for k, val in zip(order, np.split(w, np.cumsum([1, 1, 3, 5, 7, 3, 1, 1])))}
    for k in ("q_O", "q_H", "thole_O", "thole_H", "default_thole"):
        p2[k] = p2[k].reshape(())
    arrays = expand_water_parameters(p2, flat["n"] // 3)
    for name, t in zip(("charges", "dipoles", "quadrupoles", "octopoles", "polarity", "tholes"), arrays):
        np.testing.assert_allclose(t.numpy().reshape(-1), f2[name], rtol=1e-15, atol=1e-18, err_msg=name)


def test_traceless_components():
    q = quadrupole_from_free(torch.tensor([1.0, 2.0, 3.0, 4.0, 5.0]))
    assert float(q[0] + q[2] + q[5]) == 0.0
    o = octopole_from_free(torch.arange(1.0, 8.0))
    # XXX + XYY + XZZ, XXY + YYY + YZZ, XXZ + YYZ + ZZZ (API order)
    assert float(o[0] + o[2] + o[7]) == 0.0
    assert float(o[1] + o[3] + o[8]) == 0.0
    assert float(o[4] + o[6] + o[9]) == 0.0


def test_chain_rule_of_the_expansion():
    """d(sum_i w_i . x_i)/d(type parameter) for a random per-particle weight
    w equals the sum of w over the sites the parameter reaches."""
    force, pos, box = make_waterbox("water216")
    n = force.getNumMultipoles()
    p = water_type_parameters(force, dtype=torch.float64)
    charges, dipoles, quads, octs, pol, tholes = expand_water_parameters(p, n // 3)
    rng = np.random.default_rng(1)
    wq = torch.tensor(rng.normal(size=n))
    wd = torch.tensor(rng.normal(size=(n, 3)))
    wQ = torch.tensor(rng.normal(size=(n, 6)))
    (charges * wq).sum().add((dipoles * wd).sum()).add((quads * wQ).sum()).backward()
    assert np.isclose(float(p["q_O"].grad), wq[0::3].sum().item())
    assert np.isclose(float(p["q_H"].grad), (wq[1::3].sum() + wq[2::3].sum()).item())
    np.testing.assert_allclose(p["d_O"].grad.numpy(), wd[0::3].sum(0).numpy())
    wO = wQ[0::3].sum(0).numpy()  # XX XY YY XZ YZ ZZ
    np.testing.assert_allclose(p["Q_O"].grad.numpy(), [wO[0] - wO[5], wO[1], wO[2] - wO[5], wO[3], wO[4]])
