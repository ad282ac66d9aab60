"""The MPIDForce energy as a differentiable PyTorch function.

``MPIDEnergy`` evaluates the energy of an ``MPIDForce`` with the gfx950
library as a function of the positions and the per-particle parameters
(charges, molecular-frame dipoles, quadrupoles, octopoles, polarisabilities,
Thole factors) and the default Thole width, and returns a float64 scalar
tensor.  When any input requires a gradient, the evaluation goes through
``mpid_energy_gradients``: the forward solve, one adjoint solve with the
solver's symmetric operator, and the gradient kernels (pair, reciprocal and
per-atom) in one call; ``backward`` then scales the stored gradients.
Without gradients it is a plain ``execute``.

The parameter layouts are those of ``MPIDForce.flatten()`` (the reference's
``MPIDForce::addMultipole`` orderings, MPIDForce.h:236-281), one row per
particle: charges [n], dipoles [n,3], quadrupoles [n,6] (XX XY YY XZ YZ ZZ
order of the API), octopoles [n,10], polarity [n,3], tholes [n].  Frames,
covalent maps and the Ewald settings stay those of the force.

``water_type_parameters`` / ``expand_water_parameters`` give the per-type
parameterisation of the synthetic water boxes (two charges; the O dipole,
its 5 independent traceless quadrupole and 7 independent traceless octopole
components; 3 O polarisabilities; the Thole factors and the default width),
expanded to per-particle tensors in torch, so that autograd sums the
per-particle gradients into the per-type ones.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from .kernel import HipCalcMPIDForceKernel
from .mpidforce import MPIDForce

PARAMETERS = ("charges", "dipoles", "quadrupoles", "octopoles", "polarity", "tholes")


class MPIDEnergy:
    """Differentiable energy of one MPIDForce (mutual or direct polarization).

    precision: "mixed" or "double".  all_sites: also form the gradients of
    the higher moments and polarisabilities of sites that are bare charges
    (every pair then takes the general formula; slower)."""

    def __init__(self, force: MPIDForce, box, precision="mixed", all_sites=False):
        self.force = copy.deepcopy(force)
        self.flat = self.force.flatten()
        self.box = box
        self.precision = precision
        self.all_sites = all_sites
        self._kernel = None
        self._width = None
        self.last = None  # the last gradient record (numpy), for inspection

    def kernel(self, default_thole):
        """The context for this default Thole width (a global setting of the
        force: a new width creates a new context)."""
        if self._kernel is None or default_thole != self._width:
            self.force.setDefaultTholeWidth(float(default_thole))
            k = HipCalcMPIDForceKernel(self.precision)
            k.initialize(self.force.getNumMultipoles(), self.force, default_box=self.box)
            self._kernel, self._width = k, float(default_thole)
            self.flat = self.force.flatten()
        return self._kernel

    def __call__(self, positions, charges, dipoles, quadrupoles, octopoles, polarity, tholes, default_thole):
        dt = default_thole if torch.is_tensor(default_thole) else torch.tensor(float(default_thole),
                                                                               dtype=torch.float64)
        return _MPIDEnergyFunction.apply(self, positions, charges, dipoles, quadrupoles, octopoles, polarity, tholes,
                                         dt)


class _MPIDEnergyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, positions, charges, dipoles, quadrupoles, octopoles, polarity, tholes, default_thole):
        tensors = (charges, dipoles, quadrupoles, octopoles, polarity, tholes)
        kernel = model.kernel(float(default_thole))
        flat = dict(model.flat)
        for name, t in zip(PARAMETERS, tensors):
            flat[name] = np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy().reshape(-1))
        kernel.updateParameterArrays(flat)
        pos = positions.detach().to("cpu", torch.float64).numpy().reshape(-1, 3)
        inputs = (positions,) + tensors + (default_thole,)
        if any(ctx.needs_input_grad[1:]):
            g = kernel.energyGradients(pos, model.box, all_sites=model.all_sites)
            model.last = g
            names = ("positions",) + PARAMETERS
            grads = [torch.from_numpy(np.asarray(g[k], dtype=np.float64)).reshape(t.shape).to(t.device, t.dtype)
                     for k, t in zip(names, inputs[:-1])]
            grads.append(torch.tensor(g["default_thole"], dtype=default_thole.dtype, device=default_thole.device))
            ctx.grads = grads
            energy = g["energy"]
        else:
            energy, _ = kernel.execute(pos, model.box)
            ctx.grads = None
        return torch.tensor(energy, dtype=torch.float64, device=positions.device)

    @staticmethod
    def backward(ctx, grad_energy):
        out = [None]
        for need, g in zip(ctx.needs_input_grad[1:], ctx.grads):
            out.append(g * grad_energy.to(g.dtype) if need else None)
        return tuple(out)


# ---- per-type parameterisation of the synthetic water boxes -------------------------
# Independent components of a traceless quadrupole (API order XX XY YY XZ YZ
# ZZ: ZZ = -(XX + YY)) and octopole (API order XXX XXY XYY YYY XXZ XYZ YYZ
# XZZ YZZ ZZZ; traceless in every index pair: XXX = -(XYY + XZZ),
# YYY = -(XXY + YZZ), ZZZ = -(XXZ + YYZ)).
QUAD_FREE = (0, 1, 2, 3, 4)          # XX XY YY XZ YZ
OCT_FREE = (1, 2, 4, 5, 6, 7, 8)     # XXY XYY XXZ XYZ YYZ XZZ YZZ


def quadrupole_from_free(q5):
    xx, xy, yy, xz, yz = (q5[..., k] for k in range(5))
    return torch.stack([xx, xy, yy, xz, yz, -(xx + yy)], dim=-1)


def octopole_from_free(o7):
    xxy, xyy, xxz, xyz, yyz, xzz, yzz = (o7[..., k] for k in range(7))
    return torch.stack([-(xyy + xzz), xxy, xyy, -(xxy + yzz), xxz, xyz, yyz, xzz, yzz, -(xxz + yyz)], dim=-1)


def water_type_parameters(force: MPIDForce, dtype=torch.float32):
    """The per-type parameters of a synthetic water box (atoms O, H, H per
    molecule), as leaf tensors: q_O, q_H, d_O [3], Q_O [5], O_O [7],
    alpha_O [3], thole_O, thole_H, default_thole."""
    flat = force.flatten()
    q = flat["quadrupoles"].reshape(-1, 6)[0]
    o = flat["octopoles"].reshape(-1, 10)[0]

    def leaf(v):
        return torch.tensor(np.asarray(v, dtype=np.float64), dtype=dtype).requires_grad_(True)

    return dict(q_O=leaf(flat["charges"][0]), q_H=leaf(flat["charges"][1]),
                d_O=leaf(flat["dipoles"].reshape(-1, 3)[0]), Q_O=leaf(q[list(QUAD_FREE)]),
                O_O=leaf(o[list(OCT_FREE)]), alpha_O=leaf(flat["polarity"].reshape(-1, 3)[0]),
                thole_O=leaf(flat["tholes"][0]), thole_H=leaf(flat["tholes"][1]),
                default_thole=leaf(flat["default_thole"]))


def expand_water_parameters(p, nwaters):
    """Per-particle tensors (charges, dipoles, quadrupoles, octopoles,
    polarity, tholes) of nwaters molecules O, H, H from the per-type
    parameters; hydrogens carry a charge and a Thole factor only."""
    z = lambda k: torch.zeros(k, dtype=p["q_O"].dtype)
    def per_site(o_val, h_val):
        return torch.stack([o_val, h_val, h_val], dim=0).repeat((nwaters,) + (1,) * (o_val.dim()))
    charges = per_site(p["q_O"], p["q_H"])
    dipoles = per_site(p["d_O"], z(3))
    quads = per_site(quadrupole_from_free(p["Q_O"]), z(6))
    octs = per_site(octopole_from_free(p["O_O"]), z(10))
    pol = per_site(p["alpha_O"], z(3))
    tholes = per_site(p["thole_O"], p["thole_H"])
    return charges, dipoles, quads, octs, pol, tholes
