"""PyTorch autograd binding of the HIP structure-factor engine.

``StructureFactors`` is a callable holding a calculator and the fixed inputs (reflections, local
frames, ADP kinds); calling it with parameter tensors returns F(h) as a complex128 tensor whose
backward pass runs the engine's fused adjoint kernel with dT/dF = the incoming gradient.  PyTorch
passes dL/dA + i dL/dB for a complex output, which is exactly the reference's dT/dF convention
(HansenCoppensStructureFactorCalculator.h:141-143), so no conversion is needed.

Parameters (all float64, any device; results follow the device of ``xyz``):
  xyz (n, 3) fractional, adp (n, 6) U_cif (U_iso in column 0 for isotropic atoms), occ (n,),
  plm (n_types, 25) local-frame P_lm in the l*l + l + m layout, pval, kappa, kappa_prime (n_types,).
"""
from __future__ import annotations

import numpy as np
import torch

from .calculator import HansenCoppensStructureFactorCalculator
from .model import AtomSet, DerivativesSelector


def _np(t):
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


class _SF(torch.autograd.Function):
    @staticmethod
    def forward(ctx, owner, xyz, adp, occ, plm, pval, kappa, kappa_prime):
        owner._set(xyz, adp, occ, plm, pval, kappa, kappa_prime)
        f = owner.calc.calculateStructureFactors(owner._atoms, owner.lcs, owner.hkl, owner.include)
        ctx.owner = owner
        ctx.save_for_backward(xyz, adp, occ, plm, pval, kappa, kappa_prime)
        return torch.from_numpy(f).to(xyz.device)

    @staticmethod
    def backward(ctx, grad_f):
        owner = ctx.owner
        xyz, adp, occ, plm, pval, kappa, kappa_prime = ctx.saved_tensors
        owner._set(xyz, adp, occ, plm, pval, kappa, kappa_prime)
        g = grad_f.detach().to("cpu", torch.complex128).numpy()
        want_mp = any(ctx.needs_input_grad[4:8])
        sel = DerivativesSelector(True, True, True, False, want_mp)
        _, d = owner.calc.calculateStructureFactorsAndDerivatives(owner._atoms, owner.lcs, owner.hkl, g,
                                                                  owner.include, sel)
        dev = xyz.device
        out = [None,
               torch.from_numpy(d.atomic_position_derivatives.copy()).to(dev),
               torch.from_numpy(d.adp_derivatives.copy()).to(dev),
               torch.from_numpy(d.occupancy_derivatives.copy()).to(dev)]
        if want_mp:
            td = owner.calc.typeDerivatives()
            out += [torch.from_numpy(td.p_lm.copy()).to(dev), torch.from_numpy(td.p_val.copy()).to(dev),
                    torch.from_numpy(td.kappa.copy()).to(dev), torch.from_numpy(td.kappa_prime.copy()).to(dev)]
        else:
            out += [None, None, None, None]
        return tuple(out)


class StructureFactors:
    """F = StructureFactors(crystal, params, lcs, hkl)(xyz, adp, occ, plm, pval, kappa, kappa_prime)"""

    def __init__(self, crystal, params, lcs, hkl, mode="HIP", device=0, include=None):
        self.calc = HansenCoppensStructureFactorCalculator(crystal, params, device=device)
        self.calc.setCalculationMode(mode)
        self.lcs = np.asarray(lcs, dtype=np.float64)
        self.hkl = np.ascontiguousarray(np.asarray(hkl, dtype=np.int32).reshape(-1, 3))
        self.include = include
        a = crystal.atoms
        self._atoms = AtomSet(a.xyz.copy(), a.adp.copy(), a.adp_size.copy(), a.occupancy.copy(),
                              a.multiplicity.copy())
        self._types_key = None

    def initial(self, params, atoms=None, device="cpu"):
        """Parameter tensors (requires_grad) initialised from a model and an atom set."""
        a = atoms or self._atoms
        t = lambda v: torch.tensor(np.asarray(v, dtype=np.float64), device=device, requires_grad=True)
        tp = params.type_parameters
        return (t(a.xyz), t(a.adp), t(a.occupancy), t(np.array([x.plm25() for x in tp])),
                t([x.p_val for x in tp]), t([x.kappa_spherical_valence for x in tp]),
                t([x.kappa_deformation_valence for x in tp]))

    def _set(self, xyz, adp, occ, plm, pval, kappa, kappa_prime):
        self._atoms.xyz = np.ascontiguousarray(_np(xyz))
        self._atoms.adp = np.ascontiguousarray(_np(adp))
        self._atoms.occupancy = np.ascontiguousarray(_np(occ))
        types = tuple(_np(v).tobytes() for v in (plm, pval, kappa, kappa_prime))
        if types != self._types_key:
            self.calc.setTypeParameters(_np(plm), _np(pval), _np(kappa), _np(kappa_prime))
            self._types_key = types

    def __call__(self, xyz, adp, occ, plm, pval, kappa, kappa_prime):
        return _SF.apply(self, xyz, adp, occ, plm, pval, kappa, kappa_prime)
