"""PyTorch autograd binding of the HIP structure-factor engine.

``StructureFactors`` is a callable holding a calculator and the fixed inputs (reflections, local
frames, ADP kinds); calling it with parameter tensors returns F(h) as a complex128 tensor whose
backward pass runs the engine's fused adjoint kernel with dT/dF = the incoming gradient.  PyTorch
passes dL/dA + i dL/dB for a complex output, which is exactly the reference's dT/dF convention
(HansenCoppensStructureFactorCalculator.h:141-143), so no conversion is needed.

Parameters (all float64; results follow the device of ``xyz``):
  xyz (n, 3) fractional, adp (n, 6) U_cif (U_iso in column 0 for isotropic atoms), occ (n,),
  plm (n_types, 25) local-frame P_lm in the l*l + l + m layout, pval, kappa, kappa_prime (n_types,).

With xyz on the calculator's GPU the forward and the backward run as device-resident calls on
torch's current stream, without host copies or host synchronisation: dh_hc_sf_and_derivs_dev for
the structural (xyz / ADP / occupancy) gradients, dh_hc_sf_and_derivs_mp_dev when the multipole
parameters (P_lm, Pval, kappa, kappa') want gradients too (per-type sums formed on the device).
With xyz on the host both take the host call.  Type parameters are sent to the engine only when
their tensors change (tracked by tensor version counters, so an unchanged model costs no
device-to-host copy).
"""
from __future__ import annotations

import numpy as np
import torch

from .calculator import HansenCoppensStructureFactorCalculator
from .model import AtomSet, DerivativesSelector


def _np(t):
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


def _key(*ts):
    return tuple((t.data_ptr(), t._version, tuple(t.shape), t.device.type) for t in ts)


class _SF(torch.autograd.Function):
    @staticmethod
    def forward(ctx, owner, xyz, adp, occ, plm, pval, kappa, kappa_prime):
        owner._set_types(plm, pval, kappa, kappa_prime)
        ctx.owner = owner
        ctx.save_for_backward(xyz, adp, occ, plm, pval, kappa, kappa_prime)
        if owner._on_device(xyz):
            owner._ensure_device_layout()
            return owner.calc.calculateOnDevice(xyz.detach().contiguous(), owner._adp_size_dev, adp.detach().contiguous(),
                                                occ.detach().contiguous(), None, owner._lcs_dev,
                                                countAtomContribution=owner._include_dev)[0]
        owner._set_atoms(xyz, adp, occ)
        f = owner.calc.calculateStructureFactors(owner._atoms, owner.lcs, owner.hkl, owner.include)
        owner._dev_ready = False
        return torch.from_numpy(f).to(xyz.device)

    @staticmethod
    def backward(ctx, grad_f):
        owner = ctx.owner
        xyz, adp, occ, plm, pval, kappa, kappa_prime = ctx.saved_tensors
        owner._set_types(plm, pval, kappa, kappa_prime)
        want_mp = any(ctx.needs_input_grad[4:8])
        dev = xyz.device
        if owner._on_device(xyz):
            owner._ensure_device_layout()
            g = grad_f.detach().to(torch.complex128).contiguous()
            res = owner.calc.calculateOnDevice(
                xyz.detach().contiguous(), owner._adp_size_dev, adp.detach().contiguous(), occ.detach().contiguous(), g,
                owner._lcs_dev, countAtomContribution=owner._include_dev,
                derivativesSelector=DerivativesSelector(True, True, True, False, want_mp))
            out = [None, res[1], res[2], res[3]]
            if want_mp:
                td = res[4]
                out += [td.p_lm.to(plm.device), td.p_val.to(pval.device), td.kappa.to(kappa.device),
                        td.kappa_prime.to(kappa_prime.device)]
            else:
                out += [None, None, None, None]
            return tuple(out)
        owner._set_atoms(xyz, adp, occ)
        g = grad_f.detach().to("cpu", torch.complex128).numpy()
        sel = DerivativesSelector(True, True, True, False, want_mp)
        _, d = owner.calc.calculateStructureFactorsAndDerivatives(owner._atoms, owner.lcs, owner.hkl, g,
                                                                  owner.include, sel)
        owner._dev_ready = False
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
        self.device = device
        self.lcs = np.asarray(lcs, dtype=np.float64)
        self.hkl = np.ascontiguousarray(np.asarray(hkl, dtype=np.int32).reshape(-1, 3))
        self.include = include
        a = crystal.atoms
        self._atoms = AtomSet(a.xyz.copy(), a.adp.copy(), a.adp_size.copy(), a.occupancy.copy(),
                              a.multiplicity.copy())
        self._types_key = None
        self._dev_ready = False
        self._adp_size_dev = self._lcs_dev = self._include_dev = None

    def _on_device(self, t):
        return t.is_cuda and t.device.index in (None, self.device)

    def _ensure_device_layout(self):
        """Reflections and the fixed per-atom arrays on the GPU (once; again after a host call)."""
        if self._adp_size_dev is None:
            dev = torch.device("cuda", self.device)
            self._adp_size_dev = torch.tensor(self._atoms.adp_size, dtype=torch.int32, device=dev)
            self._lcs_dev = torch.tensor(self.lcs.reshape(-1, 3, 3), dtype=torch.float64, device=dev)
            if self.include is not None:
                self._include_dev = torch.tensor(np.asarray(self.include, dtype=np.uint8), device=dev)
        if not self._dev_ready:
            self.calc.setReflections(self.hkl)
            self._dev_ready = True

    def initial(self, params, atoms=None, device="cpu"):
        """Parameter tensors (requires_grad) initialised from a model and an atom set."""
        a = atoms or self._atoms
        t = lambda v: torch.tensor(np.asarray(v, dtype=np.float64), device=device, requires_grad=True)
        tp = params.type_parameters
        return (t(a.xyz), t(a.adp), t(a.occupancy), t(np.array([x.plm25() for x in tp])),
                t([x.p_val for x in tp]), t([x.kappa_spherical_valence for x in tp]),
                t([x.kappa_deformation_valence for x in tp]))

    def _set_atoms(self, xyz, adp, occ):
        self._atoms.xyz = np.ascontiguousarray(_np(xyz))
        self._atoms.adp = np.ascontiguousarray(_np(adp))
        self._atoms.occupancy = np.ascontiguousarray(_np(occ))

    def _set_types(self, plm, pval, kappa, kappa_prime):
        key = _key(plm, pval, kappa, kappa_prime)
        if key == self._types_key:
            return
        self.calc.setTypeParameters(_np(plm), _np(pval), _np(kappa), _np(kappa_prime))
        self._types_key = key

    def __call__(self, xyz, adp, occ, plm, pval, kappa, kappa_prime):
        return _SF.apply(self, xyz, adp, occ, plm, pval, kappa, kappa_prime)
