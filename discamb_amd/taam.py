"""Transferable aspherical atom model (TAAM) front end of the Hansen-Coppens path.

Mirrors the part of HcAtomBankStructureFactorCalculator (src/Scattering/
HcAtomBankStructureFactorCalculator.cpp:95-150, 440-520) that runs after atom typing: bank
parameters of the assigned atom types become an HC model (taam_utilities.cpp:44-199, 332-548),
with the valence populations scaled to the unit-cell charge (Faerman-Price,
taam_utilities.cpp:572-640); every atom with a type gets a local frame recomputed from the
positions at each call (LocalCoordinateSystemCalculator), and atoms whose element has no
Clementi-Roetti wavefunction are computed with the independent-atom model and added
(CombinedStructureFactorCalculator.cpp:128-176).

Starts from the type assignment: reading a MATTS bank file and assigning atom types
(MATTS_BankReader, MolecularAtomTypeAssigner) is not part of this package; the caller passes the
bank parameters of the types, the type index of every atom (-1 for none) and the local coordinate
system definitions the assigner produced.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .calculator import HansenCoppensStructureFactorCalculator, compute_lcs
from .model import (AtomSet, Crystal, DerivativesSelector, HC_AtomTypeParameters, HC_ModelParameters, HC_WfnType,
                    LcsDefinition, TargetFunctionAtomicParamDerivatives, bank_symbols, wfn_type_from_bank)
from .synth import cell_matrices

_ELEMENTS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
             "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
             "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf").split()


@dataclass
class AtomTypeHcParameters:
    """Multipole parameters of one bank atom type (AtomTypeHC_Parameters): kappa, kappa', P_val and
    its spread (p_val_sigma), and the non-zero P_lm as (l, m) indices with values."""
    kappa: float = 1.0
    kappa_prime: float = 1.0
    p_val: float = 0.0
    p_val_sigma: float = 0.0
    p_lm_indices: List[Tuple[int, int]] = field(default_factory=list)
    p_lms: List[float] = field(default_factory=list)


@dataclass
class TaamAtomicChargeInfo:
    """Atomic charges N_val(Z) - P_val before and after the electroneutrality scaling (0 for atoms
    outside the scaling)."""
    atomic_charges_before_scaling: List[float] = field(default_factory=list)
    atomic_charges_after_scaling: List[float] = field(default_factory=list)


def atomic_number(symbol: str) -> int:
    """Atomic number of an element symbol or an ion label ("O", "O1-", "Fe3+")."""
    s = symbol.strip()
    while s and (s[-1] in "+-" or s[-1].isdigit()):
        s = s[:-1]
    s = s[:1].upper() + s[1:].lower()
    if s not in _ELEMENTS:
        raise ValueError("unknown chemical element: %r" % symbol)
    return _ELEMENTS.index(s) + 1


def _neutral_valence_occupancy():
    """Valence electron count of every neutral atom of the Clementi-Roetti bank."""
    out = {}
    for sym in bank_symbols():
        w = wfn_type_from_bank(sym)
        if w.charge == 0:
            out[w.atomic_number] = float(sum(w.orbital_occupancy[i] for i in w.valence_orbitals_indices))
    return out


def void_atom_type():
    """Wavefunction and type of an atom with no Clementi-Roetti entry (taam_utilities.cpp:
    void_atom_type): the hydrogen wavefunction with deformation exponent 1, P_val 0, no P_lm -
    such an atom scatters with f' + i f'' only."""
    w = wfn_type_from_bank("H")
    w.deformation_valence_exponent = 1.0
    return w, HC_AtomTypeParameters([], 0.0, 1.0, 1.0)


def electroneutrality_faerman_price(type_pval, type_pval_sigma, atom_to_type, multiplicity, atomic_numbers,
                                    target_charge):
    """taam_utilities::electroneutrality_Faerman_Price (taam_utilities.cpp:572-640): the unit-cell
    charge sum_a m_a (N_val(Z_a) - P_val,t(a)) over atoms of neutral Clementi-Roetti elements is
    brought to target_charge by P_val,t -= dQ sigma_t / sum_a sigma_t(a) m_a for every type such an
    atom uses (all sigma 1 and the denominator the atom count when the sigmas sum to 0).
    multiplicity is multiplicity x occupancy per atom.  Returns (new P_val per type,
    TaamAtomicChargeInfo)."""
    pval = [float(v) for v in type_pval]
    sigma = [float(v) for v in type_pval_sigma]
    nval = _neutral_valence_occupancy()
    n = len(atom_to_type)
    used = [False] * len(pval)
    info = TaamAtomicChargeInfo()
    q = 0.0
    sigma_sum = 0.0
    counted = [atom_to_type[a] >= 0 and int(atomic_numbers[a]) in nval for a in range(n)]
    for a in range(n):
        if counted[a]:
            t = atom_to_type[a]
            dq = nval[int(atomic_numbers[a])] - pval[t]
            info.atomic_charges_before_scaling.append(dq)
            q += dq * multiplicity[a]
            sigma_sum += sigma[t] * multiplicity[a]
            used[t] = True
        else:
            info.atomic_charges_before_scaling.append(0.0)
    d_q = target_charge - q
    if d_q == 0:
        return pval, info
    if sigma_sum == 0:
        sigma_sum = float(n)
        sigma = [1.0] * len(sigma)
    for t in range(len(pval)):
        if used[t]:
            pval[t] -= d_q * sigma[t] / sigma_sum
    for a in range(n):
        info.atomic_charges_after_scaling.append(
            nval[int(atomic_numbers[a])] - pval[atom_to_type[a]] if counted[a] else 0.0)
    return pval, info


def type_assignment_to_hc_parameters(bank_parameters: Sequence[AtomTypeHcParameters], type_assignment: Sequence[int],
                                     atomic_numbers: Sequence[int], multiplicity_times_occupancy=None,
                                     target_charge: float = 0.0, scale: bool = True,
                                     not_assigned_represented_with_slaters: bool = True):
    """taam_utilities::type_assignment_to_HC_parameters (scale=True, taam_utilities.cpp:332-548)
    and type_assignment_to_unscaled_HC_parameters (scale=False, :44-199) with the Clementi-Roetti
    wavefunctions and UBDB deformation-valence parameters.

    Model layout, as the reference builds it: a void type/wavefunction first when some element has
    no Clementi-Roetti entry; then one wavefunction per present element in increasing Z and one
    type per distinct assigned bank type in increasing bank index.  Atoms without a type take type
    0, and wavefunction 0 unless their element has a Clementi-Roetti entry and
    not_assigned_represented_with_slaters; those with wavefunction 0 are the non-multipolar atoms.
    Returns (HC_ModelParameters, non-multipolar atom indices, TaamAtomicChargeInfo or None)."""
    n = len(type_assignment)
    z_all = [int(z) for z in atomic_numbers]
    if len(z_all) != n:
        raise ValueError("one atomic number per atom is needed")
    in_bank = {wfn_type_from_bank(s).atomic_number for s in bank_symbols()}
    present = sorted({z for z in z_all if z in in_bank})
    has_void = any(z not in in_bank for z in z_all)
    wfn: List[HC_WfnType] = []
    types: List[HC_AtomTypeParameters] = []
    if has_void:
        vw, vt = void_atom_type()
        wfn.append(vw)
        types.append(vt)
    z_to_wfn = {}
    for z in present:
        z_to_wfn[z] = len(wfn)
        wfn.append(wfn_type_from_bank(_ELEMENTS[z - 1], "UBDB"))
    used_types = sorted({int(t) for t in type_assignment if int(t) >= 0})
    for t in used_types:
        if t >= len(bank_parameters):
            raise ValueError("atom type index %d beyond the %d bank types" % (t, len(bank_parameters)))
    type_to_idx = {}
    for t in used_types:
        b = bank_parameters[t]
        type_to_idx[t] = len(types)
        max_l = max([l for l, _ in b.p_lm_indices], default=-1)
        p_lm = [[0.0] * (2 * l + 1) for l in range(max_l + 1)]
        for (l, m), v in zip(b.p_lm_indices, b.p_lms):
            if l > 4 or abs(m) > l:
                raise ValueError("P_lm index (%d, %d) out of range" % (l, m))
            p_lm[l][l + m] = float(v)
        types.append(HC_AtomTypeParameters(p_lm, float(b.p_val), float(b.kappa), float(b.kappa_prime)))
    if not types:
        raise ValueError("no atom has a type and every element has a wavefunction: the model has no atom type")
    atom_to_type, atom_to_wfn, non_multipolar = [], [], []
    for a in range(n):
        t = int(type_assignment[a])
        if t >= 0:
            atom_to_type.append(type_to_idx[t])
            atom_to_wfn.append(z_to_wfn.get(z_all[a], 0))
        else:
            atom_to_type.append(0)
            if not_assigned_represented_with_slaters and z_all[a] in z_to_wfn:
                atom_to_wfn.append(z_to_wfn[z_all[a]])
            else:
                atom_to_wfn.append(0)
                non_multipolar.append(a)
    info = None
    if scale:
        mult = (np.ones(n) if multiplicity_times_occupancy is None else
                np.asarray(multiplicity_times_occupancy, dtype=np.float64))
        first = 1 if has_void else 0
        pval = [tp.p_val for tp in types]
        sigma = [0.0] * len(types)
        for t in used_types:
            sigma[type_to_idx[t]] = float(bank_parameters[t].p_val_sigma)
        pval, info = electroneutrality_faerman_price(pval, sigma, atom_to_type, mult, z_all, target_charge)
        for i in range(first, len(types)):
            types[i].p_val = pval[i]
    return HC_ModelParameters(wfn, types, atom_to_wfn, atom_to_type), non_multipolar, info


_SETTINGS = {"electron scattering", "electron_scattering", "unit cell charge", "scale", "table",
             "iam electron scattering", "frozen lcs", "algorithm", "wavefunction bank", "n cores", "engine",
             "lcs chain rule", "device"}
_FILE_SETTINGS = {"bank path", "assignment info", "assignment csv", "parameters info", "multipole cif"}


def _identity_lcs(atom):
    return LcsDefinition(atom, "any_orthogonal", [], "any_orthogonal", [])


class HcAtomBankStructureFactorCalculator:
    """TAAM structure factors and derivatives from a type assignment
    (HcAtomBankStructureFactorCalculator::set, HcAtomBankStructureFactorCalculator.cpp:440-520).

    settings takes the reference's JSON keys: "electron scattering" (or "electron_scattering"),
    "unit cell charge", "scale", "table" and "iam electron scattering" (the independent-atom part),
    "frozen lcs", "algorithm" and "n cores" (accepted; one GPU path serves both algorithms), and
    "wavefunction bank" ("CR" only).  Added keys: "engine" (a calculation mode of the multipole
    part: "HIP" default, "HIP_MIXED"), "lcs chain rule" (add dT/dL : dL/dxyz to the position
    derivatives; the reference leaves it out) and "device".  Keys naming files (bank, assignment
    or parameter reports, CIF output) are refused: the front end takes arrays."""

    def __init__(self, crystal: Crystal, bank_parameters: Sequence[AtomTypeHcParameters],
                 type_assignment: Sequence[int], lcs_definitions: Sequence[Optional[LcsDefinition]],
                 settings: dict = None, atomic_numbers: Sequence[int] = None):
        settings = dict(settings or {})
        for k in settings:
            if k in _FILE_SETTINGS:
                raise ValueError("setting %r names a file; this front end takes the bank parameters and the "
                                 "type assignment as arrays" % k)
            if k not in _SETTINGS:
                raise ValueError("unknown setting %r" % k)
        if str(settings.get("wavefunction bank", "CR")).upper() != "CR":
            raise ValueError("only the Clementi-Roetti wavefunction bank (\"CR\") is available")
        atoms = crystal.atoms
        n = atoms.n
        if len(type_assignment) != n or len(lcs_definitions) != n:
            raise ValueError("one type index and one local coordinate system definition per atom are needed")
        if atomic_numbers is None:
            if atoms.elements is None:
                raise ValueError("atomic numbers are needed (AtomSet.elements or atomic_numbers)")
            atomic_numbers = [atomic_number(e) for e in atoms.elements]
        self.atomic_numbers = [int(z) for z in atomic_numbers]
        electron = bool(settings.get("electron scattering", settings.get("electron_scattering", False)))
        mult_occ = np.asarray(atoms.multiplicity, dtype=np.float64) * np.asarray(atoms.occupancy, dtype=np.float64)
        self.parameters, self.non_multipolar_atoms, self.charge_info = type_assignment_to_hc_parameters(
            bank_parameters, type_assignment, self.atomic_numbers, mult_occ, float(settings.get("unit cell charge", 0.0)),
            bool(settings.get("scale", True)), True)
        self.type_assignment = [int(t) for t in type_assignment]
        engine = str(settings.get("engine", "HIP"))
        device = int(settings.get("device", 0))
        mixed = engine.upper().endswith("MIXED")
        self._crystal = crystal
        self._hc = HansenCoppensStructureFactorCalculator(crystal, self.parameters, device)
        self._hc.setCalculationMode(engine)
        if electron:
            self._hc.setElectronScattering(True)
        # atoms without a type keep the identity frame (LocalCoordinateSystemCalculator())
        self._defs = [d if (d is not None and self.type_assignment[i] >= 0) else _identity_lcs(i)
                      for i, d in enumerate(lcs_definitions)]
        self._frozen = None
        if bool(settings.get("frozen lcs", False)):
            self._frozen = compute_lcs(crystal.cell, self._fractional(atoms.xyz), self._defs)
        else:
            self._hc.setLocalCoordinateSystems(self._defs, bool(settings.get("lcs chain rule", False)))
        self.multipolar = np.ones(n, dtype=bool)
        self.multipolar[self.non_multipolar_atoms] = False
        self._iam = None
        if self.non_multipolar_atoms:
            self._iam = self._iam_calculator(crystal, settings, device, mixed)
        self._anomalous = None

    def _fractional(self, xyz):
        xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        if self._crystal.xyz_convention == "cartesian":
            return xyz @ cell_matrices(self._crystal.cell)[1].T
        return xyz

    def _iam_calculator(self, crystal, settings, device, mixed):
        """AnyIamCalculator of the non-multipolar atoms (AnyIamCalculator.cpp:49-135): the IT92 table
        by default, "electron-IT" for "iam electron scattering"; each atom's form factor is that of its
        type label (its element symbol or ion label)."""
        table = str(settings.get("table", ""))
        if bool(settings.get("iam electron scattering", False)):
            if table not in ("", "electron-IT"):
                raise ValueError("the Mott-Bethe transform of an X-ray IAM table is not available; use \"electron-IT\"")
            table = "electron-IT"
        elif table == "":
            table = "IT92"
        labels = list(crystal.atoms.elements) if crystal.atoms.elements is not None else \
            [_ELEMENTS[z - 1] for z in self.atomic_numbers]
        distinct = sorted(set(labels), key=labels.index)
        wfn = []
        for lab in distinct:
            w, _ = void_atom_type()
            w.label = lab
            w.atomic_number = atomic_number(lab)
            wfn.append(w)
        a2w = [distinct.index(lab) for lab in labels]
        model = HC_ModelParameters(wfn, [HC_AtomTypeParameters([], 0.0, 1.0, 1.0) for _ in wfn], a2w, a2w)
        c = HansenCoppensStructureFactorCalculator(crystal, model, device)
        c.setIamTable(table)
        c.setCalculationMode("HIP_IAM_MIXED" if mixed else "CPU_IAM")
        return c

    # -- reference API -----------------------------------------------------------------------
    def setAnomalous(self, anomalous):
        """Per-atom f' + i f'' for both parts (HcAtomBankStructureFactorCalculator::setAnomalous)."""
        self._anomalous = None if anomalous is None else np.asarray(anomalous, dtype=np.complex128)

    def lastLocalCoordinateSystems(self):
        return self._frozen.copy() if self._frozen is not None else self._hc.lastLocalCoordinateSystems()

    def _masks(self, count):
        n = self._crystal.atoms.n
        c = np.ones(n, dtype=bool) if count is None else np.asarray(count, dtype=bool)
        return c & self.multipolar, c & ~self.multipolar

    def calculateStructureFactorsAndDerivatives(self, atoms: AtomSet, hkl, dTarget_dF, countAtomContribution=None,
                                                derivativesSelector: DerivativesSelector = None):
        """CombinedStructureFactorCalculator::calculateStructureFactorsAndDerivatives over the
        multipole and IAM parts: F summed, each atom's derivatives from its own part."""
        hc_mask, iam_mask = self._masks(countAtomContribution)
        f, d = self._hc.calculateStructureFactorsAndDerivatives(atoms, self._frozen, hkl, dTarget_dF, hc_mask,
                                                               derivativesSelector, self._anomalous)
        if self._iam is None:
            return f, d
        f2, d2 = self._iam.calculateStructureFactorsAndDerivatives(atoms, None, hkl, dTarget_dF, iam_mask,
                                                                  derivativesSelector, self._anomalous)
        nm = ~self.multipolar
        xyz = d.atomic_position_derivatives.copy()
        adp = d.adp_derivatives.copy()
        occ = d.occupancy_derivatives.copy()
        xyz[nm], adp[nm], occ[nm] = d2.atomic_position_derivatives[nm], d2.adp_derivatives[nm], \
            d2.occupancy_derivatives[nm]
        return f + f2, TargetFunctionAtomicParamDerivatives(xyz, adp, occ)

    def calculateStructureFactors(self, atoms: AtomSet, hkl, countAtomContribution=None):
        hc_mask, iam_mask = self._masks(countAtomContribution)
        f = self._hc.calculateStructureFactors(atoms, self._frozen, hkl, hc_mask, self._anomalous)
        if self._iam is not None:
            f = f + self._iam.calculateStructureFactors(atoms, None, hkl, iam_mask, self._anomalous)
        return f
