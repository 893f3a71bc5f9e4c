"""Input containers mirroring DiSCaMB's hot-path types, plus the built-in Clementi-Roetti bank.

* ``Crystal`` / ``AtomSet``      - Crystal, AtomInCrystal (include/discamb/CrystalStructure/Crystal.h)
* ``HC_WfnType``                 - HC_WfnType (include/discamb/HC_Model/HC_WfnData.h:126-163)
* ``HC_AtomTypeParameters``      - HC_AtomTypeParameters (include/discamb/HC_Model/HC_AtomTypeParameters.h:19-50)
* ``HC_ModelParameters``         - HC_ModelParameters (include/discamb/HC_Model/HC_ModelParameters.h:22-28)
* ``TargetFunctionAtomicParamDerivatives`` - SF_CalcDataTypes.h:36-41 (as arrays over atoms)
* ``DerivativesSelector``        - SF_CalcDataTypes.h:44-56

Atom data are held as numpy arrays over atoms rather than as a list of structs.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from fractions import Fraction
from typing import List, Optional, Sequence

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "wfn_tables.json")
_BANK = None


def parse_symop(text: str):
    """Parse a CIF-style operation such as ``-x+1/2,y,-z+1/2`` into (R int 3x3, t float 3)."""
    parts = text.replace(" ", "").lower().split(",")
    if len(parts) != 3:
        raise ValueError("symmetry operation needs three components: %r" % text)
    rot = np.zeros((3, 3), dtype=np.int32)
    trans = np.zeros(3)
    for i, p in enumerate(parts):
        t = Fraction(0)
        p = p.replace("-", "+-")
        for term in [s for s in p.split("+") if s]:
            neg = term.startswith("-")
            term = term.lstrip("-")
            sgn = -1 if neg else 1
            if term in ("x", "y", "z"):
                rot[i, "xyz".index(term)] += sgn
            elif term and term[-1] in "xyz":
                rot[i, "xyz".index(term[-1])] += sgn * int(term[:-1] or 1)
            else:
                t += sgn * Fraction(term)
        trans[i] = float(t)
    return rot, trans


@dataclass
class AtomSet:
    """vector<AtomInCrystal> as arrays: xyz (n,3), adp (n,6) with adp_size (n,) in {0,1,6},
    occupancy (n,), multiplicity (n,)."""
    xyz: np.ndarray
    adp: np.ndarray
    adp_size: np.ndarray
    occupancy: np.ndarray
    multiplicity: np.ndarray
    labels: Optional[List[str]] = None
    elements: Optional[List[str]] = None

    def __post_init__(self):
        self.xyz = np.ascontiguousarray(self.xyz, dtype=np.float64).reshape(-1, 3)
        n = self.xyz.shape[0]
        self.adp = np.ascontiguousarray(self.adp, dtype=np.float64).reshape(n, 6)
        self.adp_size = np.ascontiguousarray(self.adp_size, dtype=np.int32).reshape(n)
        self.occupancy = np.ascontiguousarray(self.occupancy, dtype=np.float64).reshape(n)
        self.multiplicity = np.ascontiguousarray(self.multiplicity, dtype=np.float64).reshape(n)

    @property
    def n(self):
        return self.xyz.shape[0]

    def copy(self):
        return AtomSet(self.xyz.copy(), self.adp.copy(), self.adp_size.copy(), self.occupancy.copy(),
                       self.multiplicity.copy(), None if self.labels is None else list(self.labels),
                       None if self.elements is None else list(self.elements))


@dataclass
class Crystal:
    cell: Sequence[float]                 # a, b, c, alpha, beta, gamma
    symops: List[str]                     # full space group
    atoms: AtomSet
    xyz_convention: str = "fractional"    # or "cartesian"
    adp_convention: str = "U_cif"         # or "U_cart"


@dataclass
class SlaterOrbital:
    n: int
    l: int
    coefficients: List[float]
    exponents: List[float]
    powers: List[int]


@dataclass
class HC_WfnType:
    label: str
    orbitals: List[SlaterOrbital]
    orbital_occupancy: List[int]
    core_orbitals_indices: List[int]
    valence_orbitals_indices: List[int]
    deformation_valence_exponent: float = 1.0
    deformation_valence_power: List[int] = field(default_factory=list)
    anomalous_scattering: complex = 0j
    atomic_number: int = 0
    charge: int = 0


@dataclass
class HC_AtomTypeParameters:
    p_lm: List[List[float]]               # p_lm[l][l + m]
    p_val: float
    kappa_spherical_valence: float = 1.0
    kappa_deformation_valence: float = 1.0

    def plm25(self):
        out = np.zeros(25)
        for l, row in enumerate(self.p_lm[:5]):
            for i, v in enumerate(row):
                out[l * l + i] = v
        return out


@dataclass
class HC_ModelParameters:
    wfn_parameters: List[HC_WfnType]
    type_parameters: List[HC_AtomTypeParameters]
    atom_to_wfn_map: List[int]
    atom_to_type_map: List[int]


@dataclass
class DerivativesSelector:
    d_xyz: bool = True
    d_adp: bool = True
    d_occ: bool = True
    d_anom: bool = False
    d_multipole: bool = False      # P_lm, Pval, kappa, kappa' of the atom types (no reference counterpart)

    def bits(self):
        return ((1 if self.d_xyz else 0) | (2 if self.d_adp else 0) | (4 if self.d_occ else 0) |
                (8 if self.d_multipole else 0))


@dataclass
class TypeParameterDerivatives:
    """dT/d(atom-type parameter), summed over the atoms of each type: p_lm (n_types, 25) in the
    l*l + l + m layout (local frame, density normalised), p_val, kappa, kappa_prime (n_types,)."""
    p_lm: np.ndarray
    p_val: np.ndarray
    kappa: np.ndarray
    kappa_prime: np.ndarray


@dataclass
class TargetFunctionAtomicParamDerivatives:
    """dT/dp for all atoms: xyz (n,3), adp (n,6) (first adp_size entries used), occ (n,)."""
    atomic_position_derivatives: np.ndarray
    adp_derivatives: np.ndarray
    occupancy_derivatives: np.ndarray


@dataclass
class LcsDefinition:
    """Definition of one atom's local coordinate system from atomic positions
    (LocalCoordinateSystem<AtomInCrystalID>, LocalCoordinateSystem.h:20-55).  Every reference atom is
    (index in the asymmetric unit, symmetry operation applied to it).  direction types:
    "average_direction" (mean of the unit vectors from the central atom to the reference atoms),
    "average_position" (from the central atom to the reference atoms' centroid), "any_orthogonal".
    coordinate_1 / coordinate_2: axis (0 X, 1 Y, 2 Z) along direction 1 / in the plane of both
    directions; is_r: right-handed; chirality: three atoms whose triple product fixes the sign of
    the third axis (or empty)."""
    central: int
    direction1_type: str
    ref1: List[tuple]
    direction2_type: str
    ref2: List[tuple]
    coordinate_1: int = 0
    coordinate_2: int = 1
    is_r: bool = True
    chirality: List[tuple] = field(default_factory=list)
    central_op: str = "x,y,z"


class SfDerivativesAtHkl:
    """dF/dp of one reflection for every atom (SF_CalcDataTypes.h:49-55), complex:
    atomicPostionDerivatives (n, 3), adpDerivatives (n, 6) (U_iso in column 0, unused columns 0),
    occupancyDerivatives (n,), anomalousScatteringDerivatives (empty: the per-hkl engine leaves it
    unset).  The field names keep the reference's spelling.  Instances handed to a per-hkl
    processor view library memory that the next reflection overwrites: copy what must outlive
    the callback."""

    __slots__ = ("row", "anomalousScatteringDerivatives")

    def __init__(self, row):
        self.row = row                               # (10, n) complex
        self.anomalousScatteringDerivatives = np.zeros(0, dtype=np.complex128)

    @property
    def atomicPostionDerivatives(self):
        return self.row[0:3].T

    atomicPositionDerivatives = atomicPostionDerivatives

    @property
    def adpDerivatives(self):
        return self.row[3:9].T

    @property
    def occupancyDerivatives(self):
        return self.row[9]


def _bank():
    global _BANK
    if _BANK is None:
        with open(_DATA) as f:
            _BANK = json.load(f)["entries"]
    return _BANK


def wfn_type_from_bank(symbol: str, def_val: str = "UBDB") -> HC_WfnType:
    """Clementi-Roetti wavefunction + default deformation-valence parameters for an element
    (the TAAM route, taam_utilities.cpp:86-104: SlaterOrbitalWfnData CR + DeformationValenceParameters
    UBDB_CR)."""
    e = _bank()[symbol]
    orbs = [SlaterOrbital(o["n"], o["l"], list(o["coefficients"]), list(o["exponents"]), list(o["powers"]))
            for o in e["orbitals"]]
    pw = e["def_val_powers_ubdb"] if def_val.upper() == "UBDB" else e["def_val_powers_standard"]
    return HC_WfnType(symbol, orbs, list(e["occupancy"]), list(e["core"]), list(e["valence"]),
                      float(e["def_val_exponent"]), list(pw), 0j, int(e["Z"]), int(e["charge"]))


def bank_symbols():
    return list(_bank().keys())
