"""discamb_amd: Hansen-Coppens structure factors and derivatives on AMD MI355X (gfx950).

A from-scratch HIP implementation of DiSCaMB's
``HansenCoppensStructureFactorCalculator::calculateStructureFactorsAndDerivatives`` behind the
same calculator interface; see include/discamb_amd.h for the C ABI.
"""
from .model import (AtomSet, Crystal, DerivativesSelector, HC_AtomTypeParameters, HC_ModelParameters,
                    HC_WfnType, LcsDefinition, SfDerivativesAtHkl, SlaterOrbital, TargetFunctionAtomicParamDerivatives,
                    TypeParameterDerivatives, parse_symop, wfn_type_from_bank)
from .calculator import BatchCalculator, HansenCoppensStructureFactorCalculator, compute_lcs
from .taam import AtomTypeHcParameters, HcAtomBankStructureFactorCalculator, type_assignment_to_hc_parameters

__all__ = ["AtomSet", "Crystal", "DerivativesSelector", "HC_AtomTypeParameters", "HC_ModelParameters",
           "HC_WfnType", "LcsDefinition", "SfDerivativesAtHkl", "SlaterOrbital", "TargetFunctionAtomicParamDerivatives", "TypeParameterDerivatives",
           "parse_symop",
           "wfn_type_from_bank", "HansenCoppensStructureFactorCalculator", "BatchCalculator", "compute_lcs",
           "AtomTypeHcParameters", "HcAtomBankStructureFactorCalculator", "type_assignment_to_hc_parameters"]
