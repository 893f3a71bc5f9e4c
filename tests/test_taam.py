"""TAAM front end on the host: bank parameters + type assignment -> HC model
(taam_utilities.cpp:44-199, 332-548) and the Faerman-Price electroneutrality scaling (:572-640).

taam_utilities.cpp cannot be compiled into oracle/_ref: it pulls in the atom-typing and
structural-properties sources and vflib.  The restatement is pinned instead by what fixes it:
  * the scaling: the cell charge after scaling equals the target, the P_val shifts are
    proportional to the types' sigmas, unused types are untouched - together these determine the
    result - plus the closed forms of the zero-sigma and zero-dQ branches;
  * the model layout: a worked example whose expected arrays follow from the reference's
    ordering rules (void type first, wavefunctions by increasing Z, types by increasing bank
    index, P_lm at [l][l + m], unassigned atoms on type 0), checked entry by entry.
"""
import numpy as np
import pytest

from discamb_amd import Crystal, AtomSet, LcsDefinition
from discamb_amd.taam import (AtomTypeHcParameters, HcAtomBankStructureFactorCalculator, atomic_number,
                              electroneutrality_faerman_price, type_assignment_to_hc_parameters)
from discamb_amd.model import wfn_type_from_bank


def _nval(z):
    el = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S"][z - 1]
    w = wfn_type_from_bank(el)
    return sum(w.orbital_occupancy[i] for i in w.valence_orbitals_indices)


def test_valence_counts_are_the_chemical_ones():
    assert [_nval(z) for z in (1, 6, 7, 8, 16)] == [1, 4, 5, 6, 6]


def test_electroneutrality_reaches_the_target_with_sigma_proportional_shifts():
    rng = np.random.default_rng(5)
    for trial in range(20):
        n_types, n = 7, 40
        pval = rng.uniform(0.7, 6.5, n_types)
        sigma = rng.uniform(0.01, 0.3, n_types)
        a2t = rng.integers(0, n_types - 2, n)              # the last two types stay unused
        z = rng.choice([1, 6, 7, 8, 16, 53], n)            # iodine (no neutral CR entry) is not counted
        mult = rng.uniform(0.5, 4.0, n)
        target = float(rng.uniform(-2, 2))
        new, info = electroneutrality_faerman_price(pval, sigma, list(a2t), mult, z, target)
        counted = z != 53
        q = sum(m * (_nval(int(zz)) - new[t]) for m, zz, t, c in zip(mult, z, a2t, counted) if c)
        assert abs(q - target) < 1e-10
        used = sorted({int(t) for t, c in zip(a2t, counted) if c})
        ratio = [(new[t] - pval[t]) / sigma[t] for t in used]
        assert np.ptp(ratio) < 1e-12
        assert all(new[t] == pval[t] for t in range(n_types) if t not in used)
        after = [(_nval(int(zz)) - new[t]) if c else 0.0 for zz, t, c in zip(z, a2t, counted)]
        assert np.allclose(info.atomic_charges_after_scaling, after, atol=1e-12)


def test_electroneutrality_zero_sigma_and_zero_charge_branches():
    # all sigmas 0: every used type moves by -dQ / (number of atoms, counted or not)
    pval, sigma = [4.2, 5.1, 3.0], [0.0, 0.0, 0.0]
    a2t, z, mult = [0, 1, 0, 1], [6, 8, 6, 53], [1.0, 2.0, 1.0, 1.0]
    q = (4 - 4.2) * 1 + (6 - 5.1) * 2 + (4 - 4.2) * 1
    new, _ = electroneutrality_faerman_price(pval, sigma, a2t, mult, z, 0.5)
    dq = 0.5 - q
    assert np.allclose(new, [4.2 - dq / 4, 5.1 - dq / 4, 3.0], atol=1e-14)
    # already at the target: nothing changes and no after-charges are reported
    new, info = electroneutrality_faerman_price([4.0, 6.0], [0.1, 0.1], [0, 1], [1.0, 1.0], [6, 8], 0.0)
    assert new == [4.0, 6.0] and info.atomic_charges_after_scaling == []


def _bank():
    return [AtomTypeHcParameters(1.0 + 0.01 * i, 0.9 + 0.02 * i, 3.8 + 0.1 * i, 0.05 + 0.01 * i,
                                 [(0, 0), (1, 1), (2, -1), (3, 2)], [0.01 * i, 0.1, -0.2, 0.3])
            for i in range(9)]


def test_model_layout_worked_example():
    bank = _bank()
    bank[2].p_lm_indices, bank[2].p_lms = [], []
    types = [7, 2, 7, -1, 4, -1]
    z = [6, 8, 6, 53, 7, 8]                     # C O C I N O
    p, nonmp, info = type_assignment_to_hc_parameters(bank, types, z, scale=False)
    assert [w.label for w in p.wfn_parameters] == ["H", "C", "N", "O"]      # void, then increasing Z
    assert p.wfn_parameters[0].deformation_valence_exponent == 1.0
    assert p.type_parameters[0].p_val == 0.0 and p.type_parameters[0].p_lm == []
    assert [t.p_val for t in p.type_parameters[1:]] == [bank[2].p_val, bank[4].p_val, bank[7].p_val]
    assert list(p.atom_to_type_map) == [3, 1, 3, 0, 2, 0]
    assert list(p.atom_to_wfn_map) == [1, 3, 1, 0, 2, 3]   # unassigned O keeps its wavefunction
    assert nonmp == [3] and info is None
    t7 = p.type_parameters[3]
    assert t7.kappa_spherical_valence == bank[7].kappa and t7.kappa_deformation_valence == bank[7].kappa_prime
    assert [len(r) for r in t7.p_lm] == [1, 3, 5, 7]
    assert t7.p_lm[0][0] == bank[7].p_lms[0] and t7.p_lm[1][2] == 0.1 and t7.p_lm[2][1] == -0.2 and t7.p_lm[3][5] == 0.3
    assert sum(abs(v) for r in t7.p_lm for v in r) == pytest.approx(abs(bank[7].p_lms[0]) + 0.6)
    assert p.type_parameters[1].p_lm == []      # a bank type without P_lm
    # no element without a wavefunction: no void entries, unassigned atoms on the first bank type
    p2, nonmp2, _ = type_assignment_to_hc_parameters(bank, [5, -1, 3], [6, 8, 7], scale=False)
    assert [w.label for w in p2.wfn_parameters] == ["C", "N", "O"]
    assert list(p2.atom_to_type_map) == [1, 0, 0] and nonmp2 == []


def test_scaled_model_counts_unassigned_atoms_with_type_zero():
    """Unassigned atoms sit on type 0 and enter the charge balance with its P_val (the reference's
    atom-to-type map has no -1 entries by then)."""
    bank = [AtomTypeHcParameters(p_val=4.2, p_val_sigma=0.1) for _ in range(6)]
    p, _, info = type_assignment_to_hc_parameters(bank, [5, -1], [6, 8], np.ones(2), 0.0, True)
    # Q = (4 - 4.2) + (6 - 4.2) = 1.6; both atoms on the one type: P_val = 4.2 + 1.6 * 0.1 / 0.2
    assert p.type_parameters[0].p_val == pytest.approx(5.0, abs=1e-14)
    assert info.atomic_charges_before_scaling == pytest.approx([-0.2, 1.8])
    assert info.atomic_charges_after_scaling == pytest.approx([-1.0, 1.0])
    # with a void type, its P_val stays 0 even where unassigned atoms are counted on it
    p, _, _ = type_assignment_to_hc_parameters(bank, [5, -1, -1], [6, 8, 53], np.ones(3), 0.0, True)
    assert p.type_parameters[0].p_val == 0.0
    assert p.type_parameters[1].p_val == pytest.approx(4.2 + ((4 - 4.2) + (6 - 0.0)) * 0.1 / 0.1)


def test_atomic_numbers_and_settings_are_validated():
    assert [atomic_number(s) for s in ("H", "O1-", "Fe3+", "cl", "I")] == [1, 8, 26, 17, 53]
    with pytest.raises(ValueError):
        atomic_number("Xx")
    atoms = AtomSet(np.zeros((1, 3)), np.zeros((1, 6)), np.zeros(1, dtype=np.int32), np.ones(1), np.ones(1), None, ["C"])
    cr = Crystal([8, 8, 8, 90, 90, 90], ["x,y,z"], atoms)
    lcs = [LcsDefinition(0, "any_orthogonal", [], "any_orthogonal", [])]
    for bad in ({"bank path": "x.bnk"}, {"assignment info": "a.log"}, {"wavefunction bank": "SCM"},
                {"no such key": 1}):
        with pytest.raises(ValueError):
            HcAtomBankStructureFactorCalculator(cr, _bank(), [0], lcs, bad)
    with pytest.raises(ValueError):
        HcAtomBankStructureFactorCalculator(cr, _bank(), [0, 1], lcs)
    with pytest.raises(ValueError):
        type_assignment_to_hc_parameters(_bank(), [12], [6], scale=False)
