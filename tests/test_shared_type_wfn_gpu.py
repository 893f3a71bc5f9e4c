"""Atom types shared by atoms of different wavefunctions (as TAAM models have: one bank type, or
type 0 for unassigned atoms, across elements).

The reference's engines take each atom's core term (and f', f'') from the atom's own wavefunction
but the valence and deformation terms from the wavefunction of the type's last atom
(HansenCoppens_SF_Engine3.cpp:218-224, 1295-1300; HansenCoppens_SF_Engine.h:639); the form-factor
route takes the core term from the type's wavefunction too (HansenCoppensStructureFactorCalculator
.cpp:1292-1300).  The oracle follows both; here every GPU route is checked against it on a model
where three types each mix wavefunctions.  Tolerances: FP64 1e-10, mixed 1e-6 / 1e-5.
"""
import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import DerivativesSelector, HansenCoppensStructureFactorCalculator, synth

pytestmark = pytest.mark.gpu


def _mixed_model(sg, seed=21):
    w = synth.small_structure(sg, n_atoms=12, d_min=1.1, seed=seed)
    p = w.params
    # types 0..3 only; wavefunctions stay per element (C N O S H H ...)
    p.atom_to_type_map = [0, 1, 2, 0, 3, 1, 2, 0, 1, 3, 3, 2]
    p.type_parameters = p.type_parameters[:4]
    return w


def _errs(f, d, ref):
    return [rel_err(f, ref[0]), rel_err(d.atomic_position_derivatives, ref[1]), rel_err(d.adp_derivatives, ref[2]),
            rel_err(d.occupancy_derivatives, ref[3])]


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
@pytest.mark.parametrize("sg", ["P21/c", "P212121", "P-1"])
def test_summed_path(sg, mode):
    w = _mixed_model(sg)
    print("case shared-type summed %s %s" % (sg, mode), flush=True)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setCalculationMode(mode)
    f, d = c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    e = _errs(f, d, hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf))
    print("  ", ["%.1e" % v for v in e], flush=True)
    assert (max(e) < 1e-10) if mode == "HIP" else (e[0] < 1e-6 and max(e) < 1e-5)


def test_rows_form_factors_and_type_derivatives():
    w = _mixed_model("P21/c")
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    print("case shared-type rows / form factors / type derivatives", flush=True)
    F, R = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl)
    rF, rR = hc_oracle.per_hkl(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl)
    assert rel_err(F, rF) < 1e-10
    for sl in (slice(0, 3), slice(3, 9), slice(9, 10)):
        assert rel_err(R[:, sl], rR[:, sl]) < 1e-10
    c2f = np.linalg.inv(hc_oracle.cell_f2c(w.crystal.cell))
    ff = c.calculateFormFactors(w.lcs, w.hkl)
    assert rel_err(ff, hc_oracle.form_factors(w.crystal, w.params, w.lcs, w.hkl.astype(np.float64) @ c2f)) < 1e-10
    # type derivatives against central differences of the oracle's target
    sel = DerivativesSelector(True, True, True, False, True)
    c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf, None, sel)
    td = c.typeDerivatives()
    assert td.p_val.shape == (4,)

    def T(params):
        F = hc_oracle.sf_and_derivs(w.crystal, params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf)[0]
        return float(np.sum((np.conj(w.dtdf) * F).real))

    import copy
    h = 1e-5
    got, want = [], []
    for t, tp in enumerate(w.params.type_parameters):
        for name, arr in (("p_val", td.p_val), ("kappa_spherical_valence", td.kappa),
                          ("kappa_deformation_valence", td.kappa_prime)):
            p1, p2 = copy.deepcopy(w.params), copy.deepcopy(w.params)
            setattr(p1.type_parameters[t], name, getattr(tp, name) + h)
            setattr(p2.type_parameters[t], name, getattr(tp, name) - h)
            want.append((T(p1) - T(p2)) / (2 * h))
            got.append(arr[t])
    assert rel_err(np.array(got), np.array(want)) < 1e-6
    # new type parameters reach the split engine types too
    tp = w.params.type_parameters
    for t in tp:
        t.p_val *= 0.9
        t.kappa_spherical_valence *= 1.02
    c.setTypeParameters(np.array([t.plm25() for t in tp]), [t.p_val for t in tp],
                        [t.kappa_spherical_valence for t in tp], [t.kappa_deformation_valence for t in tp])
    f, d = c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    assert max(_errs(f, d, hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl,
                                                   w.dtdf))) < 1e-10


def test_batch_with_shared_types():
    from discamb_amd import BatchCalculator
    ws = [_mixed_model("P21/c", 30), _mixed_model("P212121", 31)]
    calcs = [HansenCoppensStructureFactorCalculator(w.crystal, w.params) for w in ws]
    print("case shared-type batch", flush=True)
    b = BatchCalculator(calcs)
    for w, c in zip(ws, calcs):
        c.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    b.run(graph=False)
    for w, c in zip(ws, calcs):
        f, d = c.fetch()
        assert max(_errs(f, d, hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl,
                                                       w.dtdf))) < 1e-10
