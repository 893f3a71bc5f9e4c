"""Independent-atom model (CPU_IAM) and electron scattering on the GPU against the oracle.

Tolerances (normwise max|d - d_ref| / max|d_ref| per output class): FP64 1e-10; mixed precision
1e-6 for F and 1e-5 for derivatives; multipole derivatives with electron scattering against
central differences of the oracle's electron F (FP64 1e-6).
"""
import copy

import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import DerivativesSelector, HansenCoppensStructureFactorCalculator, synth

pytestmark = pytest.mark.gpu

I222 = ["x,y,z", "-x,-y,z", "-x,y,-z", "x,-y,-z", "x+1/2,y+1/2,z+1/2", "-x+1/2,-y+1/2,z+1/2",
        "-x+1/2,y+1/2,-z+1/2", "x+1/2,-y+1/2,-z+1/2"]


def _structure(case, n_atoms=9, d_min=1.0, seed=31):
    if case == "P-1_shifted":
        w = synth.small_structure("P-1", n_atoms=n_atoms, d_min=d_min, seed=seed)
        w.crystal.symops = ["x,y,z", "-x+1/2,-y,-z+1/2"]
    elif case == "I222":
        w = synth.small_structure("P212121", n_atoms=n_atoms, d_min=d_min, seed=seed)
        w.crystal.symops = I222
        w.crystal.atoms.multiplicity[:] = 8
    else:
        w = synth.small_structure(case, n_atoms=n_atoms, d_min=d_min, seed=seed)
    return w


def _errs(f, d, rf, rx, ru, ro):
    return {"F": rel_err(f, rf), "xyz": rel_err(d.atomic_position_derivatives, rx),
            "adp": rel_err(d.adp_derivatives, ru), "occ": rel_err(d.occupancy_derivatives, ro)}


CASES = ["P1", "P-1", "P21/c", "C2/c", "P212121", "P43212", "P-1_shifted", "I222"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("mode", ["CPU_IAM", "HIP_IAM_MIXED"])
def test_iam_matches_oracle(case, mode):
    w = _structure(case)
    rng = np.random.default_rng(3)
    an = rng.uniform(-0.4, 0.4, w.crystal.atoms.n) + 1j * rng.uniform(0.1, 1.5, w.crystal.atoms.n)
    print("case iam %s %s hkl=%d" % (case, mode, len(w.hkl)), flush=True)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setCalculationMode(mode)
    f, d = c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf, anomalous=an)
    ref = hc_oracle.iam_sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.hkl, w.dtdf, anomalous=an,
                                      n_threads=8)
    e = _errs(f, d, *ref)
    print("  ", e, flush=True)
    if mode == "CPU_IAM":
        assert max(e.values()) < 1e-10
    else:
        assert e["F"] < 1e-6 and max(e.values()) < 1e-5


def test_iam_rows_mode_switch_and_form_factor_refusal():
    w = _structure("P21/c", n_atoms=8)
    an = 0.1 + 0.7j * np.ones(8)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    fx = c.calculateStructureFactors(w.crystal.atoms, w.lcs, w.hkl)
    c.setCalculationMode("CPU_IAM")
    F, R = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl, anomalous=an)
    rF, rR = hc_oracle.iam_per_hkl(w.crystal, w.params, w.crystal.atoms, w.hkl, anomalous=an)
    assert rel_err(F, rF) < 1e-10
    for sl in (slice(0, 3), slice(3, 9), slice(9, 10)):
        assert rel_err(R[:, sl], rR[:, sl]) < 1e-10
    with pytest.raises(RuntimeError):
        c.calculateFormFactors(w.lcs, w.hkl[:3])
    with pytest.raises(RuntimeError):
        c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf, None,
                                                  DerivativesSelector(d_multipole=True))
    c.setCalculationMode("HIP")
    assert rel_err(c.calculateStructureFactors(w.crystal.atoms, w.lcs, w.hkl), fx) == 0.0


@pytest.mark.parametrize("case", ["P1", "P-1", "P21/c", "P212121", "P43212", "P-1_shifted", "I222"])
@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_electron_matches_oracle(case, mode):
    w = _structure(case, seed=17)
    hkl = np.vstack([np.zeros((1, 3), dtype=w.hkl.dtype), w.hkl])      # (000) included
    dtdf = np.concatenate([[0.5 - 0.2j], w.dtdf])
    print("case electron %s %s" % (case, mode), flush=True)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setCalculationMode(mode)
    c.setElectronScattering(True)
    f, d = c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, hkl, dtdf)
    ref = hc_oracle.electron_sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, hkl, dtdf, n_threads=8)
    e = _errs(f, d, *ref)
    print("  ", e, flush=True)
    if mode == "HIP":
        assert max(e.values()) < 1e-10
    else:
        assert e["F"] < 1e-6 and max(e.values()) < 1e-5
    # switching off returns the X-ray result
    c.setElectronScattering(False)
    fx = c.calculateStructureFactors(w.crystal.atoms, w.lcs, hkl)
    rx = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, hkl, dtdf)[0]
    assert rel_err(fx, rx) < (1e-10 if mode == "HIP" else 1e-6)


def test_electron_multipole_derivatives_vs_finite_differences():
    w = _structure("P21/c", n_atoms=6, d_min=1.3, seed=5)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setElectronScattering(True)
    sel = DerivativesSelector(True, True, True, False, True)
    f, d = c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf, None, sel)
    td = c.typeDerivatives()

    def T(params):
        F = hc_oracle.electron_sf_and_derivs(w.crystal, params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf)[0]
        return float(np.sum((np.conj(w.dtdf) * F).real))

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
        for l, row in enumerate(tp.p_lm):
            for m in range(len(row)):
                p1, p2 = copy.deepcopy(w.params), copy.deepcopy(w.params)
                p1.type_parameters[t].p_lm[l][m] += h
                p2.type_parameters[t].p_lm[l][m] -= h
                want.append((T(p1) - T(p2)) / (2 * h))
                got.append(td.p_lm[t, l * l + m])
    assert rel_err(np.array(got), np.array(want)) < 1e-6


def test_electron_and_iam_are_refused_in_batches():
    from discamb_amd import BatchCalculator
    w = _structure("P212121", n_atoms=6)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setElectronScattering(True)
    with pytest.raises(RuntimeError):
        BatchCalculator([c])
    c.setElectronScattering(False)
    c.setCalculationMode("CPU_IAM")
    with pytest.raises(RuntimeError):
        BatchCalculator([c])


def test_iam_tables_by_name():
    """setIamTable: IT92 (AnyIamCalculator's table) and electron-IT against the oracle with the same
    table; an unknown name is refused."""
    w = _structure("P21/c", n_atoms=8)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setCalculationMode("CPU_IAM")
    for table in ("IT92", "electron-IT", "Waasmeier-Kirfel"):
        print("case iam table %s" % table, flush=True)
        c.setIamTable(table)
        f, d = c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
        ref = hc_oracle.iam_sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.hkl, w.dtdf, table=table)
        assert max(_errs(f, d, *ref).values()) < 1e-10
    with pytest.raises(RuntimeError):
        c.setIamTable("no such table")
