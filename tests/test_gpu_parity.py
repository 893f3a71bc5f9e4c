"""GPU parity: the HIP engine through the C ABI against the FP64 oracle on the same seeded inputs.

Tolerances (normwise max|d - d_ref| / max|d_ref| per output class):
  FP64 drop-in ("HIP")        1e-10 for F and every derivative class
  mixed precision ("HIP_MIXED") 1e-6 for F, 1e-5 for derivatives
"""
import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import (AtomSet, DerivativesSelector, HansenCoppensStructureFactorCalculator, synth)

pytestmark = pytest.mark.gpu

TOL = {"HIP": (1e-10, 1e-10), "HIP_MIXED": (1e-6, 1e-5)}


def _compare(w, mode, atoms=None, include=None, anomalous=None, hkl=None, dtdf=None, tol=None, general=False,
             radial="auto"):
    atoms = atoms or w.crystal.atoms
    hkl = w.hkl if hkl is None else hkl
    dtdf = w.dtdf if dtdf is None else dtdf
    print("case %s mode=%s general=%s radial=%s atoms=%d hkl=%d" % (w.name, mode, general, radial, atoms.n, len(hkl)),
          flush=True)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    calc.setKernelOptions(general_rotations=general, radial=radial)
    f, d = calc.calculateStructureFactorsAndDerivatives(atoms, w.lcs, hkl, dtdf, include, anomalous=anomalous)
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, atoms, w.lcs, hkl, dtdf, include=include,
                                             n_threads=16, anomalous=anomalous)
    tf, td = tol or TOL[mode]
    errs = {"F": rel_err(f, rf), "dxyz": rel_err(d.atomic_position_derivatives, rx),
            "dadp": rel_err(d.adp_derivatives, ru), "docc": rel_err(d.occupancy_derivatives, ro)}
    print("  errors", errs, flush=True)
    assert errs["F"] <= tf
    for k in ("dxyz", "dadp", "docc"):
        assert errs[k] <= td, k
    return f, d


# "general" forces the general-rotation kernels; triclinic, monoclinic and orthorhombic groups
# otherwise run the parity-folded kernels for diagonal operations (P43212 and P31 always general).
# radial: per-(reflection, type) records or per-atom lookups in per-wavefunction tables (the
# default for these one-type-per-atom structures)
@pytest.mark.parametrize("radial", ["auto", "records"])
@pytest.mark.parametrize("general", [False, True])
@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
@pytest.mark.parametrize("sg", list(synth.SPACE_GROUPS))
def test_space_groups(sg, mode, general, radial):
    _compare(synth.small_structure(sg, n_atoms=14, d_min=0.9), mode, general=general, radial=radial)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_shifted_inversion_centre(mode):
    w = synth.small_structure("P-1", n_atoms=10, d_min=1.0)
    w.crystal.symops = ["x,y,z", "-x+1/2,-y,-z+1/2"]
    _compare(w, mode)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_body_centred(mode):
    w = synth.small_structure("P212121", n_atoms=8, d_min=1.1)
    w.crystal.symops = ["x,y,z", "-x,-y,z", "-x,y,-z", "x,-y,-z", "x+1/2,y+1/2,z+1/2", "-x+1/2,-y+1/2,z+1/2",
                        "-x+1/2,y+1/2,-z+1/2", "x+1/2,-y+1/2,-z+1/2"]
    w.crystal.atoms.multiplicity[:] = 8
    _compare(w, mode)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_config0_full(mode):
    """60-atom P2_1/c structure, all 15 300 unique reflections to 0.5 A."""
    _compare(synth.config0(), mode)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
@pytest.mark.parametrize("sg,ops", [("Pna21", ["x,y,z", "-x,-y,z+1/2", "x+1/2,-y+1/2,z", "-x+1/2,y+1/2,z+1/2"]),
                                    ("Pm2m", ["x,y,z", "-x,y,-z", "x,y,-z", "-x,y,z"]),
                                    ("Pc", ["x,y,z", "x,-y,z+1/2"]),
                                    ("P212121-reordered", ["x,y,z", "-x,y+1/2,-z+1/2", "x+1/2,-y+1/2,-z",
                                                           "-x+1/2,-y,z+1/2"])])
def test_compiled_sign_sets_table_lookups(sg, ops, mode):
    """Mirror and mm2 sign sets and a reordered 222 listing on the per-atom table path (the mixed
    parity-folded kernels of the compiled sets with every term direct; FP64: the general table
    kernel) against the oracle."""
    w = synth.small_structure("P212121", n_atoms=16, d_min=0.9, seed=29)
    w.crystal.symops = ops
    w.crystal.atoms.multiplicity[:] = len(ops)
    w.name = "small_" + sg
    _compare(w, mode, radial="tables")


@pytest.mark.parametrize("radial", ["auto", "tables"])
@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_config1_reflection_subset(mode, radial):
    """5000-atom P2_1 2_1 2_1 structure, all atoms, a 3000-reflection subset spanning resolution."""
    w = synth.config1()
    rng = np.random.default_rng(5)
    sel = np.sort(rng.choice(len(w.hkl), 3000, replace=False))
    _compare(w, mode, hkl=w.hkl[sel], dtdf=w.dtdf[sel], radial=radial)


def test_include_mask_anomalous_and_h000():
    w = synth.small_structure("P21/c", n_atoms=12, d_min=1.2)
    inc = np.ones(12, dtype=bool)
    inc[[1, 5]] = False
    rng = np.random.default_rng(8)
    anom = rng.normal(size=12) * 0.3 + 1j * np.abs(rng.normal(size=12)) * 0.2
    hkl = np.vstack([[[0, 0, 0]], w.hkl])
    dtdf = np.concatenate([[0.7 - 0.2j], w.dtdf])
    f, d = _compare(w, "HIP", include=inc, anomalous=anom, hkl=hkl, dtdf=dtdf)
    assert np.all(d.atomic_position_derivatives[~inc] == 0)
    assert np.all(d.occupancy_derivatives[~inc] == 0)


def test_acentric_anomalous():
    w = synth.small_structure("P212121", n_atoms=9, d_min=1.2)
    anom = np.full(9, 0.25 + 0.4j)
    _compare(w, "HIP", anomalous=anom)


def test_cartesian_and_u_cart_conventions():
    w = synth.small_structure("P21/c", n_atoms=10, d_min=1.2)
    f2c, c2f = synth.cell_matrices(w.crystal.cell)
    a = w.crystal.atoms.copy()
    a.xyz = a.xyz @ f2c.T
    J = hc_oracle.u_cif_to_cart_matrix(w.crystal.cell)
    for i in range(a.n):
        if a.adp_size[i] == 6:
            a.adp[i] = J @ a.adp[i]
    w.crystal.atoms = a
    w.crystal.xyz_convention = "cartesian"
    w.crystal.adp_convention = "U_cart"
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    f, d = calc.calculateStructureFactorsAndDerivatives(a, w.lcs, w.hkl, w.dtdf)
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, a, w.lcs, w.hkl, w.dtdf)
    assert rel_err(f, rf) < 1e-10
    assert rel_err(d.atomic_position_derivatives, rx) < 1e-10
    assert rel_err(d.adp_derivatives, ru) < 1e-10
    assert rel_err(d.occupancy_derivatives, ro) < 1e-10


def test_f_only_and_selector():
    w = synth.small_structure("P1", n_atoms=6, d_min=1.2)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    f = calc.calculateStructureFactors(w.crystal.atoms, w.lcs, w.hkl)
    rf = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, np.ones(len(w.hkl)))[0]
    assert rel_err(f, rf) < 1e-10
    f2, d = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf, None,
                                                         DerivativesSelector(True, False, False))
    assert np.all(d.adp_derivatives == 0) and np.all(d.occupancy_derivatives == 0)
    assert np.any(d.atomic_position_derivatives != 0)


def test_empty_reflections_and_errors():
    w = synth.small_structure("P1", n_atoms=4, d_min=1.5)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    f, d = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, np.zeros((0, 3), int),
                                                        np.zeros(0, complex))
    assert f.shape == (0,)
    assert np.all(d.atomic_position_derivatives == 0)
    with pytest.raises(RuntimeError):
        calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf[:-1])
    with pytest.raises(RuntimeError):
        calc.setCalculationMode("NOPE")


def test_determinism_full_config1():
    """Full 5000-atom / 66 592-reflection workload: two runs are bitwise identical (fixed-order,
    atomic-free reductions) and FP64 / mixed agree to the mixed tolerance."""
    w = synth.config1()
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    f1, d1 = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    f2, d2 = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    assert np.array_equal(f1, f2)
    assert np.array_equal(d1.adp_derivatives, d2.adp_derivatives)
    calc.setCalculationMode("HIP_MIXED")
    fm, dm = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    e = rel_err(fm, f1)
    print("full config1 mixed vs fp64: F %.3e dxyz %.3e dadp %.3e" % (
        e, rel_err(dm.atomic_position_derivatives, d1.atomic_position_derivatives),
        rel_err(dm.adp_derivatives, d1.adp_derivatives)), flush=True)
    assert e < 1e-6


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_autotune_before_first_run_and_launch_shapes(mode):
    """The autotuner runs on freshly uploaded inputs (before any run), leaves no result to fetch,
    and every reflection-chunk size (ragged last chunk, one chunk for everything) gives the
    oracle's answer."""
    w = synth.small_structure("P212121", n_atoms=150, d_min=0.8)
    print("case autotune mode=%s atoms=%d hkl=%d" % (mode, w.crystal.atoms.n, len(w.hkl)), flush=True)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    calc.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    assert calc.autotune() > 0
    with pytest.raises(RuntimeError):
        calc.fetch()
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, n_threads=16)
    tf, td = TOL[mode]
    for chunk in (0, 8, 13, 64, len(w.hkl)):
        calc.setTuning(chunk)
        calc.run()
        f, d = calc.fetch()
        errs = (rel_err(f, rf), rel_err(d.atomic_position_derivatives, rx), rel_err(d.adp_derivatives, ru),
                rel_err(d.occupancy_derivatives, ro))
        print("  chunk %d errors %s" % (chunk, errs), flush=True)
        assert errs[0] <= tf and max(errs[1:]) <= td


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_one_calculator_across_changing_reflection_lists(mode):
    """One calculator called with reflection lists that grow, shrink, reorder, empty and repeat
    (its host staging, |h| order and reflection tables are reused or rebuilt between calls) gives
    bitwise the results of a fresh calculator for each list."""
    w = synth.small_structure("P212121", n_atoms=40, d_min=0.9)
    rng = np.random.default_rng(7)
    n = len(w.hkl)
    sub = rng.permutation(n)[: n // 3]
    lists = [(w.hkl, w.dtdf), (w.hkl[sub], w.dtdf[sub]), (w.hkl, w.dtdf), (w.hkl, w.dtdf * 2.0),
             (w.hkl[:0], w.dtdf[:0]), (w.hkl[::-1], w.dtdf[::-1]), (w.hkl, w.dtdf * 2.0)]
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    for i, (h, g) in enumerate(lists):
        print("case reflection-list sequence mode=%s call=%d hkl=%d" % (mode, i, len(h)), flush=True)
        f, d = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, h, g)
        fresh = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
        fresh.setCalculationMode(mode)
        rf, rd = fresh.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, h, g)
        assert f.shape == (len(h),)
        assert np.array_equal(f, rf)
        assert np.array_equal(d.atomic_position_derivatives, rd.atomic_position_derivatives)
        assert np.array_equal(d.adp_derivatives, rd.adp_derivatives)
        assert np.array_equal(d.occupancy_derivatives, rd.occupancy_derivatives)
    # the last list against the oracle as well
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf * 2.0,
                                             n_threads=16)
    tf, td = TOL[mode]
    assert rel_err(f, rf) <= tf and rel_err(d.adp_derivatives, ru) <= td


def _type_arrays(params):
    tp = params.type_parameters
    return (np.array([t.plm25() for t in tp]), [t.p_val for t in tp], [t.kappa_spherical_valence for t in tp],
            [t.kappa_deformation_valence for t in tp])


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_one_calculator_across_changing_model_state(mode):
    """With the reflection list held fixed (so the reflection records, radial tables and radial
    records are kept between calls), every change of the model state the cached prep depends on -
    new type parameters (kappa, kappa', Pval, P_lm), electron scattering on and off, forced kernel
    paths - gives bitwise the result of a fresh calculator set up the same way."""
    import copy
    w = synth.small_structure("P212121", n_atoms=40, d_min=0.9)
    base = _type_arrays(w.params)
    p2 = copy.deepcopy(w.params)
    rng = np.random.default_rng(12)
    for t in p2.type_parameters:
        t.kappa_spherical_valence *= 1.03
        t.kappa_deformation_valence *= 0.97
        t.p_val *= 0.98
        for row in t.p_lm:
            for m in range(len(row)):
                row[m] += rng.normal() * 0.02
    changed = _type_arrays(p2)
    # (type parameters, electron scattering, kernel options) before each call
    states = [(base, False, {}), (changed, False, {}), (changed, True, {}), (changed, False, {}),
              (changed, False, {"general_rotations": True}), (changed, False, {"radial": "records"}),
              (changed, False, {"radial": "tables"}), (base, False, {}), (base, True, {"radial": "records"})]
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    for i, (types, electron, opts) in enumerate(states):
        print("case model-state sequence mode=%s call=%d electron=%s options=%s" % (mode, i, electron, opts),
              flush=True)
        for c in (calc,):
            c.setTypeParameters(*types)
            c.setElectronScattering(electron)
            c.setKernelOptions(**opts)
        f, d = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
        fresh = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
        fresh.setCalculationMode(mode)
        fresh.setTypeParameters(*types)
        fresh.setElectronScattering(electron)
        fresh.setKernelOptions(**opts)
        rf, rd = fresh.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
        assert np.array_equal(f, rf), i
        assert np.array_equal(d.atomic_position_derivatives, rd.atomic_position_derivatives), i
        assert np.array_equal(d.adp_derivatives, rd.adp_derivatives), i
        assert np.array_equal(d.occupancy_derivatives, rd.occupancy_derivatives), i
    # the changed parameters against the oracle as well (X-ray, default kernels)
    calc.setTypeParameters(*changed)
    calc.setElectronScattering(False)
    calc.setKernelOptions()
    f, d = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, p2, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, n_threads=16)
    tf, td = TOL[mode]
    assert rel_err(f, rf) <= tf and rel_err(d.adp_derivatives, ru) <= td


def test_failed_upload_leaves_no_stale_inputs():
    """An upload that fails (here: too few dT/dF values) leaves the calculator without inputs:
    run() and fetch() refuse until an upload succeeds, and the next good call matches a fresh
    calculator bitwise."""
    w = synth.small_structure("P21/c", n_atoms=12, d_min=1.0)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    f0, d0 = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    bad = w.crystal.atoms.copy()
    bad.adp_size = bad.adp_size.copy()
    bad.adp_size[0] = 3
    with pytest.raises(RuntimeError):
        calc.upload(bad, w.lcs, w.hkl[:10], w.dtdf[:10])
    with pytest.raises(RuntimeError):
        calc.run()
    with pytest.raises(RuntimeError):
        calc.fetch()
    f1, d1 = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    assert np.array_equal(f0, f1)
    assert np.array_equal(d0.adp_derivatives, d1.adp_derivatives)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
@pytest.mark.parametrize("sg", ["P1", "P-1", "P21/c", "C2/c", "P212121", "P31", "P-1_shifted", "I222", "P2_1-c",
                                "P2_1-a", "Pmm2", "Pna21", "P2mm", "Pm2m", "Pc", "Pm-x", "Pm-z", "P212121-reordered",
                                "P2/c-shifted"])
def test_line_recurrences_vs_direct_and_oracle(sg, mode):
    """Terms carried along hkl lines (groups with <= 4 explicit operations, radial records) against
    direct evaluation of every phase and Debye-Waller factor and against the oracle, with
    reflection lists that break lines (a random subset) and a short reflection chunk (many chunk
    starts).  FP64: line vs direct 1e-12, oracle 1e-10.  Mixed precision (FP32 products along
    the lines, direct evaluation every fourth reflection): both against the oracle at 1e-6 (F) and
    1e-5 (derivatives), and the line path no further from the oracle than twice the direct path's
    error plus 2e-7."""
    # every compiled sign-pattern set: 2-fold screw axes along c and along a, mirrors normal to y
    # (Pc), x and z, mm2 with the 2-fold along z (Pmm2, and Pna2_1 with translations), x and y, the
    # 222 set listed in another order (the engine sorts the operations); and a diagonal
    # four-operation set without compiled kernels (2/m with the inversion centre off the origin)
    extra = {"P2_1-c": ["x,y,z", "-x,-y,z+1/2"], "P2_1-a": ["x,y,z", "x+1/2,-y,-z"],
             "Pmm2": ["x,y,z", "-x,-y,z", "x,-y,z", "-x,y,z"],
             "Pna21": ["x,y,z", "-x,-y,z+1/2", "x+1/2,-y+1/2,z", "-x+1/2,y+1/2,z+1/2"],
             "P2mm": ["x,y,z", "x,-y,-z", "x,y,-z", "x,-y,z"], "Pm2m": ["x,y,z", "-x,y,-z", "x,y,-z", "-x,y,z"],
             "Pc": ["x,y,z", "x,-y,z+1/2"], "Pm-x": ["x,y,z", "-x,y,z"], "Pm-z": ["x,y,z", "x,y,-z"],
             "P212121-reordered": ["x,y,z", "x+1/2,-y+1/2,-z", "-x,y+1/2,-z+1/2", "-x+1/2,-y,z+1/2"],
             "P2/c-shifted": ["x,y,z", "-x,y,-z", "-x+1/2,-y,-z+1/2", "x+1/2,-y,z+1/2"]}
    base = "P212121" if sg in extra or sg == "I222" else sg.split("_")[0]
    w = synth.small_structure(base, n_atoms=20, d_min=0.8, seed=23)
    if sg in extra:
        w.crystal.symops = extra[sg]
        w.crystal.atoms.multiplicity[:] = len(extra[sg])
    if sg == "P-1_shifted":
        w.crystal.symops = ["x,y,z", "-x+1/2,-y,-z+1/2"]
    if sg == "I222":
        w.crystal.symops = ["x,y,z", "-x,-y,z", "-x,y,-z", "x,-y,-z", "x+1/2,y+1/2,z+1/2", "-x+1/2,-y+1/2,z+1/2",
                            "-x+1/2,y+1/2,-z+1/2", "x+1/2,-y+1/2,-z+1/2"]
        w.crystal.atoms.multiplicity[:] = 8
    rng = np.random.default_rng(31)
    sub = np.sort(rng.choice(len(w.hkl), int(0.8 * len(w.hkl)), replace=False))
    for hkl, g, chunk in ((w.hkl, w.dtdf, 0), (w.hkl[sub], w.dtdf[sub], 0), (w.hkl, w.dtdf, 13)):
        print("case line recurrences %s mode=%s hkl=%d chunk=%d" % (sg, mode, len(hkl), chunk), flush=True)
        out = []
        for direct in (False, True):
            calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
            calc.setCalculationMode(mode)
            calc.setKernelOptions(radial="records", direct_phases=direct)
            calc.setTuning(chunk)
            out.append(calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, hkl, g))
        rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, hkl, g, n_threads=16)

        def errs(fd):
            f, d = fd
            return (rel_err(f, rf), rel_err(d.atomic_position_derivatives, rx), rel_err(d.adp_derivatives, ru),
                    rel_err(d.occupancy_derivatives, ro))
        (f, d), (fd, dd) = out
        el, ed = errs(out[0]), errs(out[1])
        print("  vs oracle: line", el, "direct", ed, flush=True)
        if mode == "HIP":
            e = (rel_err(f, fd), rel_err(d.atomic_position_derivatives, dd.atomic_position_derivatives),
                 rel_err(d.adp_derivatives, dd.adp_derivatives),
                 rel_err(d.occupancy_derivatives, dd.occupancy_derivatives))
            print("  line vs direct", e, flush=True)
            assert max(e) <= 1e-12, e
            assert max(el) <= 1e-10, el
        else:
            for e in (el, ed):
                assert e[0] <= 1e-6 and max(e[1:]) <= 1e-5, e
            for a, b in zip(el, ed):
                assert a <= 2 * b + 2e-7, (el, ed)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_ionic_wavefunctions(mode):
    """Ionic Clementi-Roetti entries of the built-in bank (Fe3+, Cu+ without valence orbitals, O-,
    Se-) in place of the neutral atoms' wavefunctions, against the oracle; also with electron
    scattering (atomic numbers of the ions' elements)."""
    from discamb_amd import model
    w = synth.small_structure("P21/c", n_atoms=12, d_min=0.9, seed=41)
    for i, label in enumerate(["Fe3+", "Cu+", "O-", "Se-"]):
        w.params.wfn_parameters[i] = model.wfn_type_from_bank(label)
    _compare(w, mode)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    calc.setElectronScattering(True)
    f, d = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    rf = hc_oracle.electron_sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, n_threads=16)[0]
    assert rel_err(f, rf) <= TOL[mode][0]
