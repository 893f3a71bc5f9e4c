"""GPU parity at the full size of every benchmark workload (HansenCoppensStructureFactorCalculator.cpp:380-508).

* 5000-atom P2_1 2_1 2_1 structure, all 66 592 reflections to 1.2 A: FP64 and mixed precision
  against the FP64 oracle on every reflection.
* 20 000-atom P2_1 2_1 2_1 structure, 80 x 90 x 110 A, all 127 455 reflections to 1.5 A: mixed
  against FP64 on every reflection (max|dF|/max|F|, dR and derivative errors printed), and both
  precisions against the oracle on a 4000-reflection subset with all 20 000 atoms.
* 256 uneven structures (20-400 atoms, P1 / P-1 / P2_1/c / C2/c / P2_1 2_1 2_1, d_min 0.5-0.9 A):
  one batched step (eager and hipGraph-replayed) against one-at-a-time calls for all 256, in both
  precisions, and against the oracle on a stratified subset (the smallest and the largest
  structure of every group plus eight spread over the term counts).

Tolerances (normwise max|d - d_ref| / max|d_ref| per output class):
  FP64 vs oracle                  1e-10
  mixed vs oracle                 1e-6 (F), 1e-5 (derivatives)
  FP64 batch vs single calls      1e-12 (only the summation order of the chunks differs)
  mixed batch vs mixed single     1e-6
  mixed vs FP64 (20 000 atoms)    1e-6 (F), 1e-5 (derivatives)
"""
import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import BatchCalculator, HansenCoppensStructureFactorCalculator, synth

pytestmark = pytest.mark.gpu

TOL = {"HIP": (1e-10, 1e-10), "HIP_MIXED": (1e-6, 1e-5)}


def _errs(f, d, rf, rx, ru, ro):
    return {"F": rel_err(f, rf), "dxyz": rel_err(d.atomic_position_derivatives, rx),
            "dadp": rel_err(d.adp_derivatives, ru), "docc": rel_err(d.occupancy_derivatives, ro)}


def _check(errs, mode):
    tf, td = TOL[mode]
    assert errs["F"] <= tf, errs
    for k in ("dxyz", "dadp", "docc"):
        assert errs[k] <= td, (k, errs)


def _run(w, mode, hkl=None, dtdf=None):
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    hkl = w.hkl if hkl is None else hkl
    dtdf = w.dtdf if dtdf is None else dtdf
    return calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, hkl, dtdf)


@pytest.fixture(scope="module")
def config1_oracle():
    w = synth.config1()
    print("case config1 oracle: %d atoms, %d reflections" % (w.crystal.atoms.n, len(w.hkl)), flush=True)
    ref = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, n_threads=16)
    return w, ref


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_config1_all_reflections_vs_oracle(config1_oracle, mode):
    w, ref = config1_oracle
    print("case config1 full mode=%s atoms=%d hkl=%d" % (mode, w.crystal.atoms.n, len(w.hkl)), flush=True)
    f, d = _run(w, mode)
    errs = _errs(f, d, *ref)
    print("  errors", errs, flush=True)
    _check(errs, mode)


@pytest.fixture(scope="module")
def config2_large():
    w = synth.config2_large()
    print("case config2 20000-atom structure: %d atoms, %d reflections" % (w.crystal.atoms.n, len(w.hkl)), flush=True)
    return w


def test_config2_20000_atoms_mixed_vs_fp64_all_reflections(config2_large):
    w = config2_large
    f64, d64 = _run(w, "HIP")
    fm, dm = _run(w, "HIP_MIXED")
    a64, am = np.abs(f64), np.abs(fm)
    stats = {"maxdF/maxF": rel_err(fm, f64), "dR": float(np.abs(am - a64).sum() / a64.sum()),
             "dxyz": rel_err(dm.atomic_position_derivatives, d64.atomic_position_derivatives),
             "dadp": rel_err(dm.adp_derivatives, d64.adp_derivatives),
             "docc": rel_err(dm.occupancy_derivatives, d64.occupancy_derivatives)}
    print("  20000 atoms, %d reflections, mixed vs FP64: %s" % (len(w.hkl), stats), flush=True)
    assert stats["maxdF/maxF"] <= 1e-6
    assert stats["dR"] <= 1e-6
    for k in ("dxyz", "dadp", "docc"):
        assert stats[k] <= 1e-5, k


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_config2_20000_atoms_vs_oracle_on_reflection_subset(config2_large, mode):
    w = config2_large
    rng = np.random.default_rng(21)
    sel = np.sort(rng.choice(len(w.hkl), 4000, replace=False))
    print("case config2 subset mode=%s atoms=%d hkl=%d" % (mode, w.crystal.atoms.n, len(sel)), flush=True)
    f, d = _run(w, mode, w.hkl[sel], w.dtdf[sel])
    ref = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl[sel], w.dtdf[sel], n_threads=16)
    errs = _errs(f, d, *ref)
    print("  errors", errs, flush=True)
    _check(errs, mode)


@pytest.fixture(scope="module")
def config4_structures():
    ws = synth.config4(256)
    print("case config4: %d structures, %.3g terms" % (len(ws), sum(w.terms for w in ws)), flush=True)
    return ws


def _stratified(ws):
    """The smallest and largest structure (by terms) of every space group plus eight spread over the
    term-count order of the whole batch."""
    idx = set()
    groups = {}
    for i, w in enumerate(ws):
        groups.setdefault(w.name.split("_")[1], []).append(i)
    for g, members in groups.items():
        members = sorted(members, key=lambda i: ws[i].terms)
        idx.update([members[0], members[-1]])
    order = sorted(range(len(ws)), key=lambda i: ws[i].terms)
    for q in np.linspace(0, len(order) - 1, 8):
        idx.add(order[int(round(q))])
    return sorted(idx)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_config4_batch256_vs_single_calls_and_oracle(config4_structures, mode):
    ws = config4_structures
    calcs = []
    single = []
    for w in ws:
        c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
        c.setCalculationMode(mode)
        single.append(c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf))
        calcs.append(c)
    batch = BatchCalculator(calcs)
    for w, c in zip(ws, calcs):
        c.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    tol = 1e-12 if mode == "HIP" else 1e-6
    print("case config4 batch eager mode=%s structures=%d" % (mode, len(ws)), flush=True)
    batch.run(graph=False)
    eager = [c.fetch() for c in calcs]
    worst = 0.0
    for i, w in enumerate(ws):
        f, d = eager[i]
        rf, rd = single[i]
        e = max(_errs(f, d, rf, rd.atomic_position_derivatives, rd.adp_derivatives, rd.occupancy_derivatives).values())
        worst = max(worst, e)
        assert e <= tol, (w.name, e)
    print("  batch vs single: worst normwise error %.3e over %d structures" % (worst, len(ws)), flush=True)
    # captured step: the first call captures and replays, the second replays again
    for rep in range(2):
        print("case config4 batch graph replay %d mode=%s" % (rep, mode), flush=True)
        batch.run(graph=True)
        for i, c in enumerate(calcs):
            f, d = c.fetch()
            ef, ed = eager[i]
            assert np.array_equal(f, ef), (rep, ws[i].name)
            assert np.array_equal(d.atomic_position_derivatives, ed.atomic_position_derivatives), (rep, ws[i].name)
            assert np.array_equal(d.adp_derivatives, ed.adp_derivatives), (rep, ws[i].name)
            assert np.array_equal(d.occupancy_derivatives, ed.occupancy_derivatives), (rep, ws[i].name)
    sub = _stratified(ws)
    print("case config4 oracle subset mode=%s structures=%s" % (mode, sub), flush=True)
    for i in sub:
        w = ws[i]
        ref = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, n_threads=16)
        f, d = eager[i]
        errs = _errs(f, d, *ref)
        print("  %s (%d reflections) vs oracle %s" % (w.name, len(w.hkl), errs), flush=True)
        _check(errs, mode)
