"""Per-reflection derivative rows (the reference's per-hkl path) on the GPU against the oracle.

Tolerances (normwise max|d - d_ref| / max|d_ref| per quantity): F and every derivative class
1e-10 against the FP64 oracle; sub-ranges, the callback driver and the device-pointer entry must
reproduce the full host result exactly (the rows do not depend on the blocking).  In the mixed
calculation mode the terms are FP32 (FP64 phase reduction, FP64 rows): 1e-6 for F, 1e-5 for the
derivative classes.
"""
import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import HansenCoppensStructureFactorCalculator, synth
from test_oracle_per_hkl import CASES, _anomalous, _structure

pytestmark = pytest.mark.gpu


def _check_rows(F, R, rF, rR, adp_size, tol=1e-10):
    errs = {"F": rel_err(F, rF), "xyz": rel_err(R[:, 0:3], rR[:, 0:3]), "adp": rel_err(R[:, 3:9], rR[:, 3:9]),
            "occ": rel_err(R[:, 9], rR[:, 9])}
    assert max(errs.values()) <= tol, errs
    for i, na in enumerate(adp_size):
        assert np.all(R[:, 3 + na:9, i] == 0)
    return errs


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("cartesian", [False, True])
def test_rows_match_oracle(case, cartesian):
    w = _structure(case, n_atoms=9, d_min=1.0, seed=21, cartesian=cartesian)
    an = _anomalous(w.crystal.atoms.n, seed=4)
    print("case rows %s cartesian=%s atoms=%d hkl=%d" % (case, cartesian, w.crystal.atoms.n, len(w.hkl)),
          flush=True)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    F, R = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl, anomalous=an)
    rF, rR = hc_oracle.per_hkl(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, anomalous=an, n_threads=8)
    print("  ", _check_rows(F, R, rF, rR, w.crystal.atoms.adp_size), flush=True)


def _rows_on_device(c, k0, k1):
    """dh_hc_jacobian_rows_dev into device memory from the HIP runtime the library itself uses."""
    import ctypes
    from discamb_amd import _lib
    hip = ctypes.CDLL("libamdhip64.so.7")
    nh, n = c._jac_shape
    m = k1 - k0
    fb, rb = 16 * m, 16 * m * 10 * n
    pf, pr = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(pf), ctypes.c_size_t(fb)) == 0
    assert hip.hipMalloc(ctypes.byref(pr), ctypes.c_size_t(rb)) == 0
    try:
        _lib.check(c._lib.dh_hc_jacobian_rows_dev(c._h, k0, k1, pf, pr, None))
        f = np.zeros(m, dtype=np.complex128)
        r = np.zeros((m, 10, n), dtype=np.complex128)
        # hipMemcpy on the null stream, which waits for the rows
        assert hip.hipMemcpy(ctypes.c_void_p(f.ctypes.data), pf, ctypes.c_size_t(fb), 2) == 0
        assert hip.hipMemcpy(ctypes.c_void_p(r.ctypes.data), pr, ctypes.c_size_t(rb), 2) == 0
    finally:
        hip.hipFree(pf)
        hip.hipFree(pr)
    return f, r


def test_rows_excluded_atoms_blocks_and_device_entry():
    w = _structure("P21/c", n_atoms=20, d_min=0.9, seed=8)
    inc = np.ones(20, dtype=bool)
    inc[[2, 7, 19]] = False
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    print("case rows excluded atoms / blocks / device", flush=True)
    F, R = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl, countAtomContribution=inc)
    rF, rR = hc_oracle.per_hkl(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, include=inc, n_threads=8)
    _check_rows(F, R, rF, rR, w.crystal.atoms.adp_size)
    assert np.all(R[:, :, ~inc] == 0)
    m = len(w.hkl)
    for k0, k1 in [(0, 1), (5, 6), (13, 400), (m - 7, m), (0, m)]:
        f2, r2 = c.jacobianRows(k0, k1)
        assert np.array_equal(f2, F[k0:k1]) and np.array_equal(r2, R[k0:k1])
    f3, r3 = _rows_on_device(c, 17, m - 3)
    assert np.array_equal(f3, F[17:m - 3]) and np.array_equal(r3, R[17:m - 3])
    f4, r4 = c.jacobianRows(3, 3)
    assert f4.shape == (0,) and r4.shape == (0, 10, 20)
    with pytest.raises(RuntimeError):
        c.jacobianRows(0, m + 1)


def _mixed_errs(F, R, rF, rR):
    return {"F": rel_err(F, rF), "xyz": rel_err(R[:, 0:3], rR[:, 0:3]), "adp": rel_err(R[:, 3:9], rR[:, 3:9]),
            "occ": rel_err(R[:, 9], rR[:, 9])}


@pytest.mark.parametrize("case", CASES)
def test_mixed_mode_rows_match_oracle(case):
    w = _structure(case, n_atoms=12, d_min=1.0, seed=7)
    an = _anomalous(w.crystal.atoms.n, 3)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setCalculationMode("HIP_MIXED")
    print("case mixed rows %s" % case, flush=True)
    F, R = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl, anomalous=an)
    rF, rR = hc_oracle.per_hkl(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, anomalous=an)
    e = _mixed_errs(F, R, rF, rR)
    print("  ", e, flush=True)
    assert e["F"] < 1e-6 and max(e.values()) < 1e-5


def test_mixed_rows_follow_mode_switches_and_new_inputs():
    w = _structure("P212121", n_atoms=12, d_min=1.0, seed=2)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    F, R = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl)
    c.setCalculationMode("HIP_MIXED")
    # a summed mixed-precision run in between rewrites the engine's tables (parity-folded records)
    c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    F2, R2 = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl)
    e = _mixed_errs(F2, R2, F, R)
    assert e["F"] < 1e-6 and max(e.values()) < 1e-5 and e["F"] > 0
    a = w.crystal.atoms.copy()
    a.xyz = a.xyz + 0.01
    F3, R3 = c.calculateJacobian(a, w.lcs, w.hkl)
    rF, rR = hc_oracle.per_hkl(w.crystal, w.params, a, w.lcs, w.hkl)
    e = _mixed_errs(F3, R3, rF, rR)
    assert e["F"] < 1e-6 and max(e.values()) < 1e-5
    # back to FP64: the FP64 rows again, exactly
    c.setCalculationMode("HIP")
    F4, R4 = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl)
    assert np.array_equal(F, F4) and np.array_equal(R, R4)


class _Collector:
    def __init__(self, n, m, dtdf):
        self.f = np.zeros(m, dtype=np.complex128)
        self.g = np.zeros((n, 10))
        self.order = []
        self.dtdf = dtdf
        self.first = None

    def onPerHklCalculation(self, k, f, d):
        self.order.append(k)
        self.f[k] = f
        if self.first is None:
            self.first = (d.atomicPostionDerivatives.copy(), d.adpDerivatives.copy(), d.occupancyDerivatives.copy())
        c = np.conj(self.dtdf[k])
        self.g[:, 0:3] += np.real(c * d.atomicPostionDerivatives)
        self.g[:, 3:9] += np.real(c * d.adpDerivatives)
        self.g[:, 9] += np.real(c * d.occupancyDerivatives)


@pytest.mark.parametrize("case,n_atoms,n_hkl", [("C2/c", 15, None), ("P212121", 600, 1500)])
def test_per_hkl_processor_contracts_to_the_summed_derivatives(case, n_atoms, n_hkl):
    """The reference's per-hkl overload with a gradient-collecting processor.  The 600-atom case
    needs several staging blocks, so the double-buffered driver is exercised."""
    if n_atoms < 100:
        w = _structure(case, n_atoms=n_atoms, d_min=1.0, seed=5)
    else:
        w = synth.small_structure(case, n_atoms=n_atoms, d_min=1.6, seed=5, cell=(30.0, 32.0, 34.0, 90.0, 90.0, 90.0))
    hkl = w.hkl if n_hkl is None else w.hkl[:n_hkl]
    dtdf = w.dtdf[:len(hkl)]
    n, m = w.crystal.atoms.n, len(hkl)
    print("case per-hkl processor %s atoms=%d hkl=%d" % (case, n, m), flush=True)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    col = _Collector(n, m, dtdf)
    assert c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, hkl, col) is None
    assert col.order == list(range(m))
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, hkl, dtdf, n_threads=8)
    assert rel_err(col.f, rf) < 1e-10
    assert rel_err(col.g[:, 0:3], rx) < 1e-10 and rel_err(col.g[:, 3:9], ru) < 1e-10
    assert rel_err(col.g[:, 9], ro) < 1e-10
    F, R = c.calculateJacobian(w.crystal.atoms, w.lcs, hkl)
    assert np.array_equal(col.first[0], R[0, 0:3].T) and np.array_equal(col.first[2], R[0, 9])


def test_per_hkl_processor_exception_stops_the_calculation():
    w = _structure("P-1", n_atoms=6, d_min=1.2)

    class Stop(Exception):
        pass

    class P:
        calls = 0

        def onPerHklCalculation(self, k, f, d):
            P.calls += 1
            if k == 4:
                raise Stop()

    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    with pytest.raises(Stop):
        c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, P())
    assert P.calls == 5
    # the calculator stays usable
    F, R = c.calculateJacobian(w.crystal.atoms, w.lcs, w.hkl)
    assert np.isfinite(R).all()
