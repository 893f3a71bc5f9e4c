"""Pins of the independent-atom (CPU_IAM) and electron-scattering restatements in the oracle (CPU).

  * the N-Gaussian tables (data file) and lookup rule against the reference's own
    NGaussianFormFactorsTable compiled into oracle/_ref, every entry of all four tables;
  * IAM: F against the explicit full-group sum with f'' != 0, and every derivative against central
    differences of the full-group target;
  * electron: the Mott-Bethe identity F_e = 0.023934 / s^2 (F_Z - F_x) between the electron F, the
    pinned X-ray F and an independent numpy sum F_Z over the group with form factor Z, plus the
    derivatives against central differences.
"""
import ctypes

import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import synth

TABLES = ["Waasmeier-Kirfel", "IT92", "electron-IT", "electron-cctbx"]


def test_ngaussian_tables_match_the_reference(ref_lib):
    h = np.array([0.0, 0.31, 1.1, 2.7, 6.0])
    out = np.zeros(len(h))
    P = ctypes.c_void_p
    n = 0
    for table in TABLES:
        labels = list(hc_oracle.ngauss_tables()[table])
        for label in labels + ["O-", "Cl-", "Fe+", "Xx"]:
            ok = ref_lib.ref_ngauss(label.encode(), table.encode(), P(h.ctypes.data), ctypes.c_int(len(h)),
                                    P(out.ctypes.data))
            mine = hc_oracle.ngauss(label, table, h)
            if not ok:
                assert mine is None, (table, label)
                continue
            assert mine is not None, (table, label)
            assert np.abs(mine - out).max() <= 1e-13 * max(1.0, np.abs(out).max()), (table, label)
            n += 1
    assert n > 600


def _target(w, atoms, dtdf, an, fn):
    return float(np.sum((np.conj(dtdf) * fn(atoms)).real))


@pytest.mark.parametrize("sg", ["P1", "P-1", "P21/c", "C2/c", "P212121", "P-1_shifted"])
def test_iam_oracle_matches_full_group_and_finite_differences(sg):
    shifted = sg == "P-1_shifted"
    w = synth.small_structure("P-1" if shifted else sg, n_atoms=4, d_min=1.5, seed=2)
    if shifted:
        w.crystal.symops = ["x,y,z", "-x+1/2,-y,-z+1/2"]
    an = np.random.default_rng(7).uniform(-0.3, 0.3, 4) + 1j * np.random.default_rng(8).uniform(0.2, 1.5, 4)
    atoms = w.crystal.atoms
    F, dx, du, do = hc_oracle.iam_sf_and_derivs(w.crystal, w.params, atoms, w.hkl, w.dtdf, anomalous=an)
    fg = lambda a: hc_oracle.iam_full_group_f(w.crystal, w.params, a, w.hkl, anomalous=an)
    assert rel_err(F, fg(atoms)) < 1e-13
    h = 1e-6
    for i in range(atoms.n):
        cases = [("xyz", j, dx[i, j]) for j in range(3)] + [("adp", j, du[i, j]) for j in range(atoms.adp_size[i])]
        cases.append(("occupancy", None, do[i]))
        for field, j, val in cases:
            a1, a2 = atoms.copy(), atoms.copy()
            if j is None:
                getattr(a1, field)[i] += h
                getattr(a2, field)[i] -= h
            else:
                getattr(a1, field)[i, j] += h
                getattr(a2, field)[i, j] -= h
            fd = (_target(w, a1, w.dtdf, an, fg) - _target(w, a2, w.dtdf, an, fg)) / (2 * h)
            scale = max(np.abs(dx).max(), np.abs(du).max(), np.abs(do).max())
            assert abs(fd - val) < 2e-6 * scale, (sg, i, field, j, fd, val)


@pytest.mark.parametrize("sg", ["P1", "P-1", "P21/c", "P212121", "P-1_shifted"])
def test_electron_oracle_mott_bethe_identity(sg):
    shifted = sg == "P-1_shifted"
    w = synth.small_structure("P-1" if shifted else sg, n_atoms=5, d_min=1.3, seed=4)
    if shifted:
        w.crystal.symops = ["x,y,z", "-x+1/2,-y,-z+1/2"]
    cr, a = w.crystal, w.crystal.atoms
    Fe = hc_oracle.electron_sf_and_derivs(cr, w.params, a, w.lcs, w.hkl, w.dtdf)[0]
    Fx = hc_oracle.sf_and_derivs(cr, w.params, a, w.lcs, w.hkl, w.dtdf)[0]
    # F_Z: constant form factor Z over every operation of the group (numpy, no folding)
    f2c = hc_oracle.cell_f2c(cr.cell)
    c2f = np.linalg.inv(f2c)
    J = hc_oracle.u_cif_to_cart_matrix(cr.cell)
    hkl = w.hkl.astype(np.float64)
    hc = hkl @ c2f
    ops = [hc_oracle.parse_op(s) for s in cr.symops]
    FZ = np.zeros(len(hkl), dtype=np.complex128)
    for R, t in ops:
        Rc = f2c @ R @ c2f
        hs = hc @ Rc
        for i in range(a.n):
            Z = w.params.wfn_parameters[w.params.atom_to_wfn_map[i]].atomic_number
            if a.adp_size[i] == 6:
                U = J @ a.adp[i]
                Uc = np.array([[U[0], U[3], U[4]], [U[3], U[1], U[5]], [U[4], U[5], U[2]]])
                T = np.exp(-2 * np.pi ** 2 * np.einsum("ki,ij,kj->k", hs, Uc, hs))
            else:
                T = np.exp(-2 * np.pi ** 2 * a.adp[i, 0] * (hc * hc).sum(1))
            FZ += a.occupancy[i] * a.multiplicity[i] / len(ops) * Z * T * np.exp(2j * np.pi * (hkl @ R @ a.xyz[i] + hkl @ t))
    s2 = 0.25 * (hc * hc).sum(1)
    assert rel_err(Fe, 0.023934 / s2 * (FZ - Fx)) < 1e-12
    # h = 0: sum of w f_e(0)
    F0 = hc_oracle.electron_sf_and_derivs(cr, w.params, a, w.lcs, np.zeros((1, 3), dtype=int), np.ones(1))[0]
    f0 = [hc_oracle.ngauss(hc_oracle.ELEMENTS[w.params.wfn_parameters[m].atomic_number - 1], "electron-IT", 0.0)
          for m in w.params.atom_to_wfn_map]
    assert abs(F0[0] - np.dot(a.occupancy * a.multiplicity, f0)) < 1e-10 * abs(F0[0])


@pytest.mark.parametrize("sg", ["P21/c", "P212121"])
def test_electron_oracle_derivatives_match_finite_differences(sg):
    w = synth.small_structure(sg, n_atoms=4, d_min=1.6, seed=5)
    atoms = w.crystal.atoms
    F, dx, du, do = hc_oracle.electron_sf_and_derivs(w.crystal, w.params, atoms, w.lcs, w.hkl, w.dtdf)
    tgt = lambda a: float(np.sum((np.conj(w.dtdf) * hc_oracle.electron_sf_and_derivs(
        w.crystal, w.params, a, w.lcs, w.hkl, w.dtdf)[0]).real))
    h = 1e-6
    scale = max(np.abs(dx).max(), np.abs(du).max(), np.abs(do).max())
    for i in range(atoms.n):
        for j in range(3):
            a1, a2 = atoms.copy(), atoms.copy()
            a1.xyz[i, j] += h
            a2.xyz[i, j] -= h
            assert abs((tgt(a1) - tgt(a2)) / (2 * h) - dx[i, j]) < 2e-6 * scale
        a1, a2 = atoms.copy(), atoms.copy()
        a1.occupancy[i] += h
        a2.occupancy[i] -= h
        assert abs((tgt(a1) - tgt(a2)) / (2 * h) - do[i]) < 2e-6 * scale
        for j in range(atoms.adp_size[i]):
            a1, a2 = atoms.copy(), atoms.copy()
            a1.adp[i, j] += h
            a2.adp[i, j] -= h
            assert abs((tgt(a1) - tgt(a2)) / (2 * h) - du[i, j]) < 2e-6 * scale
