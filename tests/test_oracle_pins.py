"""Pins of the FP64 oracle (CPU only).

Each oracle function is checked against an independent source: the reference's own compiled
sources (oracle/_ref, built from /root/reference where it exists), closed forms, numerical
quadrature or brute force, chosen so that a dropped term, a wrong sign or index or a transposed
operand fails.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import integrate, special

import hc_oracle
from conftest import rel_err
from discamb_amd import model, synth

# valid (l, n) range of the reference closed forms (SlaterTypeOrbitalScattering.h:93-180)
LN = [(l, n) for l in range(5) for n in range(l + 2, 11)]


def test_g_function_vs_reference(ref_lib, oracle_lib):
    rng = np.random.default_rng(1)
    worst = 0.0
    for l, n in LN:
        for _ in range(20):
            h = rng.uniform(0.0, 2.5)
            z = rng.uniform(1.0, 25.0)
            a = oracle_lib.hc_oracle_g(l, n, h, z)
            b = ref_lib.ref_g_function(l, n, h, z)
            scale = abs(oracle_lib.hc_oracle_g(l, n, 0.0, z)) + abs(b) + 1e-300
            worst = max(worst, abs(a - b) / scale)
    assert worst < 1e-12


def test_g_function_closed_forms_and_quadrature(oracle_lib):
    # g(0, n, 0, Z) = n! / Z^(n+1); g(l>0, n, 0, Z) = 0
    for n in range(1, 11):
        for z in (1.3, 4.0, 11.0):
            assert oracle_lib.hc_oracle_g(0, n, 0.0, z) == pytest.approx(math.factorial(n) / z ** (n + 1), rel=1e-13)
            for l in range(1, min(n, 5)):
                assert oracle_lib.hc_oracle_g(l, n, 0.0, z) == 0.0
    # direct quadrature of int r^n exp(-Z r) j_l(2 pi h r) dr
    for l, n, h, z in [(0, 2, 0.3, 3.0), (1, 4, 0.7, 5.5), (2, 5, 0.45, 4.2), (3, 7, 1.1, 8.0), (4, 8, 0.9, 6.3),
                       (4, 10, 1.6, 9.1), (2, 9, 1.9, 12.0)]:
        k = 2 * np.pi * h
        q, _ = integrate.quad(lambda r: r ** n * np.exp(-z * r) * special.spherical_jn(l, k * r), 0, 80 / z,
                              limit=400, epsabs=0, epsrel=1e-13)
        assert oracle_lib.hc_oracle_g(l, n, h, z) == pytest.approx(q, rel=1e-9, abs=1e-14 * abs(q) + 1e-300)


def _oracle_ylm(lib, u):
    out = (ctypes.c_double * 25)()
    lib.hc_oracle_ylm((ctypes.c_double * 3)(*u), out)
    return np.array(out[:])


def test_harmonics_vs_reference(ref_lib, oracle_lib):
    rng = np.random.default_rng(2)
    for _ in range(50):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        ref = (ctypes.c_double * 25)()
        ref_lib.ref_density_normalized_ylm((ctypes.c_double * 3)(*u), ref)
        assert np.allclose(_oracle_ylm(oracle_lib, u), np.array(ref[:]), rtol=1e-13, atol=1e-14)


def test_harmonics_density_normalisation(oracle_lib):
    # int |d_lm| dOmega = 1 (l = 0) or 2 (l > 0); midpoint grid on the sphere
    nt, nphi = 600, 1200
    t = (np.arange(nt) + 0.5) * np.pi / nt
    p = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, Pp = np.meshgrid(t, p, indexing="ij")
    u = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1).reshape(-1, 3)
    w = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nphi)).reshape(-1)
    # vectorised evaluation through the oracle on a subsample is slow; use the oracle on the grid
    vals = np.array([_oracle_ylm(oracle_lib, v) for v in u[:: 37]])
    ws = w[:: 37] * 37
    integ = np.abs(vals).T @ ws
    target = np.array([1.0] + [2.0] * 24)
    assert np.allclose(integ, target, rtol=2e-3)
    # orthogonality of different (l, m)
    gram = (vals * ws[:, None]).T @ vals
    off = gram - np.diag(np.diag(gram))
    assert np.abs(off).max() < 2e-3 * np.abs(np.diag(gram)).max()


def test_global_plm_rotation_equivalence(ref_lib, oracle_lib):
    """The reference evaluates def-val with P_lm rotated to the global frame
    (Engine3.cpp:594-654 via SphConverter); the oracle evaluates it in the local frame with
    u_local = L^T u.  Both must agree for LCS given with local axes as columns."""
    rng = np.random.default_rng(3)
    for _ in range(10):
        L = synth.random_rotation(rng)
        plm = rng.uniform(-0.3, 0.3, 25)
        pg = (ctypes.c_double * 25)()
        ref_lib.ref_global_plm((ctypes.c_double * 9)(*L.reshape(-1)), (ctypes.c_double * 25)(*plm), pg)
        pg = np.array(pg[:])
        for _ in range(5):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            d_glob = _oracle_ylm(oracle_lib, u)
            d_loc = _oracle_ylm(oracle_lib, L.T @ u)
            for l in range(5):
                s = slice(l * l, (l + 1) * (l + 1))
                assert np.dot(pg[s], d_glob[s]) == pytest.approx(np.dot(plm[s], d_loc[s]), rel=1e-10, abs=1e-12)


def _ref_density(ref_lib, label, kind):
    c = (ctypes.c_double * 4000)()
    e = (ctypes.c_double * 4000)()
    p = (ctypes.c_int * 4000)()
    n = ref_lib.ref_wfn_density(label.encode(), kind, 4000, c, e, p)
    return np.array(c[:n]), np.array(e[:n]), np.array(p[:n])


@pytest.mark.parametrize("sym", model.bank_symbols())
def test_wavefunction_bank_vs_reference(ref_lib, sym):
    """Built-in Clementi-Roetti table + density products + valence normalisation against the
    reference bank (SlaterOrbitalWfnData / atomicStoWfnToSphericalDensity / setModel)."""
    w = model.wfn_type_from_bank(sym)
    for kind, valence in ((0, False), (1, True)):
        if valence and not w.valence_orbitals_indices:
            continue
        c, e, p = hc_oracle.sphere_density(w, valence)
        rc, re_, rp = _ref_density(ref_lib, sym, kind)
        assert len(c) == len(rc)
        if len(rc) == 0:
            continue
        assert np.array_equal(np.array(p), rp)
        assert np.allclose(e, re_, rtol=1e-14)
        assert np.allclose(c, rc, rtol=1e-11, atol=1e-14 * np.abs(rc).max())


@pytest.mark.parametrize("sym", model.bank_symbols())
def test_def_val_parameters_vs_reference(ref_lib, sym):
    for ubdb, key in ((1, "UBDB"), (0, "STANDARD")):
        z = ctypes.c_double()
        pw = (ctypes.c_int * 5)()
        n = ref_lib.ref_def_val(sym.encode(), ubdb, ctypes.byref(z), pw)
        assert n == 5
        w = model.wfn_type_from_bank(sym, key)
        assert w.deformation_valence_power == list(pw[:5])
        assert w.deformation_valence_exponent == pytest.approx(z.value, rel=1e-13)


def test_electron_count_of_densities():
    """f(0) of a Slater density = its electron count: core -> core electrons, valence -> 1."""
    for sym in ("H", "C", "N", "O", "S", "Fe", "Br", "Fe3+", "Cu+", "Zn2+", "O-", "Se-", "Al3+", "Ti3+"):
        w = model.wfn_type_from_bank(sym)
        for valence in (False, True):
            c, e, p = hc_oracle.sphere_density(w, valence)
            n_el = sum(4 * np.pi * ci * math.factorial(pi + 2) / ei ** (pi + 3) for ci, ei, pi in zip(c, e, p))
            idx = w.valence_orbitals_indices if valence else w.core_orbitals_indices
            target = 1.0 if valence and idx else sum(w.orbital_occupancy[i] for i in idx)
            assert n_el == pytest.approx(target, rel=1e-4)


def test_cell_and_conventions_vs_reference(ref_lib):
    rng = np.random.default_rng(4)
    for cell in [(10, 12, 15, 90, 100, 90), (7.1, 8.3, 9.2, 85, 95, 101), (50, 60, 70, 90, 90, 90)]:
        f2c = (ctypes.c_double * 9)()
        c2f = (ctypes.c_double * 9)()
        ref_lib.ref_unit_cell((ctypes.c_double * 6)(*cell), f2c, c2f)
        F = hc_oracle.cell_f2c(cell)
        assert np.allclose(F, np.array(f2c[:]).reshape(3, 3), rtol=1e-14, atol=1e-14)
        assert np.allclose(np.linalg.inv(F), np.array(c2f[:]).reshape(3, 3), rtol=1e-12, atol=1e-14)
        J = hc_oracle.u_cif_to_cart_matrix(cell)
        for _ in range(5):
            ucif = rng.uniform(-0.02, 0.05, 6)
            dcart = rng.normal(size=6)
            dxc = rng.normal(size=3)
            uc, dcif, dxf = (ctypes.c_double * 6)(), (ctypes.c_double * 6)(), (ctypes.c_double * 3)()
            ref_lib.ref_convert((ctypes.c_double * 6)(*cell), (ctypes.c_double * 6)(*ucif), uc,
                                (ctypes.c_double * 6)(*dcart), dcif, (ctypes.c_double * 3)(*dxc), dxf)
            assert np.allclose(J @ ucif, uc[:], rtol=1e-12, atol=1e-15)
            assert np.allclose(dcart @ J, dcif[:], rtol=1e-11, atol=1e-13)
            assert np.allclose(dxc @ F, dxf[:], rtol=1e-12, atol=1e-13)


GROUPS = dict(synth.SPACE_GROUPS)
GROUPS["I222"] = ["x,y,z", "-x,-y,z", "-x,y,-z", "x,-y,-z", "x+1/2,y+1/2,z+1/2", "-x+1/2,-y+1/2,z+1/2",
                  "-x+1/2,y+1/2,-z+1/2", "x+1/2,-y+1/2,-z+1/2"]
GROUPS["P-1_shifted"] = ["x,y,z", "-x+1/2,-y,-z"]


@pytest.mark.parametrize("name", list(GROUPS))
def test_space_group_decomposition_vs_reference(ref_lib, name):
    ops = GROUPS[name]
    n, ns, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    inv = (ctypes.c_double * 3)()
    ce, ob = ctypes.c_char(), ctypes.c_int()
    rot, tr = (ctypes.c_int * (9 * 48))(), (ctypes.c_double * (3 * 48))()
    assert ref_lib.ref_space_group(";".join(ops).encode(), ctypes.byref(n), ctypes.byref(ns), ctypes.byref(c), inv,
                                   ctypes.byref(ce), ctypes.byref(ob), rot, tr, 48) == 0
    order, subset, centro, t_inv, letter, obverse = hc_oracle.decompose_group(ops)
    assert order == n.value and len(subset) == ns.value and int(centro) == c.value
    assert letter == ce.value.decode()
    if centro:
        # any inversion translation differing by a centering vector is equivalent
        assert len(subset) * (2 if centro else 1) * {"P": 1, "C": 2, "I": 2}[letter] == order
    ref_rots = {tuple(rot[9 * k:9 * k + 9]) for k in range(ns.value)}
    mine = {tuple(int(v) for v in R.reshape(-1)) for R, t in subset}
    if not centro:
        assert mine == ref_rots


def test_centering_multipliers():
    hkl = np.array([[1, 0, 0], [1, 1, 0], [0, 1, 1], [1, 1, 1], [2, 0, 0], [-1, 2, 0], [1, 0, 1]])
    assert list(hc_oracle.centering_mult("C", True, hkl)) == [0, 2, 0, 2, 2, 0, 0]
    assert list(hc_oracle.centering_mult("I", True, hkl)) == [0, 2, 2, 0, 2, 0, 2]
    assert list(hc_oracle.centering_mult("F", True, hkl)) == [0, 0, 0, 4, 4, 0, 0]
    assert list(hc_oracle.centering_mult("A", True, hkl)) == [2, 0, 2, 2, 2, 2, 0]
