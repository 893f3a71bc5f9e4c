"""The oracle's derivatives against central finite differences of the oracle's own F (CPU).

dT/dp for the linear target T(p) = sum_h Re(conj(dT/dF_h) F_h(p)) is exactly what
calculateStructureFactorsAndDerivatives returns (dT/dF = dT/dA + i dT/dB), so finite differences
of F pin the adjoint independently of its formula.  Local frames are held fixed."""
import copy
import math

import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import synth


def target(w, atoms, dtdf):
    F = hc_oracle.sf_and_derivs(w.crystal, w.params, atoms, w.lcs, w.hkl, dtdf)[0]
    return float(np.sum((np.conj(dtdf) * F).real))


@pytest.mark.parametrize("sg", ["P1", "P-1", "P21/c", "C2/c", "P212121", "P-1_shifted", "P43212", "P31"])
def test_oracle_derivatives_match_finite_differences(sg):
    shifted = sg == "P-1_shifted"
    w = synth.small_structure("P-1" if shifted else sg, n_atoms=5, d_min=1.6)
    if shifted:
        w.crystal.symops = ["x,y,z", "-x+1/2,-y,-z+1/2"]
    atoms = w.crystal.atoms
    F, dx, du, do = hc_oracle.sf_and_derivs(w.crystal, w.params, atoms, w.lcs, w.hkl, w.dtdf)
    h = 1e-6
    for i in range(atoms.n):
        for j in range(3):
            a1, a2 = atoms.copy(), atoms.copy()
            a1.xyz[i, j] += h
            a2.xyz[i, j] -= h
            fd = (target(w, a1, w.dtdf) - target(w, a2, w.dtdf)) / (2 * h)
            assert fd == pytest.approx(dx[i, j], rel=1e-5, abs=1e-6 * np.abs(dx).max())
        for j in range(atoms.adp_size[i]):
            a1, a2 = atoms.copy(), atoms.copy()
            a1.adp[i, j] += h
            a2.adp[i, j] -= h
            fd = (target(w, a1, w.dtdf) - target(w, a2, w.dtdf)) / (2 * h)
            assert fd == pytest.approx(du[i, j], rel=1e-5, abs=1e-6 * np.abs(du).max())
        a1, a2 = atoms.copy(), atoms.copy()
        a1.occupancy[i] += h
        a2.occupancy[i] -= h
        fd = (target(w, a1, w.dtdf) - target(w, a2, w.dtdf)) / (2 * h)
        assert fd == pytest.approx(do[i], rel=1e-6, abs=1e-8 * np.abs(do).max())


def test_symmetry_invariants_of_the_oracle():
    # F(-h) = conj F(h) without anomalous scattering (acentric group)
    w = synth.small_structure("P212121", n_atoms=6, d_min=1.4)
    hkl = np.vstack([w.hkl[:50], -w.hkl[:50]])
    F = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, hkl, np.ones(100))[0]
    assert rel_err(F[50:], np.conj(F[:50])) < 1e-12
    # systematic absences of P2_1 2_1 2_1: h00, 0k0, 00l with odd index
    ab = np.array([[3, 0, 0], [0, 5, 0], [0, 0, 7], [1, 0, 0]])
    Fa = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, ab, np.ones(4))[0]
    assert np.abs(Fa).max() < 1e-12 * np.abs(F).max()
    # centrosymmetric P2_1/c: F real; h0l with l odd and 0k0 with k odd absent
    w = synth.small_structure("P21/c", n_atoms=6, d_min=1.4)
    F = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, np.ones(len(w.hkl)))[0]
    assert np.all(F.imag == 0)
    ab = np.array([[1, 0, 1], [2, 0, 3], [0, 3, 0]])
    Fa = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, ab, np.ones(3))[0]
    assert np.abs(Fa).max() < 1e-12 * np.abs(F).max()
    # a lattice translation of every atom leaves F unchanged
    a = w.crystal.atoms.copy()
    a.xyz += np.array([1.0, -2.0, 3.0])
    F2 = hc_oracle.sf_and_derivs(w.crystal, w.params, a, w.lcs, w.hkl, np.ones(len(w.hkl)))[0]
    assert rel_err(F2, F) < 1e-11


def test_deformation_density_fourier_transform():
    """The deformation-valence term of one atom equals the numerical 3-D Fourier transform of
    kappa'^3 N_l r^n_l exp(-kappa' zeta r) sum_m P_lm d_lm(L^T r_hat)  (pins 4 pi, i^l, N_l, the
    harmonic normalisation, the g-functions and the local-frame convention)."""
    from discamb_amd import HC_AtomTypeParameters, HC_ModelParameters, HC_WfnType, AtomSet, Crystal
    rng = np.random.default_rng(3)
    wf = HC_WfnType("X", [], [], [], [], 4.3, [2, 2, 2, 3, 4])
    for l in range(5):
        plm = [[0.0] * (2 * q + 1) for q in range(5)]
        for m in range(2 * l + 1):
            plm[l][m] = rng.uniform(-1, 1)
        tp = HC_AtomTypeParameters(plm, 0.0, 1.0, 1.13)
        params = HC_ModelParameters([wf], [tp], [0], [0])
        atoms = AtomSet(np.zeros((1, 3)), np.zeros((1, 6)), np.zeros(1, dtype=np.int32), np.ones(1), np.ones(1))
        cryst = Crystal([12.0, 12.0, 12.0, 90, 90, 90], ["x,y,z"], atoms)
        L = synth.random_rotation(rng)
        hkl = np.array([[1, 0, 0], [2, 1, 0], [1, 2, 3], [3, 0, 2]])
        F = hc_oracle.sf_and_derivs(cryst, params, atoms, L[None], hkl, np.ones(4))[0]
        # numerical transform: Gauss-Laguerre-like radial grid x product angular grid
        zeta, kp, n = 4.3, 1.13, [2, 2, 2, 3, 4][l]
        norm = zeta ** (n + 3) / math.factorial(n + 2)
        r = np.linspace(1e-4, 12.0, 3000)
        dr = r[1] - r[0]
        ct, wt = np.polynomial.legendre.leggauss(40)
        ph = np.linspace(0, 2 * np.pi, 80, endpoint=False)
        st = np.sqrt(1 - ct ** 2)
        u = np.stack([np.outer(st, np.cos(ph)), np.outer(st, np.sin(ph)), np.outer(ct, np.ones_like(ph))], -1)
        u = u.reshape(-1, 3)
        wa = np.repeat(wt, len(ph)) * (2 * np.pi / len(ph))
        ul = u @ L
        import ctypes
        lib = hc_oracle.lib()
        dl = np.array([[0.0] * 25 for _ in ul])
        for i, v in enumerate(ul):
            out = (ctypes.c_double * 25)()
            lib.hc_oracle_ylm((ctypes.c_double * 3)(*v), out)
            dl[i] = out[:]
        ang = dl[:, l * l:(l + 1) * (l + 1)] @ np.array(plm[l])
        radial = kp ** 3 * norm * (kp * r) ** n * np.exp(-zeta * kp * r)
        f2c = np.eye(3) * 12.0
        for k, hv in enumerate(hkl):
            hc = np.linalg.inv(f2c).T @ hv
            phase = np.exp(2j * np.pi * np.outer(r, u @ hc))           # (r, angle)
            integrand = (radial[:, None] * r[:, None] ** 2) * phase * ang[None, :]
            val = np.sum(integrand * wa[None, :]) * dr
            assert abs(val - F[k]) < 2e-5 * max(1e-3, abs(F).max()), (l, k, val, F[k])


@pytest.mark.parametrize("sg", list(synth.SPACE_GROUPS))
def test_oracle_symmetry_folding_matches_the_full_group(sg):
    """The reference's folding (coset representatives, 2 Re for centric-at-origin, centering
    multipliers) against an explicit sum over every operation of the group."""
    w = synth.small_structure(sg, n_atoms=6, d_min=1.3)
    F, dx, du, do = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    G, gx, gu, go = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf,
                                            mode="full_group")
    assert rel_err(F, G) < 1e-13
    assert rel_err(dx, gx) < 1e-12 and rel_err(du, gu) < 1e-12 and rel_err(do, go) < 1e-12
