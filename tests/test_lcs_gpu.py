"""Frames recomputed from the positions at every call, and the chain rule through them, on the GPU.

  * without the chain rule: F and the derivatives equal the oracle's with the frames the oracle's
    restatement computes from the same positions (1e-10, FP64);
  * with it: every xyz derivative equals the central difference of the oracle's target
    T(x) = sum_h Re(conj(dT/dF_h) F_h(x, L(x))) with the frames recomputed at each displaced
    position (FP64 1e-6, mixed 2e-5 normwise).
"""
import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import DerivativesSelector, HansenCoppensStructureFactorCalculator, LcsDefinition, synth

pytestmark = pytest.mark.gpu

OPS = ["x,y,z", "-x,y+1/2,-z+1/2", "x+1,y,z"]
KINDS = ["average_direction", "average_position"]


def definitions(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        others = rng.permutation([j for j in range(n) if j != i])
        r1 = [(int(j), OPS[int(rng.integers(3))]) for j in others[:int(rng.integers(1, 3))]]
        r2 = [(int(j), OPS[int(rng.integers(3))]) for j in others[2:2 + int(rng.integers(1, 3))]]
        t2 = KINDS[int(rng.integers(2))] if rng.random() < 0.8 else "any_orthogonal"
        c1 = int(rng.integers(3))
        c2 = int((c1 + rng.integers(1, 3)) % 3)
        ch = [(int(j), "x,y,z") for j in others[-3:]] if rng.random() < 0.4 else []
        out.append(LcsDefinition(i, KINDS[int(rng.integers(2))], r1, t2, r2, c1, c2, bool(rng.random() < 0.8), ch))
    return out


def _frames(w, atoms, defs):
    return np.array([hc_oracle.lcs_matrix(w.crystal.cell, atoms.xyz, d) for d in defs])


@pytest.mark.parametrize("sg", ["P21/c", "P212121"])
def test_frames_follow_positions(sg):
    w = synth.small_structure(sg, n_atoms=8, d_min=1.2, seed=3)
    defs = definitions(8, 1)
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setLocalCoordinateSystems(defs)
    a = w.crystal.atoms.copy()
    for step in range(2):
        print("case lcs frames %s step %d" % (sg, step), flush=True)
        f, d = c.calculateStructureFactorsAndDerivatives(a, None, w.hkl, w.dtdf)
        L = _frames(w, a, defs)
        assert np.abs(c.lastLocalCoordinateSystems() - L).max() < 1e-13
        rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, a, L, w.hkl, w.dtdf)
        assert rel_err(f, rf) < 1e-10 and rel_err(d.atomic_position_derivatives, rx) < 1e-10
        a.xyz = a.xyz + np.random.default_rng(step).normal(0, 0.01, a.xyz.shape)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
@pytest.mark.parametrize("sg", ["P21/c", "P212121", "P43212"])
def test_chain_rule_matches_finite_differences(sg, mode):
    w = synth.small_structure(sg, n_atoms=6, d_min=1.3, seed=8)
    defs = definitions(6, 2)
    a = w.crystal.atoms
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setCalculationMode(mode)
    c.setLocalCoordinateSystems(defs, chainRule=True)
    print("case lcs chain rule %s %s" % (sg, mode), flush=True)
    f, d = c.calculateStructureFactorsAndDerivatives(a, None, w.hkl, w.dtdf)

    def T(atoms):
        F = hc_oracle.sf_and_derivs(w.crystal, w.params, atoms, _frames(w, atoms, defs), w.hkl, w.dtdf)[0]
        return float(np.sum((np.conj(w.dtdf) * F).real))

    h = 1e-6
    fd = np.zeros((a.n, 3))
    for i in range(a.n):
        for j in range(3):
            a1, a2 = a.copy(), a.copy()
            a1.xyz[i, j] += h
            a2.xyz[i, j] -= h
            fd[i, j] = (T(a1) - T(a2)) / (2 * h)
    frozen = hc_oracle.sf_and_derivs(w.crystal, w.params, a, _frames(w, a, defs), w.hkl, w.dtdf)[1]
    assert rel_err(fd, frozen) > 1e-3          # the frames matter for these derivatives
    e = rel_err(d.atomic_position_derivatives, fd)
    print("   chain rule vs finite differences %.2e" % e, flush=True)
    assert e < (1e-6 if mode == "HIP" else 2e-5)
    # ADP and occupancy derivatives are unaffected by the frames' dependence on positions
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, a, _frames(w, a, defs), w.hkl, w.dtdf)
    tol = 1e-10 if mode == "HIP" else 1e-5
    assert rel_err(d.adp_derivatives, ru) < tol and rel_err(d.occupancy_derivatives, ro) < tol
    # selector without xyz: no chain-rule work, same ADP derivatives
    f2, d2 = c.calculateStructureFactorsAndDerivatives(a, None, w.hkl, w.dtdf, None,
                                                      DerivativesSelector(False, True, True))
    assert rel_err(d2.adp_derivatives, ru) < tol


def test_chain_rule_with_cartesian_conventions():
    """Cartesian input coordinates and U_cart: frames from the converted positions, position
    derivatives in Cartesian coordinates, against central differences in Cartesian coordinates."""
    w = synth.small_structure("P21/c", n_atoms=6, d_min=1.4, seed=11)
    f2c, c2f = synth.cell_matrices(w.crystal.cell)
    a = w.crystal.atoms.copy()
    a.xyz = a.xyz @ f2c.T
    for i in range(a.n):
        if a.adp_size[i] == 6:
            u = a.adp[i]
            ucif = np.array([[u[0], u[3], u[4]], [u[3], u[1], u[5]], [u[4], u[5], u[2]]])
            n = np.diag(np.linalg.norm(c2f, axis=1))     # lengths of the reciprocal axes
            uc = f2c @ n @ ucif @ n @ f2c.T
            a.adp[i] = [uc[0, 0], uc[1, 1], uc[2, 2], uc[0, 1], uc[0, 2], uc[1, 2]]
    cr = synth.Crystal(w.crystal.cell, w.crystal.symops, a, "cartesian", "U_cart")
    defs = definitions(6, 4)
    c = HansenCoppensStructureFactorCalculator(cr, w.params)
    c.setLocalCoordinateSystems(defs, chainRule=True)
    print("case lcs chain rule cartesian", flush=True)
    f, d = c.calculateStructureFactorsAndDerivatives(a, None, w.hkl, w.dtdf)

    def frames(atoms):
        frac = atoms.xyz @ c2f.T
        return np.array([hc_oracle.lcs_matrix(cr.cell, frac, dd) for dd in defs])

    rf = hc_oracle.sf_and_derivs(cr, w.params, a, frames(a), w.hkl, w.dtdf)[0]
    assert rel_err(f, rf) < 1e-10

    def T(atoms):
        F = hc_oracle.sf_and_derivs(cr, w.params, atoms, frames(atoms), w.hkl, w.dtdf)[0]
        return float(np.sum((np.conj(w.dtdf) * F).real))

    h = 1e-6
    fd = np.zeros((a.n, 3))
    for i in range(a.n):
        for j in range(3):
            a1, a2 = a.copy(), a.copy()
            a1.xyz[i, j] += h
            a2.xyz[i, j] -= h
            fd[i, j] = (T(a1) - T(a2)) / (2 * h)
    e = rel_err(d.atomic_position_derivatives, fd)
    print("   chain rule vs finite differences (Cartesian) %.2e" % e, flush=True)
    assert e < 1e-6
