"""The CPU-baseline restatement of Engine3 with hkl-line recurrences (hc_oracle_engine_lines:
global-frame P_lm by projection quadrature, per-reflection tables, phase and Debye-Waller factors
carried along lines) against the direct oracle: 1e-12 normwise for F and every derivative class,
with isotropic atoms, anomalous terms, excluded atoms and reflection lists that break lines."""
import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import synth


@pytest.mark.parametrize("sg", ["P1", "P212121", "P31", "P43212"])
def test_line_restatement_matches_the_direct_oracle(sg):
    w = synth.small_structure(sg, n_atoms=16, d_min=0.9, seed=13)
    n = w.crystal.atoms.n
    rng = np.random.default_rng(2)
    inc = np.ones(n, dtype=bool)
    inc[3] = False
    anom = rng.normal(size=n) * 0.2 + 1j * np.abs(rng.normal(size=n)) * 0.1
    sub = np.sort(rng.choice(len(w.hkl), int(0.7 * len(w.hkl)), replace=False))
    for hkl, g in ((w.hkl, w.dtdf), (w.hkl[sub], w.dtdf[sub])):
        a = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, hkl, g, include=inc, anomalous=anom,
                                    n_threads=4)
        b = hc_oracle.sf_and_derivs_lines(w.crystal, w.params, w.crystal.atoms, w.lcs, hkl, g, include=inc,
                                          anomalous=anom, n_threads=4)
        e = [rel_err(x, y) for x, y in zip(b, a)]
        assert max(e) <= 1e-12, e


def test_line_order_covers_every_reflection_once():
    w = synth.small_structure("P1", n_atoms=4, d_min=1.0)
    perm, off, ax = hc_oracle.line_order(w.hkl)
    assert sorted(perm.tolist()) == list(range(len(w.hkl)))
    h = w.hkl[perm]
    for i in range(len(off) - 1):
        seg = h[off[i]:off[i + 1]]
        assert np.all(np.diff(seg[:, ax]) == 1)
        assert np.all(seg[:, (ax + 1) % 3] == seg[0, (ax + 1) % 3])


def test_line_restatement_refuses_centrosymmetric_groups():
    w = synth.small_structure("P21/c", n_atoms=4, d_min=1.5)
    with pytest.raises(ValueError):
        hc_oracle.sf_and_derivs_lines(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
