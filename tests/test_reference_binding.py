"""The reference-side binding (include/discamb_amd/discamb_binding.hpp) compiled against DiSCaMB's
own headers and types: a discamb::Crystal + discamb::HC_ModelParameters built the way a DiSCaMB
program builds them (operation strings, the reference's Clementi-Roetti bank and UBDB
deformation-valence defaults) flatten to the descriptor the Python path builds from the same
structure (HansenCoppensStructureFactorCalculator.cpp:157-290, HC_ModelParameters.h:22-28).

Integer arrays, conventions, labels and the per-call arrays agree exactly; Slater coefficients and
exponents (the reference bank against this package's table) to 1e-13 relative; the space-group
operations as a set (the reference SpaceGroup may list them in another order).  Host only: the
driver is built in oracle/_ref from the reference checkout.
"""
import ctypes
import os
import subprocess
import types

import numpy as np
import pytest

from discamb_amd import model, synth
from discamb_amd.calculator import HansenCoppensStructureFactorCalculator, _model_desc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "binding_check")


def _input(w):
    c = w.crystal
    a = c.atoms
    out = [" ".join("%.17g" % v for v in c.cell) + " %s %s" % (c.xyz_convention, c.adp_convention), str(len(c.symops))]
    out += list(c.symops)
    out.append(str(len(w.params.wfn_parameters)))
    out += [wf.label for wf in w.params.wfn_parameters]
    out.append(str(len(w.params.type_parameters)))
    for t in w.params.type_parameters:
        vals = [len(t.p_lm), t.p_val, t.kappa_spherical_valence, t.kappa_deformation_valence]
        vals += [v for row in t.p_lm for v in row]
        out.append(" ".join("%.17g" % v for v in vals))
    out.append(str(a.n))
    for i in range(a.n):
        vals = [w.params.atom_to_wfn_map[i], w.params.atom_to_type_map[i], int(a.multiplicity[i]), a.occupancy[i]]
        vals += list(a.xyz[i]) + [int(a.adp_size[i])] + list(a.adp[i][: a.adp_size[i]]) + list(np.asarray(w.lcs[i]).reshape(9))
        out.append(" ".join("%.17g" % v for v in vals))
    out.append(str(len(w.hkl)))
    for h, g in zip(w.hkl, w.dtdf):
        out.append("%d %d %d %.17g %.17g" % (h[0], h[1], h[2], g.real, g.imag))
    return "\n".join(out) + "\n"


def _run(w):
    if not os.path.exists(DRIVER):
        pytest.skip("oracle/_ref/binding_check not built (reference checkout absent)")
    r = subprocess.run([DRIVER], input=_input(w), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in r.stdout.splitlines():
        name, n, *vals = line.split()
        got[name] = vals if name == "wfn_label" else np.array([float(v) for v in vals])
        assert len(vals) == int(n), name
    return got


def _arr(p, n):
    return np.array([p[i] for i in range(n)], dtype=float)


def _python_desc(w):
    d, keep = _model_desc(w.crystal, w.params, w.crystal.xyz_convention, w.crystal.adp_convention)
    n_orb = d.wfn_orbital_offset[d.n_wfn]
    n_terms = d.orbital_term_offset[n_orb]
    out = {"cell": np.array(list(d.cell)),
           "conventions": np.array([d.xyz_convention, d.adp_convention, d.d_xyz_convention, d.d_adp_convention])}
    sizes = {"wfn_orbital_offset": d.n_wfn + 1, "orbital_term_offset": n_orb + 1, "orbital_nl": 2 * n_orb,
             "term_coeff": n_terms, "term_exp": n_terms, "term_pow": n_terms, "orbital_occupancy": n_orb,
             "orbital_kind": n_orb, "def_val_exponent": d.n_wfn, "def_val_pow_offset": d.n_wfn + 1,
             "def_val_pow": d.def_val_pow_offset[d.n_wfn], "anomalous": 2 * d.n_wfn, "wfn_atomic_number": d.n_wfn,
             "plm_n_l": d.n_types, "plm": 25 * d.n_types, "p_val": d.n_types, "kappa": d.n_types,
             "kappa_prime": d.n_types, "atom_to_wfn": d.n_atoms, "atom_to_type": d.n_atoms,
             "atom_multiplicity": d.n_atoms, "symop_rot": 9 * d.n_symops, "symop_trans": 3 * d.n_symops}
    for k, n in sizes.items():
        out[k] = _arr(getattr(d, k), n)
    out["wfn_label"] = [d.wfn_label[i].decode() for i in range(d.n_wfn)]
    a = w.crystal.atoms
    at, keep2 = HansenCoppensStructureFactorCalculator._atoms(types.SimpleNamespace(n_atoms=a.n), a, w.lcs, None, None)
    hkl, dd = HansenCoppensStructureFactorCalculator._hkl(w.hkl, w.dtdf)
    out.update({"xyz": _arr(at.xyz, 3 * a.n), "adp_size": _arr(at.adp_size, a.n), "adp": _arr(at.adp, 6 * a.n),
                "occupancy": _arr(at.occupancy, a.n), "lcs": _arr(at.lcs, 9 * a.n), "include": np.ones(a.n),
                "hkl": hkl.reshape(-1).astype(float), "dtdf": dd.astype(float)})
    return out


def _ops(rot, trans):
    return sorted((tuple(int(v) for v in rot[9 * s: 9 * s + 9]), tuple(np.round(np.mod(trans[3 * s: 3 * s + 3], 1.0), 12)))
                  for s in range(len(rot) // 9))


CLOSE = ("term_coeff", "term_exp", "def_val_exponent")


def _compare(w):
    got, mine = _run(w), _python_desc(w)
    assert _ops(got["symop_rot"], got["symop_trans"]) == _ops(mine["symop_rot"], mine["symop_trans"])
    assert got["wfn_label"] == mine["wfn_label"]
    for k, v in mine.items():
        if k in ("symop_rot", "symop_trans", "wfn_label"):
            continue
        g = got[k]
        assert g.shape == v.shape, k
        if k in CLOSE:
            assert np.allclose(g, v, rtol=1e-13, atol=0), (k, np.abs(g - v).max())
        else:
            assert np.array_equal(g, v), k


def test_config0_flattens_to_the_python_descriptor():
    _compare(synth.config0())


def test_ions_and_cartesian_conventions_flatten_alike():
    w = synth.small_structure("P21/c", n_atoms=12, d_min=1.2, seed=41)
    for i, label in enumerate(["Fe3+", "Cu+", "O-", "Se-"]):
        w.params.wfn_parameters[i] = model.wfn_type_from_bank(label)
    f2c, _ = synth.cell_matrices(w.crystal.cell)
    w.crystal.atoms.xyz = w.crystal.atoms.xyz @ f2c.T
    w.crystal.xyz_convention = "cartesian"
    w.crystal.adp_convention = "U_cart"
    _compare(w)
