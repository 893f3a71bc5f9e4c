"""The C++ drop-in class (include/discamb_amd/HansenCoppensStructureFactorCalculator.hpp) end to end:
a compiled driver builds Crystal / HC_ModelParameters / AtomInCrystal values, calls
calculateStructureFactorsAndDerivatives on the GPU, and the result is compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import hc_oracle
from conftest import ROOT, rel_err
from discamb_amd import synth

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "tests", "cpp", "hc_cpp_driver")


def write_input(w, path, include):
    tok = []
    tok += [repr(float(v)) for v in w.crystal.cell]
    tok += [len(w.crystal.symops)] + [s.replace(" ", "") for s in w.crystal.symops]
    tok += [len(w.params.wfn_parameters)]
    for wf in w.params.wfn_parameters:
        tok += [wf.label, len(wf.orbitals)]
        for o in wf.orbitals:
            tok += [o.n, o.l, len(o.coefficients)]
            for c, e, p in zip(o.coefficients, o.exponents, o.powers):
                tok += [repr(float(c)), repr(float(e)), p]
        tok += list(wf.orbital_occupancy)
        tok += [len(wf.core_orbitals_indices)] + list(wf.core_orbitals_indices)
        tok += [len(wf.valence_orbitals_indices)] + list(wf.valence_orbitals_indices)
        tok += [repr(float(wf.deformation_valence_exponent)), len(wf.deformation_valence_power)]
        tok += list(wf.deformation_valence_power)
        tok += [0.0, 0.0]
    tok += [len(w.params.type_parameters)]
    for t in w.params.type_parameters:
        tok += [len(t.p_lm)]
        for row in t.p_lm:
            tok += [repr(float(v)) for v in row]
        tok += [repr(float(t.p_val)), repr(float(t.kappa_spherical_valence)), repr(float(t.kappa_deformation_valence))]
    a = w.crystal.atoms
    tok += [a.n]
    for i in range(a.n):
        tok += [w.params.atom_to_wfn_map[i], w.params.atom_to_type_map[i]]
        tok += [repr(float(v)) for v in a.xyz[i]]
        tok += [int(a.adp_size[i])] + [repr(float(v)) for v in a.adp[i, :a.adp_size[i]]]
        tok += [repr(float(a.occupancy[i])), repr(float(a.multiplicity[i]))]
        tok += [repr(float(v)) for v in w.lcs[i].reshape(-1)]
        tok += [1 if include[i] else 0]
    tok += [len(w.hkl)]
    for h, d in zip(w.hkl, w.dtdf):
        tok += [int(h[0]), int(h[1]), int(h[2]), repr(float(d.real)), repr(float(d.imag))]
    with open(path, "w") as f:
        f.write(" ".join(str(t) for t in tok))


@pytest.mark.parametrize("sg,mode", [("P21/c", "HIP"), ("P212121", "HIP"), ("P21/c", "per_hkl"),
                                     ("P212121", "per_hkl")])
def test_cpp_drop_in_matches_oracle(tmp_path, sg, mode):
    """mode per_hkl: the per-hkl template overload with a processor that contracts every
    reflection's SfDerivativesAtHkl with dT/dF; without f'' the contraction equals the summed
    derivatives."""
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/hc_cpp_driver not built (run __graft_entry__.build())")
    w = synth.small_structure(sg, n_atoms=9, d_min=1.1)
    include = np.ones(9, dtype=bool)
    include[3] = False
    inp = tmp_path / "in.txt"
    write_input(w, str(inp), include)
    print("case cpp driver", sg, mode, flush=True)
    out = subprocess.run([DRIVER, str(inp), mode], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    if mode == "per_hkl":
        assert "order ok" in out.stderr and "rethrow ok" in out.stderr
    else:
        assert "size check ok" in out.stderr
    lines = out.stdout.strip().split("\n")
    m = len(w.hkl)
    f = np.array([complex(*map(float, l.split())) for l in lines[:m]])
    dx, du, do = [], np.zeros((9, 6)), []
    for i, l in enumerate(lines[m:]):
        v = [float(x) for x in l.split()]
        dx.append(v[:3])
        na = int(v[3])
        du[i, :na] = v[4:4 + na]
        do.append(v[4 + na])
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf,
                                             include=include)
    assert rel_err(f, rf) < 1e-10
    assert rel_err(np.array(dx), rx) < 1e-10
    assert rel_err(du, ru) < 1e-10
    assert rel_err(np.array(do), ro) < 1e-10


def test_cpp_form_factors_match_oracle(tmp_path):
    """calculateFormFactors through the C++ class: one reflection per call and the batched form."""
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/hc_cpp_driver not built (run __graft_entry__.build())")
    w = synth.small_structure("P21/c", n_atoms=9, d_min=1.1)
    include = np.ones(9, dtype=bool)
    include[3] = False
    inp = tmp_path / "in.txt"
    write_input(w, str(inp), include)
    print("case cpp driver form factors", flush=True)
    out = subprocess.run([DRIVER, str(inp), "form_factors"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "single ok" in out.stderr
    v = np.array([complex(*map(float, l.split())) for l in out.stdout.strip().split("\n")])
    f = v.reshape(len(w.hkl), 9)
    c2f = np.linalg.inv(hc_oracle.cell_f2c(w.crystal.cell))
    ref = hc_oracle.form_factors(w.crystal, w.params, w.lcs, w.hkl.astype(np.float64) @ c2f, include=include)
    assert rel_err(f, ref) < 1e-10


def test_cpp_local_frames_from_positions(tmp_path):
    """setLocalCoordinateSystems through the C++ class with the chain rule: the frames equal the
    oracle's restatement, F the oracle's with those frames, and the position derivatives the
    Python binding's (whose chain rule is checked against finite differences in test_lcs_gpu.py)."""
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/hc_cpp_driver not built (run __graft_entry__.build())")
    from discamb_amd import HansenCoppensStructureFactorCalculator
    from test_lcs_gpu import definitions
    w = synth.small_structure("P21/c", n_atoms=8, d_min=1.2, seed=3)
    include = np.ones(8, dtype=bool)
    defs = definitions(8, 5)
    inp = tmp_path / "in.txt"
    write_input(w, str(inp), include)
    kind = {"any_orthogonal": 0, "average_direction": 1, "average_position": 2}
    tok = [len(defs)]
    refs = lambda lst: [len(lst)] + [x for r in lst for x in (r[0], r[1].replace(" ", ""))]
    for d in defs:
        tok += [d.central, d.central_op, kind[d.direction1_type]] + refs(d.ref1)
        tok += [kind[d.direction2_type]] + refs(d.ref2)
        tok += [d.coordinate_1, d.coordinate_2, 1 if d.is_r else 0] + refs(d.chirality)
    with open(inp, "a") as f:
        f.write(" " + " ".join(str(t) for t in tok))
    print("case cpp driver lcs", flush=True)
    out = subprocess.run([DRIVER, str(inp), "lcs"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().split("\n")
    n, m = 8, len(w.hkl)
    L = np.array([[float(x) for x in l.split()] for l in lines[:n]]).reshape(n, 3, 3)
    f = np.array([complex(*map(float, l.split())) for l in lines[n:n + m]])
    dx = np.array([[float(x) for x in l.split()] for l in lines[n + m:]])
    ref_L = np.array([hc_oracle.lcs_matrix(w.crystal.cell, w.crystal.atoms.xyz, d) for d in defs])
    assert np.abs(L - ref_L).max() < 1e-13
    assert rel_err(f, hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, ref_L, w.hkl, w.dtdf)[0]) < 1e-10
    c = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    c.setLocalCoordinateSystems(defs, chainRule=True)
    _, d = c.calculateStructureFactorsAndDerivatives(w.crystal.atoms, None, w.hkl, w.dtdf)
    assert rel_err(dx, d.atomic_position_derivatives) < 1e-12
