"""Multipole-parameter adjoint (P_lm, Pval, kappa, kappa') and the torch autograd binding.

The reference has no such derivatives, so they are pinned by central finite differences: of the
FP64 oracle's F on small structures, and (config-3 shape) of the GPU FP64 F on the 5000-atom
structure with a least-squares target against perturbed observations."""
import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import DerivativesSelector, HansenCoppensStructureFactorCalculator, synth

pytestmark = pytest.mark.gpu


def _types_arrays(params):
    tp = params.type_parameters
    return (np.array([t.plm25() for t in tp]), np.array([t.p_val for t in tp]),
            np.array([t.kappa_spherical_valence for t in tp]), np.array([t.kappa_deformation_valence for t in tp]))


def _oracle_T(w, params, dtdf):
    F = hc_oracle.sf_and_derivs(w.crystal, params, w.crystal.atoms, w.lcs, w.hkl, dtdf, n_threads=16)[0]
    return float(np.sum((np.conj(dtdf) * F).real))


@pytest.mark.parametrize("radial", ["auto", "records"])
@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
@pytest.mark.parametrize("sg", ["P21/c", "P212121", "P43212", "P63"])
def test_multipole_derivatives_vs_oracle_finite_differences(sg, mode, radial):
    """Multipole-parameter derivatives against central differences of the oracle.  radial
    "records" orders the reflections along hkl lines, so both multipole passes run the line
    kernels (2/m and 222 sign sets); "auto" takes per-atom table lookups in |h| order for the
    first pass' F and the direct kernels.  P4_3 2_1 2 runs the 222 kernels over the atoms' images
    under the coset representatives of its 222 subgroup and folds their derivatives back; P6_3
    runs the 2/m-set kernels of its {1, 2_1} subgroup over three images in hexagonal axes."""
    import copy
    w = synth.small_structure(sg, n_atoms=8, d_min=1.2)
    print("case multipole fd", sg, mode, radial, flush=True)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    calc.setKernelOptions(radial=radial)
    sel = DerivativesSelector(True, True, True, False, True)
    f, d = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf, None, sel)
    td = calc.typeDerivatives()
    # basic derivatives still match the oracle
    rf, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    tol = 1e-10 if mode == "HIP" else 1e-5
    assert rel_err(d.atomic_position_derivatives, rx) < tol
    assert rel_err(d.adp_derivatives, ru) < tol
    fd_tol = 1e-6 if mode == "HIP" else 2e-5
    h = 1e-5
    got, want = {}, {}
    for t, tp in enumerate(w.params.type_parameters):
        for name, key in (("p_val", "pval"), ("kappa_spherical_valence", "kappa"),
                          ("kappa_deformation_valence", "kappa_prime")):
            p1, p2 = copy.deepcopy(w.params), copy.deepcopy(w.params)
            setattr(p1.type_parameters[t], name, getattr(tp, name) + h)
            setattr(p2.type_parameters[t], name, getattr(tp, name) - h)
            want.setdefault(key, []).append((_oracle_T(w, p1, w.dtdf) - _oracle_T(w, p2, w.dtdf)) / (2 * h))
            got.setdefault(key, []).append({"pval": td.p_val, "kappa": td.kappa, "kappa_prime": td.kappa_prime}[key][t])
        for l, row in enumerate(tp.p_lm):
            for m in range(len(row)):
                p1, p2 = copy.deepcopy(w.params), copy.deepcopy(w.params)
                p1.type_parameters[t].p_lm[l][m] += h
                p2.type_parameters[t].p_lm[l][m] -= h
                want.setdefault("plm", []).append((_oracle_T(w, p1, w.dtdf) - _oracle_T(w, p2, w.dtdf)) / (2 * h))
                got.setdefault("plm", []).append(td.p_lm[t, l * l + m])
    for key in want:
        e = rel_err(np.array(got[key]), np.array(want[key]))
        print("  %s: normwise |d - fd| / max|fd| = %.2e" % (key, e), flush=True)
        assert e < fd_tol, key


def test_set_type_parameters_changes_f_like_the_oracle():
    import copy
    w = synth.small_structure("P212121", n_atoms=7, d_min=1.3)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    p2 = copy.deepcopy(w.params)
    rng = np.random.default_rng(4)
    for t in p2.type_parameters:
        t.p_val *= 1.1
        t.kappa_spherical_valence *= 0.97
        t.kappa_deformation_valence *= 1.05
        for row in t.p_lm:
            for m in range(len(row)):
                row[m] += rng.normal() * 0.05
    calc.setTypeParameters(*_types_arrays(p2))
    f = calc.calculateStructureFactors(w.crystal.atoms, w.lcs, w.hkl)
    rf = hc_oracle.sf_and_derivs(w.crystal, p2, w.crystal.atoms, w.lcs, w.hkl, np.ones(len(w.hkl)))[0]
    assert rel_err(f, rf) < 1e-10


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_autograd_least_squares_config1(mode):
    """5000-atom P2_1 2_1 2_1 structure, T = sum (|F| - |Fobs|)^2 with Fobs from a perturbed copy
    (xyz + N(0, 0.02 A), U x U(0.9, 1.1)); the torch gradient on 200 random parameters against
    central differences (step 1e-5) of the GPU FP64 F."""
    import torch
    from discamb_amd.torch_ops import StructureFactors
    w = synth.config1()
    rng = np.random.default_rng(33)
    sf = StructureFactors(w.crystal, w.params, w.lcs, w.hkl, mode=mode)
    ref = StructureFactors(w.crystal, w.params, w.lcs, w.hkl, mode="HIP")
    params = sf.initial(w.params)
    # synthetic observations from a perturbed copy
    f2c, c2f = synth.cell_matrices(w.crystal.cell)
    pert = [p.detach().clone() for p in params]
    pert[0] += torch.from_numpy((rng.normal(0, 0.02, (w.crystal.atoms.n, 3))) @ c2f.T)
    pert[1] *= torch.from_numpy(rng.uniform(0.9, 1.1, (w.crystal.atoms.n, 1)))
    with torch.no_grad():
        fobs = ref(*pert).abs()
    print("case autograd config1", mode, flush=True)
    F = sf(*params)
    T = ((F.abs() - fobs) ** 2).sum()
    T.backward()
    grads = [p.grad.numpy().copy() for p in params]
    names = ["xyz", "adp", "occ", "plm", "pval", "kappa", "kappa_prime"]
    # 200 random parameters across the classes
    picks = []
    n = w.crystal.atoms.n
    adp_size = w.crystal.atoms.adp_size
    for _ in range(200):
        c = rng.integers(7)
        if c == 0:
            picks.append((0, (rng.integers(n), rng.integers(3))))
        elif c == 1:
            i = rng.integers(n)
            picks.append((1, (i, rng.integers(adp_size[i]))))
        elif c == 2:
            picks.append((2, (rng.integers(n),)))
        elif c == 3:
            t = rng.integers(len(w.params.type_parameters))
            L = len(w.params.type_parameters[t].p_lm) - 1
            picks.append((3, (t, rng.integers((L + 1) ** 2))))
        else:
            picks.append((c, (rng.integers(len(w.params.type_parameters)),)))
    h = 1e-5
    base = [p.detach().clone() for p in params]

    def T_of(vals):
        with torch.no_grad():
            return float(((ref(*vals).abs() - fobs) ** 2).sum())

    worst = {}
    for c, idx in picks:
        v1 = [b.clone() for b in base]
        v2 = [b.clone() for b in base]
        v1[c][idx] += h
        v2[c][idx] -= h
        fd = (T_of(v1) - T_of(v2)) / (2 * h)
        g = grads[c][idx]
        scale = np.abs(grads[c]).max()
        worst[names[c]] = max(worst.get(names[c], 0.0), abs(g - fd) / scale)
    print("  worst |g - fd| / max|g| per class:", worst, flush=True)
    tol = 1e-5 if mode == "HIP" else 1e-4
    for k, v in worst.items():
        assert v < tol, (k, v)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_autograd_device_path_config1_vs_oracle(mode):
    """Parameters as GPU tensors (the device-resident calls, no host round trip) on the 5000-atom
    structure: T = sum (|F| - |Fobs|)^2; the xyz / U / occupancy gradient equals the oracle's
    derivatives for dT/dF = 2 (|F| - |Fobs|) F / |F| (FP64 1e-10, mixed 1e-5 normwise) and the
    host path's gradient; with multipole-parameter gradients requested as well the device call
    takes the multipole adjoint and gives the same structural gradient."""
    import torch
    from discamb_amd.torch_ops import StructureFactors
    w = synth.config1()
    rng = np.random.default_rng(35)
    f2c, c2f = synth.cell_matrices(w.crystal.cell)
    a = w.crystal.atoms
    xyz_obs = a.xyz + rng.normal(0, 0.02, (a.n, 3)) @ c2f.T
    adp_obs = a.adp * rng.uniform(0.9, 1.1, (a.n, 1))
    obs_atoms = a.copy()
    obs_atoms.xyz, obs_atoms.adp = xyz_obs, adp_obs
    fobs = np.abs(hc_oracle.sf_and_derivs(w.crystal, w.params, obs_atoms, w.lcs, w.hkl, np.ones(len(w.hkl)),
                                          n_threads=16)[0])
    f0 = hc_oracle.sf_and_derivs(w.crystal, w.params, a, w.lcs, w.hkl, np.ones(len(w.hkl)), n_threads=16)[0]
    g_ref = 2 * (np.abs(f0) - fobs) * f0 / np.abs(f0)
    _, rx, ru, ro = hc_oracle.sf_and_derivs(w.crystal, w.params, a, w.lcs, w.hkl, g_ref, n_threads=16)
    sf = StructureFactors(w.crystal, w.params, w.lcs, w.hkl, mode=mode)
    fo = torch.tensor(fobs, device="cuda")
    grads = {}
    for where in ("cuda", "cpu"):
        params = sf.initial(w.params, device=where)
        for p in params[3:]:
            p.requires_grad_(False)
        print("case autograd device path mode=%s params on %s" % (mode, where), flush=True)
        F = sf(*params)
        assert F.device.type == where
        T = ((F.abs() - fo.to(where)) ** 2).sum()
        T.backward()
        grads[where] = [params[i].grad.detach().cpu().numpy() for i in range(3)]
    tol = 1e-10 if mode == "HIP" else 1e-5
    e = [rel_err(g, r) for g, r in zip(grads["cuda"], (rx, ru, ro))]
    print("  device path vs oracle derivatives", e, flush=True)
    assert max(e) <= tol, e
    e = [rel_err(g, r) for g, r in zip(grads["cuda"], grads["cpu"])]
    print("  device path vs host path", e, flush=True)
    assert max(e) <= (1e-12 if mode == "HIP" else 1e-6), e
    params = sf.initial(w.params, device="cuda")
    F = sf(*params)
    ((F.abs() - fo) ** 2).sum().backward()
    assert params[3].grad is not None and params[3].grad.device.type == "cuda"
    assert rel_err(params[0].grad.cpu().numpy(), grads["cuda"][0]) <= (1e-12 if mode == "HIP" else 1e-6)


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
@pytest.mark.parametrize("sg", ["P1", "P21/c", "P212121", "P43212", "P63"])
def test_autograd_device_multipole_gradients_vs_host_path(sg, mode):
    """Every parameter as a GPU tensor: the backward runs dh_hc_sf_and_derivs_mp_dev (multipole
    adjoint and per-type sums on the device) and gives the gradients of the host path (parameters
    on the host: dh_hc_sf_and_derivs + dh_hc_fetch_type_derivatives) for xyz, U, occupancy, P_lm,
    Pval, kappa and kappa' (FP64 1e-12, mixed 1e-6 normwise); an excluded atom and the P_lm
    entries beyond a type's multipole order get no gradient."""
    import torch
    from discamb_amd.torch_ops import StructureFactors
    w = synth.small_structure(sg, n_atoms=10, d_min=1.0, seed=21)
    rng = np.random.default_rng(5)
    include = np.ones(w.crystal.atoms.n, dtype=bool)
    include[3] = False
    sf = StructureFactors(w.crystal, w.params, w.lcs, w.hkl, mode=mode, include=include)
    fobs = np.abs(rng.normal(size=len(w.hkl)) + 1j * rng.normal(size=len(w.hkl))) * 10.0
    grads = {}
    for where in ("cuda", "cpu"):
        params = sf.initial(w.params, device=where)
        print("case device multipole gradients sg=%s mode=%s params on %s" % (sg, mode, where), flush=True)
        F = sf(*params)
        T = ((F.abs() - torch.tensor(fobs, device=where)) ** 2).sum()
        T.backward()
        for p in params:
            assert p.grad is not None and p.grad.device.type == where
        grads[where] = [p.grad.detach().cpu().numpy() for p in params]
    tol = 1e-12 if mode == "HIP" else 1e-6
    names = ["xyz", "adp", "occ", "plm", "pval", "kappa", "kappa_prime"]
    e = {n: rel_err(g, r) for n, g, r in zip(names, grads["cuda"], grads["cpu"])}
    print("  device vs host path", e, flush=True)
    assert max(e.values()) <= tol, e
    assert np.all(grads["cuda"][0][3] == 0) and grads["cuda"][2][3] == 0
    for t, tp in enumerate(w.params.type_parameters):
        n_l = len(tp.p_lm)
        assert np.all(grads["cuda"][3][t, n_l * n_l:] == 0)
