"""Device-resident calls (dh_hc_set_reflections + dh_hc_sf_and_derivs_dev) against the host call and
the oracle.

The device call converts the caller's conventions on the GPU (xyz, ADPs, dT/dF, derivative
conventions), so it equals the host call to rounding, not bitwise.  Tolerances (normwise
max|d - d_ref| / max|d_ref| per output class):
  device vs host call, FP64            1e-13
  device vs host call, mixed           1e-6 (F), 1e-5 (derivatives): the FP32 rounding of an input
                                       can differ when its FP64 conversion differs by an ulp
  FP64 device call vs oracle            1e-10
  mixed device call vs oracle           1e-6 (F), 1e-5 (derivatives)
  graph replay vs eager device call     bitwise
"""
import ctypes

import numpy as np
import pytest

import hc_oracle
from conftest import rel_err
from discamb_amd import DerivativesSelector, HansenCoppensStructureFactorCalculator, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = {"HIP": (1e-10, 1e-10), "HIP_MIXED": (1e-6, 1e-5), "CPU_IAM": (1e-10, 1e-10)}

I222 = ["x,y,z", "-x,-y,z", "-x,y,-z", "x,-y,-z", "x+1/2,y+1/2,z+1/2", "-x+1/2,-y+1/2,z+1/2",
        "-x+1/2,y+1/2,-z+1/2", "x+1/2,-y+1/2,-z+1/2"]


def _structure(case, n_atoms=14, d_min=0.95, seed=17):
    if case == "P-1_shifted":
        w = synth.small_structure("P-1", n_atoms=n_atoms, d_min=d_min, seed=seed)
        w.crystal.symops = ["x,y,z", "-x+1/2,-y,-z+1/2"]
    elif case == "I222":
        w = synth.small_structure("P212121", n_atoms=n_atoms, d_min=d_min, seed=seed)
        w.crystal.symops = I222
        w.crystal.atoms.multiplicity[:] = 8
    else:
        w = synth.small_structure(case, n_atoms=n_atoms, d_min=d_min, seed=seed)
    return w


def _dev(a, lcs=None, anomalous=None, include=None):
    t = lambda x, dt: torch.tensor(np.ascontiguousarray(x), dtype=dt, device="cuda")
    out = {"xyz": t(a.xyz, torch.float64), "adp_size": t(a.adp_size, torch.int32), "adp": t(a.adp, torch.float64),
           "occupancy": t(a.occupancy, torch.float64)}
    if lcs is not None:
        out["localCoordinateSystems"] = t(np.asarray(lcs).reshape(-1, 3, 3), torch.float64)
    if anomalous is not None:
        out["anomalous"] = t(anomalous, torch.complex128)
    if include is not None:
        out["countAtomContribution"] = t(np.asarray(include, dtype=np.uint8), torch.uint8)
    return out


def _np(x):
    return None if x is None else x.detach().cpu().numpy()


def _errs(f, dx, du, do, rf, rx, ru, ro):
    return {"F": rel_err(f, rf), "dxyz": rel_err(dx, rx), "dadp": rel_err(du, ru), "docc": rel_err(do, ro)}


def _host_bar(e, mode):
    if mode == "HIP_MIXED":
        assert e["F"] <= 1e-6 and max(e.values()) <= 1e-5, e
    else:
        assert max(e.values()) <= 1e-13, e


def _check_case(w, mode, include=None, anomalous=None, selector=None):
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    a = w.crystal.atoms
    hf, hd = calc.calculateStructureFactorsAndDerivatives(a, w.lcs, w.hkl, w.dtdf, include, selector,
                                                          anomalous=anomalous)
    calc.setReflections(w.hkl)
    g = torch.tensor(w.dtdf, dtype=torch.complex128, device="cuda")
    f, dx, du, do = calc.calculateOnDevice(dTarget_dF=g, derivativesSelector=selector,
                                           **_dev(a, w.lcs, anomalous, include))
    torch.cuda.synchronize()
    f, dx, du, do = _np(f), _np(dx), _np(du), _np(do)
    e_host = _errs(f, dx, du, do, hf, hd.atomic_position_derivatives, hd.adp_derivatives, hd.occupancy_derivatives)
    print("  device vs host %s" % e_host, flush=True)
    _host_bar(e_host, mode)
    return calc, f, dx, du, do


CASES = ["P1", "P-1", "P21/c", "C2/c", "P212121", "P43212", "P31", "P63", "P4/m", "P-1_shifted", "I222"]


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
@pytest.mark.parametrize("case", CASES)
def test_device_call_matches_host_call_and_oracle(case, mode):
    w = _structure(case)
    print("case device call %s mode=%s atoms=%d hkl=%d" % (case, mode, w.crystal.atoms.n, len(w.hkl)), flush=True)
    _, f, dx, du, do = _check_case(w, mode)
    ref = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, n_threads=16)
    e = _errs(f, dx, du, do, *ref)
    print("  device vs oracle %s" % e, flush=True)
    tf, td = TOL[mode]
    assert e["F"] <= tf and max(e["dxyz"], e["dadp"], e["docc"]) <= td, e


@pytest.mark.parametrize("case", ["P21/c", "P212121"])
def test_device_call_include_anomalous_and_selector(case):
    w = _structure(case, n_atoms=12)
    n = w.crystal.atoms.n
    inc = np.ones(n, dtype=bool)
    inc[[2, 7]] = False
    rng = np.random.default_rng(4)
    anom = rng.normal(size=n) * 0.3 + 1j * np.abs(rng.normal(size=n)) * 0.2
    print("case device call include/anomalous %s" % case, flush=True)
    _, f, dx, du, do = _check_case(w, "HIP", include=inc, anomalous=anom)
    assert np.all(dx[~inc] == 0) and np.all(du[~inc] == 0) and np.all(do[~inc] == 0)
    ref = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, include=inc,
                                  anomalous=anom, n_threads=16)
    assert max(_errs(f, dx, du, do, *ref).values()) <= 1e-10
    print("case device call selector %s" % case, flush=True)
    _, f, dx, du, do = _check_case(w, "HIP", selector=DerivativesSelector(True, False, True))
    assert np.all(du == 0) and np.any(dx != 0) and np.any(do != 0)


def test_device_call_cartesian_and_u_cart_conventions():
    w = _structure("P21/c", n_atoms=10)
    f2c, _ = synth.cell_matrices(w.crystal.cell)
    a = w.crystal.atoms.copy()
    a.xyz = a.xyz @ f2c.T
    J = hc_oracle.u_cif_to_cart_matrix(w.crystal.cell)
    for i in range(a.n):
        if a.adp_size[i] == 6:
            a.adp[i] = J @ a.adp[i]
    w.crystal.atoms = a
    w.crystal.xyz_convention = "cartesian"
    w.crystal.adp_convention = "U_cart"
    print("case device call Cartesian / U_cart", flush=True)
    _, f, dx, du, do = _check_case(w, "HIP")
    ref = hc_oracle.sf_and_derivs(w.crystal, w.params, a, w.lcs, w.hkl, w.dtdf)
    assert max(_errs(f, dx, du, do, *ref).values()) <= 1e-10


def test_device_call_f_only_iam_and_electron():
    w = _structure("P-1", n_atoms=9)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setReflections(w.hkl)
    args = _dev(w.crystal.atoms, w.lcs)
    f, dx, du, do = calc.calculateOnDevice(**args)
    assert dx is None
    rf = calc.calculateStructureFactors(w.crystal.atoms, w.lcs, w.hkl)
    assert rel_err(_np(f), rf) <= 1e-13
    for mode, electron in (("CPU_IAM", False), ("HIP", True), ("HIP_MIXED", True)):
        print("case device call mode=%s electron=%s" % (mode, electron), flush=True)
        calc.setCalculationMode(mode)
        calc.setElectronScattering(electron)
        hf, hd = calc.calculateStructureFactorsAndDerivatives(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
        calc.setReflections(w.hkl)
        g = torch.tensor(w.dtdf, dtype=torch.complex128, device="cuda")
        f, dx, du, do = calc.calculateOnDevice(dTarget_dF=g, **args)
        e = _errs(_np(f), _np(dx), _np(du), _np(do), hf, hd.atomic_position_derivatives, hd.adp_derivatives,
                  hd.occupancy_derivatives)
        print("  device vs host %s" % e, flush=True)
        _host_bar(e, mode)
    calc.setElectronScattering(False)


def test_device_call_errors():
    w = _structure("P1", n_atoms=5, d_min=1.4)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    args = _dev(w.crystal.atoms, w.lcs)
    with pytest.raises(RuntimeError):          # no reflections set
        calc.calculateOnDevice(**args)
    calc.setReflections(w.hkl)
    g = torch.tensor(w.dtdf, dtype=torch.complex128, device="cuda")
    mp = DerivativesSelector(d_multipole=True)
    with pytest.raises(ValueError):            # no allocation-free form of the multipole call
        calc.calculateOnDevice(dTarget_dF=g, derivativesSelector=mp, no_alloc=True, **args)
    # and the C entry points without the multipole outputs refuse the DH_D_MULTIPOLE bit
    a, keep, _ = calc._dev_args(args["xyz"], args["adp_size"], args["adp"], args["occupancy"],
                                args.get("localCoordinateSystems"), None, None)
    f = torch.empty(len(w.hkl), dtype=torch.complex128, device="cuda")
    gr = torch.view_as_real(g)
    for fn in (calc._lib.dh_hc_sf_and_derivs_dev, calc._lib.dh_hc_sf_and_derivs_dev_noalloc):
        rc = fn(calc._h, ctypes.byref(a), gr.data_ptr(), mp.bits(), torch.view_as_real(f).data_ptr(), None, None,
                None, None)
        assert rc != 0
    torch.cuda.synchronize()
    # with the outputs: accepted
    res = calc.calculateOnDevice(dTarget_dF=g, derivativesSelector=mp, **args)
    assert len(res) == 5 and res[4].p_lm.shape == (calc.n_types, 25)
    bad = dict(args)
    bad["xyz"] = args["xyz"].float()
    with pytest.raises(RuntimeError):          # wrong dtype
        calc.calculateOnDevice(**bad)
    calc.calculateStructureFactors(w.crystal.atoms, w.lcs, w.hkl)
    with pytest.raises(RuntimeError):          # a host call replaced the device layout
        calc.calculateOnDevice(**args)
    calc.setReflections(w.hkl)
    calc.calculateOnDevice(dTarget_dF=g, **args)
    with pytest.raises(RuntimeError):          # device results are not fetched
        calc.fetch()


def test_device_call_reflections_after_kernel_option_changes():
    """Only the engine prepared by the latest setReflections serves device-resident calls.  In
    P4_3 2_1 2 the general-rotation option switches between the coset engine and the general one;
    a list set on the coset engine and then replaced while the option pointed at the other engine
    is never used: switching back needs the reflections set again."""
    w = _structure("P43212", n_atoms=10, d_min=1.0)
    print("case device call after option changes P43212 atoms=%d hkl=%d" % (w.crystal.atoms.n, len(w.hkl)), flush=True)
    short = np.ascontiguousarray(w.hkl[: len(w.hkl) // 2])
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    args = _dev(w.crystal.atoms, w.lcs)
    calc.setReflections(short)                       # coset engine, shorter list
    calc.setKernelOptions(general_rotations=True)
    calc.setReflections(w.hkl)                       # general engine, full list
    f_general = _np(calc.calculateOnDevice(**args)[0])
    calc.setKernelOptions(general_rotations=False)
    with pytest.raises(RuntimeError):                # the coset engine's shorter list was dropped
        calc.calculateOnDevice(**args)
    calc.setReflections(w.hkl)
    f_coset = _np(calc.calculateOnDevice(**args)[0])
    torch.cuda.synchronize()
    ref = hc_oracle.sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, n_threads=16)[0]
    for f in (f_general, f_coset):
        assert f.shape == ref.shape and rel_err(f, ref) <= 1e-10
    # the split host calls likewise: an upload replaces the inputs of every engine
    print("case host upload/run after option changes P43212", flush=True)
    a = w.crystal.atoms
    calc.upload(a, w.lcs, short, w.dtdf[: len(short)])      # coset engine, shorter list
    calc.run()
    calc.fetch()
    calc.setKernelOptions(general_rotations=True)
    calc.upload(a, w.lcs, w.hkl, w.dtdf)                     # general engine, full list
    calc.setKernelOptions(general_rotations=False)
    with pytest.raises(RuntimeError):                        # the coset engine's upload was dropped
        calc.run()
    calc.upload(a, w.lcs, w.hkl, w.dtdf)
    calc.run()
    f_host = calc.fetch()[0]
    assert rel_err(f_host, ref) <= 1e-10


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_device_call_in_a_captured_graph(mode):
    """The device call captured into a torch CUDA graph and replayed after new coordinates are
    copied into its input tensors: every replay equals an eager device call on the same inputs,
    with eager work and synchronisations between the replays."""
    w = synth.small_structure("P212121", n_atoms=150, d_min=0.8, seed=3)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    calc.setReflections(w.hkl)
    args = _dev(w.crystal.atoms, w.lcs)
    g = torch.tensor(w.dtdf, dtype=torch.complex128, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        calc.calculateOnDevice(dTarget_dF=g, **args)        # warm-up off the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    print("case graph capture mode=%s" % mode, flush=True)
    with torch.cuda.graph(graph):
        out = calc.calculateOnDevice(dTarget_dF=g, **args)
    host = HansenCoppensStructureFactorCalculator(w.crystal, w.params)    # host reference, own buffers
    host.setCalculationMode(mode)
    rng = np.random.default_rng(9)
    x0 = w.crystal.atoms.xyz.copy()
    for rep in range(4):
        x = x0 + rng.normal(0, 0.003, x0.shape)
        args["xyz"].copy_(torch.tensor(x, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = [_np(t).copy() for t in out]
        eager = calc.calculateOnDevice(dTarget_dF=g, **args)
        torch.cuda.synchronize()
        for a_, b_ in zip(got, eager):
            assert np.array_equal(a_, _np(b_)), rep
        a = w.crystal.atoms.copy()
        a.xyz = x
        hf, hd = host.calculateStructureFactorsAndDerivatives(a, w.lcs, w.hkl, w.dtdf)
        _host_bar({"F": rel_err(got[0], hf), "dxyz": rel_err(got[1], hd.atomic_position_derivatives)}, mode)
        print("  replay %d ok" % rep, flush=True)
