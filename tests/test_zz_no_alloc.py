"""The allocation-free device call (dh_hc_sf_and_derivs_dev_noalloc) against the allocating device
call it varies (dh_hc_sf_and_derivs_dev), through the C ABI.

Both run the same kernels; the allocation-free call takes every per-call buffer from the
library's static per-GPU scratch.  Where the launch shape is the same the results are bitwise
equal; the bar is the path's parity tolerance (normwise max|d - d_ref| / max|d_ref| per output
class): FP64 1e-10, mixed 1e-6 (F) / 1e-5 (derivatives).
"""
import numpy as np
import pytest

from conftest import rel_err
from discamb_amd import HansenCoppensStructureFactorCalculator, synth
from discamb_amd._lib import DhError

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = {"HIP": (1e-10, 1e-10), "HIP_MIXED": (1e-6, 1e-5)}


def _dev(a, lcs):
    t = lambda x, dt: torch.tensor(np.ascontiguousarray(x), dtype=dt, device="cuda")
    return {"xyz": t(a.xyz, torch.float64), "adp_size": t(a.adp_size, torch.int32), "adp": t(a.adp, torch.float64),
            "occupancy": t(a.occupancy, torch.float64),
            "localCoordinateSystems": t(np.asarray(lcs).reshape(-1, 3, 3), torch.float64)}


def _outputs(n, nh):
    z = lambda *s: torch.full(s, np.nan, dtype=torch.float64, device="cuda")
    return (torch.full((nh,), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda"), z(n, 3), z(n, 6), z(n))


def _compare(w, mode, calc=None):
    calc = calc or HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setCalculationMode(mode)
    calc.setReflections(w.hkl)
    args = _dev(w.crystal.atoms, w.lcs)
    g = torch.tensor(w.dtdf, dtype=torch.complex128, device="cuda")
    ref = calc.calculateOnDevice(dTarget_dF=g, **args)
    out = _outputs(w.crystal.atoms.n, len(w.hkl))      # NaN-filled: every entry must be written
    got = calc.calculateOnDevice(dTarget_dF=g, no_alloc=True, out=out, **args)
    torch.cuda.synchronize()
    ref = [t.cpu().numpy() for t in ref]
    got = [t.cpu().numpy() for t in got]
    errs = {k: rel_err(a, b) for k, a, b in zip(("F", "dxyz", "dadp", "docc"), got, ref)}
    bitwise = all(np.array_equal(a, b) for a, b in zip(got, ref))
    print("  %s %d atoms %d reflections: no-alloc vs allocating %s bitwise=%s" % (
        mode, w.crystal.atoms.n, len(w.hkl), errs, bitwise), flush=True)
    tf, td = TOL[mode]
    assert errs["F"] <= tf and max(errs["dxyz"], errs["dadp"], errs["docc"]) <= td, errs
    return calc


@pytest.mark.parametrize("mode", ["HIP", "HIP_MIXED"])
def test_noalloc_matches_the_allocating_call_twice_second_smaller(mode):
    big = synth.small_structure("P212121", n_atoms=300, d_min=0.7, seed=11, cell=(16.0, 18.0, 20.0, 90.0, 90.0, 90.0))
    small = synth.small_structure("P21/c", n_atoms=20, d_min=1.0, seed=12)
    print("case no-alloc mode=%s first call (larger)" % mode, flush=True)
    _compare(big, mode)
    print("case no-alloc mode=%s second call (smaller, another calculator)" % mode, flush=True)
    _compare(small, mode)


def test_noalloc_headline_size_and_refusals():
    """The 5000-atom / 66 592-reflection workload fits the scratch in FP64; 224 000 reflections do
    not (refused, nothing computed); records of another calculation mode are refused until the
    allocating call or dh_hc_set_reflections prepared them."""
    w = synth.config1()
    print("case no-alloc 5000 atoms FP64", flush=True)
    calc = _compare(w, "HIP")
    # mixed records not prepared: refused, then prepared by an allocating call
    calc.setCalculationMode("HIP_MIXED")
    args = _dev(w.crystal.atoms, w.lcs)
    g = torch.tensor(w.dtdf, dtype=torch.complex128, device="cuda")
    out = _outputs(w.crystal.atoms.n, len(w.hkl))
    with pytest.raises(DhError) as e:
        calc.calculateOnDevice(dTarget_dF=g, no_alloc=True, out=out, **args)
    assert e.value.code == 4
    print("case no-alloc 5000 atoms mixed after the allocating call", flush=True)
    ref = calc.calculateOnDevice(dTarget_dF=g, **args)
    got = calc.calculateOnDevice(dTarget_dF=g, no_alloc=True, out=out, **args)
    torch.cuda.synchronize()
    assert rel_err(got[0].cpu().numpy(), ref[0].cpu().numpy()) <= 1e-6
    big = synth.config1(d_min=0.8)
    print("case no-alloc refusal: %d atoms %d reflections" % (big.crystal.atoms.n, len(big.hkl)), flush=True)
    c2 = HansenCoppensStructureFactorCalculator(big.crystal, big.params)
    c2.setReflections(big.hkl)
    g2 = torch.tensor(big.dtdf, dtype=torch.complex128, device="cuda")
    out2 = _outputs(big.crystal.atoms.n, len(big.hkl))
    with pytest.raises(DhError) as e:
        c2.calculateOnDevice(dTarget_dF=g2, no_alloc=True, out=out2, **_dev(big.crystal.atoms, big.lcs))
    assert e.value.code == 4 and "scratch" in str(e.value)
    torch.cuda.synchronize()
    assert torch.isnan(out2[1]).all()


def test_noalloc_call_in_a_captured_graph():
    """Captured with preallocated outputs (nothing allocated inside the capture), replayed after
    new coordinates: every replay equals an eager allocation-free call on the same inputs."""
    w = synth.small_structure("P-1", n_atoms=80, d_min=0.8, seed=5)
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params)
    calc.setReflections(w.hkl)
    args = _dev(w.crystal.atoms, w.lcs)
    g = torch.tensor(w.dtdf, dtype=torch.complex128, device="cuda")
    out = _outputs(w.crystal.atoms.n, len(w.hkl))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        calc.calculateOnDevice(dTarget_dF=g, no_alloc=True, out=out, **args)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    print("case no-alloc graph capture", flush=True)
    with torch.cuda.graph(graph):
        calc.calculateOnDevice(dTarget_dF=g, no_alloc=True, out=out, **args)
    rng = np.random.default_rng(2)
    x0 = w.crystal.atoms.xyz.copy()
    eager_out = _outputs(w.crystal.atoms.n, len(w.hkl))
    for rep in range(3):
        args["xyz"].copy_(torch.tensor(x0 + rng.normal(0, 0.003, x0.shape), dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = [t.cpu().numpy().copy() for t in out]
        calc.calculateOnDevice(dTarget_dF=g, no_alloc=True, out=eager_out, **args)
        torch.cuda.synchronize()
        for a, b in zip(got, eager_out):
            assert np.array_equal(a, b.cpu().numpy()), rep
        print("  replay %d ok" % rep, flush=True)
