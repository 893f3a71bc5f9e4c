#!/usr/bin/env python3
"""Benchmark: Hansen-Coppens F + derivatives throughput on one MI355X.

Workload: 5000 protein-like atoms in P2_1 2_1 2_1 (50 x 60 x 70 A), 40 shared multipole types,
all 66 592 unique reflections to 1.2 A, dT/dF complex N(0,1) (synthetic, seeded).  One step =
one calculateStructureFactorsAndDerivatives pass (per-call prep kernels + fused F/adjoint kernel +
stage-2 reductions) on inputs already resident in HBM.  The headline value is the mixed-precision
path; the FP64 drop-in path is measured the same way and reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK = 157.3e12      # MI355X published FP32 vector peak (FLOP/s)
FP64_PEAK = 78.6e12
# HBM bytes per launch of the main kernels on this workload, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes scaled by a plain-copy calibration (profiles/*_hbm_traffic.md, "bench-traffic" lines)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_main_p212121_5000at_1.2A_hbm_traffic.md")
TRAFFIC_FILE_ROWS = os.path.join(ROOT, "profiles", "k_jac_rows_hbm_traffic.md")
BATCH_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "batch256_hbm_traffic.md")


def measured_traffic(kernel, path=TRAFFIC_FILE):
    """HBM bytes per launch recorded for `kernel` at this workload, or None."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("bench-traffic ") and line.split()[1] == kernel:
                    return float(line.split()[2])
    except OSError:
        pass
    return None


def batch_traffic(chunk, path=BATCH_TRAFFIC_FILE):
    """(chunk, HBM bytes per batched step) recorded for the 256-structure workload: at this mixed
    chunk if recorded, else at the recorded chunk closest to it; (None, None) without a record."""
    rec = {}
    try:
        with open(path) as fh:
            for line in fh:
                f = line.split()
                if line.startswith("batch-traffic "):
                    rec[int(f[1])] = float(f[2])
    except OSError:
        pass
    if not rec:
        return None, None
    c = min(rec, key=lambda x: abs(x - chunk))
    return c, rec[c]


def flops_per_term(L):
    """F + xyz/U/occ derivative flops per (atom, op, reflection); FMA = 2."""
    return 66 + 2 * (L + 1) ** 2 + 2 * (L + 1)


def algorithmic_flops(w):
    n_ops = w.n_explicit_ops
    tot = 0.0
    for i in range(w.crystal.atoms.n):
        t = w.params.type_parameters[w.params.atom_to_type_map[i]]
        L = len(t.p_lm) - 1
        tot += flops_per_term(4 if L > 2 else 2)
    return tot * n_ops * w.hkl.shape[0]


def time_mode(calc, mode, steps, warmup, tune=True):
    calc.setCalculationMode(mode)
    calc.setTiming(False)
    if tune:
        calc.autotune()
    for _ in range(warmup):
        calc.run()
    if warmup:
        calc.fetch()
    calc.setTiming(True)
    main_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        calc.run()
        main_ms.append(calc.lastTiming()["main_ms"])
    f, d = calc.fetch()          # synchronises the stream
    t1 = time.perf_counter()
    info = calc.lastTiming()
    return (t1 - t0) / steps, float(np.mean(main_ms)), info, f, d


def time_host_calls(calc, w, calls=10, retype=False):
    """Whole calls from host buffers: upload (atoms and dT/dF; the unchanged reflection list stays on the
    device with its order and the reflection records), run, download of F and derivatives.  retype:
    new atom-type parameters (kappa, kappa' scaled by turns) before every call, which rebuilds the
    radial records and tables the calls otherwise reuse.  The mode and tuning are whatever `calc`
    last used."""
    tp = w.params.type_parameters
    plm = np.array([t.plm25() for t in tp])
    pv = [t.p_val for t in tp]
    ka = np.array([t.kappa_spherical_valence for t in tp])
    kp = np.array([t.kappa_deformation_valence for t in tp])
    calc.setTiming(False)
    calc.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    calc.run()
    calc.fetch()
    t0 = time.perf_counter()
    for i in range(calls):
        if retype:
            sc = 1.0 + 1e-3 * (i % 2)
            calc.setTypeParameters(plm, pv, ka * sc, kp / sc)
        calc.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
        calc.run()
        calc.fetch()
    dt = (time.perf_counter() - t0) / calls
    if retype:
        calc.setTypeParameters(plm, pv, ka, kp)
    return dt


# Calibration of the CPU restatement against the reference itself, measured in the build container
# (Intel Xeon @ 2.70 GHz, 8 threads) on one shape: 500 atoms of one type each, P2_1 2_1 2_1,
# 50 x 60 x 70 A, 66 592 reflections to 1.2 A.  The reference's Engine3 (built from its sources with
# its own CMake build, engine "CPU") ran it at 1.46e8 terms/s; the line-recurrence restatement
# (oracle hc_oracle_engine_lines) runs it at 6.6e7 terms/s.  The reference cannot be built from the
# repository, so the box times the restatement and reports this ratio beside it.
CPU_CALIBRATION = {"shape": "500 atoms (one type each), P212121 50x60x70 A, 66592 hkl to 1.2 A, 8 threads of an "
                            "Intel Xeon @ 2.70 GHz",
                   "reference_engine3_terms_per_s": 1.46e8, "restatement_terms_per_s": 6.6e7,
                   "restatement_over_reference_rate": 6.6e7 / 1.46e8}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, budget_refl=66592, label="the same 5000-atom P212121 structure"):
    """The FP64 OpenMP CPU restatement on the host cores: for acentric groups the Engine3-shaped one
    with hkl-line recurrences (global P_lm, per-reflection tables, products along lines), otherwise
    the direct oracle (sin/cos/exp per term, deformation term in each atom's frame)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import hc_oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(11)
    sel = np.sort(rng.choice(len(w.hkl), min(budget_refl, len(w.hkl)), replace=False))
    centro = hc_oracle.decompose_group(w.crystal.symops)[2]
    fn = hc_oracle.sf_and_derivs if centro else hc_oracle.sf_and_derivs_lines
    t0 = time.perf_counter()
    fn(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl[sel], w.dtdf[sel], n_threads=threads)
    dt = time.perf_counter() - t0
    terms = w.crystal.atoms.n * w.n_explicit_ops * len(sel)
    how = ("direct FP64 OpenMP oracle (sin/cos/exp per term)" if centro else
           "FP64 OpenMP restatement of Engine3 with hkl-line recurrences")
    return {"value": terms / dt, "unit": "terms/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": "%s on %s, %d of its %d reflections, F + xyz/U/occ derivatives, %.1f s wall" % (
                how, label, len(sel), len(w.hkl), dt),
            "calibration": CPU_CALIBRATION if not centro else None}


def bench_batch(args):
    """256 uneven structures per step through one batched launch sequence replayed from a
    hipGraph; per-structure step times from one-structure batches of the same structures."""
    from discamb_amd import BatchCalculator, HansenCoppensStructureFactorCalculator, synth

    t0 = time.perf_counter()
    ws = synth.config4(256)
    calcs = [HansenCoppensStructureFactorCalculator(w.crystal, w.params, device=0) for w in ws]
    for w, c in zip(ws, calcs):
        c.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    terms = float(sum(w.terms for w in ws))
    terms_g = float(sum(w.crystal.atoms.n * len(w.crystal.symops) * len(w.hkl) for w in ws))
    flops = float(sum(algorithmic_flops(w) for w in ws))
    print("batch workload: %d structures, %.3e terms per step, built in %.1f s" % (len(ws), terms,
          time.perf_counter() - t0), file=sys.stderr, flush=True)
    batch = BatchCalculator(calcs)
    res = {}
    for mode in ("HIP_MIXED", "HIP"):
        for c in calcs:
            c.setCalculationMode(mode)
        fixed = (args.chunk_fp64 or args.chunk) if mode == "HIP" else args.chunk
        batch.setTuning(fixed)
        chunk = batch.autotune() if fixed == 0 else fixed
        for _ in range(args.warmup):
            batch.run(graph=True)
        dev = []
        t1 = time.perf_counter()
        for _ in range(args.steps):
            batch.run(graph=True)
            dev.append(batch.lastTiming()["step_ms"])
        t2 = time.perf_counter()
        out = [c.fetch() for c in calcs]
        # each structure alone: a one-structure batch (same kernels, graph replay, its own tuned
        # reflection chunk), device time
        single = []
        for c in calcs if args.singles else []:
            b1 = BatchCalculator([c])
            if fixed:
                b1.setTuning(fixed)
            else:
                b1.autotune()
            b1.run(graph=True)
            t = []
            for _ in range(3):
                b1.run(graph=True)
                t.append(b1.lastTiming()["step_ms"])
            single.append(min(t))
            del b1
        res[mode] = {"wall_ms": (t2 - t1) / args.steps * 1e3, "device_ms": float(np.median(dev)), "chunk": chunk,
                     "out": out, "single_ms": np.array(single), "units": batch.lastTiming()["units"],
                     "grid": batch.lastTiming()["grid"]}
        print("%s: batched step %.3f ms wall, %.3f ms device, chunk %d%s" % (
            mode, res[mode]["wall_ms"], res[mode]["device_ms"], chunk, "; one structure at a time: sum %.3f ms, p50 %.3f "
            "ms, p99 %.3f ms" % (np.sum(single), np.percentile(single, 50), np.percentile(single, 99)) if single else ""),
            file=sys.stderr, flush=True)
    mx, fp = res["HIP_MIXED"], res["HIP"]
    errs = [float(np.abs(a[0] - b[0]).max() / np.abs(b[0]).max()) for a, b in zip(mx["out"], fp["out"])]

    def pct(r):
        if len(r["single_ms"]) == 0:
            return None
        return {"p50_ms": float(np.percentile(r["single_ms"], 50)), "p99_ms": float(np.percentile(r["single_ms"], 99)),
                "sum_ms": float(np.sum(r["single_ms"]))}

    achieved = flops / (mx["device_ms"] * 1e-3)
    line = {
        "metric": "atom*symop*hkl terms/s for F+derivs (mixed precision), 256-structure batched step",
        "value": terms / (mx["device_ms"] * 1e-3),
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mx["device_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 terms, f64 phase reduction and accumulation",
        "data": "synthetic, seeded (numpy PCG64 seed 4000)",
        "config": {"workload": "256 uneven structures: 20-400 atoms log-uniform, P1/P-1/P21/c/C2/c/P212121, "
                               "cell edges 6-25 A, d_min 0.5-0.9 A (2k-40k reflections), one type per atom",
                   "structures": len(ws), "terms_per_step": terms,
                   "terms_per_step_full_group": terms_g,
                   "term_count": "terms_per_step (the value's count) = sum of atoms x explicit operations x "
                                 "reflections, the work done (inversion and centring folded analytically); "
                                 "terms_per_step_full_group counts atoms x |G| x reflections",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "roof": "FP32 vector ALU; whole batched step (prep + main + reductions)",
                     "achieved": achieved / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK, "traffic": batch_traffic(mx["chunk"])[1],
                     "traffic_chunk": batch_traffic(mx["chunk"])[0],
                     "traffic_note": "HBM bytes per batched step, summed over the step's kernels, from the calibrated "
                                     "FETCH_SIZE + WRITE_SIZE counters recorded for this workload at traffic_chunk "
                                     "(profiles/batch256_hbm_traffic.md)"},
        "step_time": {"device_ms_mixed": mx["device_ms"], "wall_ms_mixed": mx["wall_ms"],
                      "device_ms_fp64": fp["device_ms"], "wall_ms_fp64": fp["wall_ms"],
                      "note": "device = events around the graph launch; wall adds descriptor upload and sync"},
        "per_structure_mixed": pct(mx),
        "per_structure_fp64": pct(fp),
        "fp64": {"value": terms / (fp["device_ms"] * 1e-3), "unit": "terms/s", "ms_per_step": fp["device_ms"]},
        "speedup_mixed_vs_fp64": fp["device_ms"] / mx["device_ms"],
        "batching_speedup_mixed": pct(mx)["sum_ms"] / mx["device_ms"] if pct(mx) else None,
        "accuracy_mixed_vs_fp64": {"max_over_structures_max_dF_over_maxF": max(errs)},
        "launch": {"chunk_mixed": mx["chunk"], "chunk_fp64": fp["chunk"], "units": mx["units"],
                   "persistent_grid": mx["grid"], "graph": True},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import hc_oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sample = ws[:8]
        t0 = time.perf_counter()
        for w in sample:
            centro = hc_oracle.decompose_group(w.crystal.symops)[2]
            fn = hc_oracle.sf_and_derivs if centro else hc_oracle.sf_and_derivs_lines
            fn(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, w.dtdf, n_threads=threads)
        dt = time.perf_counter() - t0
        st = sum(w.terms for w in sample)
        line["cpu_baseline"] = {"value": st / dt, "unit": "terms/s", "cores": threads, "kind": "port",
                                "cpu_model": cpu_model(),
                                "sample": "FP64 OpenMP restatement (hkl-line recurrences for the acentric groups, the "
                                          "direct oracle for the centrosymmetric ones) on the first 8 structures of "
                                          "the batch (%s; %.3e terms), %.1f s wall" % (
                                              ", ".join(w.name for w in sample), st, dt),
                                "calibration": CPU_CALIBRATION}
    print(json.dumps(line), flush=True)


def bench_per_hkl(args):
    """Per-reflection derivative rows (the per-hkl path): F and dF/d(xyz, ADP, occ) of every atom
    for every reflection, FP64 rows, written by k_jac into device memory (dh_hc_jacobian_rows_dev),
    with FP64 terms (HIP, the value) and with mixed-precision terms (HIP_MIXED, reported beside).
    Two structures: the 60-atom P2_1/c small molecule (all its reflections in one launch) and the
    5000-atom P2_1 2_1 2_1 structure (a block of 2400 reflections, 1.9 GB of rows, per launch)."""
    import ctypes

    from discamb_amd import HansenCoppensStructureFactorCalculator, _lib, synth

    hip = ctypes.CDLL("libamdhip64.so.7")      # the runtime libdiscamb_amd.so itself uses
    lib = _lib.load()
    out = {}
    for key, w, m, mode in (("p21c_60", synth.config0(), None, "HIP"), ("p212121_5000", synth.config1(), 2400, "HIP"),
                            ("p21c_60_mixed", synth.config0(), None, "HIP_MIXED"),
                            ("p212121_5000_mixed", synth.config1(), 2400, "HIP_MIXED")):
        n = w.crystal.atoms.n
        m = len(w.hkl) if m is None else m
        c = HansenCoppensStructureFactorCalculator(w.crystal, w.params, device=0)
        c.setCalculationMode(mode)
        c.jacobianUpload(w.crystal.atoms, w.lcs, w.hkl)
        fb, rb = 16 * m, 16 * m * 10 * n
        pf, pr = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(pf), ctypes.c_size_t(fb)) or hip.hipMalloc(ctypes.byref(pr), ctypes.c_size_t(rb)):
            raise RuntimeError("hipMalloc failed")
        for _ in range(args.warmup):
            _lib.check(lib.dh_hc_jacobian_rows_dev(c._h, 0, m, pf, pr, None))
        hip.hipDeviceSynchronize()
        c.setTiming(True)
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _lib.check(lib.dh_hc_jacobian_rows_dev(c._h, 0, m, pf, pr, None))
            kms.append(c.lastTiming()["main_ms"])
        hip.hipDeviceSynchronize()
        wall = (time.perf_counter() - t0) / args.steps
        c.setTiming(False)
        info = c.lastTiming()
        hip.hipFree(pf)
        hip.hipFree(pr)
        # host delivery: the same rows copied block by block into pageable host memory
        th = time.perf_counter()
        c.jacobianRows(0, m)
        host = time.perf_counter() - th
        kms = float(np.mean(kms))
        out[key] = {"atoms": n, "reflections": m, "explicit_ops": w.n_explicit_ops, "kernel_ms": kms,
                    "wall_ms": wall * 1e3, "bytes_written": rb + fb, "pairs": n * m,
                    "hbm_GBps": (rb + fb) / (kms * 1e-3) / 1e9, "host_rows_s": host,
                    "host_GBps": (rb + fb) / host / 1e9, "chunk": info["chunk"], "grid": info["grid"][0]}
        print("per-hkl %s: %d atoms x %d reflections, k_jac %.3f ms (%.0f GB/s of rows), wall %.3f ms, host "
              "copy-out %.1f ms" % (key, n, m, kms, out[key]["hbm_GBps"], wall * 1e3, host * 1e3),
              file=sys.stderr, flush=True)
    big, small = out["p212121_5000"], out["p21c_60"]
    line = {
        "metric": "per-reflection derivative rows: (atom, reflection) pairs/s, F + dF/d(xyz, U, occ)",
        "value": big["pairs"] / (big["kernel_ms"] * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": big["kernel_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic, seeded (numpy PCG64 seeds 1000 and 0)",
        "config": {"workload": "P212121 5000 atoms (the headline structure), rows of 2400 of its reflections per "
                               "launch (10 complex FP64 derivatives per atom and reflection, 1.92 GB)",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "roof": "HBM writes of the rows (algorithmic bytes = 160 B per (atom, "
                                             "reflection) + 16 B of F per reflection)",
                     "achieved": big["hbm_GBps"], "peak": 8000.0, "unit": "GB/s",
                     "frac": big["hbm_GBps"] / 8000.0,
                     "traffic": measured_traffic("k_jac<acentric>", TRAFFIC_FILE_ROWS), "kernel": "k_jac<acentric>",
                     "kernel_ms": big["kernel_ms"], "bytes_per_launch": big["bytes_written"]},
        "small_molecule": small,
        "large_block": big,
        "mixed": {"note": "HIP_MIXED: FP32 terms with FP64 phase reduction, FP64 rows",
                  "small_molecule": out["p21c_60_mixed"], "large_block": out["p212121_5000_mixed"],
                  "hbm_frac_large_block": out["p212121_5000_mixed"]["hbm_GBps"] / 8000.0,
                  "traffic_large_block": measured_traffic("k_jac<acentric,mixed>", TRAFFIC_FILE_ROWS)},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import hc_oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        w = synth.config0()
        t0 = time.perf_counter()
        hc_oracle.per_hkl(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl, n_threads=threads)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": w.crystal.atoms.n * len(w.hkl) / dt, "unit": "pairs/s", "cores": threads,
                                "kind": "port",
                                "sample": "FP64 OpenMP per-hkl oracle on the 60-atom P21/c structure, all %d "
                                          "reflections, %.2f s wall" % (len(w.hkl), dt)}
    print(json.dumps(line), flush=True)


def bench_variant(args):
    """The 5000-atom P2_1 2_1 2_1 structure in the independent-atom model (CPU_IAM / HIP_IAM_MIXED)
    or with electron scattering (Mott-Bethe), mixed and FP64, or its form factors for every
    reflection (calculateFormFactors)."""
    from discamb_amd import HansenCoppensStructureFactorCalculator, synth

    w = synth.config1()
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params, device=0)
    terms = w.terms
    wl = ("P212121 5000 atoms (C1575 N425 O475 S50 H2475), cell 50x60x70 A, 66592 unique hkl to 1.2 A, "
          "4 explicit ops")
    if args.workload == "form_factors":
        f = calc.calculateFormFactors(w.lcs, w.hkl[:200])          # splines and set-up
        calc.setTiming(True)
        t0 = time.perf_counter()
        f = calc.calculateFormFactors(w.lcs, w.hkl)
        wall = time.perf_counter() - t0
        kms = calc.lastTiming()["main_ms"]
        n, m = w.crystal.atoms.n, len(w.hkl)
        bytes_out = 16.0 * n * m
        line = {"metric": "atomic form factors per second (form-factor route, FP64)",
                "value": n * m / (kms * 1e-3), "unit": "form factors/s", "n_gpus": 1, "steps": 1, "warmup": 1,
                "ms_per_step": kms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f64", "data": "synthetic, seeded (numpy PCG64 seed 1000)",
                "config": {"workload": wl + "; f of every atom for every reflection (%.2f GB)" % (bytes_out / 1e9),
                           "parallelism": "single GPU"},
                "roofline": {"bound": "hbm", "roof": "HBM writes of the form factors (16 B per atom and reflection)",
                             "achieved": bytes_out / (kms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                             "frac": bytes_out / (kms * 1e-3) / 8e12, "traffic": None, "kernel": "k_ff",
                             "kernel_ms": kms},
                "host_delivery": {"wall_s": wall, "GBps": bytes_out / wall / 1e9,
                                  "note": "256 MB device blocks copied to pageable host memory"}}
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import hc_oracle
            c2f = np.linalg.inv(hc_oracle.cell_f2c(w.crystal.cell))
            sel = w.hkl[:2000].astype(np.float64) @ c2f
            t0 = time.perf_counter()
            hc_oracle.form_factors(w.crystal, w.params, w.lcs, sel)
            dt = time.perf_counter() - t0
            line["cpu_baseline"] = {"value": n * len(sel) / dt, "unit": "form factors/s", "cores": 1, "kind": "port",
                                    "sample": "oracle (numpy splines + C deformation term) on 2000 reflections x "
                                              "5000 atoms, %.1f s" % dt}
        print(json.dumps(line), flush=True)
        return
    calc.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    if args.workload == "iam":
        modes = ("HIP_IAM_MIXED", "CPU_IAM")
        label = "independent-atom model (Waasmaier-Kirfel form factors)"
    else:
        calc.setElectronScattering(True)
        calc.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
        modes = ("HIP_MIXED", "HIP")
        label = "electron scattering (Mott-Bethe on the multipole model)"
    res = {}
    for mode in modes:
        calc.setCalculationMode(mode)
        calc.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
        res[mode] = time_mode(calc, mode, args.steps, args.warmup, True)
        print("%s %s: %.3f ms/step, kernel %.3f ms" % (args.workload, mode, res[mode][0] * 1e3, res[mode][1]),
              file=sys.stderr, flush=True)
    mx, fp = res[modes[0]], res[modes[1]]
    line = {"metric": "atom*symop*hkl terms/s for F+derivs (mixed precision), " + label,
            "value": terms / mx[0], "unit": "terms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": mx[0] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 terms, f64 phase reduction and accumulation",
            "data": "synthetic, seeded (numpy PCG64 seed 1000)",
            "config": {"workload": wl, "terms_per_step": terms, "parallelism": "single GPU"},
            "kernel_ms": mx[1], "launch": mx[2],
            "fp64": {"value": terms / fp[0], "ms_per_step": fp[0] * 1e3, "kernel_ms": fp[1]},
            "accuracy_mixed_vs_fp64": {"max_dF_over_maxF": float(np.abs(mx[3] - fp[3]).max() / np.abs(fp[3]).max())}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import hc_oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        k = 300 if args.workload == "iam" else 4000
        t0 = time.perf_counter()
        if args.workload == "iam":
            hc_oracle.iam_sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.hkl[:k], w.dtdf[:k], n_threads=threads)
        else:
            hc_oracle.electron_sf_and_derivs(w.crystal, w.params, w.crystal.atoms, w.lcs, w.hkl[:k], w.dtdf[:k],
                                             n_threads=threads)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": w.crystal.atoms.n * w.n_explicit_ops * k / dt, "unit": "terms/s",
                                "cores": threads, "kind": "port",
                                "sample": "FP64 OpenMP oracle, same structure, first %d reflections, %.1f s" % (k, dt)}
    print(json.dumps(line), flush=True)


def neighbour_frames(w):
    """One local coordinate system per atom of the structure, defined from its two nearest
    neighbours in the asymmetric unit (x along the nearest, y towards the second nearest), the
    kind of definition a TAAM assignment produces."""
    from scipy.spatial import cKDTree
    from discamb_amd import LcsDefinition
    from discamb_amd.synth import cell_matrices

    cart = w.crystal.atoms.xyz @ cell_matrices(w.crystal.cell)[0].T
    _, nb = cKDTree(cart).query(cart, k=3)
    return [LcsDefinition(i, "average_direction", [(int(nb[i, 1]), "x,y,z")], "average_direction",
                          [(int(nb[i, 2]), "x,y,z")], 0, 1) for i in range(len(cart))]


def bench_taam(args):
    """Whole calculateStructureFactorsAndDerivatives calls on the 5000-atom structure (atoms, 66 592
    reflections and dT/dF uploaded, results downloaded, every call) with (a) frames given, (b) frames
    recomputed from the positions on the host at every call (AnyHcCalculator without frozen frames),
    (c) (b) plus the chain rule dT/dL : dL/dxyz (multipole adjoint kernel, per-frame forward-mode
    derivatives on host threads).  Positions move by 1e-4 between calls."""
    from discamb_amd import HansenCoppensStructureFactorCalculator, synth

    w = synth.config1()
    defs = neighbour_frames(w)
    terms = w.terms
    res = {}
    for mode in ("HIP_MIXED", "HIP"):
        calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params, device=0)
        calc.setCalculationMode(mode)
        calc.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
        calc.autotune()
        for variant in ("given", "recomputed", "chain"):
            if variant == "recomputed":
                calc.setLocalCoordinateSystems(defs)
            elif variant == "chain":
                calc.setLocalCoordinateSystems(defs, chainRule=True)
            a = w.crystal.atoms.copy()
            rng = np.random.default_rng(5)
            for _ in range(args.warmup):
                calc.calculateStructureFactorsAndDerivatives(a, w.lcs, w.hkl, w.dtdf)
            calc.setTiming(True)
            ms, kms = [], []
            for _ in range(args.steps):
                a.xyz = a.xyz + rng.normal(0, 1e-4, a.xyz.shape)
                t0 = time.perf_counter()
                calc.calculateStructureFactorsAndDerivatives(a, w.lcs, w.hkl, w.dtdf)
                ms.append((time.perf_counter() - t0) * 1e3)
                kms.append(calc.lastTiming()["main_ms"])
            calc.setTiming(False)
            res[(mode, variant)] = (float(np.median(ms)), float(np.mean(kms)))
            print("taam %s %s: %.3f ms/call (median), main kernel %.3f ms" % (mode, variant, res[(mode, variant)][0],
                                                                          res[(mode, variant)][1]),
                  file=sys.stderr, flush=True)
        calc.setLocalCoordinateSystems(None)
    mx = res[("HIP_MIXED", "chain")]
    wl = ("P212121 5000 atoms (C1575 N425 O475 S50 H2475), cell 50x60x70 A, 66592 unique hkl to 1.2 A, "
          "4 explicit ops; one two-neighbour local frame per atom")
    line = {"metric": "atom*symop*hkl terms/s per whole call (F + derivs, frames recomputed, chain rule; mixed)",
            "value": terms / (mx[0] * 1e-3), "unit": "terms/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": mx[0], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 terms, f64 phase reduction and accumulation",
            "data": "synthetic, seeded (numpy PCG64 seed 1000)",
            "config": {"workload": wl, "terms_per_step": terms, "parallelism": "single GPU",
                       "timed": "host call incl. uploads of atoms/hkl/dT/dF and downloads"},
            "calls_ms": {"%s %s" % k: {"call_ms": v[0], "main_kernel_ms": v[1]} for k, v in res.items()}}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, budget_refl=4000)
    print(json.dumps(line), flush=True)


def bench_autograd(args):
    """Gradient steps of T = sum (|F| - |Fobs|)^2 through the torch.autograd binding on the
    5000-atom structure: Fobs from a perturbed copy (xyz + N(0, 0.02 A), U x U(0.9, 1.1)); a step
    is the forward F (F-only kernel), the loss and the backward pass (fused adjoint with the
    multipole-parameter derivatives), gradients w.r.t. xyz, U, occupancy, P_lm, Pval, kappa and
    kappa'.  Three steps per precision: every parameter as a GPU tensor (device-resident forward and
    backward, the multipole adjoint and its per-type sums on the device: the line's value), the same
    with host tensors (host calls), and the structural parameters alone as GPU tensors."""
    import torch
    from discamb_amd import synth
    from discamb_amd.torch_ops import StructureFactors

    w = synth.config1()
    terms = w.terms
    rng = np.random.default_rng(33)
    f2c, c2f = synth.cell_matrices(w.crystal.cell)
    ref = StructureFactors(w.crystal, w.params, w.lcs, w.hkl, mode="HIP")
    base = ref.initial(w.params)
    pert = [p.detach().clone() for p in base]
    pert[0] += torch.from_numpy(rng.normal(0, 0.02, (w.crystal.atoms.n, 3)) @ c2f.T)
    pert[1] *= torch.from_numpy(rng.uniform(0.9, 1.1, (w.crystal.atoms.n, 1)))
    with torch.no_grad():
        fobs = ref(*pert).abs()
    del ref
    res, gpu_all, dev = {}, {}, {}
    fobs_dev = fobs.to("cuda")

    def timed(run, sync):
        for _ in range(args.warmup):
            run()
        if sync:
            torch.cuda.synchronize()
        ms = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            run()
            if sync:
                torch.cuda.synchronize()
            ms.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ms))

    for mode in ("HIP_MIXED", "HIP"):
        sf = StructureFactors(w.crystal, w.params, w.lcs, w.hkl, mode=mode)

        def backward_kernel_ms(step):
            # main kernels of the backward call (the step's last engine call), from one extra step
            # with HIP-event timing (host calls only: device-resident calls are not event-timed)
            sf.calc.setTiming(True)
            step()
            torch.cuda.synchronize()
            kms = sf.calc.lastTiming()["main_ms"]
            sf.calc.setTiming(False)
            return kms

        def stepper(params, fo):
            def step():
                for p in params:
                    p.grad = None
                T = ((sf(*params).abs() - fo) ** 2).sum()
                T.backward()
                return T
            return step

        # every parameter as a GPU tensor: device-resident forward and backward, multipole
        # gradients (per-type sums) formed on the device
        gparams = sf.initial(w.params, device="cuda")
        gpu_all[mode] = timed(stepper(gparams, fobs_dev), True)
        print("autograd %s, GPU tensors, all parameters: %.3f ms/step (median)" % (mode, gpu_all[mode]),
              file=sys.stderr, flush=True)
        # the same step with the parameters on the host: host calls for forward and backward
        params = sf.initial(w.params)
        res[mode] = (timed(stepper(params, fobs), False), backward_kernel_ms(stepper(params, fobs)))
        print("autograd %s, host tensors, all parameters: %.3f ms/step (median), backward main kernel %.3f ms" % (
            mode, *res[mode]), file=sys.stderr, flush=True)
        # structural parameters as GPU tensors: the plain adjoint, no multipole passes
        dparams = sf.initial(w.params, device="cuda")
        for p in dparams[3:]:
            p.requires_grad_(False)
        dev[mode] = timed(stepper(dparams, fobs_dev), True)
        print("autograd %s, GPU tensors, xyz/U/occ: %.3f ms/step (median)" % (mode, dev[mode]), file=sys.stderr,
              flush=True)
    mx, fp = gpu_all["HIP_MIXED"], gpu_all["HIP"]
    kmx, kfp = res["HIP_MIXED"][1], res["HIP"][1]
    wl = ("P212121 5000 atoms (C1575 N425 O475 S50 H2475), 40 multipole types, cell 50x60x70 A, 66592 unique "
          "hkl to 1.2 A, 4 explicit ops; least-squares target against a perturbed copy")
    line = {"metric": "atom*symop*hkl terms/s per torch gradient step (forward F + backward incl. multipole "
                      "parameters; mixed)",
            "value": terms / (mx * 1e-3), "unit": "terms/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": mx, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 terms, f64 phase reduction and accumulation",
            "data": "synthetic, seeded (numpy PCG64 seed 1000; perturbation seed 33)",
            "config": {"workload": wl, "terms_per_step": terms, "parallelism": "single GPU",
                       "timed": "parameters as GPU tensors (xyz, U, occupancy, P_lm, Pval, kappa, kappa'): "
                                "device-resident forward (F only) and backward (both multipole-adjoint passes, "
                                "per-type sums on the device) on torch's stream; synchronised at the end of "
                                "each step"},
            "backward_main_kernel_ms": kmx,
            "backward_main_kernel_note": "both multipole-adjoint passes of the backward call, HIP events, measured "
                                         "on the host-tensor step (the same kernels)",
            "fp64": {"value": terms / (fp * 1e-3), "ms_per_step": fp, "backward_main_kernel_ms": kfp},
            "host_tensors": {
                "timed": "the same step with the parameters as host tensors: host calls for forward and backward, "
                         "gradients copied back",
                "mixed_ms_per_step": res["HIP_MIXED"][0], "fp64_ms_per_step": res["HIP"][0]},
            "gpu_tensors_structural": {
                "timed": "parameters as GPU tensors, gradients of xyz / U / occupancy only: device-resident "
                         "forward (F only) and backward (F + derivatives) on torch's stream, no host copies; "
                         "synchronised at the end of each step",
                "mixed_ms_per_step": dev["HIP_MIXED"], "fp64_ms_per_step": dev["HIP"],
                "mixed_terms_per_s": terms / (dev["HIP_MIXED"] * 1e-3),
                "fp64_terms_per_s": terms / (dev["HIP"] * 1e-3)}}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, budget_refl=4000)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=0, help="fixed reflection chunk (disables the autotuner)")
    ap.add_argument("--chunk-fp64", type=int, default=0, help="fixed reflection chunk for the FP64 path only")
    ap.add_argument("--workload", default="p212121_5000",
                    choices=["p212121_5000", "p212121_20000", "p43212_5000", "batch256", "per_hkl", "iam",
                             "electron", "form_factors", "taam", "autograd"],
                    help="p212121_5000: one 5000-atom structure at 1.2 A (headline); p212121_20000: 20 000 atoms "
                         "at 1.5 A; p43212_5000: 5000 atoms in P4_3 2_1 2 at 1.4 A (coset images over the 222 "
                         "subgroup); batch256: 256 uneven structures per batched step; per_hkl: per-reflection "
                         "derivative rows; iam / electron: the headline structure in those modes; form_factors: "
                         "its atomic form factors; taam: whole calls with local frames recomputed from the "
                         "positions and the chain rule through them; autograd: torch least-squares gradient "
                         "steps (forward F, backward with the multipole adjoint)")
    ap.add_argument("--no-singles", dest="singles", action="store_false",
                    help="batch256: skip the one-structure-at-a-time timings")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain read+write elementwise kernel over MIB MiB (PMC byte calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the path would shard by reflection); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch256":
        return bench_batch(args)
    if args.workload == "per_hkl":
        return bench_per_hkl(args)
    if args.workload in ("iam", "electron", "form_factors"):
        return bench_variant(args)
    if args.workload == "taam":
        return bench_taam(args)
    if args.workload == "autograd":
        return bench_autograd(args)
    from discamb_amd import HansenCoppensStructureFactorCalculator, synth

    if args.workload == "p212121_20000":
        w = synth.config2_large()
        wl_text = ("P212121 20000 atoms (protein composition), 40 multipole types, cell 80x90x110 A, %d unique hkl "
                   "to 1.5 A, 4 explicit ops" % len(w.hkl))
        seed_text = "synthetic, seeded (numpy PCG64 seed 2000)"
        cpu_label, cpu_refl = "the same 20000-atom P212121 structure", 16648
    elif args.workload == "p43212_5000":
        rng = np.random.Generator(np.random.PCG64(1000))
        comp = synth.protein_composition(5000, rng)
        w = synth.build_workload("p43212_5000at_1.4A", rng, (60.0, 60.0, 70.0, 90.0, 90.0, 90.0), "P43212", comp,
                                 1.4, synth.PROTEIN_TYPES)
        wl_text = ("P43212 5000 atoms (protein composition), cell 60x60x70 A, %d unique hkl to 1.4 A, 8 explicit ops "
                   "(222 subgroup x 2 coset representatives)" % len(w.hkl))
        seed_text = "synthetic, seeded (numpy PCG64 seed 1000)"
        cpu_label, cpu_refl = "the same 5000-atom P43212 structure", 4000
    else:
        w = synth.config1()
        wl_text = ("P212121 5000 atoms (C1575 N425 O475 S50 H2475), 40 multipole types, cell 50x60x70 A, 66592 "
                   "unique hkl to 1.2 A, 4 explicit ops")
        seed_text = "synthetic, seeded (numpy PCG64 seed 1000)"
        cpu_label, cpu_refl = "the same 5000-atom P212121 structure", 66592
    calc = HansenCoppensStructureFactorCalculator(w.crystal, w.params, device=0)
    calc.upload(w.crystal.atoms, w.lcs, w.hkl, w.dtdf)
    terms = w.terms
    flops = algorithmic_flops(w)
    print("workload %s: %d atoms x %d ops x %d hkl = %.3e terms" % (
        w.name, w.crystal.atoms.n, w.n_explicit_ops, w.hkl.shape[0], terms), file=sys.stderr, flush=True)
    chunk64 = args.chunk_fp64 or args.chunk
    calc.setTuning(chunk64)            # 0 = autotuned
    s64, k64, info64, f64, d64 = time_mode(calc, "HIP", args.steps, args.warmup, chunk64 == 0)
    calc.setTuning(args.chunk)
    print("fp64: %.3f ms/step, main kernel %.3f ms, grid %s" % (s64 * 1e3, k64, info64["grid"]), file=sys.stderr,
          flush=True)
    tune = args.chunk == 0
    smx, kmx, infomx, fmx, dmx = time_mode(calc, "HIP_MIXED", args.steps, args.warmup, tune)
    print("mixed: %.3f ms/step, main kernel %.3f ms, grid %s" % (smx * 1e3, kmx, infomx["grid"]),
          file=sys.stderr, flush=True)
    host_mx = time_host_calls(calc, w)
    host_mx_rt = time_host_calls(calc, w, retype=True)
    calc.setCalculationMode("HIP")
    calc.setTuning(chunk64)
    host_64 = time_host_calls(calc, w)
    host_64_rt = time_host_calls(calc, w, retype=True)
    calc.setTuning(args.chunk)
    calc.setCalculationMode("HIP_MIXED")
    print("whole calls from host buffers: mixed %.3f ms (new type parameters %.3f ms), FP64 %.3f ms (%.3f ms)" % (
        host_mx * 1e3, host_mx_rt * 1e3, host_64 * 1e3, host_64_rt * 1e3), file=sys.stderr, flush=True)

    def rel(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())

    fo = np.abs(f64)
    r_fp64 = float(np.sum(np.abs(np.abs(fmx) - fo)) / np.sum(fo))
    acc = {"max_dF_over_maxF": rel(fmx, f64), "dR_abs_mixed_vs_fp64": r_fp64,
           "d_xyz": rel(dmx.atomic_position_derivatives, d64.atomic_position_derivatives),
           "d_adp": rel(dmx.adp_derivatives, d64.adp_derivatives),
           "d_occ": rel(dmx.occupancy_derivatives, d64.occupancy_derivatives)}
    achieved = flops / (kmx * 1e-3)
    kname = "k_main<mixed,deriv,line,diag222>"    # P2_1 2_1 2_1: diagonal sign matrices, hkl-line recurrences
    line = {
        "metric": "atom*symop*hkl terms/s for F+derivs (mixed precision)",
        "value": terms / smx,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": smx * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 terms, f64 phase reduction and accumulation",
        "data": seed_text,
        "config": {"workload": wl_text, "terms_per_step": terms, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "roof": "FP32 vector ALU (the path uses VALU, not matrix cores)",
                     "achieved": achieved / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK, "traffic": measured_traffic(kname) if args.workload == "p212121_5000" else None,
                     "frac_of_unpacked_fp32": achieved / (FP32_PEAK / 2),
                     "kernel": kname, "kernel_ms": kmx,
                     "flops_per_launch": flops,
                     "flops_per_term": "66 + 2(L+1)^2 + 2(L+1): 126 (L=4), 90 (L=2)"},
        "fp64": {"value": terms / s64, "unit": "terms/s", "ms_per_step": s64 * 1e3, "kernel_ms": k64,
                 "achieved_tflops": flops / (k64 * 1e-3) / 1e12, "frac_of_fp64_peak": flops / (k64 * 1e-3) / FP64_PEAK,
                 "frac_of_fp32_peak": flops / (k64 * 1e-3) / FP32_PEAK,
                 "kernel": "k_main<fp64,deriv,line,diag222>",
                 "traffic": measured_traffic("k_main<fp64,deriv,line,diag222>")
                 if args.workload == "p212121_5000" else None},
        "host_call": {"ms": host_mx * 1e3, "terms_per_s": terms / host_mx,
                      "fp64_ms": host_64 * 1e3, "fp64_terms_per_s": terms / host_64,
                      "after_new_type_parameters_ms": host_mx_rt * 1e3,
                      "fp64_after_new_type_parameters_ms": host_64_rt * 1e3,
                      "timed": "whole calls from host buffers: upload of atoms and dT/dF, run, download of F and "
                               "derivatives; the reflection list is unchanged between calls, so its order, device "
                               "copy, reflection records and radial records/tables are reused (the *_after_new_"
                               "type_parameters figures change kappa, kappa' before every call, which rebuilds the "
                               "radial records and tables)"},
        "timed_step": "device-resident step: run() on uploaded inputs; the reflection-dependent prep (records, "
                      "radial records and tables) is built once and reused, the per-atom prep runs every step",
        "speedup_mixed_vs_fp64": s64 / smx,
        "accuracy_mixed_vs_fp64": acc,
        "launch": {"grid": infomx["grid"], "chunk": infomx["chunk"], "fp64_chunk": info64["chunk"],
                   "autotuned": tune, "fp64_autotuned": chunk64 == 0},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, cpu_refl, cpu_label)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
