"""Python mirror of DiSCaMB's HansenCoppensStructureFactorCalculator on the HIP engine.

Reference interface: include/discamb/Scattering/HansenCoppensStructureFactorCalculator.h:36-375
(constructor ``(Crystal, HC_ModelParameters)``, ``setCalculationMode``,
``calculateStructureFactorsAndDerivatives(atoms, lcs, hkl, f, dTarget_dparam, dTarget_df,
countAtomContribution[, selector])``, ``calculateStructureFactors``).  Errors raise ``DhError``
(the reference throws discamb::Exception).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .model import (AtomSet, Crystal, DerivativesSelector, HC_ModelParameters, SfDerivativesAtHkl,
                    TargetFunctionAtomicParamDerivatives, TypeParameterDerivatives, parse_symop)

XYZ = {"fractional": 0, "cartesian": 1}
ADP = {"U_cif": 0, "U_cart": 1}


class _Keep:
    """Holds numpy arrays alive while ctypes structs point into them."""

    def __init__(self):
        self.items = []

    def arr(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.items.append(a)
        return a


def _model_desc(crystal: Crystal, params: HC_ModelParameters, d_xyz="fractional", d_adp="U_cif"):
    keep = _Keep()
    d = _lib.DhModelDesc()
    d.cell = (ctypes.c_double * 6)(*[float(v) for v in crystal.cell])
    rots, trans = [], []
    for s in crystal.symops:
        r, t = parse_symop(s)
        rots.append(r)
        trans.append(t)
    rot = keep.arr(np.array(rots).reshape(-1), np.int32)
    tr = keep.arr(np.array(trans).reshape(-1), np.float64)
    d.n_symops = len(crystal.symops)
    d.symop_rot = _lib.ptr(rot, ctypes.c_int)
    d.symop_trans = _lib.ptr(tr, ctypes.c_double)
    d.xyz_convention = XYZ[crystal.xyz_convention]
    d.adp_convention = ADP[crystal.adp_convention]
    d.d_xyz_convention = XYZ[d_xyz]
    d.d_adp_convention = ADP[d_adp]
    wo, to, nl, cc, ce, cp, occ, kind, dve, dvo, dvp, an = [0], [0], [], [], [], [], [], [], [], [0], [], []
    for w in params.wfn_parameters:
        for i, o in enumerate(w.orbitals):
            nl += [o.n, o.l]
            cc += list(o.coefficients)
            ce += list(o.exponents)
            cp += list(o.powers)
            to.append(len(cc))
            occ.append(w.orbital_occupancy[i])
            kind.append(0 if i in w.core_orbitals_indices else (1 if i in w.valence_orbitals_indices else -1))
        wo.append(len(occ))
        dve.append(w.deformation_valence_exponent)
        dvp += list(w.deformation_valence_power)
        dvo.append(len(dvp))
        an += [complex(w.anomalous_scattering).real, complex(w.anomalous_scattering).imag]
    d.n_wfn = len(params.wfn_parameters)
    for name, vals, dt, ct in [("wfn_orbital_offset", wo, np.int32, ctypes.c_int),
                               ("orbital_term_offset", to, np.int32, ctypes.c_int),
                               ("orbital_nl", nl + [0], np.int32, ctypes.c_int),
                               ("term_coeff", cc + [0.0], np.float64, ctypes.c_double),
                               ("term_exp", ce + [0.0], np.float64, ctypes.c_double),
                               ("term_pow", cp + [0], np.int32, ctypes.c_int),
                               ("orbital_occupancy", occ + [0], np.int32, ctypes.c_int),
                               ("orbital_kind", kind + [0], np.int32, ctypes.c_int),
                               ("def_val_exponent", dve + [0.0], np.float64, ctypes.c_double),
                               ("def_val_pow_offset", dvo, np.int32, ctypes.c_int),
                               ("def_val_pow", dvp + [0], np.int32, ctypes.c_int),
                               ("anomalous", an + [0.0, 0.0], np.float64, ctypes.c_double)]:
        setattr(d, name, _lib.ptr(keep.arr(vals, dt), ct))
    tp = params.type_parameters
    d.n_types = len(tp)
    d.plm_n_l = _lib.ptr(keep.arr([len(t.p_lm) for t in tp] + [0], np.int32), ctypes.c_int)
    d.plm = _lib.ptr(keep.arr(np.concatenate([t.plm25() for t in tp] + [np.zeros(25)]), np.float64), ctypes.c_double)
    d.p_val = _lib.ptr(keep.arr([t.p_val for t in tp] + [0.0], np.float64), ctypes.c_double)
    d.kappa = _lib.ptr(keep.arr([t.kappa_spherical_valence for t in tp] + [1.0], np.float64), ctypes.c_double)
    d.kappa_prime = _lib.ptr(keep.arr([t.kappa_deformation_valence for t in tp] + [1.0], np.float64),
                             ctypes.c_double)
    d.n_atoms = len(params.atom_to_type_map)
    d.atom_to_wfn = _lib.ptr(keep.arr(list(params.atom_to_wfn_map) + [0], np.int32), ctypes.c_int)
    d.atom_to_type = _lib.ptr(keep.arr(list(params.atom_to_type_map) + [0], np.int32), ctypes.c_int)
    d.atom_multiplicity = _lib.ptr(keep.arr(np.append(crystal.atoms.multiplicity, 0.0), np.float64),
                                   ctypes.c_double)
    labels = (ctypes.c_char_p * (len(params.wfn_parameters) + 1))(
        *[w.label.encode() for w in params.wfn_parameters], None)
    keep.items.append(labels)
    d.wfn_label = ctypes.cast(labels, ctypes.POINTER(ctypes.c_char_p))
    d.wfn_atomic_number = _lib.ptr(keep.arr([int(w.atomic_number) for w in params.wfn_parameters] + [0], np.int32),
                                   ctypes.c_int)
    return d, keep


_LCS_KIND = {"any_orthogonal": 0, "average_direction": 1, "average_position": 2}


def _lcs_defs(defs):
    """ctypes array of dh_lcs_def (and the arrays it points into) from LcsDefinition objects."""
    keep = []

    def ref(r):
        idx, op = (r, "x,y,z") if isinstance(r, (int, np.integer)) else r
        R, t = parse_symop(op)
        a = _lib.DhAtomRef()
        a.atom = int(idx)
        a.rot = (ctypes.c_int * 9)(*[int(v) for v in np.asarray(R).reshape(-1)])
        a.trans = (ctypes.c_double * 3)(*[float(v) for v in t])
        return a

    def refs(lst):
        arr = (_lib.DhAtomRef * max(len(lst), 1))(*[ref(r) for r in lst])
        keep.append(arr)
        return ctypes.cast(arr, ctypes.POINTER(_lib.DhAtomRef))

    out = (_lib.DhLcsDef * max(len(defs), 1))()
    for i, d in enumerate(defs):
        o = out[i]
        o.central = ref((d.central, d.central_op))
        o.direction1_type = _LCS_KIND[d.direction1_type]
        o.direction2_type = _LCS_KIND[d.direction2_type]
        o.n_ref1, o.n_ref2, o.n_chirality = len(d.ref1), len(d.ref2), len(d.chirality)
        o.ref1, o.ref2, o.chirality = refs(d.ref1), refs(d.ref2), refs(d.chirality)
        o.coordinate_1, o.coordinate_2 = int(d.coordinate_1), int(d.coordinate_2)
        o.is_r = 1 if d.is_r else 0
    return out, keep


def compute_lcs(cell, xyz_frac, definitions):
    """Local frames (n, 3, 3), axes as columns, of the definitions for atoms at fractional xyz
    (LocalCoordinateSystemCalculator::calculate); host code, no GPU needed."""
    lib = _lib.load()
    arr, keep = _lcs_defs(definitions)
    x = np.ascontiguousarray(np.asarray(xyz_frac, dtype=np.float64).reshape(-1, 3))
    c6 = np.ascontiguousarray(np.asarray(cell, dtype=np.float64))
    out = np.zeros((max(len(definitions), 1), 9))
    _lib.check(lib.dh_lcs_compute(_lib.ptr(c6, ctypes.c_double), x.shape[0], _lib.ptr(x, ctypes.c_double),
                                  len(definitions), arr, _lib.ptr(out, ctypes.c_double)))
    return out[:len(definitions)].reshape(-1, 3, 3)


class HansenCoppensStructureFactorCalculator:
    """F(h) and dT/dp for the Hansen-Coppens multipole model on one GPU."""

    def __init__(self, crystal: Crystal, parameters: HC_ModelParameters, device: int = 0,
                 d_xyz_convention: str = None, d_adp_convention: str = None):
        """Derivative conventions default to the crystal's input conventions, as the reference
        constructor sets them (HansenCoppensStructureFactorCalculator.cpp:57-73)."""
        self._lib = _lib.load()
        d_xyz_convention = d_xyz_convention or crystal.xyz_convention
        d_adp_convention = d_adp_convention or crystal.adp_convention
        desc, keep = _model_desc(crystal, parameters, d_xyz_convention, d_adp_convention)
        h = ctypes.c_void_p()
        _lib.check(self._lib.dh_hc_create(ctypes.byref(desc), int(device), ctypes.byref(h)))
        self._h = h
        self.n_atoms = len(parameters.atom_to_type_map)
        self.n_types = len(parameters.type_parameters)
        self._keep = None
        self._nh = self._n = 0          # sizes of the last upload (fetch reads them)
        self._dev_nh = 0                # reflections of the device-resident calls

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.dh_hc_destroy(h)
            self._h = None

    # -- reference API -----------------------------------------------------------------------
    def setCalculationMode(self, mode: str):
        _lib.check(self._lib.dh_hc_set_calculation_mode(self._h, mode.encode()))

    def setLocalCoordinateSystems(self, definitions, chainRule: bool = False):
        """One LcsDefinition per atom (or None to detach): the frames are then recomputed from the
        positions at every call and the localCoordinateSystems argument is ignored (TAAM,
        AnyHcCalculator with unfrozen frames).  chainRule adds dT/dL : dL/dxyz to the xyz
        derivatives."""
        defs = definitions or []
        arr, keep = _lcs_defs(defs)
        _lib.check(self._lib.dh_hc_set_lcs_definitions(self._h, len(defs), arr, 1 if chainRule else 0))

    def lastLocalCoordinateSystems(self):
        out = np.zeros((max(self.n_atoms, 1), 9))
        _lib.check(self._lib.dh_hc_last_lcs(self._h, _lib.ptr(out, ctypes.c_double)))
        return out[:self.n_atoms].reshape(-1, 3, 3)

    def setElectronScattering(self, electronScattering: bool):
        """Mott-Bethe electron structure factors for the summed path (needs HC_WfnType.atomic_number)."""
        _lib.check(self._lib.dh_hc_set_electron_scattering(self._h, 1 if electronScattering else 0))

    def setIamTable(self, table: str):
        """N-Gaussian table of the CPU_IAM mode: "Waasmeier-Kirfel" (default), "IT92" (the table of
        AnyIamCalculator), "electron-IT" or "electron-cctbx"."""
        _lib.check(self._lib.dh_hc_set_iam_table(self._h, table.encode()))

    def setTypeParameters(self, plm, p_val, kappa, kappa_prime):
        """Replace the atom-type parameters: plm (n_types, 25) in the l*l + l + m layout, p_val,
        kappa, kappa_prime (n_types,)."""
        nt = self.n_types
        plm = np.ascontiguousarray(np.asarray(plm, dtype=np.float64).reshape(nt, 25))
        pv, ka, kp = (np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(nt)) for v in (p_val, kappa, kappa_prime))
        _lib.check(self._lib.dh_hc_set_type_parameters(self._h, nt, _lib.ptr(plm, ctypes.c_double),
                                                       _lib.ptr(pv, ctypes.c_double), _lib.ptr(ka, ctypes.c_double),
                                                       _lib.ptr(kp, ctypes.c_double)))

    def typeDerivatives(self):
        """Per-type multipole derivatives of the last call made with d_multipole."""
        nt = self.n_types
        plm = np.zeros((max(nt, 1), 25))
        pv, ka, kp = np.zeros(max(nt, 1)), np.zeros(max(nt, 1)), np.zeros(max(nt, 1))
        _lib.check(self._lib.dh_hc_fetch_type_derivatives(self._h, _lib.ptr(plm, ctypes.c_double),
                                                          _lib.ptr(pv, ctypes.c_double), _lib.ptr(ka, ctypes.c_double),
                                                          _lib.ptr(kp, ctypes.c_double)))
        return TypeParameterDerivatives(plm[:nt], pv[:nt], ka[:nt], kp[:nt])

    def setTuning(self, hkl_chunk: int):
        _lib.check(self._lib.dh_hc_set_tuning(self._h, int(hkl_chunk)))

    def setKernelOptions(self, general_rotations: bool = False, radial: str = "auto", direct_phases: bool = False):
        """general_rotations: run the general-rotation kernels even for groups whose operations
        are all diagonal sign matrices (those otherwise use the parity-folded kernels).
        radial: "records" (per-(reflection, type) tables), "tables" (per-(atom, reflection)
        lookups in per-wavefunction tables) or "auto".
        direct_phases: every phase and Debye-Waller factor is evaluated directly instead of carried
        along hkl lines (reflections are then ordered by |h|); takes effect at the next upload."""
        flags = (1 if general_rotations else 0) | {"auto": 0, "records": 2, "tables": 4}[radial] | \
            (8 if direct_phases else 0)
        _lib.check(self._lib.dh_hc_set_options(self._h, flags))

    def engineInfo(self):
        n, o, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ce = ctypes.create_string_buffer(2)
        _lib.check(self._lib.dh_hc_engine_info(self._h, ctypes.byref(n), ctypes.byref(o), ctypes.byref(c), ce))
        return {"n_explicit_ops": n.value, "order": o.value, "centrosymmetric": bool(c.value),
                "centering": ce.value.decode()}

    def _atoms(self, atoms: AtomSet, lcs, include, anomalous):
        keep = _Keep()
        a = _lib.DhAtoms()
        n = atoms.n
        if n != self.n_atoms:
            raise _lib.DhError(1, "number of atoms differs from the model's atom-to-type map")
        a.n_atoms = n
        a.xyz = _lib.ptr(keep.arr(atoms.xyz, np.float64), ctypes.c_double)
        a.adp_size = _lib.ptr(keep.arr(atoms.adp_size, np.int32), ctypes.c_int)
        a.adp = _lib.ptr(keep.arr(atoms.adp, np.float64), ctypes.c_double)
        a.occupancy = _lib.ptr(keep.arr(atoms.occupancy, np.float64), ctypes.c_double)
        if lcs is None:
            lcs = np.broadcast_to(np.eye(3), (n, 3, 3))
        a.lcs = _lib.ptr(keep.arr(np.asarray(lcs).reshape(n, 9), np.float64), ctypes.c_double)
        if anomalous is not None:
            an = np.asarray(anomalous, dtype=np.complex128).reshape(n)
            a.anomalous = _lib.ptr(keep.arr(np.stack([an.real, an.imag], 1), np.float64), ctypes.c_double)
        if include is not None:
            a.include = _lib.ptr(keep.arr(np.asarray(include, dtype=bool).astype(np.uint8), np.uint8),
                                 ctypes.c_ubyte)
        return a, keep

    @staticmethod
    def _hkl(hkl, dtdf):
        hkl = np.ascontiguousarray(np.asarray(hkl).reshape(-1, 3), dtype=np.int32)
        d = np.ascontiguousarray(np.asarray(dtdf, dtype=np.complex128).reshape(-1))
        if d.shape[0] != hkl.shape[0]:
            raise _lib.DhError(1, "Inconsistent size of hkl vectors set and (d target/ d F) table when "
                                  "calculating multipolar structure factors.")
        return hkl, d.view(np.float64)          # (re, im) interleaved, no copy

    def calculateStructureFactorsAndDerivatives(self, atoms: AtomSet, localCoordinateSystems, hkl,
                                                dTarget_dF, countAtomContribution=None,
                                                derivativesSelector: DerivativesSelector = None,
                                                anomalous=None):
        """Returns (f (n_hkl,) complex, TargetFunctionAtomicParamDerivatives).

        With an object that has ``onPerHklCalculation(hkl_idx, f, SfDerivativesAtHkl)`` in place of
        dTarget_dF this is the reference's per-hkl overload (HansenCoppensStructureFactorCalculator
        .h:206-242): the processor is called once per reflection, in order, and nothing is
        returned."""
        if hasattr(dTarget_dF, "onPerHklCalculation"):
            return self._per_hkl(atoms, localCoordinateSystems, hkl, dTarget_dF, countAtomContribution, anomalous)
        sel = (derivativesSelector or DerivativesSelector()).bits()
        a, keep = self._atoms(atoms, localCoordinateSystems, countAtomContribution, anomalous)
        hkl_i, dd = self._hkl(hkl, dTarget_dF)
        nh, n = hkl_i.shape[0], atoms.n
        f = np.zeros(max(nh, 1), np.complex128)       # written through its (re, im) float64 view
        dx = np.zeros((max(n, 1), 3))
        du = np.zeros((max(n, 1), 6))
        do = np.zeros(max(n, 1))
        _lib.check(self._lib.dh_hc_sf_and_derivs(
            self._h, ctypes.byref(a), nh, _lib.ptr(hkl_i, ctypes.c_int), _lib.ptr(dd, ctypes.c_double), sel,
            _lib.ptr(f.view(np.float64), ctypes.c_double), _lib.ptr(dx, ctypes.c_double),
            _lib.ptr(du, ctypes.c_double), _lib.ptr(do, ctypes.c_double)))
        return f[:nh], TargetFunctionAtomicParamDerivatives(dx[:n], du[:n], do[:n])

    # -- device-resident calls (torch tensors on the calculator's GPU) ---------------------------
    def setReflections(self, hkl):
        """Reflection list (host array) of the device-resident calls; kept until set again (a host
        call on this calculator replaces it)."""
        hkl_i = np.ascontiguousarray(np.asarray(hkl).reshape(-1, 3), dtype=np.int32)
        _lib.check(self._lib.dh_hc_set_reflections(self._h, hkl_i.shape[0], _lib.ptr(hkl_i, ctypes.c_int)))
        self._dev_nh = hkl_i.shape[0]

    def _dev_args(self, xyz, adp_size, adp, occupancy, localCoordinateSystems, anomalous, countAtomContribution):
        import torch
        n = self.n_atoms
        keep = []

        def dev(t, dtype, shape, name):
            if t is None:
                return None
            if not isinstance(t, torch.Tensor) or not t.is_cuda:
                raise _lib.DhError(1, "%s: a torch tensor on the GPU is required" % name)
            if t.dtype != dtype or tuple(t.shape) != shape or not t.is_contiguous():
                raise _lib.DhError(1, "%s: contiguous %s of shape %s required" % (name, dtype, shape))
            keep.append(t)
            return t.data_ptr()

        a = _lib.DhAtomsDev()
        a.n_atoms = n
        a.xyz = dev(xyz, torch.float64, (n, 3), "xyz")
        a.adp_size = dev(adp_size, torch.int32, (n,), "adp_size")
        a.adp = dev(adp, torch.float64, (n, 6), "adp")
        a.occupancy = dev(occupancy, torch.float64, (n,), "occupancy")
        a.lcs = dev(localCoordinateSystems, torch.float64, (n, 3, 3), "localCoordinateSystems")
        if anomalous is not None:
            a.anomalous = dev(torch.view_as_real(anomalous) if anomalous.is_complex() else anomalous, torch.float64,
                              (n, 2), "anomalous")
        if countAtomContribution is not None:
            a.include = dev(countAtomContribution, torch.uint8, (n,), "countAtomContribution")
        return a, keep, dev

    def calculateOnDevice(self, xyz, adp_size, adp, occupancy, dTarget_dF=None, localCoordinateSystems=None,
                          anomalous=None, countAtomContribution=None, derivativesSelector: DerivativesSelector = None,
                          stream=None, no_alloc=False, out=None):
        """calculateStructureFactorsAndDerivatives on torch tensors resident on the calculator's GPU,
        ordered on `stream` (default: torch's current stream) without host synchronisation; the
        reflections are those of setReflections.  xyz (n,3) / adp (n,6) float64 in the model's
        conventions, adp_size (n,) int32, occupancy (n,) float64, dTarget_dF (n_hkl,) complex128
        (None for F only), optional local frames (n,3,3), anomalous (n,) complex128 and
        countAtomContribution (n,) uint8.  Returns (F, dxyz, dadp, docc) as new GPU tensors (the
        derivatives None for F only), or writes them into `out` = (F, dxyz, dadp, docc) tensors.
        no_alloc: the allocation-free entry (dh_hc_sf_and_derivs_dev_noalloc; with `out` given
        nothing is allocated on the Python side either).  A selector with d_multipole runs
        dh_hc_sf_and_derivs_mp_dev and appends the per-type multipole derivatives
        (TypeParameterDerivatives of GPU tensors: p_lm (n_types, 25), p_val, kappa, kappa_prime);
        it takes neither no_alloc nor `out`."""
        import torch
        sel = (derivativesSelector or DerivativesSelector()).bits() if dTarget_dF is not None else 0
        mp = bool(sel & 8)
        if mp and (no_alloc or out is not None):
            raise ValueError("multipole derivatives on the device: no allocation-free form and no `out` tensors")
        a, keep, dev = self._dev_args(xyz, adp_size, adp, occupancy, localCoordinateSystems, anomalous,
                                      countAtomContribution)
        nh, n = getattr(self, "_dev_nh", 0), self.n_atoms
        device = xyz.device
        if out is not None:
            f, outs = out[0], list(out[1:])
            dev(torch.view_as_real(f), torch.float64, (nh, 2), "F output")
            for o, shape in zip(outs, ((n, 3), (n, 6), (n,))):
                dev(o, torch.float64, shape, "derivative output")
        else:
            f = torch.empty(nh, dtype=torch.complex128, device=device)
            outs = [torch.empty(s, dtype=torch.float64, device=device) for s in ((n, 3), (n, 6), (n,))] if sel \
                else [None] * 3
        dd = None
        if dTarget_dF is not None:
            dd = dev(torch.view_as_real(dTarget_dF), torch.float64, (nh, 2), "dTarget_dF")
        st = stream if stream is not None else torch.cuda.current_stream(device)
        fptr = torch.view_as_real(f).data_ptr() if nh else None
        optrs = [o.data_ptr() if o is not None else None for o in outs]
        if mp:
            nt = max(self.n_types, 1)
            td = TypeParameterDerivatives(*(torch.zeros(s, dtype=torch.float64, device=device)
                                            for s in ((nt, 25), (nt,), (nt,), (nt,))))
            _lib.check(self._lib.dh_hc_sf_and_derivs_mp_dev(
                self._h, ctypes.byref(a), dd, sel, fptr, *optrs, td.p_lm.data_ptr(), td.p_val.data_ptr(),
                td.kappa.data_ptr(), td.kappa_prime.data_ptr(), ctypes.c_void_p(st.cuda_stream)))
            return (f,) + tuple(outs) + (td,)
        fn = self._lib.dh_hc_sf_and_derivs_dev_noalloc if no_alloc else self._lib.dh_hc_sf_and_derivs_dev
        _lib.check(fn(self._h, ctypes.byref(a), dd, sel, fptr, *optrs, ctypes.c_void_p(st.cuda_stream)))
        return (f,) + tuple(outs)

    # -- per-reflection derivatives (the per-hkl path) -------------------------------------------
    def _per_hkl(self, atoms, lcs, hkl, processor, include, anomalous):
        a, keep = self._atoms(atoms, lcs, include, anomalous)
        hkl_i = np.ascontiguousarray(np.asarray(hkl).reshape(-1, 3), dtype=np.int32)
        n = atoms.n
        failure = []

        def cb(user, k, f, row):
            try:
                fv = complex(f[0], f[1])
                r = np.ctypeslib.as_array(row, shape=(_lib.JAC_PLANES * n * 2,)).view(np.complex128)
                processor.onPerHklCalculation(int(k), fv, SfDerivativesAtHkl(r.reshape(_lib.JAC_PLANES, n)))
                return 0
            except BaseException as e:      # re-raised after the library has stopped
                failure.append(e)
                return 1

        fn = _lib.PER_HKL_FN(cb)
        rc = self._lib.dh_hc_per_hkl(self._h, ctypes.byref(a), hkl_i.shape[0], _lib.ptr(hkl_i, ctypes.c_int), fn,
                                     None)
        if failure:
            raise failure[0]
        _lib.check(rc)

    def jacobianUpload(self, atoms: AtomSet, localCoordinateSystems, hkl, countAtomContribution=None,
                       anomalous=None):
        """Inputs of the per-reflection rows (no dT/dF); reflections stay in the given order."""
        a, keep = self._atoms(atoms, localCoordinateSystems, countAtomContribution, anomalous)
        hkl_i = np.ascontiguousarray(np.asarray(hkl).reshape(-1, 3), dtype=np.int32)
        _lib.check(self._lib.dh_hc_jacobian_upload(self._h, ctypes.byref(a), hkl_i.shape[0],
                                                   _lib.ptr(hkl_i, ctypes.c_int)))
        self._jac_shape = (hkl_i.shape[0], atoms.n)

    def jacobianRows(self, k0: int = 0, k1: int = None):
        """(F (k1-k0,) complex, rows (k1-k0, 10, n) complex) of the uploaded reflections [k0, k1)."""
        nh, n = self._jac_shape
        k1 = nh if k1 is None else k1
        m = max(k1 - k0, 0)
        f = np.zeros(max(m, 1), dtype=np.complex128)
        rows = np.zeros((max(m, 1), _lib.JAC_PLANES, max(n, 1)), dtype=np.complex128)
        _lib.check(self._lib.dh_hc_jacobian_rows(self._h, int(k0), int(k1),
                                                 f.ctypes.data_as(_lib.c_dbl_p), rows.ctypes.data_as(_lib.c_dbl_p)))
        return f[:m], rows[:m, :, :n]

    def calculateJacobian(self, atoms: AtomSet, localCoordinateSystems, hkl, countAtomContribution=None,
                          anomalous=None):
        """All per-reflection rows at once: (F (n_hkl,), rows (n_hkl, 10, n)) complex, planes
        dF/dxyz (3), dF/dADP (6; U_iso in plane 3), dF/docc, in the calculator's conventions."""
        self.jacobianUpload(atoms, localCoordinateSystems, hkl, countAtomContribution, anomalous)
        return self.jacobianRows()

    # -- form-factor route -------------------------------------------------------------------------
    def _form_factors(self, fn, lcs, h, dtype, countAtom):
        n = self.n_atoms
        h = np.asarray(h)
        single = h.ndim == 1
        h = np.ascontiguousarray(h.reshape(-1, 3), dtype=dtype)
        m = h.shape[0]
        lcs_a = np.ascontiguousarray(np.broadcast_to(np.eye(3), (n, 3, 3)) if lcs is None else
                                     np.asarray(lcs, dtype=np.float64).reshape(n, 3, 3))
        inc = None if countAtom is None else np.ascontiguousarray(np.asarray(countAtom, dtype=bool).astype(np.uint8))
        f = np.zeros((max(m, 1), max(n, 1)), dtype=np.complex128)
        _lib.check(fn(self._h, lcs_a.ctypes.data_as(_lib.c_dbl_p), None if inc is None else _lib.ptr(inc, ctypes.c_ubyte),
                      m, h.ctypes.data_as(_lib.c_int_p if dtype == np.int32 else _lib.c_dbl_p),
                      f.ctypes.data_as(_lib.c_dbl_p)))
        f = f[:m, :n]
        return f[0] if single else f

    def calculateFormFactors(self, localCoordinateSystems, hkl, countAtomContribution=None):
        """Atomic form factors (HansenCoppensStructureFactorCalculator.h:171-175): hkl (3,) gives
        f (n_atoms,) complex as the reference; hkl (m, 3) gives (m, n_atoms) for m reflections."""
        return self._form_factors(self._lib.dh_hc_form_factors, localCoordinateSystems, hkl, np.int32,
                                  countAtomContribution)

    def calculateFormFactorsCart(self, localCoordinateSystems, h_cart, countAtomContribution=None):
        """As calculateFormFactors for Cartesian reciprocal vectors (1/Angstrom)."""
        return self._form_factors(self._lib.dh_hc_form_factors_cart, localCoordinateSystems, h_cart, np.float64,
                                  countAtomContribution)

    def calculateStructureFactors(self, atoms: AtomSet, localCoordinateSystems, hkl, countAtomContribution=None,
                                  anomalous=None):
        nh = np.asarray(hkl).reshape(-1, 3).shape[0]
        f, _ = self.calculateStructureFactorsAndDerivatives(
            atoms, localCoordinateSystems, hkl, np.ones(nh, dtype=np.complex128), countAtomContribution,
            DerivativesSelector(False, False, False, False), anomalous)
        return f

    # -- device-resident split (benchmarks, refinement loops) -----------------------------------
    def upload(self, atoms: AtomSet, localCoordinateSystems, hkl, dTarget_dF, countAtomContribution=None,
               anomalous=None):
        a, keep = self._atoms(atoms, localCoordinateSystems, countAtomContribution, anomalous)
        hkl_i, dd = self._hkl(hkl, dTarget_dF)
        self._keep = (keep, hkl_i, dd)
        self._nh, self._n = hkl_i.shape[0], atoms.n
        _lib.check(self._lib.dh_hc_upload(self._h, ctypes.byref(a), self._nh, _lib.ptr(hkl_i, ctypes.c_int),
                                          _lib.ptr(dd, ctypes.c_double)))

    def run(self, derivativesSelector: DerivativesSelector = None):
        sel = (derivativesSelector or DerivativesSelector()).bits()
        _lib.check(self._lib.dh_hc_run(self._h, sel))

    def fetch(self):
        nh, n = self._nh, self._n
        f = np.zeros(max(nh, 1), np.complex128)       # written through its (re, im) float64 view
        dx = np.zeros((max(n, 1), 3))
        du = np.zeros((max(n, 1), 6))
        do = np.zeros(max(n, 1))
        _lib.check(self._lib.dh_hc_fetch(self._h, _lib.ptr(f.view(np.float64), ctypes.c_double),
                                         _lib.ptr(dx, ctypes.c_double), _lib.ptr(du, ctypes.c_double),
                                         _lib.ptr(do, ctypes.c_double)))
        return f[:nh], TargetFunctionAtomicParamDerivatives(dx[:n], du[:n], do[:n])

    def autotune(self, derivativesSelector: DerivativesSelector = None):
        """Pick the fastest launch shape for the uploaded inputs in the current mode."""
        sel = (derivativesSelector or DerivativesSelector()).bits()
        c = ctypes.c_int()
        _lib.check(self._lib.dh_hc_autotune(self._h, sel, ctypes.byref(c)))
        return c.value

    def setTiming(self, on=True):
        _lib.check(self._lib.dh_hc_set_timing(self._h, 1 if on else 0))

    def lastTiming(self):
        ms = ctypes.c_double()
        gx, gy, ch = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.dh_hc_last_timing(self._h, ctypes.byref(ms), ctypes.byref(gx), ctypes.byref(gy),
                                               ctypes.byref(ch)))
        return {"main_ms": ms.value, "grid": (gx.value, gy.value), "chunk": ch.value}


class BatchCalculator:
    """Many calculators computed as one step: the per-structure prep, a single persistent F +
    adjoint kernel over every structure and the stage-2 reductions, optionally replayed from a
    captured hipGraph.  Inputs are uploaded per calculator (``calc.upload``), results read per
    calculator (``calc.fetch``).  No reference counterpart; it replaces a loop of
    calculateStructureFactorsAndDerivatives calls over small structures."""

    def __init__(self, calculators):
        self._calcs = list(calculators)           # the batch borrows them
        self._lib = _lib.load()
        arr = (ctypes.c_void_p * len(self._calcs))(*[c._h.value for c in self._calcs])
        h = ctypes.c_void_p()
        _lib.check(self._lib.dh_batch_create(arr, len(self._calcs), ctypes.byref(h)))
        self._b = h

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            self._lib.dh_batch_destroy(b)
            self._b = None

    def run(self, derivativesSelector: DerivativesSelector = None, graph: bool = True):
        sel = (derivativesSelector or DerivativesSelector()).bits()
        _lib.check(self._lib.dh_batch_run(self._b, sel, 1 if graph else 0))

    def autotune(self, derivativesSelector: DerivativesSelector = None):
        sel = (derivativesSelector or DerivativesSelector()).bits()
        c = ctypes.c_int()
        _lib.check(self._lib.dh_batch_autotune(self._b, sel, ctypes.byref(c)))
        return c.value

    def setTuning(self, hkl_chunk: int):
        _lib.check(self._lib.dh_batch_set_tuning(self._b, int(hkl_chunk)))

    def lastTiming(self):
        ms = ctypes.c_double()
        ch, un, gr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.dh_batch_last_timing(self._b, ctypes.byref(ms), ctypes.byref(ch), ctypes.byref(un),
                                                  ctypes.byref(gr)))
        return {"step_ms": ms.value, "chunk": ch.value, "units": un.value, "grid": gr.value}
