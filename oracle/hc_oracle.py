"""TEST INFRASTRUCTURE ONLY - calculator-level FP64 oracle.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module; the product
(discamb_amd) never imports it.  It restates, in numpy, the pre/post-processing of
HansenCoppensStructureFactorCalculator (src/Scattering/HansenCoppensStructureFactorCalculator.cpp):

  setModel                         .cpp:157-290   Slater density products, valence / occupancy,
                                                  multiplicity weights m/|G|
  setEngineSpecificSpaceGroupRep.  .cpp:107-155   Cartesian explicit-subset operations
  calculateStructureFactors...     .cpp:380-508   centering multipliers, conj(dT/dF) * mult, F * mult
  prepareDataForEngine_part        .cpp:897-1011  U_cif -> U_cart * 2 pi^2, frac -> Cartesian,
                                                  h_cart = h a* + k b* + l c*
  convertDerivatives               .cpp:634-664   Cartesian -> fractional, U_cart -> U_cif

and calls the C engine restatement in oracle/liboracle.so (hc_oracle.c).  ``mode="full_group"``
instead sums every operation of the space group explicitly (no inversion / centering folding,
no multipliers) - an independent check of the folding logic.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Model(ctypes.Structure):
    _fields_ = [("n_wfn", ctypes.c_int),
                ("core_off", ctypes.c_void_p), ("val_off", ctypes.c_void_p),
                ("core_c", ctypes.c_void_p), ("core_e", ctypes.c_void_p), ("core_p", ctypes.c_void_p),
                ("val_c", ctypes.c_void_p), ("val_e", ctypes.c_void_p), ("val_p", ctypes.c_void_p),
                ("dv_zeta", ctypes.c_void_p), ("dv_n", ctypes.c_void_p), ("dv_nl", ctypes.c_void_p),
                ("anom", ctypes.c_void_p),
                ("n_types", ctypes.c_int),
                ("pval", ctypes.c_void_p), ("kappa", ctypes.c_void_p), ("kappa_p", ctypes.c_void_p),
                ("plm", ctypes.c_void_p), ("plm_nl", ctypes.c_void_p)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _LIB = ctypes.CDLL(path)
        _LIB.hc_oracle_engine.restype = ctypes.c_int
    return _LIB


def parse_op(text):
    from fractions import Fraction
    rot = np.zeros((3, 3))
    tr = np.zeros(3)
    for i, comp in enumerate(text.replace(" ", "").lower().split(",")):
        comp = comp.replace("-", "+-")
        for term in [t for t in comp.split("+") if t]:
            s = -1.0 if term.startswith("-") else 1.0
            term = term.lstrip("-")
            if term[-1:] in ("x", "y", "z"):
                rot[i, "xyz".index(term[-1])] += s * (float(term[:-1]) if term[:-1] else 1.0)
            else:
                tr[i] += s * float(Fraction(term))
    return rot, tr


def cell_f2c(cell):
    a, b, c, al, be, ga = cell
    f = np.pi / 180.0
    ca, cb, cg, sg = np.cos(al * f), np.cos(be * f), np.cos(ga * f), np.sin(ga * f)
    v = np.sqrt(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg)
    return np.array([[a, b * cg, c * cb], [0.0, b * sg, c * (ca - cb * cg) / sg], [0.0, 0.0, c * v / sg]])


def decompose_group(symops):
    """(order, subset [(R, t)], centro, t_inv, centering letter, obverse).  Coset representatives
    are chosen by rotation matrix: one operation per rotation R, and of each (R, -R) pair only the
    first listed when the group is centrosymmetric."""
    ops = [parse_op(s) for s in symops]
    order = len(ops)
    cen = [t % 1.0 for R, t in ops if np.allclose(R, np.eye(3))]
    inv = [t % 1.0 for R, t in ops if np.allclose(R, -np.eye(3))]
    centro = len(inv) > 0
    t_inv = inv[0] if centro else np.zeros(3)
    subset, seen = [], []
    for R, t in ops:
        if any(np.allclose(R, S) for S in seen) or (centro and any(np.allclose(R, -S) for S in seen)):
            continue
        seen.append(R)
        subset.append((R, t % 1.0))

    def has(v):
        return any(np.allclose(((c - np.array(v)) + 0.5) % 1.0 - 0.5, 0, atol=1e-8) for c in cen)

    letter, obverse = "P", True
    if len(cen) == 4:
        letter = "F"
    elif len(cen) == 2:
        for L, v in (("A", (0, .5, .5)), ("B", (.5, 0, .5)), ("C", (.5, .5, 0)), ("I", (.5, .5, .5))):
            if has(v):
                letter = L
    elif len(cen) == 3:
        letter = "R"
        obverse = has((2 / 3, 1 / 3, 1 / 3))
    return order, subset, centro, t_inv, letter, obverse


def centering_mult(letter, obverse, hkl):
    h, k, l = hkl[:, 0], hkl[:, 1], hkl[:, 2]
    if letter == "P":
        return np.ones(len(hkl))
    if letter == "C":
        return np.where((h + k) % 2 == 0, 2.0, 0.0)
    if letter == "A":
        return np.where((k + l) % 2 == 0, 2.0, 0.0)
    if letter == "B":
        return np.where((h + l) % 2 == 0, 2.0, 0.0)
    if letter == "I":
        return np.where((h + k + l) % 2 == 0, 2.0, 0.0)
    if letter == "F":
        return np.where(((h % 2) == (k % 2)) & ((k % 2) == (l % 2)), 4.0, 0.0)
    if letter == "R":
        if obverse:
            return np.where(np.abs(-h + k + l) % 3 == 0, 3.0, 0.0)
        return np.where(np.abs(h - k + l) % 2 == 0, 3.0, 0.0)   # scattering_utilities.cpp:130
    raise ValueError(letter)


def sphere_density(w, valence):
    """atomicStoWfnToSphericalDensity (HC_WfnData.cpp:11-73)."""
    c, e, p = [], [], []
    idx = w.valence_orbitals_indices if valence else w.core_orbitals_indices
    for oi in idx:
        occ = w.orbital_occupancy[oi]
        if occ == 0:
            continue
        o = w.orbitals[oi]
        for i in range(len(o.coefficients)):
            for j in range(i + 1):
                v = o.coefficients[i] * o.coefficients[j] * occ / (4 * np.pi)
                if i != j:
                    v *= 2
                c.append(v)
                e.append(o.exponents[i] + o.exponents[j])
                p.append(o.powers[i] + o.powers[j])
    if valence:
        tot = sum(w.orbital_occupancy[i] for i in w.valence_orbitals_indices)
        c = [v / tot for v in c]
    return c, e, p


def u_cif_to_cart_matrix(cell):
    """6x6 linear map U_cart(6) = J U_cif(6) in the (11 22 33 12 13 23) parameter order."""
    f2c = cell_f2c(cell)
    c2f = np.linalg.inv(f2c)
    A = f2c * np.linalg.norm(c2f, axis=1)[None, :]
    J = np.zeros((6, 6))
    idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for j, (a, b) in enumerate(idx):
        E = np.zeros((3, 3))
        E[a, b] = E[b, a] = 1.0
        M = A @ E @ A.T
        J[:, j] = [M[x, y] for x, y in idx]
    return J


def _engine_inputs(crystal, params, atoms, lcs, hkl, include, mode, anomalous):
    """Everything the C restatement takes, in the engine's units (prepareDataForEngine_part)."""
    hkl = np.ascontiguousarray(np.asarray(hkl, dtype=np.int64).reshape(-1, 3))
    m = hkl.shape[0]
    n = atoms.n
    cell = crystal.cell
    f2c = cell_f2c(cell)
    c2f = np.linalg.inv(f2c)
    order, subset, centro, t_inv, letter, obverse = decompose_group(crystal.symops)
    if mode == "full_group":
        ops = [(R, t) for R, t in (parse_op(s) for s in crystal.symops)]
        centro_flag, tinv, mult = 0, np.zeros(3), np.ones(m)
    else:
        ops = subset
        centro_flag, tinv = (1 if centro else 0), t_inv
        mult = centering_mult(letter, obverse, hkl)
    x = {}
    x["rot"] = np.ascontiguousarray(np.array([f2c @ R @ c2f for R, t in ops]).reshape(-1))
    x["trans"] = np.ascontiguousarray(np.array([f2c @ t for R, t in ops]).reshape(-1))
    x["n_ops"], x["centro"] = len(ops), centro_flag
    x["inv_cart"] = np.ascontiguousarray(f2c @ tinv)
    # model
    wfn = params.wfn_parameters
    core, val = [sphere_density(w, False) for w in wfn], [sphere_density(w, True) for w in wfn]

    def flat(parts):
        off = np.cumsum([0] + [len(q[0]) for q in parts]).astype(np.int32)
        c = np.array(sum([q[0] for q in parts], []) + [0.0])
        e = np.array(sum([q[1] for q in parts], []) + [1.0])
        p = np.array(sum([q[2] for q in parts], []) + [0], dtype=np.int32)
        return off, c, e, p

    co, cc, ce, cp = flat(core)
    vo, vc, ve, vp = flat(val)
    dv_zeta = np.array([w.deformation_valence_exponent for w in wfn])
    dv_n = np.zeros((len(wfn), 5), dtype=np.int32)
    dv_nl = np.zeros(len(wfn), dtype=np.int32)
    for i, w in enumerate(wfn):
        k = min(5, len(w.deformation_valence_power))
        dv_nl[i] = k
        dv_n[i, :k] = w.deformation_valence_power[:k]
    anom = np.array([[complex(w.anomalous_scattering).real, complex(w.anomalous_scattering).imag] for w in wfn])
    tp = params.type_parameters
    pval = np.array([t.p_val for t in tp])
    kap = np.array([t.kappa_spherical_valence for t in tp])
    kapp = np.array([t.kappa_deformation_valence for t in tp])
    plm = np.ascontiguousarray(np.array([t.plm25() for t in tp]))
    plm_nl = np.array([len(t.p_lm) for t in tp], dtype=np.int32)
    x["keep"] = [co, cc, ce, cp, vo, vc, ve, vp, dv_zeta, dv_n, dv_nl, anom, pval, kap, kapp, plm, plm_nl]
    x["model"] = _Model(len(wfn), co.ctypes.data, vo.ctypes.data, cc.ctypes.data, ce.ctypes.data, cp.ctypes.data,
                        vc.ctypes.data, ve.ctypes.data, vp.ctypes.data, dv_zeta.ctypes.data, dv_n.ctypes.data,
                        dv_nl.ctypes.data, anom.ctypes.data, len(tp), pval.ctypes.data, kap.ctypes.data,
                        kapp.ctypes.data, plm.ctypes.data, plm_nl.ctypes.data)
    # atoms (prepareDataForEngine_part)
    xyz = atoms.xyz if crystal.xyz_convention == "fractional" else atoms.xyz @ c2f.T
    x["xyz_cart"] = np.ascontiguousarray(xyz @ f2c.T)
    J = u_cif_to_cart_matrix(cell)
    adp_n = np.ascontiguousarray(atoms.adp_size.astype(np.int32))
    adp_flat = []
    for i in range(n):
        if adp_n[i] == 6:
            u = atoms.adp[i] if crystal.adp_convention == "U_cart" else J @ atoms.adp[i]
            adp_flat += list(2 * np.pi ** 2 * u)
        elif adp_n[i] == 1:
            adp_flat.append(2 * np.pi ** 2 * atoms.adp[i, 0])
    x["adp_n"], x["adp_flat"] = adp_n, np.array(adp_flat + [0.0])
    x["occ"] = np.ascontiguousarray(atoms.occupancy, dtype=np.float64)
    x["mult_w"] = np.ascontiguousarray(crystal.atoms.multiplicity / float(order))
    x["lcs"] = np.ascontiguousarray(np.asarray(lcs, dtype=np.float64).reshape(n, 9))
    x["anom_atom"] = None
    if anomalous is not None:
        a = np.asarray(anomalous, dtype=np.complex128)
        x["anom_atom"] = np.ascontiguousarray(np.stack([a.real, a.imag], 1))
    x["include"] = np.ones(n, dtype=np.uint8) if include is None else np.ascontiguousarray(include, dtype=np.uint8)
    x["a_wfn"] = np.ascontiguousarray(np.array(params.atom_to_wfn_map, dtype=np.int32))
    x["a_type"] = np.ascontiguousarray(np.array(params.atom_to_type_map, dtype=np.int32))
    x["h_cart"] = np.ascontiguousarray(hkl.astype(np.float64) @ c2f)
    x["mult"], x["m"], x["n"], x["f2c"], x["J"] = np.ascontiguousarray(mult, dtype=np.float64), m, n, f2c, J
    return x


def _atom_args(x):
    P = ctypes.c_void_p
    return [ctypes.byref(x["model"]), ctypes.c_int(x["n"]), P(x["a_wfn"].ctypes.data), P(x["a_type"].ctypes.data),
            P(x["xyz_cart"].ctypes.data), P(x["adp_n"].ctypes.data), P(x["adp_flat"].ctypes.data),
            P(x["occ"].ctypes.data), P(x["mult_w"].ctypes.data), P(x["lcs"].ctypes.data),
            P(x["anom_atom"].ctypes.data) if x["anom_atom"] is not None else P(0), P(x["include"].ctypes.data),
            ctypes.c_int(x["n_ops"]), P(x["rot"].ctypes.data), P(x["trans"].ctypes.data), ctypes.c_int(x["centro"]),
            P(x["inv_cart"].ctypes.data)]


def sf_and_derivs(crystal, params, atoms, lcs, hkl, dtdf, include=None, n_threads=1, mode="reference",
                  anomalous=None):
    """Returns (F complex (m,), d_xyz (n,3), d_adp (n,6) (or U_iso in column 0), d_occ (n,)); derivative
    conventions follow the crystal's input conventions."""
    x = _engine_inputs(crystal, params, atoms, lcs, hkl, include, mode, anomalous)
    m, n, f2c, J, adp_n, mult = x["m"], x["n"], x["f2c"], x["J"], x["adp_n"], x["mult"]
    d = np.asarray(dtdf, dtype=np.complex128) * 1.0
    d = np.conj(d) * mult
    dd = np.ascontiguousarray(np.stack([d.real, d.imag], 1))
    fo = np.zeros((m, 2))
    dx = np.zeros((n, 3))
    du = np.zeros(max(1, int(adp_n.sum())))
    do = np.zeros(n)
    P = ctypes.c_void_p
    rc = lib().hc_oracle_engine(*_atom_args(x), ctypes.c_int(m), P(x["h_cart"].ctypes.data), P(dd.ctypes.data),
                                P(fo.ctypes.data), P(dx.ctypes.data), P(du.ctypes.data), P(do.ctypes.data),
                                ctypes.c_int(n_threads))
    if rc != 0:
        raise RuntimeError("oracle engine failed")
    F = (fo[:, 0] + 1j * fo[:, 1]) * mult
    # convertDerivatives: Cartesian -> fractional, U_cart -> U_cif
    d_xyz = dx @ f2c if crystal.xyz_convention == "fractional" else dx
    d_adp = np.zeros((n, 6))
    off = 0
    for i in range(n):
        if adp_n[i] == 6:
            g = du[off:off + 6]
            d_adp[i] = g @ J if crystal.adp_convention == "U_cif" else g
            off += 6
        elif adp_n[i] == 1:
            d_adp[i, 0] = du[off]
            off += 1
    return F, d_xyz, d_adp, do


def line_order(hkl):
    """Reflections along lines: the hkl axis with the fewest lines (findPreferredHklOrderingDirection,
    scattering_utilities.cpp:371-435), lines in lexicographic order of the other two indices, each by
    increasing index along the axis and split where a step is not +1 (splitHklLines, :463-517).
    Returns (permutation, line offsets, axis)."""
    hkl = np.asarray(hkl, dtype=np.int64).reshape(-1, 3)
    best, best_n = 2, None
    for ax in range(3):
        o = [(ax + 1) % 3, (ax + 2) % 3]
        n = len({(a, b) for a, b in hkl[:, o]})
        if best_n is None or n < best_n:
            best, best_n = ax, n
    o1, o2 = (best + 1) % 3, (best + 2) % 3
    perm = np.lexsort((hkl[:, best], hkl[:, o2], hkl[:, o1]))
    hs = hkl[perm]
    brk = np.ones(len(hs), dtype=bool)
    if len(hs) > 1:
        same = (hs[1:, o1] == hs[:-1, o1]) & (hs[1:, o2] == hs[:-1, o2]) & (hs[1:, best] == hs[:-1, best] + 1)
        brk[1:] = ~same
    off = np.append(np.nonzero(brk)[0], len(hs)).astype(np.int32)
    return perm, off, best


def sf_and_derivs_lines(crystal, params, atoms, lcs, hkl, dtdf, include=None, n_threads=1, anomalous=None):
    """sf_and_derivs restated the way the reference's Engine3 computes it (global-frame P_lm, per
    reflection tables, phase and Debye-Waller factors carried along hkl lines: hc_oracle_engine_lines)
    - the CPU baseline of the benchmark.  Acentric groups only.  Same results as sf_and_derivs to
    rounding (the recurrence error grows by a few ulp per step)."""
    x = _engine_inputs(crystal, params, atoms, lcs, hkl, include, "reference", anomalous)
    if x["centro"]:
        raise ValueError("the line-recurrence restatement covers acentric groups")
    m, n, f2c, J, adp_n, mult = x["m"], x["n"], x["f2c"], x["J"], x["adp_n"], x["mult"]
    perm, off, ax = line_order(hkl)
    c2f = np.linalg.inv(f2c)
    d = np.conj(np.asarray(dtdf, dtype=np.complex128)) * mult
    dd = np.ascontiguousarray(np.stack([d.real, d.imag], 1)[perm])
    h_lines = np.ascontiguousarray(x["h_cart"][perm])
    step = np.ascontiguousarray(c2f[ax].astype(np.float64))
    fo = np.zeros((m, 2))
    dx = np.zeros((n, 3))
    du = np.zeros(max(1, int(adp_n.sum())))
    do = np.zeros(n)
    P = ctypes.c_void_p
    rc = lib().hc_oracle_engine_lines(*_atom_args(x), ctypes.c_int(m), P(h_lines.ctypes.data), ctypes.c_int(len(off) - 1),
                                      P(off.ctypes.data), P(step.ctypes.data), P(dd.ctypes.data), P(fo.ctypes.data),
                                      P(dx.ctypes.data), P(du.ctypes.data), P(do.ctypes.data), ctypes.c_int(n_threads))
    if rc != 0:
        raise RuntimeError("oracle line engine failed")
    F = np.zeros(m, dtype=np.complex128)
    F[perm] = (fo[:, 0] + 1j * fo[:, 1])
    F *= mult
    d_xyz = dx @ f2c if crystal.xyz_convention == "fractional" else dx
    d_adp = np.zeros((n, 6))
    o = 0
    for i in range(n):
        if adp_n[i] == 6:
            g = du[o:o + 6]
            d_adp[i] = g @ J if crystal.adp_convention == "U_cif" else g
            o += 6
        elif adp_n[i] == 1:
            d_adp[i, 0] = du[o]
            o += 1
    return F, d_xyz, d_adp, do


def per_hkl(crystal, params, atoms, lcs, hkl, include=None, n_threads=1, anomalous=None):
    """Per-reflection F and derivatives of the reference's per-hkl path (HansenCoppens_SF_Engine.h:
    512-780 behind HansenCoppensStructureFactorCalculator.h:384-545).  Returns (F (m,) complex,
    rows (m, 10, n) complex): planes dF/dxyz (3), dF/dU (6; U_iso in plane 3), dF/docc, converted to
    the crystal's conventions as WrapperForSingleHklResultProcessor::convert does (.h:512-531)."""
    x = _engine_inputs(crystal, params, atoms, lcs, hkl, include, "reference", anomalous)
    m, n, f2c, J, adp_n = x["m"], x["n"], x["f2c"], x["J"], x["adp_n"]
    fo = np.zeros((m, 2))
    rows = np.zeros((m, 10, max(n, 1), 2))
    P = ctypes.c_void_p
    rc = lib().hc_oracle_per_hkl(*_atom_args(x), ctypes.c_int(m), P(x["h_cart"].ctypes.data),
                                 P(x["mult"].ctypes.data), P(fo.ctypes.data), P(rows.ctypes.data),
                                 ctypes.c_int(n_threads))
    if rc != 0:
        raise RuntimeError("oracle per-hkl engine failed")
    F = fo[:, 0] + 1j * fo[:, 1]
    R = rows[..., 0] + 1j * rows[..., 1]
    R = R[:, :, :n]
    if crystal.xyz_convention == "fractional":
        R[:, 0:3, :] = np.einsum("ji,mjn->min", f2c, R[:, 0:3, :])
    if crystal.adp_convention == "U_cif":
        aniso = adp_n == 6
        R[:, 3:9, aniso] = np.einsum("ji,mjn->min", J, R[:, 3:9, aniso])
    return F, R


# ------------------------------------------------------------------------------ form-factor route
SPLINE_STEP, SPLINE_NODES = 0.001, 10000     # HansenCoppensStructureFactorCalculator.cpp:1250-1251


def spline_coefficients(y):
    """NaturalCubicSpline::set (src/MathUtilities/NaturalCubicSpline.cpp) restated: derivatives x
    from the tridiagonal system with rows (1, 4, 1) - 4 on the diagonal in the first and last row
    too - and right-hand side 3 (y_{i+1} - y_{i-1}) (one-sided at the ends); segment i is
    y_i + t x_i + t^2 (3 (y_{i+1} - y_i) - 2 x_i - x_{i+1}) + t^3 (2 (y_i - y_{i+1}) + x_i + x_{i+1});
    the last node's segment is zero."""
    from scipy.linalg import solve_banded
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    d = np.empty(n)
    d[0] = 3 * (y[1] - y[0])
    d[1:-1] = 3 * (y[2:] - y[:-2])
    d[-1] = 3 * (y[-1] - y[-2])
    ab = np.zeros((3, n))
    ab[0, 1:] = 1.0
    ab[1, :] = 4.0
    ab[2, :-1] = 1.0
    x = solve_banded((1, 1), ab, d)
    c = np.zeros((n, 4))
    c[:-1, 0] = y[:-1]
    c[:-1, 1] = x[:-1]
    c[:-1, 2] = 3 * (y[1:] - y[:-1]) - 2 * x[:-1] - x[1:]
    c[:-1, 3] = 2 * (y[:-1] - y[1:]) + x[:-1] + x[1:]
    return c


def spline_eval(c, x, step=SPLINE_STEP):
    """NaturalCubicSpline::evaluate: zero below the grid and from the last node on."""
    x = np.asarray(x, dtype=np.float64)
    u = x / step
    i = np.floor(u).astype(np.int64)
    ok = (u >= 0) & (i < len(c))
    i = np.clip(i, 0, len(c) - 1)
    t = (x - i * step) / step
    q = c[i]
    v = q[..., 0] + t * q[..., 1] + t * t * q[..., 2] + t * t * t * q[..., 3]
    return np.where(ok, v, 0.0)


def _sph_density(parts, h):
    c, e, p = parts
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros(len(h))
    if len(c):
        cc, ee = np.ascontiguousarray(c, dtype=np.float64), np.ascontiguousarray(e, dtype=np.float64)
        pp = np.ascontiguousarray(p, dtype=np.int32)
        P = ctypes.c_void_p
        lib().hc_oracle_sph_density(P(cc.ctypes.data), P(ee.ctypes.data), P(pp.ctypes.data), ctypes.c_int(len(cc)),
                                    P(h.ctypes.data), ctypes.c_int(len(h)), P(out.ctypes.data))
    return out


def form_factors(crystal, params, lcs, h_cart, include=None):
    """Atomic form factors of the form-factor route (HansenCoppensStructureFactorCalculator::
    calculateFormFactorsCart, .cpp:1305-1333, with the splines of setDataForCpuEngineCallsForSingleHkl,
    .cpp:1244-1272, and HansenCoppens_SF_Engine::calculateFormFactors, HansenCoppens_SF_Engine.cpp:
    383-495): f[k, a] = spline_core[wfn(type)](|h|) + spline_val[type](|h|) + f_def, zero for
    excluded atoms.  h_cart (m, 3) in 1/Angstrom."""
    h = np.ascontiguousarray(np.asarray(h_cart, dtype=np.float64).reshape(-1, 3))
    m, n = h.shape[0], len(params.atom_to_type_map)
    wfn, tp = params.wfn_parameters, params.type_parameters
    type_wfn = np.zeros(len(tp), dtype=int)
    for a in range(n):
        type_wfn[params.atom_to_type_map[a]] = params.atom_to_wfn_map[a]
    grid = np.arange(SPLINE_NODES) * SPLINE_STEP
    core = [spline_coefficients(_sph_density(sphere_density(w, False), grid)) for w in wfn]
    val = [spline_coefficients(tp[t].p_val * _sph_density(sphere_density(wfn[type_wfn[t]], True),
                                                            grid / tp[t].kappa_spherical_valence))
           for t in range(len(tp))]
    hl = np.sqrt((h * h).sum(1))
    S = np.stack([spline_eval(core[type_wfn[t]], hl) + spline_eval(val[t], hl) for t in range(len(tp))], 1)
    # deformation valence (the pinned def_val of the C restatement)
    x = _engine_inputs(crystal, params, crystal.atoms, lcs, np.zeros((1, 3), dtype=np.int64), include,
                       "reference", None)
    dv = np.zeros((m, max(n, 1), 2))
    P = ctypes.c_void_p
    lib().hc_oracle_def_val_ff(ctypes.byref(x["model"]), ctypes.c_int(n), P(x["a_wfn"].ctypes.data),
                               P(x["a_type"].ctypes.data), P(x["lcs"].ctypes.data), P(x["include"].ctypes.data),
                               ctypes.c_int(m), P(h.ctypes.data), P(dv.ctypes.data))
    f = S[:, params.atom_to_type_map] + dv[:, :n, 0] + 1j * dv[:, :n, 1]
    inc = np.ones(n, dtype=bool) if include is None else np.asarray(include, dtype=bool)
    f[:, ~inc] = 0.0
    return f


# ------------------------------------------------------------------------------ N-Gaussian tables
_NGAUSS = None


def ngauss_tables():
    """The N-Gaussian tables as data (discamb_amd/data/ngaussian_tables.json)."""
    global _NGAUSS
    if _NGAUSS is None:
        import json
        path = os.path.join(HERE, "..", "discamb_amd", "data", "ngaussian_tables.json")
        _NGAUSS = {name: {e["label"]: e for e in entries} for name, entries in json.load(open(path))["tables"].items()}
    return _NGAUSS


def ngauss(label, table, h):
    """n_gaussian_form_factors_table::getFormFactor(label, table).calculate_h(h)
    (NGaussianFormFactorsTable.cpp:3176-3217, NGaussianFormFactor.cpp:115-127): exact label, else
    the ion spelling with '1' before the sign; None if the table has neither."""
    t = ngauss_tables()[table]
    e = t.get(label)
    if e is None and len(label) > 1 and label[-1] in "+-" and not label[-2].isdigit():
        return ngauss(label[:-1] + "1" + label[-1], table, h)
    if e is None:
        return None
    h = np.asarray(h, dtype=np.float64)
    x2 = h * h / 4
    return sum(a * np.exp(-b * x2) for a, b in zip(e["a"], e["b"])) + e["c"]


ELEMENTS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
            "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
            "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf").split()


def _iam_params(params):
    """The model the reference's calculateSF_IAM passes to its engine: one type per wavefunction
    with no valence or multipole terms (HansenCoppens_SF_Engine.cpp:229-278)."""
    from discamb_amd.model import HC_AtomTypeParameters, HC_ModelParameters
    types = [HC_AtomTypeParameters([[0.0]], 0.0, 1.0, 1.0) for _ in params.wfn_parameters]
    return HC_ModelParameters(params.wfn_parameters, types, list(params.atom_to_wfn_map),
                              list(params.atom_to_wfn_map))


def _iam_core(crystal, params, hkl, table="Waasmeier-Kirfel"):
    c2f = np.linalg.inv(cell_f2c(crystal.cell))
    hl = np.sqrt(((np.asarray(hkl, dtype=np.float64).reshape(-1, 3) @ c2f) ** 2).sum(1))
    cols = []
    for w in params.wfn_parameters:
        f = ngauss(w.label, table, hl)
        if f is None:
            raise ValueError("no Gaussian form factor for " + w.label)
        cols.append(f)
    return np.ascontiguousarray(np.stack(cols, 1))


def iam_per_hkl(crystal, params, atoms, hkl, include=None, anomalous=None, n_threads=1, table="Waasmeier-Kirfel"):
    """Per-reflection F and derivative rows of the independent-atom model (the per-hkl engine with
    the Waasmaier-Kirfel form factors, calculateStructureFactorsAndDerivatives<T> in CPU_IAM mode;
    another N-Gaussian table by name)."""
    core = _iam_core(crystal, params, hkl, table)
    lcs = np.broadcast_to(np.eye(3), (atoms.n, 3, 3))
    lib().hc_oracle_set_core_override(ctypes.c_void_p(core.ctypes.data), ctypes.c_int(core.shape[1]))
    try:
        return per_hkl(crystal, _iam_params(params), atoms, lcs, hkl, include, n_threads, anomalous)
    finally:
        lib().hc_oracle_set_core_override(ctypes.c_void_p(0), ctypes.c_int(0))


def iam_sf_and_derivs(crystal, params, atoms, hkl, dtdf, include=None, anomalous=None, n_threads=1,
                      table="Waasmeier-Kirfel"):
    """CPU_IAM calculateStructureFactorsAndDerivatives (HansenCoppensStructureFactorCalculator.cpp:
    478-499 over HansenCoppens_SF_Engine::calculateSF): F and dT/dp = sum_h Re(conj(dT/dF_h) dF_h/dp)
    - the engine's summed paths are the contraction of its per-hkl derivatives."""
    F, R = iam_per_hkl(crystal, params, atoms, hkl, include, anomalous, n_threads, table)
    g = np.conj(np.asarray(dtdf, dtype=np.complex128))[:, None]
    n = atoms.n
    dx = np.stack([np.real(g * R[:, j]).sum(0) for j in range(3)], 1)
    du = np.stack([np.real(g * R[:, 3 + j]).sum(0) for j in range(6)], 1)
    do = np.real(g * R[:, 9]).sum(0)
    return F, dx[:n], du[:n], do[:n]


def iam_full_group_f(crystal, params, atoms, hkl, anomalous=None, table="Waasmeier-Kirfel"):
    """Explicit sum over every operation of the group with the IAM form factors (no folding)."""
    core = _iam_core(crystal, params, hkl, table)
    lcs = np.broadcast_to(np.eye(3), (atoms.n, 3, 3))
    lib().hc_oracle_set_core_override(ctypes.c_void_p(core.ctypes.data), ctypes.c_int(core.shape[1]))
    try:
        return sf_and_derivs(crystal, _iam_params(params), atoms, lcs, hkl, np.ones(len(core)), mode="full_group",
                             anomalous=anomalous)[0]
    finally:
        lib().hc_oracle_set_core_override(ctypes.c_void_p(0), ctypes.c_int(0))


def electron_sf_and_derivs(crystal, params, atoms, lcs, hkl, dtdf, include=None, n_threads=1, mode="reference",
                           anomalous=None):
    """sf_and_derivs with setElectronScattering(true): every term's atomic form factor through
    the Mott-Bethe formula with Z of the atom's wavefunction and, at h = 0, the "electron-IT"
    Gaussian value (HansenCoppens_SF_Engine3.cpp:952-970, 1387-1398)."""
    z = np.array([params.wfn_parameters[w].atomic_number for w in params.atom_to_wfn_map], dtype=np.float64)
    f0 = np.array([ngauss(ELEMENTS[int(v) - 1], "electron-IT", 0.0) for v in z], dtype=np.float64)
    lib().hc_oracle_set_electron(ctypes.c_void_p(z.ctypes.data), ctypes.c_void_p(f0.ctypes.data))
    try:
        return sf_and_derivs(crystal, params, atoms, lcs, hkl, dtdf, include, n_threads, mode, anomalous)
    finally:
        lib().hc_oracle_set_electron(ctypes.c_void_p(0), ctypes.c_void_p(0))


# ------------------------------------------------------------------------------ local coordinate systems
def lcs_matrix(cell, xyz_frac, d):
    """LocalCoordinateSystemCalculator::calculate restated (LocalCoordinateSystemCalculator.cpp:54-238,
    CrossProductLcs.cpp:10-62): the local axes (as matrix columns) of definition d (a
    discamb_amd.LcsDefinition) for atoms at fractional positions xyz_frac."""
    f2c = cell_f2c(cell)
    xyz_frac = np.asarray(xyz_frac, dtype=np.float64)

    def pos(ref):
        R, t = parse_op(ref[1])
        return f2c @ (R @ xyz_frac[ref[0]] + t)

    r0 = pos((d.central, d.central_op))

    def direction(refs, kind):
        if kind == "average_direction":
            v = np.zeros(3)
            for r in refs:
                u = pos(r) - r0
                v += u / np.sqrt(u @ u)
        elif kind == "average_position":
            v = np.mean([pos(r) for r in refs], axis=0) - r0
        else:
            raise ValueError(kind)
        return v / np.sqrt(v @ v)

    def any_orthogonal(v):
        rx, ry = np.cross(v, [1.0, 0, 0]), np.cross(v, [0, 1.0, 0])
        r = rx if rx @ rx > ry @ ry else ry
        return r / np.sqrt(r @ r)

    t1, t2 = d.direction1_type, d.direction2_type
    if t1 == "any_orthogonal" and t2 == "any_orthogonal":
        return np.eye(3)
    if t1 == "any_orthogonal":
        r2 = direction(d.ref2, t2)
        r1 = any_orthogonal(r2)
    else:
        r1 = direction(d.ref1, t1)
        r2 = any_orthogonal(r1) if t2 == "any_orthogonal" else direction(d.ref2, t2)
    c1, c2 = d.coordinate_1 + 1, d.coordinate_2 + 1
    c3 = 6 - c1 - c2
    s = (1 - 2 * ((2 + c2 - c1) % 3)) * (1.0 if d.is_r else -1.0)
    a0 = r1 / np.sqrt(r1 @ r1)
    a2 = s * np.cross(a0, r2)
    a2 /= np.sqrt(a2 @ a2)
    a1 = s * np.cross(a2, a0)
    axes = [None, None, None]
    axes[c1 - 1], axes[c2 - 1], axes[c3 - 1] = a0, a1, a2
    if len(d.chirality) == 3:
        p = [pos(r) - r0 for r in d.chirality]
        if np.cross(p[0], p[1]) @ p[2] < 0:
            axes[3 - d.coordinate_1 - d.coordinate_2] = -axes[3 - d.coordinate_1 - d.coordinate_2]
    return np.stack(axes, 1)
