"""Seeded synthetic Hansen-Coppens structures shaped like the benchmark workloads.

Every value derives from numpy PCG64 with seed ``config * 1000 + replicate``.  Common choices:
random orthonormal local frames (QR of a Gaussian 3x3, det +1); anisotropic U = Q diag(l) Q^T with
l ~ U(0.01, 0.05) A^2 given to the library as U_cif; isotropic H with Uiso ~ U(0.02, 0.06);
minimum separation 0.9 A over symmetry images and lattice translations; occupancy 1; general
positions (multiplicity |G|); f' = f'' = 0; dT/dF ~ complex N(0, 1); reflections: all
symmetry-unique, Friedel-merged hkl with d >= d_min.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

from .model import (AtomSet, Crystal, HC_AtomTypeParameters, HC_ModelParameters, parse_symop,
                    wfn_type_from_bank)

SPACE_GROUPS = {
    "P1": ["x,y,z"],
    "P-1": ["x,y,z", "-x,-y,-z"],
    "P21/c": ["x,y,z", "-x,y+1/2,-z+1/2", "-x,-y,-z", "x,-y+1/2,z+1/2"],
    "C2/c": ["x,y,z", "-x,y,-z+1/2", "x+1/2,y+1/2,z", "-x+1/2,y+1/2,-z+1/2", "-x,-y,-z", "x,-y,z+1/2",
             "-x+1/2,-y+1/2,-z", "x+1/2,-y+1/2,z+1/2"],
    "P212121": ["x,y,z", "-x+1/2,-y,z+1/2", "-x,y+1/2,-z+1/2", "x+1/2,-y+1/2,-z"],
    # non-diagonal rotations (general-rotation kernels)
    "P43212": ["x,y,z", "-x,-y,z+1/2", "-y+1/2,x+1/2,z+3/4", "y+1/2,-x+1/2,z+1/4", "-x+1/2,y+1/2,-z+3/4",
               "x+1/2,-y+1/2,-z+1/4", "y,x,-z", "-y,-x,-z+1/2"],
    "P31": ["x,y,z", "-y,x-y,z+1/3", "-x+y,-x,z+2/3"],
    # a two-operation diagonal subgroup {1, 2_1 along c} under three (P6_3, hexagonal axes) or two
    # (P4/m, centrosymmetric) coset representatives
    "P63": ["x,y,z", "-y,x-y,z", "-x+y,-x,z", "-x,-y,z+1/2", "y,-x+y,z+1/2", "x-y,x,z+1/2"],
    "P4/m": ["x,y,z", "-x,-y,z", "-y,x,z", "y,-x,z", "-x,-y,-z", "x,y,-z", "y,-x,-z", "-y,x,-z"],
}


@dataclass
class Workload:
    name: str
    crystal: Crystal
    params: HC_ModelParameters
    lcs: np.ndarray          # (n, 3, 3)
    hkl: np.ndarray          # (m, 3) int32
    dtdf: np.ndarray         # (m,) complex
    n_explicit_ops: int

    @property
    def terms(self):
        return self.crystal.atoms.n * self.n_explicit_ops * self.hkl.shape[0]


def cell_matrices(cell):
    a, b, c, al, be, ga = cell
    f = np.pi / 180
    ca, cb, cg, sg = np.cos(al * f), np.cos(be * f), np.cos(ga * f), np.sin(ga * f)
    phi = np.sqrt(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg)
    f2c = np.array([[a, b * cg, c * cb], [0, b * sg, c * (ca - cb * cg) / sg], [0, 0, c * phi / sg]])
    return f2c, np.linalg.inv(f2c)


def random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def u_cart_to_cif(u_cart, f2c):
    c2f = np.linalg.inv(f2c)
    rec = np.linalg.norm(c2f, axis=1)          # |a*|, |b*|, |c*|
    B = c2f / rec[:, None]
    return B @ u_cart @ B.T


def unique_hkl(cell, symops, d_min):
    """All hkl with d >= d_min, one representative per orbit of the Laue group (+ Friedel)."""
    f2c, c2f = cell_matrices(cell)
    rots = [parse_symop(s)[0] for s in symops]
    hmax = [int(np.floor(np.linalg.norm(f2c[:, i]) / d_min)) + 1 for i in range(3)]
    g = np.stack(np.meshgrid(*[np.arange(-m, m + 1) for m in hmax], indexing="ij"), -1).reshape(-1, 3)
    hc = g @ c2f                      # Cartesian reciprocal vectors (rows)
    s2 = np.einsum("ij,ij->i", hc, hc)
    g = g[(s2 <= 1.0 / d_min ** 2 + 1e-12) & (s2 > 0)]
    best = g.copy()
    key = lambda v: (v[:, 0] * 1000003 + v[:, 1]) * 1000003 + v[:, 2]
    bk = key(best)
    for R in rots:
        for sgn in (1, -1):
            img = sgn * (g @ R)       # h R  (reflection transforms by R^T acting on row vector)
            ik = key(img)
            upd = ik > bk
            best[upd] = img[upd]
            bk[upd] = ik[upd]
    _, idx = np.unique(bk, return_index=True)
    out = best[np.sort(idx)]
    return np.ascontiguousarray(out, dtype=np.int32)


def _fast_place(rng, n, cell, symops, min_sep=0.9):
    """Vectorised rejection sampling for large structures (same acceptance rule)."""
    f2c, c2f = cell_matrices(cell)
    ops = [parse_symop(s) for s in symops]
    cs = min_sep
    ext = np.array([np.linalg.norm(f2c[:, i]) for i in range(3)])
    ncell = np.maximum(1, np.floor(ext / (1.2 * cs)).astype(int))
    grid = {}
    out = []
    shifts = [(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)]
    misses = 0
    while len(out) < n:
        if misses > 20000:
            raise ValueError("cannot place %d atoms %.2f A apart in cell %s (%d placed)" % (n, min_sep, cell, len(out)))
        misses += 1
        x = rng.uniform(0, 1, 3)
        imgs = np.array([(R @ x + t) % 1.0 for R, t in ops])
        ok = True
        for i in range(len(imgs)):
            for j in range(i + 1, len(imgs)):
                d = imgs[i] - imgs[j]
                d -= np.round(d)
                if np.linalg.norm(f2c @ d) < min_sep:
                    ok = False
        if not ok:
            continue
        cells = [tuple((np.floor(y * ncell)).astype(int) % ncell) for y in imgs]
        for y, c in zip(imgs, cells):
            for s in shifts:
                k = ((c[0] + s[0]) % ncell[0], (c[1] + s[1]) % ncell[1], (c[2] + s[2]) % ncell[2])
                for z in grid.get(k, ()):
                    d = y - z
                    d -= np.round(d)
                    if np.linalg.norm(f2c @ d) < min_sep:
                        ok = False
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        out.append(x)
        misses = 0
        for y, c in zip(imgs, cells):
            grid.setdefault(c, []).append(y)
    return np.array(out)


def build_workload(name, rng, cell, sg, composition, d_min, n_types_per_element=None, lmax_heavy=4,
                   lmax_h=2):
    """composition: list of element symbols per atom (order kept)."""
    symops = SPACE_GROUPS[sg]
    n = len(composition)
    f2c, c2f = cell_matrices(cell)
    frac = _fast_place(rng, n, cell, symops)
    elements = sorted(set(composition), key=lambda e: composition.index(e))
    wfn = [wfn_type_from_bank(e) for e in elements]
    wfn_idx = {e: i for i, e in enumerate(elements)}
    types: List[HC_AtomTypeParameters] = []
    atom_type = np.zeros(n, dtype=int)

    def new_type(el):
        L = lmax_h if el == "H" else lmax_heavy
        plm = [list(rng.uniform(-0.3, 0.3, 2 * l + 1)) for l in range(L + 1)]
        pval = rng.uniform(0.8, 1.2) if el == "H" else rng.uniform(3.5, 6.5)
        return HC_AtomTypeParameters(plm, pval, rng.uniform(0.95, 1.05), rng.uniform(0.8, 1.2))

    if n_types_per_element is None:
        for i, el in enumerate(composition):
            types.append(new_type(el))
            atom_type[i] = i
    else:
        base = {}
        for el in elements:
            base[el] = len(types)
            for _ in range(n_types_per_element[el]):
                types.append(new_type(el))
        for i, el in enumerate(composition):
            atom_type[i] = base[el] + rng.integers(n_types_per_element[el])
    adp = np.zeros((n, 6))
    adp_size = np.zeros(n, dtype=np.int32)
    lcs = np.zeros((n, 3, 3))
    for i, el in enumerate(composition):
        lcs[i] = random_rotation(rng)
        if el == "H":
            adp[i, 0] = rng.uniform(0.02, 0.06)
            adp_size[i] = 1
        else:
            Q = random_rotation(rng)
            uc = Q @ np.diag(rng.uniform(0.01, 0.05, 3)) @ Q.T
            ucif = u_cart_to_cif(uc, f2c)
            adp[i] = [ucif[0, 0], ucif[1, 1], ucif[2, 2], ucif[0, 1], ucif[0, 2], ucif[1, 2]]
            adp_size[i] = 6
    order = len(symops)
    atoms = AtomSet(frac, adp, adp_size, np.ones(n), np.full(n, float(order)), None, list(composition))
    crystal = Crystal(list(cell), symops, atoms)
    params = HC_ModelParameters(wfn, types, [wfn_idx[e] for e in composition], list(atom_type))
    hkl = unique_hkl(cell, symops, d_min)
    dtdf = rng.normal(size=hkl.shape[0]) + 1j * rng.normal(size=hkl.shape[0])
    n_centering = sum(1 for s in symops if np.all(parse_symop(s)[0] == np.eye(3, dtype=int)))
    centro = any(np.all(parse_symop(s)[0] == -np.eye(3, dtype=int)) for s in symops)
    n_exp = order // (n_centering * (2 if centro else 1))
    return Workload(name, crystal, params, lcs, hkl, dtdf, n_exp)


def config0(replicate=0):
    """60 atoms C14 N4 O6 H36, P2_1/c, 10 x 12 x 15 A, beta 100, hkl to 0.5 A, one type per atom."""
    rng = np.random.Generator(np.random.PCG64(0 * 1000 + replicate))
    comp = ["C"] * 14 + ["N"] * 4 + ["O"] * 6 + ["H"] * 36
    return build_workload("cfg0_p21c_60at_0.5A", rng, (10.0, 12.0, 15.0, 90.0, 100.0, 90.0), "P21/c", comp, 0.5)


def protein_composition(n_atoms, rng):
    frac = {"C": 1575 / 5000, "N": 425 / 5000, "O": 475 / 5000, "S": 50 / 5000}
    comp = []
    for el, f in frac.items():
        comp += [el] * int(round(f * n_atoms))
    comp += ["H"] * (n_atoms - len(comp))
    return comp


PROTEIN_TYPES = {"C": 12, "N": 6, "O": 6, "S": 2, "H": 14}


def config1(replicate=0, n_atoms=5000, d_min=1.2):
    """5000 protein-like atoms, P2_1 2_1 2_1, 50 x 60 x 70 A, 40 shared types, hkl to 1.2 A."""
    rng = np.random.Generator(np.random.PCG64(1 * 1000 + replicate))
    comp = protein_composition(n_atoms, rng)
    return build_workload("cfg1_p212121_%dat_%.1fA" % (n_atoms, d_min), rng, (50.0, 60.0, 70.0, 90.0, 90.0, 90.0),
                          "P212121", comp, d_min, PROTEIN_TYPES)


def config2_large(replicate=0):
    """20 000 protein-like atoms, P2_1 2_1 2_1, 80 x 90 x 110 A, hkl to 1.5 A."""
    rng = np.random.Generator(np.random.PCG64(2 * 1000 + replicate))
    comp = protein_composition(20000, rng)
    return build_workload("cfg2_p212121_20000at_1.5A", rng, (80.0, 90.0, 110.0, 90.0, 90.0, 90.0), "P212121",
                          comp, 1.5, PROTEIN_TYPES)


def small_structure(sg, n_atoms=12, d_min=1.0, seed=7, cell=None, lmax_heavy=4):
    """Small test structure in any of SPACE_GROUPS (one type per atom)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if cell is None:
        if sg == "P1":
            cell = (7.0, 8.0, 9.0, 85.0, 95.0, 100.0)
        elif sg == "P-1":
            cell = (7.5, 8.0, 9.5, 92.0, 104.0, 88.0)
        elif sg in ("P21/c", "C2/c"):
            cell = (8.0, 9.0, 10.0, 90.0, 105.0, 90.0)
        elif sg in ("P43212", "P4/m"):
            cell = (8.5, 8.5, 10.0, 90.0, 90.0, 90.0)
        elif sg in ("P31", "P63"):
            cell = (8.5, 8.5, 10.0, 90.0, 90.0, 120.0)
        else:
            cell = (8.0, 9.0, 10.0, 90.0, 90.0, 90.0)
    els = ["C", "N", "O", "S", "H", "H"]
    comp = [els[i % len(els)] for i in range(n_atoms)]
    return build_workload("small_%s" % sg, rng, cell, sg, comp, d_min, None, lmax_heavy)


BATCH_GROUPS = ["P1", "P-1", "P21/c", "C2/c", "P212121"]
_LAUE_ORDER = {"P1": 2, "P-1": 2, "P21/c": 4, "C2/c": 4, "P212121": 8}


def config4(n_structures=256, seed=4000, max_refl=40000, min_refl=2000):
    """Uneven small-molecule structures for batched refinement steps.

    Atoms per asymmetric unit log-uniform in [20, 400] (composition C14 N4 O6 H36 scaled), space
    group uniform over BATCH_GROUPS, cell edges U(6, 25) A (triclinic angles U(80, 100), monoclinic
    beta U(95, 110)), d_min U(0.5, 0.9) A, parameters as in config0 (one type per atom).  Two
    adjustments keep the structures physical and the reflection counts in the 2k-40k range: a cell
    that cannot hold the atoms at 11 A^3 per atom is scaled up isotropically, and d_min is raised
    (lowered, not below 0.5 A) when the expected unique-reflection count exceeds max_refl (falls
    below min_refl)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for b in range(n_structures):
        n = int(round(np.exp(rng.uniform(np.log(20), np.log(400)))))
        sg = BATCH_GROUPS[rng.integers(len(BATCH_GROUPS))]
        edges = rng.uniform(6.0, 25.0, 3)
        if sg in ("P1", "P-1"):
            ang = rng.uniform(80.0, 100.0, 3)
        elif sg in ("P21/c", "C2/c"):
            ang = np.array([90.0, rng.uniform(95.0, 110.0), 90.0])
        else:
            ang = np.array([90.0, 90.0, 90.0])
        cell = tuple(edges) + tuple(ang)
        f2c, _ = cell_matrices(cell)
        vol = abs(np.linalg.det(f2c))
        need = 11.0 * n * len(SPACE_GROUPS[sg])
        if vol < need:
            edges = edges * (need / vol) ** (1.0 / 3.0)
            cell = tuple(edges) + tuple(ang)
            vol = need
        d_min = rng.uniform(0.5, 0.9)
        est = 4.18879 * vol / d_min ** 3 / _LAUE_ORDER[sg]
        if est > max_refl:
            d_min = (4.18879 * vol / (_LAUE_ORDER[sg] * max_refl)) ** (1.0 / 3.0)
        elif est < min_refl:
            d_min = max(0.5, (4.18879 * vol / (_LAUE_ORDER[sg] * min_refl)) ** (1.0 / 3.0))
        comp = []
        for el, k in (("C", 14), ("N", 4), ("O", 6)):
            comp += [el] * int(round(n * k / 60.0))
        comp += ["H"] * max(0, n - len(comp))
        comp = comp[:n]
        sub = np.random.Generator(np.random.PCG64(seed * 1000 + b))
        out.append(build_workload("batch%03d_%s_%dat_%.2fA" % (b, sg.replace("/", ""), n, d_min), sub, cell, sg, comp,
                                  d_min))
    return out
