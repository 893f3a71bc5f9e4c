"""Pin of the oracle's local-coordinate-system restatement (hc_oracle.lcs_matrix) against the
reference's own LocalCoordinateSystemCalculator + CrossProductLcs compiled into oracle/_ref, on
random definitions covering every direction type, axis pair, handedness, symmetry-related reference
atoms and chirality flips."""
import ctypes

import numpy as np

import hc_oracle
from discamb_amd import LcsDefinition

OPS = ["x,y,z", "-x,y+1/2,-z+1/2", "-x,-y,-z", "x,-y+1/2,z+1/2", "x+1,y,z-1"]
KINDS = ["any_orthogonal", "average_direction", "average_position"]


def _ref_lcs(ref_lib, cell, xyz, d):
    enc = lambda refs: ((ctypes.c_int * max(len(refs), 1))(*[r[0] for r in refs]),
                        (ctypes.c_char_p * max(len(refs), 1))(*[r[1].encode() for r in refs]))
    i1, o1 = enc(d.ref1)
    i2, o2 = enc(d.ref2)
    ic, oc = enc(d.chirality)
    m = np.zeros(9)
    c6 = np.ascontiguousarray(cell, dtype=np.float64)
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    P = ctypes.c_void_p
    ref_lib.ref_lcs(P(c6.ctypes.data), ctypes.c_int(len(xyz)), P(x.ctypes.data), ctypes.c_int(d.central),
                    d.central_op.encode(), ctypes.c_int(KINDS.index(d.direction1_type)), ctypes.c_int(len(d.ref1)), i1,
                    o1, ctypes.c_int(KINDS.index(d.direction2_type)), ctypes.c_int(len(d.ref2)), i2, o2,
                    ctypes.c_int(d.coordinate_1), ctypes.c_int(d.coordinate_2), ctypes.c_int(1 if d.is_r else 0),
                    ctypes.c_int(len(d.chirality)), ic, oc, P(m.ctypes.data))
    return m.reshape(3, 3)


def random_definitions(n_atoms, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        c = int(rng.integers(n_atoms))
        others = [i for i in range(n_atoms) if i != c]
        pick = rng.permutation(others)
        half = len(pick) // 2
        pools = [pick[:half], pick[half:]]       # disjoint reference atoms: the two directions differ
        refs = lambda k: [(int(i), OPS[int(rng.integers(len(OPS)))] if rng.random() < 0.3 else "x,y,z")
                          for i in rng.choice(pools[k], int(rng.integers(1, 4)), replace=False)]
        t1 = KINDS[int(rng.integers(1, 3))] if rng.random() < 0.9 else "any_orthogonal"
        t2 = KINDS[int(rng.integers(0, 3))]
        c1 = int(rng.integers(3))
        c2 = int((c1 + rng.integers(1, 3)) % 3)
        ch = [(int(i), "x,y,z") for i in rng.choice(others, 3, replace=False)] if rng.random() < 0.5 else []
        out.append(LcsDefinition(c, t1, refs(0), t2, refs(1), c1, c2, bool(rng.random() < 0.7), ch))
    return out


def test_lcs_restatement_matches_the_reference(ref_lib):
    rng = np.random.default_rng(1)
    cell = [8.0, 9.5, 10.5, 90.0, 104.0, 90.0]
    xyz = rng.uniform(0, 1, (12, 3))
    flips = 0
    for d in random_definitions(12, 400, 2):
        mine = hc_oracle.lcs_matrix(cell, xyz, d)
        ref = _ref_lcs(ref_lib, cell, xyz, d)
        assert np.abs(mine - ref).max() < 1e-12, d
        assert np.abs(mine.T @ mine - np.eye(3)).max() < 1e-12
        flips += int(np.linalg.det(mine) < 0)
    assert flips > 0      # left-handed frames (is_r false or chirality flips) are exercised
