"""Local frames computed by the library from positions (host code, dh_lcs_compute) against the
oracle's restatement (itself pinned to the reference calculator in test_oracle_lcs.py)."""
import numpy as np
import pytest

import hc_oracle
from discamb_amd import LcsDefinition, compute_lcs
from test_oracle_lcs import random_definitions


def test_library_frames_match_the_oracle():
    rng = np.random.default_rng(4)
    cell = [7.5, 8.0, 9.5, 92.0, 104.0, 88.0]
    xyz = rng.uniform(0, 1, (15, 3))
    defs = random_definitions(15, 500, 9)
    L = compute_lcs(cell, xyz, defs)
    for i, d in enumerate(defs):
        assert np.abs(L[i] - hc_oracle.lcs_matrix(cell, xyz, d)).max() < 1e-13


def test_invalid_definitions_are_rejected():
    cell = [8.0, 8.0, 8.0, 90.0, 90.0, 90.0]
    xyz = np.random.default_rng(1).uniform(0, 1, (4, 3))
    bad = [LcsDefinition(0, "average_direction", [(7, "x,y,z")], "average_position", [(1, "x,y,z")]),
           LcsDefinition(0, "average_direction", [(1, "x,y,z")], "average_position", [(2, "x,y,z")], 1, 1),
           LcsDefinition(0, "average_direction", [(1, "x,y,z")], "average_position", [(2, "x,y,z")],
                         chirality=[(1, "x,y,z"), (2, "x,y,z")]),
           LcsDefinition(0, "average_direction", [], "average_position", [(2, "x,y,z")])]
    for d in bad:
        with pytest.raises(RuntimeError):
            compute_lcs(cell, xyz, [d])
