"""The oracle's touching, contact and wake-column rules against an independent adaptive cubature of the pair block
(tests/cubature.py), case by case as the reference's test/assembly/test_contact.jl:9-96: a T-junction both ways, a
point contact, gaps of 1% and 0.1%, wide corners of 150 and 170 degrees, a needle along an edge; and the wake's
columns (src/assembly/assemble.jl:288-321): a body triangle meeting the first wake row along the trailing edge, a tip
filament starting at a test corner, a filament passing near.  Each case also checks that a plausible wrong rule (the
shared corner's rule alone, the near rule, the unsplit wide-corner rule) misses the block, so that the check
discriminates.  The kernel inside the cubature is the oracle's, pinned separately by the golden vectors
(tests/test_oracle_kernels.py)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
import cubature as cub

TOL = 1e-11     # the cubature's absolute error target; the blocks are O(0.01-1)
NE, NV, NS = 1512, 378, 4536   # edge, vertex and self rule lengths at the default TierRules (rules.jl:169-174)

TA, TB, TH = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.5, 0, 0])
TLONG = np.array([TA, TB, [0.45, -0.8, 0.0]])
TSHORT = np.array([TH, TA, [0.2, 0.7 * math.cos(math.radians(20)), 0.7 * math.sin(math.radians(20))]])
ID_LONG, ID_SHORT = (0, 1, 2), (3, 0, 4)   # TA shared, TH a hanging node on TLONG's edge a-b


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def mono(p, q, s):
    return 2 * math.factorial(p) * math.factorial(q) * math.factorial(s) / math.factorial(p + q + s + 2)


def ref_block(T, S, apex=0):
    blk, _, err = cub.block(T, S, O.source_velocity_many, TOL, apex=apex)
    assert err <= TOL
    return blk


def test_contact_rule_pieces():
    """test_contact.jl:29-51: the T-junction's contact rule is the edge rule along a-h plus the corner rule toward h in
    two parts (a 94 degree corner); a partition of unity with exact monomials; the reverse pair takes the edge rule
    unsplit; a point contact takes three corner rules; shared edges, shared corners and distant triangles none."""
    tier, lam, w, _, _ = O.pair_geom(TLONG, TSHORT, ID_LONG, ID_SHORT)
    assert tier == "contact" and len(w) == NE + 2 * NV
    assert abs(w.sum() - 1) < 1e-13
    assert lam.min() >= -1e-15 and np.abs(lam.sum(1) - 1).max() < 1e-13
    for p, q, s in ((1, 0, 0), (2, 1, 0), (1, 1, 1), (3, 0, 2)):
        assert abs(w @ (lam[:, 0] ** p * lam[:, 1] ** q * lam[:, 2] ** s) / mono(p, q, s) - 1) < 1e-12
    tier, lam, w, _, _ = O.pair_geom(TSHORT, TLONG, ID_SHORT, ID_LONG)
    assert tier == "contact" and len(w) == NE
    point_contact = np.array([TH, [0.3, 0.5, 0.1], [0.7, 0.5, 0.1]])
    tier, _, w, _, _ = O.pair_geom(TLONG, point_contact, ID_LONG, (3, 5, 6))
    assert tier == "contact" and len(w) == 3 * NV
    for S, sid in ((np.array([TB, TA, [0.4, 0.7, 0.2]]), (1, 0, 7)), (np.array([TB, [1.6, 0.4, 0.0], [1.2, 0.8, 0.1]]), (1, 8, 9)),
                   (np.array([[0.0, 0, 2], [1.0, 0, 2], [0.5, 1, 2]]), (10, 11, 12))):
        assert O.pair_geom(TLONG, S, ID_LONG, sid, kind="contact") is None
    # and the classification then takes the shared edge, the shared corner and the far or near tier
    assert O.pair_geom(TLONG, np.array([TB, TA, [0.4, 0.7, 0.2]]), ID_LONG, (1, 0, 7))[0] == "edge"
    assert O.pair_geom(TLONG, np.array([TB, [1.6, 0.4, 0.0], [1.2, 0.8, 0.1]]), ID_LONG, (1, 8, 9))[0] == "vertex"


@pytest.mark.parametrize("case", ["long-short", "short-long"])
def test_contact_accuracy_t_junction(case):
    """test_contact.jl:53-60: the contact rule within 1e-6 of the cubature; the shared corner's rule alone misses the
    singular segment by more than 1e-3"""
    T, S, tid, sid = (TLONG, TSHORT, ID_LONG, ID_SHORT) if case == "long-short" else (TSHORT, TLONG, ID_SHORT, ID_LONG)
    ref = ref_block(T, S)
    tier, _, _, blk, _ = O.pair_geom(T, S, tid, sid)
    assert tier == "contact"
    assert rel(blk, ref) < 1e-6, rel(blk, ref)
    k = [i for i in range(3) if tid[i] in sid][0]
    assert rel(O.pair_geom(T, S, tid, sid, kind=f"vertex{k}")[3], ref) > 1e-3


@pytest.mark.parametrize("gap", [0.01, 0.001])
def test_contact_accuracy_gap(gap):
    """test_contact.jl:61-68: S faces T's edge a-b across a slit, sharing no node: the contact rule (graded edge rule)
    within 1e-6; the near rule the distance tiers alone would take misses by more than 1e-4"""
    S = np.array([TB + [0, gap, 0], TA + [0, gap, 0], [0.5, 0.7, 0.3]])
    ref = ref_block(TLONG, S)
    tier, _, _, blk, _ = O.pair_geom(TLONG, S, ID_LONG, (5, 6, 7))
    assert tier == "contact"
    assert rel(blk, ref) < 1e-6, rel(blk, ref)
    assert rel(O.pair_geom(TLONG, S, ID_LONG, (5, 6, 7), kind="near")[3], ref) > 1e-4


def _dir(phi):
    return np.array([math.cos(math.radians(phi)), math.sin(math.radians(phi)), 0.0])


def test_wide_corners():
    """test_contact.jl:71-86: a 150 degree corner shared with S beyond a 20 degree gap takes the vertex rule in two
    equal angles (within 1e-7 of the cubature, where the stored rule misses by more than 1e-4); 90 degree corners keep
    the stored rule; 170 degrees take three pieces.  The classification gives the same rule to the shared corner."""
    T = np.array([TA, _dir(-75), _dir(75)])
    S = np.array([TA, 0.9 * _dir(95) + [0, 0, 0.1], 1.1 * _dir(145) + [0, 0, 0.2]])
    ref = ref_block(T, S)
    _, lam, w, blk, _ = O.pair_geom(T, S, (0, 1, 2), (0, 3, 4), kind="vertex_rule0")
    assert len(w) == 2 * NV and abs(w.sum() - 1) < 1e-13
    assert rel(blk, ref) < 1e-7, rel(blk, ref)
    assert rel(O.pair_geom(T, S, (0, 1, 2), (0, 3, 4), kind="vertex0")[3], ref) > 1e-4
    tier, lam2, _, blk2, _ = O.pair_geom(T, S, (0, 1, 2), (0, 3, 4))
    assert tier == "vertex" and np.array_equal(lam2, lam) and np.array_equal(blk2, blk)
    Tn = np.array([TA, _dir(-45), _dir(45)])
    stored = O.pair_geom(Tn, S, (0, 1, 2), (0, 3, 4), kind="vertex0")[1]
    assert np.array_equal(O.pair_geom(Tn, S, (0, 1, 2), (0, 3, 4), kind="vertex_rule0")[1], stored)
    assert len(O.pair_geom(np.array([TA, _dir(-85), _dir(85)]), S, (0, 1, 2), (0, 3, 4), kind="vertex_rule0")[2]) == 3 * NV


def test_several_contacts_needle():
    """test_contact.jl:88-96: a needle sharing S's corner a and lying along S's edge a-b within the tolerance takes
    the self rule (graded toward all its edges), within 1e-6 of the cubature"""
    S = np.array([TA, TB, [0.5, 0.8, 0.0]])
    T = np.array([TA, [0.3, -0.05, 0.0], [0.6, -0.03, 0.0]])
    ref = ref_block(T, S)
    tier, _, w, blk, _ = O.pair_geom(T, S, (0, 3, 4), (0, 1, 2))
    assert tier == "contact" and len(w) == NS
    assert rel(blk, ref) < 1e-6, rel(blk, ref)


# the trailing edge along y at x = 0; a body triangle upstream with its edge 2 -> 0 on it; the first wake row downstream
TE0, TE1 = np.array([0.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
TBODY = np.array([TE0, [-0.6, 0.45, 0.06], TE1])


def test_wake_sheet_at_trailing_edge():
    """wake_pair_rule (assemble.jl:288-293): a wake triangle shares no node with the body, so the first row, which meets
    the body triangle along its trailing edge, is found by the contact test and takes the edge rule there (within 1e-6
    of the cubature; the near rule misses by more than 1e-4); a wake triangle further downstream is near or far."""
    S = np.array([TE0, [0.7, 0.0, 0.0], TE1])
    ref = ref_block(TBODY, S, apex=2)
    tier, _, w, blk, _ = O.pair_geom(TBODY, S, (0, 1, 2), (10, 11, 12))
    assert tier == "contact" and len(w) == NE
    assert rel(blk, ref) < 1e-6, rel(blk, ref)
    assert rel(O.pair_geom(TBODY, S, (0, 1, 2), (10, 11, 12), kind="near")[3], ref) > 1e-4
    S2 = S + [0.9, 0, 0]
    ref2 = ref_block(TBODY, S2)
    tier2, _, _, blk2, _ = O.pair_geom(TBODY, S2, (0, 1, 2), (10, 11, 12))
    assert tier2 == "near" and rel(blk2, ref2) < 1e-6, rel(blk2, ref2)
    assert O.pair_geom(TBODY, S + [6.0, 0, 0], (0, 1, 2), (10, 11, 12))[0] == "far"


def test_tip_filament_columns():
    """segment_rule and segment_block (assemble.jl:299-321): a tip filament starting at a test corner takes that
    corner's vertex rule (its 1/r singularity there is integrable) and matches the cubature within 1e-6, where the
    near rule's points would not; one passing near the triangle takes the near rule (within its own quadrature error);
    one far away the far rule."""
    fa, fb = TE1, TE1 + [4.0, 0.0, 0.0]
    ca_ref, cb_ref, err = cub.segment_rows(TBODY, fa, fb, O.segment_velocity_many, TOL, apex=2)
    assert err <= TOL
    tier, n, ca, cb = O.segment_geom(TBODY, fa, fb)
    assert tier == "vertex" and n == NV
    ref = np.concatenate([ca_ref, cb_ref])
    assert rel(np.concatenate([ca, cb]), ref) < 1e-6, rel(np.concatenate([ca, cb]), ref)
    f2a, f2b = TE1 + [0.05, 0.1, 0.02], TE1 + [4.0, 0.1, 0.02]
    ca_ref, cb_ref, err = cub.segment_rows(TBODY, f2a, f2b, O.segment_velocity_many, TOL)
    tier, n, ca, cb = O.segment_geom(TBODY, f2a, f2b)
    # the near rule's own quadrature error at 0.05 from the corner is 2.4e-6; three far points miss by far more
    ref2 = np.concatenate([ca_ref, cb_ref])
    assert tier == "near" and rel(np.concatenate([ca, cb]), ref2) < 1e-5
    o = O.opts(far_ratio=0.0)  # every filament takes the far rule
    assert rel(np.concatenate(O.segment_geom(TBODY, f2a, f2b, o)[2:]), ref2) > 1e-3
    assert O.segment_geom(TBODY, TE1 + [9.0, 0, 0], TE1 + [12.0, 0, 0])[0] == "far"
