"""The DOFMap indexing contract (vphip/dofmap.py) on the generated NACA wings, by the reference's own checks of it
(test/dofmap/test_dofmap.jl:5-92): slots (nodes, then lower trailing edge copies in node order, build.jl:19-48), the
zero-tip constraints (build.jl:112-127), C and K rows, the corner-slot rule, the trailing edge line, a uniform doublet,
slot areas, the border rows, linear reproduction, the cap corners' copy option, the free-tip Kutta condition and the
Galerkin reduction (system.jl:17-50).  The wings are the C3 (rectangular 0012) and C4 (tapered, swept 2412) shapes at
reduced resolution."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from vphip import dofmap, mesh
from vphip.mesh import TE_CAP, TE_LOWER

SPECS = {"rect0012": mesh.WingSpec(code="0012", n_around=24, n_span=10, aspect=8.0),
         "swept2412": mesh.WingSpec(code="2412", n_around=24, n_span=12, aspect=10.0, taper=0.5, sweep_c4_deg=25.0)}


def linear_field(x):
    return 0.3 + 1.7 * x[0] - 0.9 * x[1] + 2.3 * x[2]


def tri_areas(g):
    a, b, c = g.x[g.tris[:, 0]], g.x[g.tris[:, 1]], g.x[g.tris[:, 2]]
    return np.linalg.norm(np.cross(b - a, c - a), axis=1) / 2


def check_row_sums(dm):
    assert np.abs(np.asarray(dm.C.sum(1)).ravel() - 1).max() < 1e-12
    assert dm.K.shape[0] == 0 or np.abs(np.asarray(dm.K.sum(1)).ravel()).max() < 1e-12


@pytest.fixture(scope="module", params=list(SPECS))
def wing(request):
    g = mesh.naca_wing(SPECS[request.param])
    return request.param, g, dofmap.build_dofmap(g)


def test_clean_wing(wing):
    name, g, dm = wing
    te = dm.wake_te
    nte = te.n
    assert nte == SPECS[name].n_span + 1
    # one lower copy per trailing edge node, after the nodes, in node order
    copies = np.nonzero(dm.lower_slot >= 0)[0]
    assert len(copies) == nte and dm.nslot == g.nn + nte
    assert np.array_equal(dm.lower_slot[copies], g.nn + np.arange(nte))
    assert np.array_equal(dm.slot_node[g.nn:], copies) and np.array_equal(dm.slot_node[:g.nn], np.arange(g.nn))
    # by default (tip_kutta zero) the tips' lower copies follow the upper ones; every other slot is free
    tips = sorted(int(dm.lower_slot[te.node[k]]) for k in (0, nte - 1))
    assert [s for s in range(dm.nslot) if dm.constraint[s] == "tip"] == tips and dm.nq == dm.nslot - 2
    C = dm.C.tocsr()
    for s in range(dm.nslot):
        cols, vals = C.indices[C.indptr[s]:C.indptr[s + 1]], C.data[C.indptr[s]:C.indptr[s + 1]]
        want = dm.q_of_slot[dm.slot_node[s] if s in tips else s]
        assert list(cols) == [want] and list(vals) == [1.0]
    check_row_sums(dm)
    assert dm.K.shape == (nte, dm.nq)
    # each Kutta row +1 on the trailing edge node's slot and -1 on its lower copy; at the tips the two cancel
    K = dm.K.tocsr()
    for k in range(nte):
        v = te.node[k]
        row = dict(zip(K.indices[K.indptr[k]:K.indptr[k + 1]], K.data[K.indptr[k]:K.indptr[k + 1]]))
        if k in (0, nte - 1):
            assert not row
        else:
            assert row == {dm.q_of_slot[v]: 1.0, dm.q_of_slot[dm.lower_slot[v]]: -1.0}
    # lower surface corners at trailing edge nodes take the copies, every other corner (caps too) its node
    for t in range(g.nt):
        for k in range(3):
            v, s = g.tris[t, k], dm.corner_slot[t, k]
            assert s == (dm.lower_slot[v] if dm.lower_slot[v] >= 0 and g.corner_side[t, k] == TE_LOWER else v)
    # one open trailing edge line along +y, tangents along the span
    assert te.lines == [list(range(nte))] and te.closed == [False]
    assert np.all(np.diff(te.x[:, 1]) > 0)
    if name == "rect0012":  # the symmetric section: the bisector is the chord line, the tangent the span
        assert np.abs(te.bisector - [1.0, 0, 0]).max() < 1e-12
        assert np.abs(te.tangent - [0, 1.0, 0]).max() < 1e-12
    else:
        assert np.all(te.bisector[:, 0] > 0.9) and np.abs(np.linalg.norm(te.tangent, axis=1) - 1).max() < 1e-12
    # a uniform doublet: uniform on every slot, nothing shed
    assert np.abs(dm.C @ np.ones(dm.nq) - 1).max() < 1e-14
    assert np.linalg.norm(dm.K @ np.ones(dm.nq)) < 1e-12
    assert abs(dm.slot_area.sum() / tri_areas(g).sum() - 1) < 1e-13
    rows = dofmap.bordered_rows(dm)
    assert len(rows) == 1 and np.array_equal(rows[0][1], np.ones(dm.nq))
    assert abs(rows[0][2].sum() / tri_areas(g).sum() - 1) < 1e-13
    assert dofmap.reproduction_error(dm, g, linear_field) < 1e-12


def test_cap_trailing_edge_copy(wing):
    """test_dofmap.jl:54-63: with cap_te_copy lower the cap corners at trailing edge nodes take the lower copy"""
    _, g, dm = wing
    dl = dofmap.build_dofmap(g, cap_te_copy="lower")
    ncap = 0
    for t in range(g.nt):
        for k in range(3):
            if g.corner_side[t, k] != TE_CAP:
                continue
            v = g.tris[t, k]
            ncap += 1
            assert dm.corner_slot[t, k] == v
            assert dl.corner_slot[t, k] == (dm.lower_slot[v] if dm.lower_slot[v] >= 0 else v)
    assert ncap > 0
    with pytest.raises(ValueError):
        dofmap.build_dofmap(g, cap_te_copy="both")


def test_free_tip_kutta(wing):
    """test_dofmap.jl:65-76: tip_kutta free: no constraint, C the identity, every Kutta row +1/-1, tips included"""
    _, g, _ = wing
    df = dofmap.build_dofmap(g, tip_kutta="free")
    assert df.nq == df.nslot and (df.C != sp.identity(df.nslot)).nnz == 0 and "tip" not in df.constraint
    check_row_sums(df)
    K = df.K.tocsr()
    for k in range(df.wake_te.n):
        v = df.wake_te.node[k]
        row = dict(zip(K.indices[K.indptr[k]:K.indptr[k + 1]], K.data[K.indptr[k]:K.indptr[k + 1]]))
        assert row == {df.q_of_slot[v]: 1.0, df.q_of_slot[df.lower_slot[v]]: -1.0}
    assert dofmap.reproduction_error(df, g, linear_field) < 1e-12
    with pytest.raises(ValueError):
        dofmap.build_dofmap(g, tip_kutta="none")


def test_galerkin_reduction(wing):
    """test_dofmap.jl:78-89: W = K[te_of_wake_node], and the bordered system's block is Cᵀ A [C; W]; A without its
    wake columns is rejected"""
    _, g, dm = wing
    te_of_wake_node = [0, 0, 1, dm.wake_te.n - 1]
    W = dofmap.wake_rows(dm, te_of_wake_node)
    assert (W != dm.K[te_of_wake_node, :]).nnz == 0
    rng = np.random.default_rng(1)
    A = sp.random(dm.nslot, dm.nslot + 4, density=0.01, random_state=2).toarray()
    b = rng.random(dm.nslot)
    M, rhs, _ = O.bordered_system(dm, A, b, W)
    Ar = dm.C.T @ A @ sp.vstack([dm.C, W]).toarray()
    assert np.abs(M[:dm.nq, :dm.nq] - Ar).max() < 1e-14 * max(1.0, np.abs(Ar).max())
    assert np.abs(rhs[:dm.nq] - dm.C.T @ b).max() < 1e-14
    with pytest.raises(ValueError):
        O.bordered_system(dm, A, b)
