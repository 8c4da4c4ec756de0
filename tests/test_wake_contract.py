"""Rigid wake strips (vphip/wake.py, src/wake/rigid.jl:18-169) by the reference's own checks
(test/wake/test_wake.jl:9-47) on the rectangular NACA 0012 wing at 5 degrees: row count and node/triangle counts, the
20-span length, geometric growth, the trailing edge row on the body's trailing edge nodes, strips along the
freestream, triangles facing wake_dir x tau, clockwise tip filaments, the bisector option, the semi-infinite ending,
the first row's size, rejected options and bodies without a trailing edge; and te_panel's closed far end."""
import numpy as np
import pytest

from vphip import dofmap, mesh, wake as wk

V = np.array([np.cos(np.radians(5)), 0.0, np.sin(np.radians(5))])


@pytest.fixture(scope="module")
def wing():
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=24, n_span=12, aspect=8.0))
    return g, dofmap.build_dofmap(g)


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def tri_normals(x, tris):
    return unit(np.cross(x[tris[:, 1]] - x[tris[:, 0]], x[tris[:, 2]] - x[tris[:, 0]]))


def test_strips(wing):
    g, dm = wing
    w = wk.build_wake(g, dm, V)
    n = dm.wake_te.n
    nrow = len(w.rows) - 1
    assert w.nw == n * (nrow + 1) and len(w.tris) == 2 * (n - 1) * nrow
    assert abs(w.span - 8.0) < 1e-12
    assert w.rows[-2] < 20 * w.span <= w.rows[-1]                       # 20 spans
    assert abs((w.rows[2] - w.rows[1]) / (w.rows[1] - w.rows[0]) - 1.2) < 1e-12   # geometric growth
    r0 = w.row == 0
    assert np.array_equal(w.x[r0], dm.wake_te.x[w.te[r0]])             # on the body's trailing edge nodes
    assert np.abs(unit(w.x[~r0] - dm.wake_te.x[w.te[~r0]]) - V).max() < 1e-12   # along the freestream
    assert np.all(tri_normals(w.x, w.tris)[:, 2] > 0.99)                # facing wake_dir x tau (+z here)
    # node id(k, j) = j n + k (rigid.jl:125)
    assert np.array_equal(w.te, np.tile(np.arange(n), nrow + 1)) and np.array_equal(w.row, np.repeat(np.arange(nrow + 1), n))


def test_tip_filaments(wing):
    g, dm = wing
    w = wk.build_wake(g, dm, V)
    nrow = len(w.rows) - 1
    line = dm.wake_te.lines[0]  # along tau = +y, from the tip at y = -b/2 to the tip at y = +b/2
    assert dm.wake_te.x[line[0], 1] < 0 < dm.wake_te.x[line[-1], 1]
    assert len(w.filaments) == 2 * nrow
    for a, b in w.filaments:  # clockwise about the normal: upstream at y = -b/2, downstream at y = +b/2
        assert w.te[a] == w.te[b] and abs(w.row[a] - w.row[b]) == 1
        assert w.te[a] == (line[-1] if w.row[b] > w.row[a] else line[0])


def test_options(wing):
    g, dm = wing
    w = wk.build_wake(g, dm, V)
    n = dm.wake_te.n
    wb = wk.build_wake(g, dm, V, wk.WakeOptions(direction="bisector"))
    assert np.abs(unit(wb.x[n:2 * n] - wb.x[:n]) - dm.wake_te.bisector).max() < 1e-12
    assert np.abs(unit(wb.x[2 * n:3 * n] - wb.x[n:2 * n]) - V).max() < 1e-12   # then the freestream
    ws = wk.build_wake(g, dm, V, wk.WakeOptions(ending="semi_infinite"))
    assert len(ws.rows) == len(w.rows) + 1 and abs(ws.rows[-1] / (1e3 * w.rows[-1]) - 1) < 1e-12
    assert wk.build_wake(g, dm, V, wk.WakeOptions(first_row=0.045)).rows[1] == 0.045
    assert abs(w.rows[1] - wk.te_panel_size(g)) < 1e-15
    with pytest.raises(ValueError):
        wk.build_wake(g, dm, V, wk.WakeOptions(ending="open"))
    s = mesh.icosphere(1)
    with pytest.raises(ValueError):
        wk.build_wake(s, dofmap.build_dofmap(s), V)


def test_te_panel(wing):
    """te_panel (rigid.jl:162-163): one row of the trailing edge panels' size, closed at its far end by filaments
    running against the trailing edge's direction, plus the tip filaments"""
    g, dm = wing
    p = wk.te_panel(g, dm, V)
    n = dm.wake_te.n
    assert np.array_equal(p.rows, [0.0, wk.te_panel_size(g)]) and p.nw == 2 * n and len(p.tris) == 2 * (n - 1)
    far = [(a, b) for a, b in p.filaments if p.row[a] == 1 and p.row[b] == 1]
    assert len(far) == n - 1 and len(p.filaments) == 2 + n - 1
    for a, b in far:  # from B = line[i+1] back to A = line[i]
        assert p.te[a] == p.te[b] + 1
