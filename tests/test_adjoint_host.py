"""The adjoint's host-side explicit derivatives (vphip.adjoint): the hand-derived partials of CL (pressure route) and
CDi (Trefftz) against central differences and complex steps of the same functions, and the objectives against the host loads
(vphip.sweep.wing_loads) and the oracle's."""
import numpy as np

from oracle import oracle as O
from vphip import adjoint as AD, api, dofmap, mesh, sweep


def _wing():
    g = mesh.naca_wing(mesh.WingSpec(code="2412", n_around=16, n_span=6, aspect=6.0, taper=0.5, sweep_c4_deg=25.0))
    dm = dofmap.build_dofmap(g)
    rng = np.random.default_rng(8)
    q = rng.standard_normal(dm.nq) * 0.1
    return g, dm, q


def test_cl_partials():
    g, dm, q = _wing()
    mu = dm.C @ q
    CL, dmu, dx, da = AD.cl_explicit(g, dm, g.x, mu, 4.0, 6.0)
    V, eL = sweep.wing_freestream(4.0)
    ref = sweep.wing_loads(api.build_elements(g), dm, mu, q, V, eL, 6.0)
    assert abs(CL - ref["CL"]) < 1e-14
    rng = np.random.default_rng(1)
    h = 1e-6
    for _ in range(3):
        v = rng.standard_normal(g.x.shape)
        fp = AD.cl_explicit(g, dm, g.x + h * v, mu, 4.0, 6.0)[0]
        fm = AD.cl_explicit(g, dm, g.x - h * v, mu, 4.0, 6.0)[0]
        assert abs((fp - fm) / (2 * h) - np.sum(dx * v)) < 1e-7 * max(1.0, np.abs(dx).sum())
        w = rng.standard_normal(dm.nslot)
        fp = AD.cl_explicit(g, dm, g.x, mu + h * w, 4.0, 6.0)[0]
        fm = AD.cl_explicit(g, dm, g.x, mu - h * w, 4.0, 6.0)[0]
        assert abs((fp - fm) / (2 * h) - dmu @ w) < 1e-7 * max(1.0, np.abs(dmu).sum())
    fp = AD.cl_explicit(g, dm, g.x, mu, 4.0 + 1e-4, 6.0)[0]
    fm = AD.cl_explicit(g, dm, g.x, mu, 4.0 - 1e-4, 6.0)[0]
    assert abs((fp - fm) / 2e-4 - da) < 1e-8


def test_cdi_partials():
    g, dm, q = _wing()
    mw = dm.K @ q
    CD, dmw, dx, da = AD.cdi_explicit(dm, g.x, mw, 4.0, 6.0)
    V, _ = sweep.wing_freestream(4.0)
    assert abs(CD - O.trefftz_drag(dm.wake_te, mw, V) / 6.0) < 1e-14 * max(1.0, abs(CD))
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(3):
        v = rng.standard_normal(g.x.shape)
        fp = AD.cdi_explicit(dm, g.x + h * v, mw, 4.0, 6.0)[0]
        fm = AD.cdi_explicit(dm, g.x - h * v, mw, 4.0, 6.0)[0]
        assert abs((fp - fm) / (2 * h) - np.sum(dx * v)) < 1e-7 * max(1.0, np.abs(dx).sum())
        w = rng.standard_normal(len(mw))
        fp = AD.cdi_explicit(dm, g.x, mw + h * w, 4.0, 6.0)[0]
        fm = AD.cdi_explicit(dm, g.x, mw - h * w, 4.0, 6.0)[0]
        assert abs((fp - fm) / (2 * h) - dmw @ w) < 1e-7 * max(1.0, np.abs(dmw).sum())
    fp = AD.cdi_explicit(dm, g.x, mw, 4.0 + 1e-4, 6.0)[0]
    fm = AD.cdi_explicit(dm, g.x, mw, 4.0 - 1e-4, 6.0)[0]
    assert abs((fp - fm) / 2e-4 - da) < 1e-8


def test_explicit_closed_forms_match_complex_steps():
    """The hand-derived d/dx and d/dmu of CL and CDi against complex steps of the same analytic functions
    (cl_terms, the batched Trefftz sum), one coordinate at a time"""
    g = mesh.naca_wing(mesh.WingSpec(code="2412", n_around=12, n_span=4, aspect=5.0, taper=0.6, sweep_c4_deg=20.0))
    dm = dofmap.build_dofmap(g)
    rng = np.random.default_rng(7)
    mu = rng.standard_normal(dm.nslot) * 0.2
    mw = rng.standard_normal(dm.K.shape[0]) * 0.2
    h = 1e-30
    CL, dmu, dx, _ = AD.cl_explicit(g, dm, g.x, mu, 4.0, 6.0)
    V, eL = sweep.wing_freestream(4.0)
    m = mu[dm.corner_slot]
    for node in rng.choice(len(g.x), 6, replace=False):
        for d in range(3):
            xc = g.x.astype(complex)
            xc[node, d] += 1j * h
            f, *_ = AD.cl_terms(xc[g.tris], m, V, eL, 6.0)
            assert abs(np.sum(f).imag / h - dx[node, d]) < 1e-13 * np.abs(dx).max()
    for slot in rng.choice(dm.nslot, 6, replace=False):
        mc = mu.astype(complex)
        mc[slot] += 1j * h
        f, *_ = AD.cl_terms(g.x[g.tris], mc[dm.corner_slot], V, eL, 6.0)
        assert abs(np.sum(f).imag / h - dmu[slot]) < 1e-13 * np.abs(dmu).max()
    CD, dmw, dxd, _ = AD.cdi_explicit(dm, g.x, mw, 4.0, 6.0)
    wte = dm.wake_te
    topo = AD._trefftz_topology(wte.lines, wte.closed)
    dd = V / np.linalg.norm(V)
    scale = 2 / np.dot(V, V) / 6.0
    P = g.x[wte.node]
    for j in range(len(mw)):
        mc = mw.astype(complex)
        mc[j] += 1j * h
        D = AD._trefftz_batch(P[None].astype(complex), mc[None], topo, dd)[0]
        assert abs(D.imag / h * scale - dmw[j]) < 1e-13 * np.abs(dmw).max()
    for j in range(len(P)):
        for d in range(3):
            Pc = P.astype(complex)
            Pc[j, d] += 1j * h
            D = AD._trefftz_batch(Pc[None], mw[None].astype(complex), topo, dd)[0]
            assert abs(D.imag / h * scale - dxd[wte.node[j], d]) < 1e-12 * np.abs(dxd).max()
