"""The adjoint's host-side explicit derivatives (vphip.adjoint): complex-step partials of CL (pressure route) and CDi
(Trefftz) against central differences of the same functions, and the objectives against the host loads
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
