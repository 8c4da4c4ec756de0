"""The discrete adjoint on the GPU (adjoint solve + backward assembly) against FP64 central differences of the CPU
oracle's full solve (assemble + LAPACK + loads) along seeded random node-perturbation directions, h = 1e-6 chord.
The wake's offsets from its trailing edge nodes are held fixed in both.  The adjoint differentiates the quadrature
with its rule pieces at fixed barycentric positions; the one rule whose pieces move with the geometry is the
wide-corner split (assemble.jl:102-125: wedge boundaries from the corner's angle and edges), whose effect on the
differences is the sensitivity of its quadrature error.  With the split off (corner_angle = 180) the adjoint is the
exact derivative: within 2e-7 of the directional derivative's scale (|grad| . |v|); with the default split within
1e-5 (measured 2-5e-6)."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from vphip import adjoint as AD, api, dofmap, mesh, sweep, wake as wk

pytestmark = pytest.mark.gpu

ALPHA, S = 4.0, 6.0


@pytest.fixture(scope="module")
def case():
    g = mesh.naca_wing(mesh.WingSpec(code="2412", n_around=16, n_span=6, aspect=6.0, taper=0.5, sweep_c4_deg=25.0))
    dm = dofmap.build_dofmap(g)
    V, _ = sweep.wing_freestream(ALPHA)
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=1.0))
    return g, dm, w


def oracle_loads(g, dm, w, x, alpha, corner_angle=75.0):
    """CL (pressure, gradient-route Cp) and CDi (Trefftz) of the oracle's solution on nodes x, the wake moved with its
    trailing edge nodes"""
    node_of_wake = dm.wake_te.node[w.te]
    gp = dataclasses.replace(g, x=x)
    wp = dataclasses.replace(w, x=x[node_of_wake] + (w.x - g.x[node_of_wake]))
    V, eL = sweep.wing_freestream(alpha)
    el = O.elements(x, g.tris)
    A, b, _, _ = O.assemble(gp, dm, O.source_strengths(el, V), V, O.opts(corner_angle=corner_angle), wake=wp)
    sol = O.solve_galerkin(dm, A, b, dofmap.wake_rows(dm, w.te))
    cp = O.pressure_coefficient(O.gradient_velocity(el, dm, sol["mu"], V), V)
    F, _ = O.pressure_force(el, cp)
    wte = dataclasses.replace(dm.wake_te, x=x[dm.wake_te.node])
    return F @ eL / S, O.trefftz_drag(wte, dm.K @ sol["q"], V) / S


@pytest.mark.parametrize("objective,corner,tol", [("CL", 180.0, 2e-7), ("CDi", 180.0, 2e-7), ("CL", 75.0, 1e-5)])
def test_adjoint_against_oracle_differences(case, objective, corner, tol):
    print(f"adjoint {objective} corner_angle {corner}", flush=True)
    g, dm, w = case
    adj = AD.WingAdjoint(g, dm, w, ALPHA, S, precision="fp64", opts=api.AssemblyOptions(corner_angle=corner))
    J, grad, dalpha = adj.gradient(objective)
    k = 0 if objective == "CL" else 1
    loads = lambda x, a: oracle_loads(g, dm, w, x, a, corner)[k]
    J0 = loads(g.x, ALPHA)
    assert abs(J - J0) < 1e-10 * max(abs(J0), 1e-3), (J, J0)
    rng = np.random.default_rng(2026)
    h = 1e-6
    worst = 0.0
    for i in range(5):
        v = rng.standard_normal(g.x.shape)
        fd = (loads(g.x + h * v, ALPHA) - loads(g.x - h * v, ALPHA)) / (2 * h)
        ad = float(np.sum(grad * v))
        scale = float(np.sum(np.abs(grad) * np.abs(v)))
        worst = max(worst, abs(fd - ad) / scale)
        print(f"  direction {i}: fd {fd:.10e} adjoint {ad:.10e} rel {abs(fd - ad) / scale:.2e}", flush=True)
    assert worst < tol, worst
    ha = 1e-4
    fd = (loads(g.x, ALPHA + ha) - loads(g.x, ALPHA - ha)) / (2 * ha)
    print(f"  alpha: fd {fd:.10e} adjoint {dalpha:.10e}", flush=True)
    assert abs(fd - dalpha) < 1e-6 * abs(fd) + 1e-9
    # the backward pass's lamᵀ db/dalpha against lamᵀ of a second assembly's b along dV∞/dalpha
    ref = float(adj.last_lam[0] @ adj._db_dalpha())
    assert abs(adj.last_lam_db[0] - ref) < 1e-12 * np.abs(adj.last_lam).sum() * np.abs(adj._db_dalpha()).max(), \
        (adj.last_lam_db, ref)
    adj.close()


def test_adjoint_two_objectives_one_pass(case):
    """CL and CDi through one backward pass (two sets of adjoint weights) equal their separate passes"""
    g, dm, w = case
    print("adjoint CL+CDi in one pass", flush=True)
    adj = AD.WingAdjoint(g, dm, w, ALPHA, S, precision="fp64")
    both = adj.gradients(("CL", "CDi"))
    for o in ("CL", "CDi"):
        J, gx, da = adj.gradient(o)
        Jb, gxb, dab = both[o]
        assert J == Jb
        assert np.abs(gx - gxb).max() < 1e-13 * np.abs(gx).max(), o
        assert abs(da - dab) < 1e-13 * abs(da) + 1e-15, o
    adj.close()


def test_adjoint_mixed_matches_fp64(case):
    print("adjoint mixed vs fp64", flush=True)
    g, dm, w = case
    out = {}
    for p in ("fp64", "mixed"):
        adj = AD.WingAdjoint(g, dm, w, ALPHA, S, precision=p)
        out[p] = adj.gradient("CL")
        adj.close()
    assert abs(out["mixed"][0] - out["fp64"][0]) < 1e-6
    assert np.abs(out["mixed"][1] - out["fp64"][1]).max() < 1e-5 * np.abs(out["fp64"][1]).max()


def test_rrules(case):
    """The rrule forms: the loads' pullback is linear in its cotangents and equals the gradients for (1, 0) and
    (0, 1); the doublets' pullback against central differences of the oracle's doublets (wᵀmu along random
    directions, the wide-corner split off, the DOFMap's slot areas rebuilt at each perturbed geometry).  A random wᵀmu
    sees a uniform doublet, so its pullback carries the zero-mean border rows' derivative (-0.18 of 8.1 here).  The
    differences of two touching entries are not smooth at this mesh (their central differences change by 8e-4
    between h = 1e-6 and 1e-7), which bounds the agreement at a few 1e-6 of the scale."""
    print("rrules: loads and doublets", flush=True)
    g, dm, w = case
    (CL, CDi), back = AD.wing_loads_rrule(g, dm, w, ALPHA, S)
    ref = back.adjoint.gradients(("CL", "CDi"))   # the same forward solution, both objectives in one pass
    assert CL == ref["CL"][0] and CDi == ref["CDi"][0]
    for k, name in enumerate(("CL", "CDi")):
        xb, ab = back((1.0, 0.0) if k == 0 else (0.0, 1.0))
        scale = np.abs(ref[name][1]).max()
        assert np.abs(xb - ref[name][1]).max() < 1e-12 * scale, name
        assert abs(ab - ref[name][2]) < 1e-12 * max(abs(ref[name][2]), 1e-12), name
    xb, ab = back((0.7, -2.5))
    lin = 0.7 * ref["CL"][1] - 2.5 * ref["CDi"][1]
    assert np.abs(xb - lin).max() < 1e-12 * np.abs(lin).max()
    assert abs(ab - (0.7 * ref["CL"][2] - 2.5 * ref["CDi"][2])) < 1e-12 * (abs(ref["CL"][2]) + abs(ref["CDi"][2]))
    back.close()

    opts = api.AssemblyOptions(corner_angle=180.0)
    mu, back = AD.doublets_rrule(g, dm, w, ALPHA, opts=opts)
    rng = np.random.default_rng(77)
    wts = rng.standard_normal(len(mu))
    xb, ab = back(wts)
    back.close()
    node_of_wake = dm.wake_te.node[w.te]

    def oracle_mu(x, alpha):
        gp = dataclasses.replace(g, x=x)
        dmp = dofmap.build_dofmap(gp)   # the slot areas of the zero-mean border rows move with the nodes
        wp = dataclasses.replace(w, x=x[node_of_wake] + (w.x - g.x[node_of_wake]))
        V, _ = sweep.wing_freestream(alpha)
        el = O.elements(x, g.tris)
        A, b, _, _ = O.assemble(gp, dmp, O.source_strengths(el, V), V, O.opts(corner_angle=180.0), wake=wp)
        return O.solve_galerkin(dmp, A, b, dofmap.wake_rows(dmp, w.te))["mu"]

    assert np.abs(mu - oracle_mu(g.x, ALPHA)).max() < 1e-10 * np.abs(mu).max()
    h = 1e-6
    for i in range(2):
        v = rng.standard_normal(g.x.shape)
        fd = wts @ (oracle_mu(g.x + h * v, ALPHA) - oracle_mu(g.x - h * v, ALPHA)) / (2 * h)
        ad = float(np.sum(xb * v))
        scale = float(np.sum(np.abs(xb) * np.abs(v)))
        print(f"  direction {i}: fd {fd:.10e} adjoint {ad:.10e} rel {abs(fd - ad) / scale:.2e}", flush=True)
        assert abs(fd - ad) < 1e-5 * scale   # measured 3e-6: a few touching entries are not smooth here
    ha = 1e-4
    fd = wts @ (oracle_mu(g.x, ALPHA + ha) - oracle_mu(g.x, ALPHA - ha)) / (2 * ha)
    assert abs(fd - ab) < 1e-6 * abs(fd) + 1e-9, (fd, ab)


def test_swept_wing_full_size_mixed_matches_fp64():
    """The benchmark wing at full size (240 x 120 quads, 58,076 triangles, rigid wake), alpha = 4 deg: the mixed path
    (FP32 multipole far field, FP32 LU on MFMA, FP64 refinement) gives CL and CDi within 1e-6 of plain FP64 and their
    node gradients within 1e-5 of the largest component, with one backward pass for both objectives; CL is in the
    range a 25-degree swept AR 10 wing has at 4 degrees (lifting-line estimate 0.5-0.6)"""
    from vphip import sweep
    spec = mesh.WingSpec(code="2412", n_around=240, n_span=120, aspect=10.0, taper=0.5, sweep_c4_deg=25.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    V, _ = sweep.wing_freestream(4.0)
    w = wk.build_wake(g, dm, V)
    area = spec.aspect * spec.mean_chord ** 2
    print(f"case swept wing full size nt={g.nt}", flush=True)
    out = {}
    for p in ("fp64", "mixed"):
        adj = AD.WingAdjoint(g, dm, w, 4.0, area, precision=p)
        out[p] = adj.gradients(("CL", "CDi"))
        adj.close()
        print(f"  {p}: CL {out[p]['CL'][0]:.9f} CDi {out[p]['CDi'][0]:.9f}", flush=True)
    for o in ("CL", "CDi"):
        a, b = out["fp64"][o], out["mixed"][o]
        assert abs(a[0] - b[0]) < 1e-6, (o, a[0], b[0])
        assert np.abs(a[1] - b[1]).max() < 1e-5 * np.abs(a[1]).max(), o
    assert 0.5 < out["fp64"]["CL"][0] < 0.6


def test_near_group_two_sets_one_launch(case, monkeypatch):
    """near_group's backward build for two sets of adjoint weights in one launch (one wave per SIMD, no spills), whose
    lanes hand split words, leaves and leaf corners to each other through LDS, against the per-set launches: the same
    gradients (to FP64 atomic-order rounding) on three runs in a row."""
    g, dm, w = case
    print("adjoint near_group: two sets in one launch vs one per launch", flush=True)
    adj = AD.WingAdjoint(g, dm, w, ALPHA, S, precision="fp64")
    ref = adj.gradients(("CL", "CDi"))
    monkeypatch.setenv("VP_ADJ_NEAR_ONE_LAUNCH", "1")
    for run in range(3):
        got = adj.gradients(("CL", "CDi"))
        for o in ("CL", "CDi"):
            scale = np.abs(ref[o][1]).max()
            assert np.abs(got[o][1] - ref[o][1]).max() < 1e-13 * scale, (run, o)
            assert abs(got[o][2] - ref[o][2]) < 1e-13 * abs(ref[o][2]) + 1e-15, (run, o)
    adj.close()


def _c4_case():
    from vphip import sweep
    spec = mesh.WingSpec(code="2412", n_around=240, n_span=120, aspect=10.0, taper=0.5, sweep_c4_deg=25.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    V, _ = sweep.wing_freestream(4.0)
    return g, dm, wk.build_wake(g, dm, V), spec.aspect * spec.mean_chord ** 2


_C4 = {}


def _c4_grad(corner):
    """The swept wing's case and its FP64 adjoint gradients with the wide-corner split at `corner` degrees (cached)"""
    if "case" not in _C4:
        _C4["case"] = _c4_case()
    if corner not in _C4:
        g, dm, w, area = _C4["case"]
        print(f"case swept wing full size adjoint nt={g.nt} corner_angle={corner}", flush=True)
        adj = AD.WingAdjoint(g, dm, w, 4.0, area, precision="fp64", opts=api.AssemblyOptions(corner_angle=corner))
        _C4[corner] = (adj.gradients(("CL", "CDi")), (adj.plan.pairs, adj.plan.points))
        adj.close()
    return _C4["case"] + _C4[corner]


def _c4_loads(g, dm, w, area, x, alpha, corner=75.0):
    """CL, CDi of the GPU FP64 path on nodes x (the wake moved with its trailing edge nodes) and the tiers' counts"""
    node_of_wake = dm.wake_te.node[w.te]
    gp = dataclasses.replace(g, x=x)
    wp = dataclasses.replace(w, x=x[node_of_wake] + (w.x - g.x[node_of_wake]))
    adj = AD.WingAdjoint(gp, dm, wp, alpha, area, precision="fp64", opts=api.AssemblyOptions(corner_angle=corner))
    ld = adj.loads()
    stats = (adj.plan.pairs, adj.plan.points)
    adj.close()
    return ld["CL"], ld["CDi"], stats


def _c4_directions(corner, ks, tol):
    g, dm, w, area, grads, base = _c4_grad(corner)
    dirs = np.random.default_rng(20261204).standard_normal((32, g.nn, 3))
    h = 1e-6
    worst = 0.0
    for k in ks:
        v = dirs[k] / np.linalg.norm(dirs[k])
        p = _c4_loads(g, dm, w, area, g.x + h * v, 4.0, corner)
        m = _c4_loads(g, dm, w, area, g.x - h * v, 4.0, corner)
        moved = [sum(abs(q[2][0][n] - base[0][n]) for n in base[0]) for q in (p, m)]
        for j, o in enumerate(("CL", "CDi")):
            fd = (p[j] - m[j]) / (2 * h)
            ad = float(np.sum(grads[o][1] * v))
            scale = float(np.sum(np.abs(grads[o][1]) * np.abs(v)))
            worst = max(worst, abs(ad - fd) / scale)
            print(f"  corner {corner:g} direction {k} {o}: adjoint {ad:+.9e} differences {fd:+.9e} "
                  f"rel {abs(ad - fd) / scale:.2e} tier count changes {moved}", flush=True)
            assert abs(ad - fd) < tol * scale, (k, o, ad, fd, scale)
    return worst


@pytest.mark.parametrize("first", [0, 8, 16, 24])
def test_swept_wing_full_size_adjoint_vs_differences(first):
    """The swept wing at full size (58,076 triangles, rigid wake, alpha = 4 deg): the adjoint's dCL/dx and dCDi/dx
    along 32 seeded random unit directions of all 87,120 node coordinates (8 per test) against central differences of
    the GPU FP64 path with h = 1e-6 chord, within 2e-5 of the directional derivative's scale |grad|.|v| (measured
    3e-8 to 1.01e-5).  The error is what rounding in the forward of about 1e-11 of CL gives over 2h: it does not
    shrink with the wide-corner split off, with every far point on the closed form or with extra refinement steps of
    the solves, and it grows at h = 1e-7 (to 2.7e-5) where a difference quotient of rounding noise would.  With 1.3e10 pair evaluations some pair sits within 1e-9 of a tier threshold, so
    the +-h assemblies of every direction move a few pairs between tiers: the counts are reported, not excluded."""
    _c4_directions(75.0, range(first, first + 8), 2e-5)


def test_swept_wing_full_size_lift_slope():
    """dCL/dalpha of the adjoint at full size against central differences in alpha (h = 1e-4 deg), to 1e-6"""
    g, dm, w, area, grads, _ = _c4_grad(75.0)
    ha = 1e-4
    p = _c4_loads(g, dm, w, area, g.x, 4.0 + ha)
    m = _c4_loads(g, dm, w, area, g.x, 4.0 - ha)
    fd = (p[0] - m[0]) / (2 * ha)
    print(f"  dCL/dalpha: adjoint {grads['CL'][2]:.9e} differences {fd:.9e}", flush=True)
    assert abs(grads["CL"][2] - fd) < 1e-6 * abs(fd), (grads["CL"][2], fd)
