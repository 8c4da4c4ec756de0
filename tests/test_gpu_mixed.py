"""GPU assembly over mixed meshes (isoparametric quads and flat triangles, src/assembly/mixed.jl:444-598) through the
C ABI (vp_assemble_mixed) against the FP64 oracle: tier labels and per-tier pair and point counts bit for bit, A and
b within 1e-12 of the largest entry; the cube sphere's solution against the exact flow (the quads' doublet error
0.54 of the split triangles', test/assembly/test_quads.jl:55-80); a triangle-only mesh equals vp_assemble."""
import numpy as np
import pytest

from oracle import oracle as O
from vphip import api, dofmap, geometry, mesh, mixed, solve
from vphip import wake as W
from vphip.exact import EllipsoidFlow

pytestmark = pytest.mark.gpu

VG = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])
OPT_KEYS = ("test", "far_ratio", "far_points", "near_points", "near_ratio", "near_depth", "touch_levels", "touch_ratio",
            "touch_order", "corner_angle", "contact_rel", "kernel_far")


def _compare(g, dm, V, opts=None, w=None, tol=1e-12):
    opts = opts or api.AssemblyOptions()
    m = mixed.mixed_mesh(g, dm)
    print(f"case {m} nslot={dm.nslot} wake={0 if w is None else len(w.x)} opts={opts}", flush=True)
    A, b, st = api.assemble_mixed(g, dm, m, V, opts, wake=w, tiers=True)
    oo = O.opts(**{k: getattr(opts, k) for k in OPT_KEYS})
    Ao, bo, po, qo, to = O.assemble_mixed(g.x, m, dm.nslot, V, touch=mixed.element_touch_pairs(g, m), o=oo, wake=w,
                                          quad_levels=opts.quad_touch_levels, tiers=True)
    assert np.array_equal(st["tiers"], to), f"tier labels differ at {np.argwhere(st['tiers'] != to)[:5]}"
    assert st["pairs"] == po, (st["pairs"], po)
    assert st["points"] == qo, (st["points"], qo)
    scale = np.max(np.abs(Ao))
    ea = np.max(np.abs(A - Ao)) / scale
    eb = np.max(np.abs(b - bo)) / max(np.max(np.abs(bo)), 1e-300)
    print(f"  A {ea:.2e}  b {eb:.2e}  pairs {st['pairs']}", flush=True)
    assert ea < tol, f"A differs by {ea:.3e}"
    assert eb < tol, f"b differs by {eb:.3e}"
    return m, A, b, st


@pytest.mark.parametrize("n", [2, 3])
def test_cube_sphere(n):
    X, F = mesh.cube_sphere_raw(n)
    g = geometry.build_geometry(X, F)
    _compare(g, dofmap.build_dofmap(g), VG)


@pytest.mark.parametrize("test", ["onering", "dual"])
def test_mixed_test_weights(test):
    X, F = mesh.cube_sphere_raw(2)
    g = geometry.build_geometry(X, F)
    _compare(g, dofmap.build_dofmap(g), VG, api.AssemblyOptions(test=test))


def test_wing_quads_caps_and_wake():
    """Quads on the wing, triangles on the tip caps, a rigid wake with tip filaments: every pair kind (quad and
    triangle tests against quad, triangle, wake triangle and filament sources), contact along the trailing edge"""
    X, F = mesh.naca_wing_raw(mesh.WingSpec(n_around=16, n_span=6, aspect=4.0))
    g = geometry.build_geometry(X, F)
    dm = dofmap.build_dofmap(g)
    w = W.build_wake(g, dm, np.array([1.0, 0.0, 0.0]), W.WakeOptions(length=1.0))
    m, _, _, st = _compare(g, dm, np.array([np.cos(np.radians(4)), 0.0, np.sin(np.radians(4))]), w=w)
    assert m.nq > 0 and m.ntri > 0 and st["pairs"]["contact"] > 0


def test_triangle_only_equals_assemble():
    """No quads: the mixed path's matrix is vp_assemble's (to rounding) and its tiers are the triangle path's"""
    g = mesh.icosphere(2)
    dm = dofmap.build_dofmap(g)
    m, A, b, st = _compare(g, dm, VG)
    els = api.build_elements(g)
    A2, b2, st2 = api.assemble(g, dm, els, api.source_strengths(els, VG), VG)
    assert np.max(np.abs(A - A2)) < 1e-13 * np.max(np.abs(A2)) and np.max(np.abs(b - b2)) < 1e-13 * np.max(np.abs(b2))
    assert st["pairs"] == st2["pairs"]


def test_cube_sphere_exact_flow():
    """test/assembly/test_quads.jl:55-80 on the device: 216 quads, the doublet error below 0.7 of the split
    triangles' (both solved on the device), Cp at the quad centres within 0.03 of the exact flow"""
    X, F = mesh.cube_sphere_raw(6)
    g = geometry.build_geometry(X, F)
    dm = dofmap.build_dofmap(g)
    m, Aq, bq, st = _compare(g, dm, VG)
    assert st["pairs"]["self"] == 216 and sum(st["pairs"].values()) == 216 ** 2
    flow = EllipsoidFlow((1.0, 1.0, 1.0), VG)
    els = api.build_elements(g)
    At, bt, _ = api.assemble(g, dm, els, api.source_strengths(els, VG), VG)
    ex = flow.doublet(g.x[dm.slot_node])

    def mu_error(A, b):
        s = solve.solve_galerkin(dm, A, b)
        e = s["mu"] - ex
        e -= np.sum(dm.slot_area * e) / np.sum(dm.slot_area)
        return np.sqrt(np.sum(dm.slot_area * e ** 2) / np.sum(dm.slot_area)) / flow.doublet_scale(), s

    eq, sol = mu_error(Aq, bq)
    et, _ = mu_error(At, bt)
    print(f"mu error quads {eq:.3e} triangles {et:.3e} ratio {eq / et:.3f}", flush=True)
    assert eq < 0.7 * et
    pts = mixed.surface_pressure(m, sol["mu"], VG, m=1)
    assert np.max(np.abs(pts["cp"] - flow.cp(pts["x"]))) < 0.03


def test_degenerate_quad_raises():
    X, F = mesh.cube_sphere_raw(2)
    g = geometry.build_geometry(X, F)
    dm = dofmap.build_dofmap(g)
    m = mixed.mixed_mesh(g, dm)
    q = m.quad_nodes[0]
    m.quad_nodes[0] = [q[0], q[0], q[0], q[0]]
    with pytest.raises(ValueError):
        api.assemble_mixed(g, dm, m, VG)
