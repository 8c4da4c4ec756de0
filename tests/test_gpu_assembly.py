"""GPU assembly through the C ABI against the FP64 oracle: matrix entries to 1e-12 of the largest entry, the
right hand side likewise, and tier labels and counts bit for bit (src/assembly/assemble.jl:260-447)."""
import numpy as np
import pytest

from oracle import oracle as O
from vphip import api, dofmap, geometry, mesh, solve, wake as wk
from vphip.exact import EllipsoidFlow

pytestmark = pytest.mark.gpu

VG = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])


def _compare(g, dm, V, opts=None, w=None, touch=None, tol=1e-12):
    opts = opts or api.AssemblyOptions()
    els = api.build_elements(g)
    sig = api.source_strengths(els, V)
    oo = O.opts(**{k: getattr(opts, k) for k in ("test", "far_ratio", "near_ratio", "near_depth", "touch_levels",
                                                  "touch_ratio", "touch_order", "corner_angle", "contact_rel",
                                                  "kernel_far")})
    print(f"case nt={g.nt} nslot={dm.nslot} wake={0 if w is None else w.nw} opts={opts}", flush=True)
    plan = api.AssemblyPlan(g, dm, w, opts, record_tiers=True)
    plan.assemble(sig, V)
    A, b = plan.read()
    tiers = plan.tiers()
    Ao, bo, po, qo, to = O.assemble(g, dm, sig, V, oo, wake=w, tiers=True)
    scale = np.max(np.abs(Ao))
    ea = np.max(np.abs(A - Ao)) / scale
    eb = np.max(np.abs(b - bo)) / max(np.max(np.abs(bo)), 1e-300)
    assert np.array_equal(tiers, to), f"tier labels differ at {np.argwhere(tiers != to)[:5]}"
    # pairs: the oracle counts filament pairs too
    assert plan.pairs == po, (plan.pairs, po)
    assert plan.points == qo, (plan.points, qo)
    assert ea < tol, f"A differs by {ea:.3e}"
    assert eb < tol, f"b differs by {eb:.3e}"
    plan.close()
    return ea, eb


@pytest.mark.parametrize("level", [1, 2, 3])
def test_icosphere(level):
    g = mesh.icosphere(level)
    _compare(g, dofmap.build_dofmap(g), VG)


@pytest.mark.parametrize("test", ["onering", "dual"])
def test_test_weights(test):
    g = mesh.icosphere(2)
    _compare(g, dofmap.build_dofmap(g), VG, api.AssemblyOptions(test=test))


def test_abused_ellipsoid_contact_tiers():
    g = mesh.abused_ellipsoid(nth=12, nph=24, sliver_ratio=0.02, seed=5)
    dm = dofmap.build_dofmap(g)
    _compare(g, dm, np.array([np.cos(np.radians(5)), 0, np.sin(np.radians(5))]))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_far_point_closed_form(precision):
    """kernel_far = 1e300 puts every far-rule point on source_velocity's closed form: the mixed path's far kernel then
    queues every far point of every tile (its deferral queue full on the all-far tiles) and evaluates them after its
    multipole pass.  Against the oracle with the same option: the FP64 bar, or the mixed path's (its FP32 near
    kernels, 4e-7 of the largest entry on the full-size ellipsoid)."""
    g = mesh.abused_ellipsoid(nth=24, nph=48, seed=7)
    dm = dofmap.build_dofmap(g)
    _compare(g, dm, VG, api.AssemblyOptions(precision=precision, kernel_far=1e300),
             tol=1e-12 if precision == "fp64" else 2e-6)


def test_wing_with_wake():
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=16, n_span=6, aspect=4.0))
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(np.radians(5)), 0, np.sin(np.radians(5))])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
    _compare(g, dm, V, w=w)


def test_pipeline_wing_with_wake():
    """A swept, tapered wing built by the geometry pipeline from a raw quad mesh (split trailing edge welded, tips
    capped by the ladder, diagonals by the default rule), with its wake"""
    spec = mesh.WingSpec(code="2412", n_around=16, n_span=7, aspect=5.0, taper=0.6, sweep_c4_deg=20.0)
    g = geometry.build_geometry(*mesh.naca_wing_raw(spec, split_te=True))
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(np.radians(5)), 0, np.sin(np.radians(5))])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
    _compare(g, dm, V, w=w)


def test_pipeline_notched_trailing_edge():
    """A prism whose section has a V notch in its trailing edge, through the geometry pipeline: two Kutta chains (the
    notch tips), two wake sheets, caps with the notch in them"""
    sec = [(0.0, 0.0), (0.5, 0.1), (1.0, 0.05), (0.7, 0.0), (1.0, -0.05), (0.5, -0.1)]
    n, m = 6, len(sec)
    X = np.array([(px, -1.5 + 3 * j / n, pz) for j in range(n + 1) for px, pz in sec])
    faces = [[m * j + k, m * j + (k + 1) % m, m * (j + 1) + (k + 1) % m, m * (j + 1) + k]
             for j in range(n) for k in range(m)]
    g = geometry.build_geometry(X, faces)
    assert len(g.chains) == 2
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(np.radians(3)), 0, np.sin(np.radians(3))])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=1.5))
    _compare(g, dm, V, w=w)


def test_nonfar_list_regrowth(monkeypatch):
    """Work lists budgeted too small: the far kernels' enumerate-only pass refills the grown non-far list without
    adding the far blocks twice, and the rule generator refills the grown rule-piece lists without counting its
    pairs twice (same entries, tiers and counts)"""
    monkeypatch.setenv("VP_NONFAR_CAP", "100")
    monkeypatch.setenv("VP_LEAF_CAP", "3")
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=16, n_span=6, aspect=4.0))
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(np.radians(5)), 0, np.sin(np.radians(5))])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
    _compare(g, dm, V, w=w)


def test_options_deeper_rules():
    g = mesh.icosphere(1)
    _compare(g, dofmap.build_dofmap(g), VG, api.AssemblyOptions(far_ratio=8.0, near_ratio=4.0, near_depth=5,
                                                                 touch_levels=8, touch_order=4))


def test_bad_option_raises():
    g = mesh.icosphere(1)
    dm = dofmap.build_dofmap(g)
    els = api.build_elements(g)
    with pytest.raises(ValueError):
        api.assemble(g, dm, els, api.source_strengths(els, VG), VG, opts=api.AssemblyOptions(test="box"))


def test_launch_autotuner():
    """The assembly's launch autotuner: a plan large enough to time runs each occupancy candidate of the forward
    evaluation kernels on successive assemblies, every assembly's A agrees with the first (same arithmetic, other
    register budget) and the oracle's, and the report names a chosen cap for every kernel afterwards."""
    import json
    print("launch autotuner on an abused ellipsoid 36 x 72", flush=True)
    g = mesh.abused_ellipsoid(nth=36, nph=72, axes=(2.0, 1.0, 0.5), jitter=0.25, sliver_frac=0.05, sliver_ratio=0.1,
                              seed=5)
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(0.087), 0.0, np.sin(0.087)])
    sig = api.source_strengths(api.build_elements(g), V)
    p = api.AssemblyPlan(g, dm, None, api.AssemblyOptions())
    ref = None
    for k in range(4):
        p.assemble(sig, V)
        A, b = p.read()
        if ref is None:
            ref = (A, b)
        else:
            assert np.abs(A - ref[0]).max() <= 1e-13 * np.abs(ref[0]).max(), k
            assert np.abs(b - ref[1]).max() <= 1e-13 * np.abs(ref[1]).max(), k
    p.close()
    rep = json.loads(api.default_context().lib.vp_assembly_tuning().decode())
    for name in ("touch_simple", "near_points", "near_group", "group_eval"):
        assert rep[name]["chosen_waves"] in (2, 3), (name, rep[name])
    Ao, bo, _, _ = O.assemble(g, dm, sig, V)
    assert np.abs(ref[0] - Ao).max() < 1e-12 * np.abs(Ao).max()


def test_abused_ellipsoid_full_size_tiers():
    """The tier label of every one of the 3.95e8 pairs of the full-size abused ellipsoid (19,880 triangles, 5% slivers)
    bit-exact against the oracle's classification (pair_rule, assemble.jl:260-282, without the pair blocks), and the
    pairs and rule points per tier equal"""
    g = mesh.abused_ellipsoid(nth=71, nph=142, axes=(2.0, 1.0, 0.5), jitter=0.25, sliver_frac=0.05, sliver_ratio=0.1,
                              seed=20261203)
    dm = dofmap.build_dofmap(g)
    a = np.radians(5.0)
    V = np.array([np.cos(a), 0.0, np.sin(a)])
    print(f"case abused ellipsoid full size tiers nt={g.nt}", flush=True)
    plan = api.AssemblyPlan(g, dm, None, api.AssemblyOptions(), record_tiers=True)
    plan.assemble(api.source_strengths(api.build_elements(g), V), V)
    tiers = plan.tiers()
    pairs, points = plan.pairs, plan.points
    plan.close()
    print("  oracle classification", flush=True)
    to, po, qo = O.classify(g)
    assert pairs == po and points == qo, (pairs, po, points, qo)
    diff = np.argwhere(tiers != to)
    assert len(diff) == 0, f"{len(diff)} tier labels differ, first at {diff[:5]}"


def test_abused_ellipsoid_full_size():
    """The abused ellipsoid at its full benchmark size (2:1:0.5, uv 71 x 142: 19,880 triangles, 5% slivers, jitter
    0.25): every pair is classified once (tier pair counts sum to nt²); the rows of A and b of every slot whose
    triangles all lie among test triangles 0..1135 equal the oracle's assembly of those test rows (all 9,942 columns,
    1e-12 of the largest entry); the mixed assembly (FP32 multipole far field and FP32 points of the split near pairs)
    stays within 2e-6 of the largest entry of FP64 everywhere (measured 4.0e-7); both
    bordered solves give the exact potential flow's doublet (mean removed) to 1e-3 of its scale in the area-weighted
    L2 norm (the discretization error, measured 2.8e-4) and agree with each other to 1e-6"""
    g = mesh.abused_ellipsoid(nth=71, nph=142, axes=(2.0, 1.0, 0.5), jitter=0.25, sliver_frac=0.05, sliver_ratio=0.1,
                              seed=20261203)
    dm = dofmap.build_dofmap(g)
    a = np.radians(5.0)
    V = np.array([np.cos(a), 0.0, np.sin(a)])
    print(f"case abused ellipsoid full size nt={g.nt} nslot={dm.nslot}", flush=True)
    assert g.nt == 19880
    els = api.build_elements(g)
    sig = api.source_strengths(els, V)
    plan = api.AssemblyPlan(g, dm, None, api.AssemblyOptions())
    plan.assemble(sig, V)
    A, b = plan.read()
    assert sum(plan.pairs.values()) == g.nt * g.nt
    plan.close()
    k = 1136
    _, Ao, bo = O.assemble_rows(g, dm, O.source_strengths(O.elements(g.x, g.tris), V), V, 0, k, full=True)
    cs = np.asarray(dm.corner_slot)
    complete = np.zeros(dm.nslot, bool)
    complete[cs[:k].ravel()] = True
    complete[cs[k:].ravel()] = False
    rows = np.flatnonzero(complete)
    print(f"  {len(rows)} complete slots from test triangles 0..{k - 1}", flush=True)
    assert len(rows) >= 300
    scale = np.abs(A).max()
    ea = np.abs(A[rows] - Ao[rows]).max() / scale
    eb = np.abs(b[rows] - bo[rows]).max() / np.abs(b).max()
    print(f"  sampled rows: A {ea:.2e}, b {eb:.2e}", flush=True)
    assert ea < 1e-12 and eb < 1e-12, (ea, eb)
    ef = EllipsoidFlow((2.0, 1.0, 0.5), V)
    mu_ex, mscale = ef.doublet(g.x[dm.slot_node]), ef.doublet_scale()
    mus = {}
    for m in ("fp64", "mixed"):
        plan = api.AssemblyPlan(g, dm, None, api.AssemblyOptions(precision=m))
        plan.assemble(sig, V)
        if m == "mixed":
            Am, bm = plan.read()
            em = np.abs(Am - A).max() / scale
            print(f"  mixed vs FP64 entries: {em:.2e}", flush=True)
            assert em < 2e-6, em
        F = solve.factorize_galerkin(dm, plan, mode=m)
        mus[m] = solve.solve_galerkin(F, bm if m == "mixed" else b)["mu"]
        F.close()
        plan.close()
        err = mus[m] - mu_ex
        err -= np.sum(dm.slot_area * err) / np.sum(dm.slot_area)
        l2 = np.sqrt(np.sum(dm.slot_area * err ** 2) / np.sum(dm.slot_area)) / mscale
        print(f"  {m} solve: mu L2 error against the exact flow {l2:.2e}", flush=True)
        assert l2 < 1e-3, (m, l2)
    dmu = np.abs(mus["mixed"] - mus["fp64"]).max() / np.abs(mus["fp64"]).max()
    print(f"  mixed vs FP64 mu: {dmu:.2e}", flush=True)
    assert dmu < 1e-6, dmu


def test_swept_wing_full_size():
    """The swept tapered NACA 2412 wing at its full benchmark size (240 x 120 quads: 58,076 triangles; rigid wake of
    16,560 triangles and 138 filaments): every pair is classified once (Σ pairs = nt (nt + wake triangles +
    filaments)); the rows of A (all 29,161 slot and 8,470 wake-node columns) and b of the 362 slots whose triangles
    all lie among test triangles 0..1199 equal the oracle's assembly of those rows to 1e-12 of the largest entry
    (measured 6e-13; the device's own log and atan2 and its order of FP64 sums differ from the oracle's); the mixed
    assembly's rows there stay within 2e-6 of FP64's largest entry (measured 7.5e-7)"""
    from vphip import sweep
    spec = mesh.WingSpec(code="2412", n_around=240, n_span=120, aspect=10.0, taper=0.5, sweep_c4_deg=25.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    V, _ = sweep.wing_freestream(4.0)
    w = wk.build_wake(g, dm, V)
    print(f"case swept wing full size nt={g.nt} nslot={dm.nslot} wake nodes={len(w.x)}", flush=True)
    k = 1200
    cs = np.asarray(dm.corner_slot)
    complete = np.zeros(dm.nslot, bool)
    complete[cs[:k].ravel()] = True
    complete[cs[k:].ravel()] = False
    rows = np.flatnonzero(complete)
    assert len(rows) >= 300
    sig = api.source_strengths(api.build_elements(g), V)
    got = {}
    for m in ("fp64", "mixed"):
        plan = api.AssemblyPlan(g, dm, w, api.AssemblyOptions(precision=m))
        plan.assemble(sig, V)
        A, b = plan.read()
        if m == "fp64":
            assert sum(plan.pairs.values()) == g.nt * (g.nt + len(w.tris) + len(w.filaments))
            scale = np.abs(A).max()
        got[m] = (np.array(A[rows]), b[rows].copy())
        del A, b
        plan.close()
    _, Ao, bo = O.assemble_rows(g, dm, O.source_strengths(O.elements(g.x, g.tris), V), V, 0, k, wake=w, full=True)
    Ao, bo = np.array(Ao[rows]), bo[rows]
    ea = np.abs(got["fp64"][0] - Ao).max() / scale
    eb = np.abs(got["fp64"][1] - bo).max() / np.abs(bo).max()
    em = np.abs(got["mixed"][0] - got["fp64"][0]).max() / scale
    print(f"  {len(rows)} complete slots: A {ea:.2e}, b {eb:.2e}; mixed vs FP64 {em:.2e}", flush=True)
    assert ea < 1e-12 and eb < 1e-12, (ea, eb)
    assert em < 2e-6, em
