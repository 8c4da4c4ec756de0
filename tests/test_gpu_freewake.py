"""The free wake on the GPU against the same host logic on the FP64 oracle: explicit_rhs of sheets, filaments and
particles (near and far) within 1e-12 of the largest entry, and the simulation step by step (trailing edge mu_w,
Kutta-Joukowski force, particle positions) within 1e-10 in FP64 over steps with conversions; the mixed path's mu_w
within 1e-6."""
import types

import numpy as np
import pytest

from oracle.freewake_oracle import OracleFreeWakeSim, oracle_explicit_rhs
from vphip import api, dofmap, freewake as FW, mesh, sweep, wake as wk

pytestmark = pytest.mark.gpu

V, LIFT = sweep.wing_freestream(8.0)


@pytest.fixture(scope="module")
def wing():
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=16, n_span=8, aspect=6.0))
    return g, dofmap.build_dofmap(g)


def test_explicit_rhs(wing):
    print("explicit_rhs sheets+filaments+particles", flush=True)
    g, dm = wing
    rng = np.random.default_rng(4)
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=0.5))
    e = FW.rigid_elements(w, rng.standard_normal(dm.wake_te.n))
    # particles: a cloud behind the wing and a few right next to it (near rules)
    nb = 700
    bx = np.vstack([np.array([1.5, 0, 0]) + rng.standard_normal((nb - 20, 3)) * [0.5, 2.0, 0.3],
                    g.x[rng.integers(0, len(g.x), 20)] + rng.standard_normal((20, 3)) * 0.02])
    blobs = FW.Blobs(bx, rng.standard_normal((nb, 3)) * 0.01, rng.uniform(0.02, 0.08, nb))
    got = FW.explicit_rhs(g, dm, e, blobs)
    ref = oracle_explicit_rhs(g, dm, e, blobs, api.AssemblyOptions())
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), np.abs(got - ref).max() / np.abs(ref).max()
    only = FW.explicit_rhs(g, dm, None, blobs)
    assert np.abs(only - oracle_explicit_rhs(g, dm, None, blobs, api.AssemblyOptions())).max() < 1e-12 * np.abs(only).max()


def test_explicit_rhs_moved_sheets(wing):
    """explicit_rhs through one vp_explicit handle (the C ABI the Julia binding calls): known sheets of one topology
    moved twice (the handle's plan moved in place) against the oracle at each position, then a different topology (a
    new plan), then particles alone and with the sheets (the body plan, made once)."""
    print("explicit_rhs handle, moved sheets", flush=True)
    g, dm = wing
    rng = np.random.default_rng(6)
    h = FW.ExplicitRHS(g, dm)
    for k, length in enumerate((0.5, 0.5, 0.5, 0.8)):
        w = wk.build_wake(g, dm, V, wk.WakeOptions(length=length))
        if k in (1, 2):   # same topology, nodes moved downstream and sheared
            w.x = w.x + np.array([0.05 * k, 0.0, 0.02 * k]) + 0.01 * k * np.outer(w.x[:, 1], [0.0, 0.0, 1.0])
        e = FW.rigid_elements(w, rng.standard_normal(dm.wake_te.n))
        got = h(e, None)
        ref = oracle_explicit_rhs(g, dm, e, None, api.AssemblyOptions())
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), (k, np.abs(got - ref).max() / np.abs(ref).max())
    assert h.info() == (2, 2, 4)
    n = 300
    blobs = FW.Blobs(np.column_stack([rng.uniform(1.0, 2.0, n), rng.uniform(-3.0, 3.0, n), rng.normal(0.0, 0.1, n)]),
                     1e-3 * rng.standard_normal((n, 3)), np.full(n, 0.05))
    for ee in (None, e):
        got = h(ee, blobs)
        ref = oracle_explicit_rhs(g, dm, ee, blobs, api.AssemblyOptions())
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    assert h.info() == (2, 3, 6)
    assert np.array_equal(h(None, None), np.zeros(dm.nslot))
    # the body as a triangle soup (its elements' own corners, as the Julia binding passes `els`): the same values to
    # rounding (the atomic sums' order differs; measured 1.3e-15)
    soup = types.SimpleNamespace(x=g.x[g.tris].reshape(-1, 3), tris=np.arange(3 * len(g.tris)).reshape(-1, 3),
                                 touch_pairs=[])
    hs = FW.ExplicitRHS(soup, dm)
    ref = h(e, blobs)
    assert np.abs(hs(e, blobs) - ref).max() <= 1e-13 * np.abs(ref).max()
    hs.close()
    h.close()
    with pytest.raises(ValueError):
        p = api.AssemblyPlan(g, dm, w, api.AssemblyOptions(), wake_only=True)
        try:
            p.update_wake(np.zeros((p.nw + 1, 3)))
        finally:
            p.close()


def test_free_wake_steps_match_oracle(wing):
    print("free wake 8 steps fp64 vs oracle", flush=True)
    g, dm = wing
    dt = wk.te_panel_size(g) * 2
    o = FW.FreeWakeOptions(dt=dt, max_rows=3)
    sim = FW.FreeWakeSim(g, dm, V, o)
    ref = OracleFreeWakeSim(g, dm, V, o)
    for k in range(8):
        a, b = sim.step(), ref.step()
        sc = np.abs(b["mu_w"]).max()
        assert np.abs(a["mu_w"] - b["mu_w"]).max() < 1e-10 * sc, (k, np.abs(a["mu_w"] - b["mu_w"]).max() / sc)
        assert np.abs(a["F_kj"] - b["F_kj"]).max() < 1e-10 * np.abs(b["F_kj"]).max()
        assert a["particles"] == b["particles"] and a["rows"] == b["rows"]
    assert sim.state.conversions > 0
    dx = np.abs(sim.state.particles.x - ref.state.particles.x).max()
    assert dx < 1e-10 * np.abs(ref.state.particles.x).max(), dx
    for r1, r2 in zip(sim.state.x, ref.state.x):
        assert np.abs(r1 - r2).max() < 1e-10 * np.abs(r2).max()
    sim.close()


@pytest.mark.parametrize("variant", ["merge_delete_scaled", "auto_freestream", "mixed"])
def test_native_engine_matches_host_mirror(wing, variant):
    """The native step (vp_freewake_*: rows, conversions with strip merging, loads, advection and deletion in C++)
    against the host mirror of step! (FreeWakeHost) over the same device calls, twelve steps: every record field,
    the rows, the particles with their merge tags and groups, mu and Cp, to 1e-11 of each quantity's scale (the total
    vorticity and the border multiplier, zero to rounding, to 1e-12 and 1e-15 absolute; only the host sums' order differs); the first step's impulse force is
    NaN in both."""
    print(f"free wake native engine vs host mirror {variant}", flush=True)
    g, dm = wing
    dt = wk.te_panel_size(g) * 2
    prec = "mixed" if variant == "mixed" else "fp64"
    if variant == "merge_delete_scaled":
        o = FW.FreeWakeOptions(dt=dt, max_rows=2, merge_rows=2, core="scaled", delete_beyond=0.04)
    elif variant == "auto_freestream":
        o = FW.FreeWakeOptions(dt=dt, max_rows=3, core="auto", advection="freestream")
    else:
        o = FW.FreeWakeOptions(dt=dt, max_rows=2, merge_rows="auto")
    sim = FW.FreeWakeSim(g, dm, V, o, precision=prec)
    ref = FW.FreeWakeHost(g, dm, V, o, precision=prec)
    for k in range(12):
        a, b = sim.step(), ref.step()
        assert a["n"] == b["n"] and a["rows"] == b["rows"] and a["particles"] == b["particles"], (k, a, b)
        assert a["t"] == pytest.approx(b["t"], rel=1e-15)
        for key in ("mu_w", "F_pressure", "M_pressure", "F_kj", "F_impulse", "Omega", "I"):
            x, y = np.asarray(a[key]), np.asarray(b[key])
            if k == 0 and key == "F_impulse":
                assert np.all(np.isnan(x)) and np.all(np.isnan(y))
                continue
            # the total vorticity is zero to rounding (Kelvin): compared in absolute terms
            tol = 1e-12 if key == "Omega" else 1e-11 * max(np.abs(y).max(), 1e-300)
            assert np.abs(x - y).max() <= tol, (k, key, x, y)
        assert abs(a["umax"] - b["umax"]) <= 1e-11 * b["umax"], (k, a["umax"], b["umax"])
        # the border multiplier is zero but for rounding (3e-8 here): compared in absolute terms
        assert abs(a["lam"] - b["lam"]) <= 1e-15, (k, a["lam"], b["lam"])
    sa, sb = sim.state, ref.state
    assert sa.conversions == sb.conversions > 0 and len(sa.x) == len(sb.x)
    for r1, r2, m1, m2 in zip(sa.x, sb.x, sa.mu, sb.mu):
        assert np.abs(r1 - r2).max() <= 1e-11 * np.abs(r2).max()
        assert np.abs(m1 - m2).max() <= 1e-11 * max(np.abs(m2).max(), 1e-300)
    assert np.array_equal(sa.tag, sb.tag) and np.array_equal(sa.group, sb.group)
    assert np.array_equal(sa.batch_rows, sb.batch_rows)
    for f in ("x", "alpha", "delta"):
        u, v = getattr(sa.particles, f), getattr(sb.particles, f)
        assert np.abs(u - v).max() <= 1e-11 * np.abs(v).max(), f
    assert np.abs(sim.mu - ref.mu).max() <= 1e-11 * np.abs(ref.mu).max()
    assert np.abs(sim.cp - ref.cp).max() <= 1e-11 * np.abs(ref.cp).max()
    if variant == "merge_delete_scaled":  # merges made, and particles beyond the limit deleted
        assert sb.conversions >= 4
        d = V / np.linalg.norm(V)
        lim = max(float(x @ d) for x in ref.edge) + o.delete_beyond * ref.te.span
        print(f"  particles {len(sb.particles)}, furthest {np.max(sb.particles.x @ d) - lim:+.3e} from the limit",
              flush=True)
    sim.close()
    ref.close()


def test_free_wake_mixed(wing):
    print("free wake 6 steps mixed vs fp64", flush=True)
    g, dm = wing
    dt = wk.te_panel_size(g) * 2
    o = FW.FreeWakeOptions(dt=dt, max_rows=3)
    runs = {}
    for p in ("fp64", "mixed"):
        s = FW.FreeWakeSim(g, dm, V, o, precision=p)
        runs[p] = s.run(6)
        s.close()
    for a, b in zip(runs["mixed"], runs["fp64"]):
        assert np.abs(a["mu_w"] - b["mu_w"]).max() < 1e-6 * np.abs(b["mu_w"]).max()


def test_free_wake_tree(wing):
    """The native engine with the wake blobs' velocities by the treecode (theta 0.5, degree 6) against direct sums
    over 30 steps: the same rows and particle counts, mu_w within 1e-7 of its largest value in FP64, within 1e-6 in
    mixed precision (FP32 kernels in both the treecode and the direct sums)"""
    print("free wake 30 steps tree vs direct", flush=True)
    g, dm = wing
    dt = wk.te_panel_size(g) * 2
    o = FW.FreeWakeOptions(dt=dt, max_rows=3)
    runs = {}
    for t in (None, (0.5, 6)):
        s = FW.FreeWakeSim(g, dm, V, o, tree=t)
        runs[t] = s.run(30)
        s.close()
    for a, b in zip(runs[(0.5, 6)], runs[None]):
        assert a["particles"] == b["particles"] and a["rows"] == b["rows"]
        assert np.abs(a["mu_w"] - b["mu_w"]).max() < 1e-7 * np.abs(b["mu_w"]).max()
    with pytest.raises(ValueError):
        FW.FreeWakeSim(g, dm, V, o, tree=(1.5, 6))
    # mixed precision: the treecode's FP32 kernels against the mixed direct sums
    mixed = {}
    for t in (None, (0.5, 6)):
        s = FW.FreeWakeSim(g, dm, V, o, precision="mixed", tree=t)
        mixed[t] = s.run(30)
        s.close()
    for a, b in zip(mixed[(0.5, 6)], mixed[None]):
        assert a["particles"] == b["particles"] and a["rows"] == b["rows"]
        assert np.abs(a["mu_w"] - b["mu_w"]).max() < 1e-6 * np.abs(b["mu_w"]).max()


def test_free_wake_full_size():
    """The benchmark's free-wake run at full size (rectangular NACA 0012 wing, AR 6, 120 x 40 quads: 9,836
    triangles; alpha 8 deg, dt 0.02, fixed core), its first 300 steps to ~25,000 particles: mixed (FP32 velocity
    kernels, FP32 LU with FP64 refinement) and the wake blobs by the treecode (theta 0.5, degree 6) against FP64 direct
    sums step by step (the same particle counts and rows, CL by Kutta-Joukowski within 1e-5), and the impulsive
    start's lift growing from one 25-step checkpoint to the next (Wagner)"""
    spec = mesh.WingSpec(code="0012", n_around=120, n_span=40, aspect=6.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    V8, L = sweep.wing_freestream(8.0)
    edges = np.concatenate([np.linalg.norm(g.x[g.tris[:, (k + 1) % 3]] - g.x[g.tris[:, k]], axis=1) for k in range(3)])
    o = FW.FreeWakeOptions(dt=0.02, core=float(0.1 * edges.mean()))
    print(f"case free wake full size nt={g.nt}", flush=True)
    runs = {}
    for p in ("fp64", "mixed", "tree"):
        s = FW.FreeWakeSim(g, dm, V8, o, precision="fp64" if p == "tree" else p, tree=(0.5, 6) if p == "tree" else None)
        runs[p] = s.run(300)
        s.close()
        print(f"  {p}: {runs[p][-1]['particles']} particles, CL {runs[p][-1]['F_kj'] @ L / spec.aspect:.6f}", flush=True)
    cl = {p: np.array([r["F_kj"] @ L / spec.aspect for r in runs[p]]) for p in runs}
    for p in ("mixed", "tree"):
        for a, b in zip(runs[p], runs["fp64"]):
            assert a["particles"] == b["particles"] and a["rows"] == b["rows"]
        d = np.abs(cl[p] - cl["fp64"]).max()
        print(f"  {p} vs fp64: max |dCL| {d:.2e}", flush=True)
        assert d < 1e-5, d
    assert runs["fp64"][-1]["particles"] > 20000
    assert np.all(np.diff(cl["fp64"][24::25]) > 0), cl["fp64"][24::25]


def test_explicit_rhs_full_size():
    """explicit_rhs at the free wake's end size: the benchmark wing (AR 6, 120 x 40 quads, 9,836 triangles) with
    trailing sheets, filaments and 84,000 particles (a cloud behind the wing and 200 next to its surface, so near rules
    run too), all 4,961 entries against the oracle to 1e-12 of the largest"""
    print("explicit_rhs full size", flush=True)
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=120, n_span=40, aspect=6.0))
    dm = dofmap.build_dofmap(g)
    rng = np.random.default_rng(5)
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=0.5))
    e = FW.rigid_elements(w, rng.standard_normal(dm.wake_te.n))
    nb = 84000
    bx = np.vstack([np.array([4.0, 0, 0]) + rng.standard_normal((nb - 200, 3)) * [2.0, 3.0, 0.5],
                    g.x[rng.integers(0, len(g.x), 200)] + rng.standard_normal((200, 3)) * 0.01])
    blobs = FW.Blobs(bx, rng.standard_normal((nb, 3)) * 1e-3, np.full(nb, 0.02))
    got = FW.explicit_rhs(g, dm, e, blobs)
    ref = oracle_explicit_rhs(g, dm, e, blobs, api.AssemblyOptions())
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  nt={g.nt} particles={nb}: {err:.2e}", flush=True)
    assert err < 1e-12, err
