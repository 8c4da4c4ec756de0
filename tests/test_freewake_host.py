"""The free wake's host logic (vphip.freewake, a restatement of src/freewake/state.jl and step.jl) run on the CPU
oracle: core radii, sheet vorticity, the trailing edge panel with explicit rows against the implicit rigid wake,
conservation of the vorticity and its first moment by conversion and merging, and an impulsive start (Kelvin's theorem,
circulation build-up, force routes, bounded velocities) -- test/freewake/test_freewake.jl:15-139 on a NACA 0012 wing."""
import math
import warnings

import numpy as np
import pytest

from oracle.freewake_oracle import OracleFreeWakeSim, oracle_explicit_rhs, oracle_opts
from oracle import oracle as O
from vphip import api, dofmap, freewake as FW, mesh, sweep, wake as wk

V, LIFT = sweep.wing_freestream(5.0)


@pytest.fixture(scope="module")
def wing():
    spec = mesh.WingSpec(code="0012", n_around=16, n_span=12, aspect=4.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    return g, dm, spec.aspect * spec.mean_chord ** 2


def test_core_radius(wing):
    o = np.zeros(3)
    auto = FW.CoreRule(FW.FreeWakeOptions(dt=0.1), np.array([1.0, 0, 0]))
    for w, l in ((0.25, 0.06), (0.25, 0.006), (0.02, 0.06)):
        assert math.isclose(auto.tri(o, o + [w, 0, 0], o + [w, l, 0]), 1.5 * math.hypot(w, l) / 2, rel_tol=1e-14)
    assert math.isclose(auto.seg(o, np.array([0.2, 0, 0])), 1.5 * 0.1, rel_tol=1e-14)
    assert FW.CoreRule(FW.FreeWakeOptions(dt=0.1, core=0.05), np.array([1.0, 0, 0])).seg(o, np.array([1.0, 0, 0])) == 0.05
    sc = FW.CoreRule(FW.FreeWakeOptions(dt=0.1, core="scaled"), np.array([1.0, 0, 0]))
    assert math.isclose(sc.tri(o, np.array([0.25, 0, 0]), np.array([0.25, 0.006, 0])), max(1.5 * math.sqrt(0.25 * 0.006), 0.1))
    g, dm, _ = wing
    with pytest.warns(UserWarning, match=r"max_rows = \d+ keeps them two cores away"):
        OracleFreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=0.06, max_rows=0))
    with pytest.raises(ValueError):
        FW.FreeWakeOptions(dt=-1.0).check()


def test_sheet_vorticity():
    p, q, r, mu = np.array([0.1, 0, 0.2]), np.array([1.0, 0.3, -0.1]), np.array([0.4, 0.9, 0.3]), np.array([0.7, -0.2, 1.1])
    g = mesh.Surface.__new__(mesh.Surface)
    els = api.build_elements(type("G", (), {"x": np.array([p, q, r]), "tris": np.array([[0, 1, 2]])})())
    gam = np.einsum("k,kj->j", mu, els.gamma[0])
    assert np.allclose(FW.sheet_vorticity(p, q, r, mu), gam * els.area[0], rtol=1e-13, atol=1e-15)


def test_te_panel_with_explicit_rows(wing):
    """the rigid wake solved implicitly equals its first row as the trailing edge panel plus the other rows explicit"""
    g, dm, _ = wing
    el = O.elements(g.x, g.tris)
    sig = O.source_strengths(el, V)
    wf = wk.build_wake(g, dm, V, wk.WakeOptions(length=1.0))
    A, b, _, _ = O.assemble(g, dm, sig, V, O.opts(), wake=wf)
    ref = O.solve_galerkin(dm, A, b, dofmap.wake_rows(dm, wf.te))
    muw = dm.K @ ref["q"]
    te = wk.te_panel(g, dm, V, length=wf.rows[1])
    At, bt, _, _ = O.assemble(g, dm, sig, V, O.opts(), wake=te)
    topo = FW.Topology.of(dm)
    n = topo.nte
    st = FW.FreeWakeState.empty(n)
    for j in range(1, len(wf.rows)):
        st.x.append(wf.x[j * n:(j + 1) * n].copy())
        st.mu.append(muw.copy())
    st.mu_line = muw.copy()
    we = FW.vortex_elements(st, topo)
    rhs = oracle_explicit_rhs(g, dm, we, None, api.AssemblyOptions())
    s = O.solve_galerkin(dm, At, bt + rhs, dofmap.wake_rows(dm, te.te))
    assert np.abs(dm.K @ s["q"] - muw).max() < 1e-12 * np.abs(muw).max()
    assert np.abs(s["q"] - ref["q"]).max() < 1e-12 * np.abs(ref["q"]).max()
    cols = At.T[:, dm.nslot:] @ muw[te.te] if At.shape[0] != dm.nslot else At[:, dm.nslot:] @ muw[te.te]
    ex = oracle_explicit_rhs(g, dm, FW.rigid_elements(te, muw), None, api.AssemblyOptions())
    assert np.abs(ex + cols).max() < 1e-13 * np.abs(cols).max()


def _blobs(sim):
    return FW.Blobs.cat([FW.free_wake_blobs(FW.vortex_elements(sim.state, sim.topo), sim.core), sim.state.particles])


def test_conversion_conserves_vorticity_and_impulse(wing):
    g, dm, _ = wing
    sim = OracleFreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=0.06, max_rows=10_000))
    sim.run(6)
    b = _blobs(sim)
    O0, I0 = FW.vorticity_moments(b)
    sO, sI = np.linalg.norm(b.alpha, axis=1).sum(), (np.linalg.norm(b.x, axis=1) * np.linalg.norm(b.alpha, axis=1)).sum()
    for _ in range(3):
        FW.convert_oldest(sim.state, sim.topo, sim.core)
        Om, I = FW.vorticity_moments(_blobs(sim))
        assert np.linalg.norm(Om - O0) < 1e-14 * sO and np.linalg.norm(I - I0) < 1e-14 * sI
    assert sim.state.conversions == 3 and len(sim.state.x) == 4


def test_conversion_distance_and_merging(wing):
    g, dm, _ = wing
    dt = wk.te_panel_size(g)
    keep = [OracleFreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=dt / f)).keep_rows for f in (1, 4)]
    assert keep[0] >= 1 and abs(keep[1] / keep[0] - 4) < 0.5
    sim = OracleFreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=dt / 4, merge_rows="auto"))
    row = np.linalg.norm(V) * dt / 4   # rows per particle pair: each strip's width over the row length, at least 1
    widths = [np.linalg.norm(sim.edge[B] - sim.edge[A]) for A, B in sim.topo.segments]
    assert sim.merge == [max(1, int(round(w / row))) for w in widths] and min(sim.merge) >= 1
    sim = OracleFreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=dt, max_rows=2, merge_rows=3))
    sim.run(8)
    st = sim.state
    merged = False
    for _ in range(3):
        O0, I0 = FW.vorticity_moments(_blobs(sim))
        pending = int((st.tag > 0).sum())
        n0 = len(st.particles)
        FW.convert_oldest(st, sim.topo, sim.core, sim.merge)
        Om, I = FW.vorticity_moments(_blobs(sim))
        assert np.linalg.norm(Om - O0) < 1e-14 * np.linalg.norm(_blobs(sim).alpha, axis=1).sum()
        assert np.linalg.norm(I - I0) < 1e-3 * np.linalg.norm(I0)
        if int((st.tag > 0).sum()) < pending:
            merged = True
            assert len(st.particles) < n0
        st.x.insert(0, sim.edge.copy())
        st.mu.insert(0, sim.mu_w.copy())
    assert merged


def test_impulsive_start(wing):
    g, dm, S = wing
    dt = wk.te_panel_size(g)
    sim = OracleFreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=dt, max_rows=4, advection="freestream"))
    hist = sim.run(20)
    st = sim.state
    for k in range(len(st.x) - 1):
        assert np.allclose(st.x[k + 1][2], sim.edge[2] + (k + 1) * dt * V, rtol=0, atol=1e-12)
    assert all(np.linalg.norm(r["Omega"]) < 1e-12 for r in hist)
    kj = [r["F_kj"] @ LIFT / S for r in hist]
    assert kj[0] < kj[4] < kj[9] < kj[19]
    clp, cli = hist[-1]["F_pressure"] @ LIFT / S, hist[-1]["F_impulse"] @ LIFT / S
    assert abs(clp - cli) < 0.05 * abs(clp), (clp, cli)
    assert len(st.particles) > 0 and st.conversions == 20 - 4
    free = OracleFreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=dt, max_rows=4))
    fh = free.run(20)
    assert all(r["umax"] < 3 * np.linalg.norm(V) for r in fh)
    assert abs((fh[-1]["F_kj"] @ LIFT / S) / kj[-1] - 1) < 0.03


def test_paraview_series(wing, tmp_path):
    """run(n, series=...) writes a ParaView frame every `every` steps and the last (src/freewake/step.jl:203-216,
    src/freewake/export.jl): the free rows, particles, body and trailing edge panel grids with their arrays, and .pvd
    collections listing each frame at its step's time."""
    import xml.etree.ElementTree as ET
    from vphip import export
    g, dm, S = wing
    dt = wk.te_panel_size(g)
    sim = OracleFreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=dt, max_rows=2))
    series = export.FreeWakeSeries(str(tmp_path), every=3)
    hist = sim.run(7, series=series)
    frames = sorted(p.name for p in tmp_path.glob("frame_*_body.vtu"))
    assert frames == ["frame_00000_body.vtu", "frame_00003_body.vtu", "frame_00006_body.vtu"]
    assert sorted(p.name for p in tmp_path.glob("*.pvd")) == ["body.pvd", "free.pvd", "particles.pvd", "te_panel.pvd"]
    times = [float(d.get("timestep")) for d in ET.parse(tmp_path / "body.pvd").getroot().iter("DataSet")]
    assert np.allclose(times, [hist[0]["t"], hist[3]["t"], hist[6]["t"]])

    def arrays(path):
        piece = ET.parse(path).getroot().find("UnstructuredGrid/Piece")
        out = {a.get("Name"): np.array(a.text.split(), float) for a in piece.iter("DataArray")}
        return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), out

    npts, ncell, a = arrays(tmp_path / "frame_00006_body.vtu")
    assert (npts, ncell) == (g.nn, g.nt) and len(a["Cp"]) == g.nt and len(a["Cp_nodes"]) == g.nn
    assert np.allclose(a["Cp"], sim.cp) and np.allclose(a["mu"], sim.mu[:g.nn])
    assert np.all(a["types"] == 5) and np.array_equal(a["connectivity"].astype(int), g.tris.ravel())
    # the frame's free rows are the state before its step moved them: rows 0..nr-1 of nte nodes
    npts, ncell, a = arrays(tmp_path / "frame_00006_free.vtu")
    nte = sim.topo.nte
    assert npts % nte == 0 and ncell == 2 * len(sim.topo.segments) * (npts // nte - 1)
    assert set(a["row"].astype(int)) == set(range(npts // nte))
    npts, ncell, a = arrays(tmp_path / "frame_00006_particles.vtu")
    assert npts == ncell and len(a["alpha"]) == 3 * npts and np.all(a["types"] == 1)
    assert np.allclose(a["strength"], np.linalg.norm(a["alpha"].reshape(-1, 3), axis=1))
    # the area-weighted nodal average reproduces a constant
    els = api.build_elements(g)
    assert np.allclose(export.nodal_average(g, els, np.full(g.nt, 2.5)), 2.5)
