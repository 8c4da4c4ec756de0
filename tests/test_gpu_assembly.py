"""GPU assembly through the C ABI against the FP64 oracle: matrix entries to 1e-12 of the largest entry, the
right hand side likewise, and tier labels and counts bit for bit (src/assembly/assemble.jl:260-447)."""
import numpy as np
import pytest

from oracle import oracle as O
from vphip import api, dofmap, mesh, wake as wk

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


def test_wing_with_wake():
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=16, n_span=6, aspect=4.0))
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(np.radians(5)), 0, np.sin(np.radians(5))])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
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
