"""GPU N-body sums (vp_surface_velocity, vp_body_velocity, vp_blob_velocities, vp_wake_field_*) against the FP64 oracle
(src/post/Post.jl:32-112, src/freewake/step.jl:176-187).  FP64: within 1e-12 of the largest velocity; FP32 kernels
with FP64 sums: within 2e-5 (source triangles) and 5e-5 (blobs) of it."""
import os
import types

import numpy as np
import pytest

from oracle import oracle as O
from vphip import api, dofmap, mesh, post, wake as wk

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "mixed": 2e-5}


def _err(u, ref):
    """largest deviation relative to the largest reference velocity (absolute where the reference is all zero)"""
    return np.abs(u - ref).max() / max(np.abs(ref).max(), 1e-300)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("level", [1, 3])
def test_surface_velocity_sphere(level, precision):
    print(f"surface_velocity icosphere {level} {precision}", flush=True)
    g = mesh.icosphere(level)
    dm = dofmap.build_dofmap(g)
    V = np.array([1.0, 0.3, -0.2])
    sig = api.source_strengths(api.build_elements(g), V)
    mu = np.random.default_rng(level).standard_normal(dm.nslot)
    u = post.surface_velocity(g, dm, sig, mu, V, precision=precision)
    ref = O.surface_velocity(g, dm, sig, mu, V)
    assert _err(u, ref) < TOL[precision], _err(u, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_surface_velocity_wing_with_wake(precision):
    print(f"surface_velocity wing+wake {precision}", flush=True)
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=24, n_span=8, aspect=4.0))
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(np.radians(5)), 0, np.sin(np.radians(5))])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
    rng = np.random.default_rng(5)
    sig = api.source_strengths(api.build_elements(g), V)
    mu = rng.standard_normal(dm.nslot)
    mu_wake = wk.wake_values(w, rng.standard_normal(dm.wake_te.n))
    u = post.surface_velocity(g, dm, sig, mu, V, wake=w, mu_wake=mu_wake, precision=precision)
    ref = O.surface_velocity(g, dm, sig, mu, V, wake=w, mu_wake=mu_wake)
    assert _err(u, ref) < TOL[precision], _err(u, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_body_velocity_off_body(precision):
    print(f"body_velocity {precision}", flush=True)
    g = mesh.abused_ellipsoid(nth=12, nph=24, axes=(2.0, 1.0, 0.5), jitter=0.25, sliver_frac=0.05,
                              sliver_ratio=0.1, seed=9)
    rng = np.random.default_rng(7)
    sig = rng.standard_normal(g.nt)
    gam = rng.standard_normal((g.nt, 3))
    d = rng.standard_normal((777, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    xs = d * rng.uniform(0.6, 6.0, 777)[:, None] * np.array([2.0, 1.0, 0.5])   # inside and outside, off the surface
    u = post.body_velocity(g, sig, gam, xs, precision=precision)
    ref = O.body_velocity(g, sig, gam, xs)
    assert _err(u, ref) < TOL[precision], _err(u, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nb,npts", [(1, 1), (300, 257), (2049, 1000)])
def test_blob_velocities(nb, npts, precision):
    print(f"blob_velocities nb={nb} np={npts} {precision}", flush=True)
    rng = np.random.default_rng(nb + npts)
    bx = rng.standard_normal((nb, 3)) + np.array([5.0, 0, 0])
    ba = rng.standard_normal((nb, 3))
    bd = rng.uniform(0.02, 0.2, nb)
    xs = np.vstack([rng.standard_normal((npts - 1, 3)) + np.array([5.0, 0, 0]), bx[:1]])  # one target on a blob
    u = post.blob_velocities(bx, ba, bd, xs, precision=precision)
    ref = O.blob_velocities(bx, ba, bd, xs)
    assert _err(u, ref) < (1e-12 if precision == "fp64" else 5e-5), _err(u, ref)


def test_empty_and_bad_inputs():
    print("field edge cases", flush=True)
    u = post.blob_velocities(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.ones((4, 3)))
    assert np.all(u == 0)
    assert post.blob_velocities(np.ones((2, 3)), np.ones((2, 3)), np.ones(2), np.zeros((0, 3))).shape == (0, 3)
    with pytest.raises(ValueError):
        post.blob_velocities(np.ones((2, 3)), np.ones((2, 3)), np.ones(2), np.ones((3, 3)), precision="fp16")
    g = mesh.icosphere(1)
    dm = dofmap.build_dofmap(g)
    with pytest.raises(ValueError):
        post.surface_velocity(g, dm, np.zeros(g.nt), np.zeros(dm.nslot + 1), [1.0, 0, 0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wake_field_graph(precision):
    """The free-wake step's velocity sums as one HIP graph (vp_wake_field_*): every replay, with the target and blob
    counts changing from call to call (and the blobs vanishing and coming back), against the eager body_velocity +
    blob_velocities of the same inputs run between the replays, and the first ones against the FP64 oracle."""
    print(f"wake field graph {precision}", flush=True)
    g = mesh.uv_ellipsoid(12, 24, (2.0, 1.0, 0.5))
    rng = np.random.default_rng(7)
    nt = len(g.tris)
    sig = rng.standard_normal(nt)
    gam = rng.standard_normal((nt, 3))
    V = np.array([1.0, 0.0, 0.1])
    field = post.WakeField(g, precision=precision)
    tol = 1e-13 if precision == "fp64" else 1e-12   # the same kernels and chunks: only the final sums' order differs
    sizes = [(300, 0), (300, 257), (2049, 1000), (500, 40), (1, 1), (7000, 3000), (64, 0), (1500, 800)]
    for k, (npts, nb) in enumerate(sizes):
        xs = 3.0 * rng.standard_normal((npts, 3))
        bx = 3.0 * rng.standard_normal((nb, 3))
        ba = rng.standard_normal((nb, 3))
        bd = rng.uniform(0.05, 0.3, nb)
        u = field(sig, gam, xs, bx, ba, bd, base=V) if nb else field(sig, gam, xs, base=V)
        ref = V + post.body_velocity(g, sig, gam, xs, precision=precision)
        if nb:
            ref = ref + post.blob_velocities(bx, ba, bd, xs, precision=precision)
        assert _err(u, ref) < tol, (k, npts, nb, _err(u, ref))
        if k < 3:
            o = V + O.body_velocity(g, sig, gam, xs) + (O.blob_velocities(bx, ba, bd, xs) if nb else 0.0)
            assert _err(u, o) < TOL[precision] * (2.5 if nb else 1.0), (k, _err(u, o))
    builds, updates, launches = field.info()
    assert launches == len(sizes)
    assert builds == 4 and updates == len(sizes) - builds   # rebuilt only when the blobs appear or vanish
    assert field(sig, gam, np.zeros((0, 3))).shape == (0, 3)
    # the body as a triangle soup (each element with its own corners, as the Julia binding passes `els`)
    soup = types.SimpleNamespace(x=g.x[g.tris].reshape(-1, 3), tris=np.arange(3 * nt).reshape(-1, 3))
    fs = post.WakeField(soup, precision=precision)
    xs = 3.0 * rng.standard_normal((400, 3))
    bx, ba, bd = 3.0 * rng.standard_normal((90, 3)), rng.standard_normal((90, 3)), rng.uniform(0.05, 0.3, 90)
    assert _err(fs(sig, gam, xs, bx, ba, bd, base=V), field(sig, gam, xs, bx, ba, bd, base=V)) <= 1e-13
    fs.close()
    # VP_GRAPHS=0: the same three kernels launched one by one, no graph launch
    launches = field.info()[2]
    os.environ["VP_GRAPHS"] = "0"
    try:
        eager = field(sig, gam, xs, bx, ba, bd, base=V)
    finally:
        del os.environ["VP_GRAPHS"]
    assert field.info()[2] == launches
    assert _err(eager, field(sig, gam, xs, bx, ba, bd, base=V)) <= 1e-15
    with pytest.raises(ValueError):
        field(sig[:-1], gam, np.ones((2, 3)))
    with pytest.raises(ValueError):
        field(sig, gam, np.ones((2, 3)), np.ones((2, 3)), np.ones((3, 3)), np.ones(2))
    field.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_blob_velocities_full_size(precision):
    """The free wake's end size: 84,000 blobs (the benchmark's spread and strengths, cores 0.01) summed at all 84,000
    blob positions on the device; 600 of the targets (every 140th) against the oracle's FP64 sums over all 84,000
    blobs, to the same tolerances as the small cases"""
    print(f"blob_velocities full size {precision}", flush=True)
    rng = np.random.default_rng(0)
    nb = 84000
    bx = rng.standard_normal((nb, 3)) * np.array([3.0, 3.0, 0.3])
    ba = rng.standard_normal((nb, 3)) * 1e-3
    bd = np.full(nb, 0.01)
    u = post.blob_velocities(bx, ba, bd, bx, precision=precision)
    pick = np.arange(0, nb, 140)
    ref = O.blob_velocities(bx, ba, bd, bx[pick])
    e = _err(u[pick], ref)
    print(f"  {len(pick)} sampled targets: {e:.2e}", flush=True)
    assert e < (1e-12 if precision == "fp64" else 5e-5), e


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_body_velocity_full_size(precision):
    """body_velocity of the abused ellipsoid at its full benchmark size (19,880 triangles, random source and sheet
    strengths) at 777 points inside and outside, off the surface, against the oracle's sums over every triangle"""
    print(f"body_velocity full size {precision}", flush=True)
    g = mesh.abused_ellipsoid(nth=71, nph=142, axes=(2.0, 1.0, 0.5), jitter=0.25, sliver_frac=0.05, sliver_ratio=0.1,
                              seed=20261203)
    rng = np.random.default_rng(11)
    sig = rng.standard_normal(g.nt)
    gam = rng.standard_normal((g.nt, 3))
    d = rng.standard_normal((777, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    xs = d * rng.uniform(0.6, 6.0, 777)[:, None] * np.array([2.0, 1.0, 0.5])
    u = post.body_velocity(g, sig, gam, xs, precision=precision)
    ref = O.body_velocity(g, sig, gam, xs)
    e = _err(u, ref)
    print(f"  {e:.2e}", flush=True)
    assert e < TOL[precision], e
