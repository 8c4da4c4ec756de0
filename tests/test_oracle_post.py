"""The oracle's N-body sums (surface_velocity, body_velocity, blob sums) and the host force formulas, pinned against
the exact sphere flow, the reference's published sum-route accuracy (README.md:195-197) and its own test values
(test/assembly/test_sphere.jl:20-38, test/post/test_forces.jl:13-57).  The host mirror's force formulas
(vphip.post) are checked against the same values."""
import numpy as np
import pytest

from oracle import oracle as O
from vphip import api, dofmap, mesh, post

VG = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])


def sphere(level):
    g = mesh.icosphere(level)
    dm = dofmap.build_dofmap(g)
    el = O.elements(g.x, g.tris)
    sig = O.source_strengths(el, VG)
    A, b, _, _ = O.assemble(g, dm, sig, VG, O.opts())
    sol = O.solve_galerkin(dm, A, b)
    u = O.surface_velocity(g, dm, sig, sol["mu"], VG)
    n, area, c = el["n"], el["area"], el["centroid"]
    rms = lambda v: np.sqrt(np.sum(area * v ** 2) / np.sum(area))
    cpex = 1 - 9 / 4 * (1 - ((c / np.linalg.norm(c, axis=1)[:, None]) @ VG) ** 2)
    ug = O.gradient_velocity(el, dm, sol["mu"], VG)
    ut = u - np.sum(u * n, 1)[:, None] * n
    return dict(g=g, dm=dm, el=el, sig=sig, mu=sol["mu"], u=u, cpL2=rms(O.pressure_coefficient(u, VG) - cpex),
                leak=rms(np.sum(u * n, 1)), agree=np.max(np.linalg.norm(ut - ug, axis=1)))


@pytest.fixture(scope="module")
def S():
    return {lev: sphere(lev) for lev in (2, 3, 4)}


def test_sum_route_published_icosphere4(S):
    """README.md:197: icosphere 4, Cp L2 (sums) 1.9e-2, leak rms 9.9e-3"""
    r = S[4]
    assert abs(r["cpL2"] - 1.9e-2) < 0.05e-2 and abs(r["leak"] - 9.9e-3) < 0.05e-3


def test_sum_route_reference_bounds(S):
    """test/assembly/test_sphere.jl:38-40: Cp L2 (sums) < 0.08, routes agree < 0.05, leak < 0.045, leak first order"""
    assert S[2]["cpL2"] < 0.08 and S[2]["agree"] < 0.05 and S[2]["leak"] < 0.045
    assert S[2]["leak"] / S[3]["leak"] > 1.7 and S[3]["leak"] / S[4]["leak"] > 1.7


def test_body_velocity_is_the_exterior_dipole(S):
    """Off the body the sources and sheets of the sphere solution induce the exact perturbation velocity
    grad((V.x)/(2 r^3)); the error is the discretisation error, falling with refinement"""
    rng = np.random.default_rng(3)
    d = rng.standard_normal((40, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    xs = d * rng.uniform(1.3, 4.0, 40)[:, None]
    r = np.linalg.norm(xs, axis=1)[:, None]
    exact = 0.5 * (VG / r ** 3 - 3 * (xs @ VG)[:, None] * xs / r ** 5)
    err = {}
    for lev in (3, 4):
        s = S[lev]
        gam = O.sheet_strengths(s["el"], s["dm"], s["mu"])
        ub = O.body_velocity(s["g"], s["sig"], gam, xs)
        err[lev] = np.max(np.linalg.norm(ub - exact, axis=1) / np.linalg.norm(exact, axis=1))
    assert err[4] < 4e-3 and err[3] / err[4] > 2.5, err


def test_blob_sum_is_the_sum_of_blobs():
    """blob_velocities adds blob_velocity (pinned in test_oracle_kernels) over the particles, and a particle's own
    position gets nothing from it"""
    rng = np.random.default_rng(11)
    bx, ba, bd = rng.standard_normal((30, 3)), rng.standard_normal((30, 3)), rng.uniform(0.05, 0.2, 30)
    xs = np.vstack([rng.standard_normal((20, 3)), bx[:5]])
    u = O.blob_velocities(bx, ba, bd, xs)
    ref = np.array([sum(O.blob_velocity(x, bx[k], ba[k], bd[k]) for k in range(30)) for x in xs])
    assert np.abs(u - ref).max() < 1e-14 * np.abs(ref).max() * 30
    own = np.array([O.blob_velocity(bx[k], bx[k], ba[k], bd[k]) for k in range(5)])
    assert np.abs(own).max() == 0.0


def straight_te(y):
    """test/post/test_forces.jl:7-11: a straight trailing edge along +y"""
    n = len(y)
    x = np.column_stack([np.zeros(n), np.asarray(y, dtype=np.float64), np.zeros(n)])
    return dofmap.WakeTE(node=np.arange(n), x=x, tangent=np.tile([0.0, 1.0, 0.0], (n, 1)),
                         bisector=np.tile([1.0, 0.0, 0.0], (n, 1)), lines=[list(range(n))], closed=[False])


@pytest.mark.parametrize("impl", ["oracle", "host"])
def test_forces_reference_values(impl):
    """test/post/test_forces.jl:13-57"""
    KJ = O.kutta_joukowski_force if impl == "oracle" else post.kutta_joukowski_force
    TD = O.trefftz_drag if impl == "oracle" else post.trefftz_drag
    # uniform pressure on a closed surface: no force, no moment
    g = mesh.icosphere(2)
    cp = np.full(g.nt, 0.7)
    if impl == "oracle":
        F, M = O.pressure_force(O.elements(g.x, g.tris), cp, origin=(0.3, -0.2, 0.1))
    else:
        F, M = post.pressure_force(api.build_elements(g), cp, origin=(0.3, -0.2, 0.1))
    assert np.linalg.norm(F) < 1e-13 and np.linalg.norm(M) < 1e-13
    # uniform circulation 0.8 over a span of 2
    f = KJ(straight_te(np.linspace(-1, 1, 9)), np.full(9, 0.8), np.array([2.0, 0, 0]))
    assert np.allclose(f, [0, 0, 2 * 0.8 * 2.0 / 2.0], atol=1e-14)
    # elliptic loading: CDi -> CL^2/(pi AR) as the cosine-spaced strips refine
    b, c = 2.0, 0.5
    V = np.array([1.0, 0, 0])

    def ratio(n):
        y = -b / 2 * np.cos(np.pi * np.arange(n + 1) / n)
        wte = straight_te(y)
        mu = np.sqrt(np.maximum(0.0, 1 - (2 * y / b) ** 2))
        CL = KJ(wte, mu, V)[2] / (b * c)
        CD = TD(wte, mu, V) / (b * c)
        return CD / (CL ** 2 / (np.pi * b / c))

    assert abs(ratio(160) - 1) < 0.01
    assert abs(ratio(160) - 1) < abs(ratio(40) - 1) < abs(ratio(10) - 1)
    y = -np.cos(np.pi * np.arange(41) / 40)
    mu = np.sqrt(np.maximum(0.0, 1 - y ** 2))
    a5 = np.radians(5)
    assert abs(TD(straight_te(y), mu, np.array([np.cos(a5), 0, np.sin(a5)])) - TD(straight_te(y), mu, V)) < 1e-13


def test_host_post_formulas_match_oracle():
    """vphip.post's per-element formulas give the oracle's numbers"""
    g = mesh.icosphere(2)
    dm = dofmap.build_dofmap(g)
    el, els = O.elements(g.x, g.tris), api.build_elements(g)
    mu = np.random.default_rng(2).standard_normal(dm.nslot)
    assert np.abs(post.sheet_strengths(els, dm, mu) - O.sheet_strengths(el, dm, mu)).max() < 1e-12
    ug = post.gradient_velocity(els, dm, mu, VG)
    assert np.abs(ug - O.gradient_velocity(el, dm, mu, VG)).max() < 1e-12
    assert np.abs(post.pressure_coefficient(ug, VG) - O.pressure_coefficient(ug, VG)).max() < 1e-14


def test_wing_loads_host_matches_oracle():
    """vphip.sweep.wing_loads (pressure CL, Trefftz CDi, KJ) gives the oracle's numbers for the same solution"""
    from vphip import sweep
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=16, n_span=6, aspect=4.0))
    dm = dofmap.build_dofmap(g)
    rng = np.random.default_rng(4)
    q = rng.standard_normal(dm.nq)
    mu = dm.C @ q
    V, eL = sweep.wing_freestream(6.0)
    got = sweep.wing_loads(api.build_elements(g), dm, mu, q, V, eL, 4.0)
    el = O.elements(g.x, g.tris)
    cp = O.pressure_coefficient(O.gradient_velocity(el, dm, mu, V), V)
    Fp, _ = O.pressure_force(el, cp)
    mu_w = dm.K @ q
    assert abs(got["CL"] - Fp @ eL / 4.0) < 1e-13 * max(1.0, abs(got["CL"]))
    assert abs(got["CDi_T"] - O.trefftz_drag(dm.wake_te, mu_w, V) / 4.0) < 1e-12 * max(1.0, abs(got["CDi_T"]))
    assert np.allclose(got["CL_KJ"], O.kutta_joukowski_force(dm.wake_te, mu_w, V) @ eL / 4.0, rtol=1e-13, atol=1e-14)
