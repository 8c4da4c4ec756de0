"""Pins of the oracle's isoparametric quads and mixed-mesh assembly (src/elements/iso_quad.jl, quad_field.jl,
quad_rules.jl, src/assembly/mixed.jl), after the reference's own checks (test/elements/test_iso_quad.jl,
test/assembly/test_quads.jl): closest points against a brute-force grid, the inner integrals against an adaptive
cubature graded at the closest point (its integrand written out here from the definition), principal values against
two-sided limits, the square rules' monomial and log integrals, wide sheared corners and T-junction contact, touching
pair blocks against cubature, the triangle-only mixed mesh against the triangle assembly, and the cube-sphere's exact
flow (the quads' error 0.54 of the triangles', Cp 0.025 at the quad centres: test/assembly/test_quads.jl:55-80)."""
import numpy as np
import pytest

import cubature
from oracle import oracle as O
from vphip import dofmap, exact, geometry, mesh, mixed
from vphip import wake as W

QUADS = {
    "square": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]],
    "trapezoid": [[0, 0, 0], [2, 0, 0], [1.4, 0.8, 0], [0.4, 0.8, 0]],
    "warped": [[0, 0, 0], [1, 0.1, 0.2], [1.2, 1.1, -0.1], [-0.1, 0.9, 0.15]],
    "sliver": [[0, 0, 0], [1, 0, 0], [1, 0.02, 0], [0, 0.02, 0]],
    "sheared sliver": [[0, 0, 0], [1, 0, 0], [1.6, 0.06, 0.01], [0.6, 0.06, 0]],  # 5.7 degree corners
}
QUADS = {k: np.array(v, dtype=np.float64) for k, v in QUADS.items()}
SQ = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
VQ = np.array([0.8, -0.3, 0.5])


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def integrand(qx, x, nx, V, p):
    """The inner integrands per unit parameter area at parameters p (n, 2), from their definition: the four unit
    nodal sheets' nx.(g_b x r)/|r|^3 with g_b = N_b,xi a2 - N_b,eta a1, and the source sheet sigma = -V.n as
    -V.(a1 x a2) nx.r/|r|^3, r = x - y(p); without 1/4pi.  Returns (n, 5)."""
    xi, eta = p[:, 0], p[:, 1]
    N = np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta], 1)
    y = N @ qx
    a1 = (1 - eta)[:, None] * (qx[1] - qx[0]) + eta[:, None] * (qx[2] - qx[3])
    a2 = (1 - xi)[:, None] * (qx[3] - qx[0]) + xi[:, None] * (qx[2] - qx[1])
    Nx = np.stack([-(1 - eta), 1 - eta, eta, -eta], 1)
    Ne = np.stack([-(1 - xi), -xi, xi, 1 - xi], 1)
    r = x - y
    r3 = np.linalg.norm(r, axis=1) ** 3
    out = np.zeros((len(p), 5))
    for b in range(4):
        g = Nx[:, b:b + 1] * a2 - Ne[:, b:b + 1] * a1
        out[:, b] = np.cross(g, r) @ nx / r3
    out[:, 4] = -(np.cross(a1, a2) @ V) * (r @ nx) / r3
    return out


def field_cubature(qx, x, nx, V, P, scale):
    """The quad's inner integrals at x by adaptive cubature: the square split at the closest point P, each part graded
    toward it by u -> P + D u^3 (test/elements/test_iso_quad.jl:24-37)"""
    total = np.zeros(5)
    for C in SQ:
        D = C - P
        if abs(D[0]) < 1e-15 or abs(D[1]) < 1e-15:
            continue

        def f(uv, D=D):
            p = P + D * uv ** 3
            jac = abs(D[0] * D[1]) * 9 * (uv[:, 0] * uv[:, 1]) ** 2
            return integrand(qx, x, nx, V, p) * jac[:, None]

        val, _ = cubature.square(f, 1e-8 * scale, max_rounds=600)
        total += val
    return total / (4 * np.pi)


def field5(qx, x, nx, V, side=0):
    c, s = O.quad_field(qx, x, nx, V, side=side)
    return np.concatenate([c, [s]])


def point(qx, p):
    return mixed.point(qx, p[0], p[1])


def normal(qx, p):
    return unit(mixed.area_normal(qx, p[0], p[1]))


# ------------------------------------------------------------------------------------------------- geometry -----

def area_cubature(qx):
    return cubature.square(lambda uv: np.linalg.norm(mixed.area_normal(qx, uv[:, 0], uv[:, 1]), axis=1)[:, None],
                           1e-13)[0][0]


@pytest.mark.parametrize("name", list(QUADS))
def test_quad_geometry(name):
    """test/elements/test_iso_quad.jl:38-62: corners, area, centroid, sheet densities, gradients of linear fields"""
    qx = QUADS[name]
    info = O.quad_info(qx)
    assert abs(info["area"] / area_cubature(qx) - 1) < 1e-6
    assert np.allclose(info["centroid"], point(qx, (0.5, 0.5)), rtol=0, atol=1e-15)
    for k in range(4):
        assert np.allclose(point(qx, SQ[k]), qx[k], atol=1e-15)
    for xi, eta in ((0.3, 0.7), (0.0, 1.0), (0.91, 0.12)):
        assert abs(mixed.shape(xi, eta).sum() - 1) < 1e-15
        Nx, Ne = mixed.shape_derivatives(xi, eta)
        assert abs(Nx.sum()) < 1e-15 and abs(Ne.sum()) < 1e-15
    assert np.linalg.norm(mixed.sheet_strength(qx, np.ones(4), 0.3, 0.6)) < 1e-14  # a uniform doublet sheds nothing
    k = np.array([0.7, -1.3, 0.4])
    mu = qx @ k  # bilinear on the quad's own surface, so interpolated exactly
    for xi, eta in ((0.3, 0.6), (0.95, 0.05), (0.0, 1.0)):
        n = normal(qx, (xi, eta))
        kt = k - (k @ n) * n
        assert np.allclose(mixed.gradient(qx, mu, xi, eta), kt, atol=1e-12)
        assert np.allclose(mixed.sheet_strength(qx, mu, xi, eta), np.cross(n, kt), atol=1e-12)


@pytest.mark.parametrize("name", list(QUADS))
def test_closest_point(name):
    """test/elements/test_iso_quad.jl:64-84: never farther than the best node of a 401 x 401 parameter grid"""
    qx = QUADS[name]
    info = O.quad_info(qx)
    rng = np.random.default_rng(3)
    g = np.linspace(0.0, 1.0, 401)
    XI, ETA = np.meshgrid(g, g, indexing="ij")
    grid = mixed.point(qx, XI.ravel(), ETA.ravel())
    for _ in range(30):
        x = info["centroid"] + info["diam"] * (rng.random(3) - 0.5) * np.array([1.6, 1.6, 0.4])
        P, d = O.quad_project(qx, x)
        assert np.all((P >= 0) & (P <= 1))
        assert abs(np.linalg.norm(point(qx, P) - x) - d) < 1e-14
        assert d <= np.min(np.linalg.norm(grid - x, axis=1)) + 1e-12


# ------------------------------------------------------------------------------------------ inner integrals -----

@pytest.mark.parametrize("name", list(QUADS))
def test_inner_integrals_vs_cubature(name):
    """test/elements/test_iso_quad.jl:86-102: above an interior point, just outside the middle of side 1 and just
    outside corner 3, at 3 down to 1e-6 diameters, with a tilted direction nx: within 1e-6 of the cubature.  The
    distances cover both the tensor Gauss tier (3 diameters) and the polar one."""
    qx = QUADS[name]
    diam = O.quad_info(qx)["diam"]
    for p0, off in (((0.37, 0.58), (0.0, 0.0)), ((0.5, 0.0), (0.0, -0.02)), ((1.0, 1.0), (0.01, 0.015))):
        p0 = np.array(p0)
        n = normal(qx, p0)
        base = point(qx, p0 + np.array(off))
        nx = unit(n + 0.4 * unit(qx[1] - qx[0]))
        for d in (3.0, 0.3, 1e-2, 1e-4, 1e-6):
            x = base + d * diam * n
            got = field5(qx, x, nx, VQ)
            P, _ = O.quad_project(qx, x)
            ref = field_cubature(qx, x, nx, VQ, P, max(np.linalg.norm(got), 1e-12))
            assert np.linalg.norm(got - ref) <= 1e-6 * np.linalg.norm(ref), (name, p0.tolist(), d)


@pytest.mark.parametrize("name", list(QUADS))
def test_principal_value_and_limits(name):
    """test/elements/test_iso_quad.jl:104-117: principal value = mean of the two-sided limits, one-sided limits =
    the fields just off the surface, and the source sheet's jump sigma (nx.n)"""
    qx = QUADS[name]
    diam = O.quad_info(qx)["diam"]
    for p0 in ((0.37, 0.58), (0.08, 0.9)):
        n = normal(qx, p0)
        x0 = point(qx, p0)
        nx = unit(n + 0.5 * unit(qx[1] - qx[0]))
        h = 1e-7 * diam
        up, down = field5(qx, x0 + h * n, nx, VQ), field5(qx, x0 - h * n, nx, VQ)
        pv, plus, minus = field5(qx, x0, nx, VQ), field5(qx, x0, nx, VQ, 1), field5(qx, x0, nx, VQ, -1)
        scale = np.linalg.norm(up) + np.linalg.norm(down)
        assert np.linalg.norm(pv - (up + down) / 2) <= 1e-5 * scale
        assert np.linalg.norm(plus - up) <= 1e-5 * scale and np.linalg.norm(minus - down) <= 1e-5 * scale
        assert abs((plus[4] - minus[4]) - (-(VQ @ n)) * (nx @ n)) < 1e-12


def test_boundary_point_is_an_error():
    """polar_field refuses a point on the quad's boundary (src/elements/quad_field.jl:108-109)"""
    qx = QUADS["square"]
    with pytest.raises(ValueError):
        O.quad_field(qx, np.array([0.5, 0.0, 0.0]), np.array([0.0, 0.0, 1.0]), VQ)


# --------------------------------------------------------------------------------------------------- rules ------

def integrate(p, w, f):
    return float(np.sum(w * f(p)))


def mono2(a, b):
    return 1 / ((a + 1) * (b + 1))


def test_square_rules():
    """test/elements/test_iso_quad.jl:119-138 (levels = 6): tensor Gauss exactness, weights and range, and the log
    integrals -1 along a side, log 2 - 3 + pi/2 at a corner, -4 along all four sides"""
    o = O.opts()
    for m in (2, 3):  # far_points 3 -> 2 x 2 and 7 -> 3 x 3 (square_tiers, src/assembly/mixed.jl:70-71)
        p, w = O.square_rule("far", O.opts(far_points=3 if m == 2 else 6), quad_levels=6)
        assert len(w) == m * m
        for a in range(2 * m):
            for b in range(2 * m):
                assert abs(integrate(p, w, lambda p: p[:, 0] ** a * p[:, 1] ** b) - mono2(a, b)) < 1e-14
    dist_side = (lambda p: p[:, 1], lambda p: 1 - p[:, 0], lambda p: 1 - p[:, 1], lambda p: p[:, 0])
    kinds = ["self"] + [f"edge{k}" for k in range(4)] + [f"vertex{k}" for k in range(4)]
    for kind in kinds:
        p, w = O.square_rule(kind, o, quad_levels=6)
        assert abs(w.sum() - 1) < 1e-13 and np.all((p >= 0) & (p <= 1))
        assert abs(integrate(p, w, lambda p: p[:, 0] ** 3 * p[:, 1] ** 2) / mono2(3, 2) - 1) < 1e-12
    for k in range(4):
        p, w = O.square_rule(f"edge{k}", o, quad_levels=6)
        assert abs(integrate(p, w, lambda p: np.log(dist_side[k](p))) / -1 - 1) < 1e-6
        p, w = O.square_rule(f"vertex{k}", o, quad_levels=6)
        ref = np.log(2) - 3 + np.pi / 2
        assert abs(integrate(p, w, lambda p: np.log(np.sum((p - SQ[k]) ** 2, 1))) / ref - 1) < 1e-6
    p, w = O.square_rule("self", o, quad_levels=6)
    assert abs(integrate(p, w, lambda p: sum(np.log(f(p)) for f in dist_side)) / -4 - 1) < 1e-6
    # the point counts at the default four levels: 25 x 24 per side, 2 x 25 x 7 per corner, 4 x 25 x 24 in all
    assert len(O.square_rule("edge0", o)[1]) == 600 and len(O.square_rule("vertex0", o)[1]) == 350
    assert len(O.square_rule("self", o)[1]) == 2400


def test_wide_corner_and_contact():
    """test/elements/test_iso_quad.jl:140-168: a corner sheared to 150 degrees split by physical angle (1e-6 against
    cubature, the stored rule > 1e-5 off); an acute corner keeps the stored rule; a T-junction cut at the hanging
    node, the half along the neighbour on the edge rule; neighbours sharing a whole side or a corner need none"""
    o = O.opts()
    a4 = 0.4 * np.array([np.cos(np.radians(150)), np.sin(np.radians(150)), 0.0])
    qx = np.array([[0.0, 0, 0], [1.0, 0, 0], np.array([1.1, 0, 0]) + a4, a4])
    _, p, w = O.quad_rule(qx, "vertex_square", corner=0)
    assert abs(w.sum() - 1) < 1e-13 and np.all((p >= -1e-14) & (p <= 1 + 1e-14))
    assert abs(integrate(p, w, lambda p: p[:, 0] ** 3 * p[:, 1] ** 2) / mono2(3, 2) - 1) < 1e-12

    def f(pp):
        x = mixed.point(qx, pp[:, 0], pp[:, 1])
        return np.log(np.linalg.norm(x - qx[0], axis=1)) * np.linalg.norm(mixed.area_normal(qx, pp[:, 0], pp[:, 1]), axis=1)

    ref = cubature.square(lambda uv: (f(uv ** 3) * 9 * (uv[:, 0] * uv[:, 1]) ** 2)[:, None], 1e-13)[0][0]
    assert abs(integrate(p, w, f) / ref - 1) < 1e-6
    ps, ws = O.square_rule("vertex0", o)
    assert abs(integrate(ps, ws, f) / ref - 1) > 1e-5
    _, p2, w2 = O.quad_rule(qx, "vertex_square", corner=1)
    ps1, ws1 = O.square_rule("vertex1", o)
    assert np.array_equal(p2, ps1) and np.array_equal(w2, ws1)
    # T-junction: S covers half of T's side 1, folded by 30 degrees
    T = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0, 1, 0], [0.0, 1, 0]])
    c30, s30 = np.cos(np.radians(30)), np.sin(np.radians(30))
    S = np.array([[0.5, 0, 0], [0.0, 0, 0], [0.0, -0.5 * c30, 0.5 * s30], [0.5, -0.5 * c30, 0.5 * s30]])
    _, p, w = O.quad_rule(T, "contact", S, tn=(0, 1, 2, 3), sn=(8, 0, 9, 10))
    assert p is not None and abs(w.sum() - 1) < 1e-13
    assert abs(integrate(p, w, lambda p: p[:, 0] ** 3 * p[:, 1] ** 2) / mono2(3, 2) - 1) < 1e-12
    assert np.sum(p[:, 0] < 0.5) == len(O.square_rule("edge0", o)[1])
    nb = np.array([T[1], T[0], [0.0, -1.0, 0.2], [1.0, -1.0, 0.2]])
    assert O.quad_rule(T, "contact", nb, tn=(0, 1, 2, 3), sn=(1, 0, 11, 12))[1] is None


def block_cubature(T, S, V, maps, self_pair=False, tol=1e-8, lo=1e-3):
    """The pair block of source quad S on test quad T (hat weights) by adaptive cubature over T's parameter square in
    graded pieces (test/assembly/test_quads.jl:9-29); the inner integrals are the oracle's quad_field (pinned above).
    Returns the 16 block entries (row-major) then the 4 right-hand-side entries."""
    total = np.zeros(20)
    for m in maps:
        def f(uv, m=m):
            u = lo + (1 - lo) * uv
            p, jac = m(u)
            jac = jac * (1 - lo) ** 2
            nJ = mixed.area_normal(T, p[:, 0], p[:, 1])
            J = np.linalg.norm(nJ, axis=1)
            x = mixed.point(T, p[:, 0], p[:, 1])
            c, s = O.quad_field_many(S, x, nJ / J[:, None], V, side=1 if self_pair else 0,
                                     P=p if self_pair else None)
            psi = mixed.shape(p[:, 0], p[:, 1]) * (J * jac)[:, None]
            return np.concatenate([(psi[:, :, None] * c[:, None, :]).reshape(-1, 16), -s[:, None] * psi], 1)

        total += cubature.square(f, tol)[0]
    return total


def block_vector(T, S, p, w, V, side=0):
    blk, rb = O.quad_block(T, S, p, w, V, side=side)
    return np.concatenate([blk.ravel(), rb])


def test_touching_pairs_vs_cubature():
    """test/assembly/test_quads.jl:37-62: a shared side folded by 25 degrees (edge tier, 1e-5), a shared corner sheared
    to 150 degrees (vertex tier, 1e-6; the stored rule > 1e-5 off), the self pair of a warped quad (2e-6)"""
    x1, x2 = np.array([0.0, 0, 0]), np.array([1.0, 0, 0])
    c25, s25 = np.cos(np.radians(25)), np.sin(np.radians(25))
    T = np.array([x1, x2, [0.9, 0.8, 0.0], [0.1, 0.7, 0.0]])
    S = np.array([x2, x1, x1 + [0.05, -0.6 * c25, 0.6 * s25], x2 + [-0.05, -0.65 * c25, 0.65 * s25 + 0.05]])
    tier, p, w = O.quad_rule(T, "quad", S, tn=(0, 1, 2, 3), sn=(1, 0, 4, 5))
    assert tier == "edge"
    halves = (lambda u: (np.stack([0.5 * u[:, 0] ** 3, u[:, 1] ** 3], 1), 4.5 * (u[:, 0] * u[:, 1]) ** 2),
              lambda u: (np.stack([1 - 0.5 * u[:, 0] ** 3, u[:, 1] ** 3], 1), 4.5 * (u[:, 0] * u[:, 1]) ** 2))
    ref = block_cubature(T, S, VQ, halves)
    assert np.linalg.norm(block_vector(T, S, p, w, VQ) - ref) < 1e-5 * np.linalg.norm(ref)

    a4 = 0.4 * np.array([np.cos(np.radians(150)), np.sin(np.radians(150)), 0.0])
    T = np.array([x1, x2, np.array([1.1, 0, 0]) + a4, a4])
    S = np.array([x1, [-0.3, -0.7, 0.25], [0.4, -0.9, 0.3], [0.7, -0.4, 0.1]])
    tier, p, w = O.quad_rule(T, "quad", S, tn=(0, 1, 2, 3), sn=(0, 6, 7, 8))
    assert tier == "vertex"
    ref = block_cubature(T, S, VQ, (lambda u: (u ** 3, 9 * (u[:, 0] * u[:, 1]) ** 2),))
    assert np.linalg.norm(block_vector(T, S, p, w, VQ) - ref) < 1e-6 * np.linalg.norm(ref)
    ps, ws = O.square_rule("vertex0", O.opts())
    assert np.linalg.norm(block_vector(T, S, ps, ws, VQ) - ref) > 1e-5 * np.linalg.norm(ref)

    W = QUADS["warped"]
    quarters = []
    for i in (0, 1):
        for j in (0, 1):
            quarters.append(lambda u, i=i, j=j: (np.stack([0.5 * u[:, 0] ** 3 if i == 0 else 1 - 0.5 * u[:, 0] ** 3,
                                                           0.5 * u[:, 1] ** 3 if j == 0 else 1 - 0.5 * u[:, 1] ** 3], 1),
                                                 2.25 * (u[:, 0] * u[:, 1]) ** 2))
    ref = block_cubature(W, W, VQ, quarters, self_pair=True, tol=1e-7)
    ps, ws = O.square_rule("self", O.opts())
    assert np.linalg.norm(block_vector(W, W, ps, ws, VQ, side=1) - ref) < 2e-6 * np.linalg.norm(ref)


# -------------------------------------------------------------------------------------------- assembly -----------

def test_triangle_only_mixed_mesh_is_the_triangle_assembly():
    """With no quads the mixed assembly's rows are assemble's (test_tri!, src/assembly/mixed.jl:563-598, restates
    assemble.jl:377-447): the generic source path of the oracle against its triangle path, with a wake"""
    g = mesh.naca_wing(mesh.WingSpec(n_around=16, n_span=6, aspect=4.0))
    dm = dofmap.build_dofmap(g)
    wk = W.build_wake(g, dm, np.array([1.0, 0.0, 0.0]), W.WakeOptions(length=1.0))
    m = mixed.mixed_mesh(g, dm)
    assert m.nq == 0 and m.ntri == g.nt
    V = unit([1.0, 0.0, 0.07])
    el = O.elements(g.x, g.tris)
    A1, b1, p1, q1 = O.assemble(g, dm, O.source_strengths(el, V), V, wake=wk)
    A2, b2, p2, q2 = O.assemble_mixed(g.x, m, dm.nslot, V, touch=mixed.element_touch_pairs(g, m), wake=wk)
    assert np.abs(A1 - A2).max() <= 1e-14 * np.abs(A1).max() and np.abs(b1 - b2).max() <= 1e-14 * np.abs(b1).max()
    assert p1 == p2 and q1 == q2


@pytest.fixture(scope="module")
def cube6():
    X, F = mesh.cube_sphere_raw(6)
    g = geometry.build_geometry(X, F)
    dm = dofmap.build_dofmap(g)
    return g, dm, mixed.mixed_mesh(g, dm)


def test_mixed_mesh_contract(cube6):
    """MixedMesh (src/assembly/mixed.jl:26-58): 216 quads and no triangles on the 6 x 6 cube sphere; every geometry
    triangle in exactly one element; a quad's slots are its triangles' corner slots"""
    g, dm, m = cube6
    assert m.nq == 216 and m.ntri == 0
    assert np.all(m.element_of >= 0) and np.array_equal(np.bincount(m.element_of), np.full(216, 2))
    for e in range(m.nq):
        for t in m.quad_tris[e]:
            for k in range(3):
                j = list(m.quad_nodes[e]).index(g.tris[t][k])
                assert m.quad_slots[e][j] == dm.corner_slot[t][k]


def test_cube_sphere_exact_flow(cube6):
    """test/assembly/test_quads.jl:55-80: the quads' doublet error below 0.7 of the split triangles' (measured
    0.537; the reference's 0.54), column defect < 1e-5, lambda ~ 0, Cp at the quad centres within 0.03 of the exact
    sphere flow (the reference's 0.025)"""
    g, dm, m = cube6
    V = unit([1.0, 0.3, -0.2])
    flow = exact.EllipsoidFlow((1.0, 1.0, 1.0), V)
    Aq, bq, pairs, _ = O.assemble_mixed(g.x, m, dm.nslot, V, touch=mixed.element_touch_pairs(g, m))
    assert pairs["self"] == 216 and sum(pairs.values()) == 216 ** 2
    el = O.elements(g.x, g.tris)
    At, bt, _, _ = O.assemble(g, dm, O.source_strengths(el, V), V)
    ex = flow.doublet(g.x[dm.slot_node])

    def mu_error(A, b):
        s = O.solve_galerkin(dm, A, b)
        e = s["mu"] - ex
        e -= np.sum(dm.slot_area * e) / np.sum(dm.slot_area)
        return np.sqrt(np.sum(dm.slot_area * e ** 2) / np.sum(dm.slot_area)) / flow.doublet_scale(), s

    eq, sol = mu_error(Aq, bq)
    et, _ = mu_error(At, bt)
    assert eq < 0.7 * et
    assert O.consistency(dm, Aq, bq)["columns"] < 1e-5 and abs(sol["lam"][0]) < 1e-12
    pts = mixed.surface_pressure(m, sol["mu"], V, m=1)
    assert np.max(np.abs(pts["cp"] - flow.cp(pts["x"]))) < 0.03
