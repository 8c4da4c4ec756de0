"""The geometry pipeline (vphip/geometry.py) by the reference's own checks of build_geometry (test/geometry/
test_bodies.jl, test_units.jl, test_wing.jl and the sharp leading edge case of test_pathologies.jl), on the same raw
meshes where the reference builds them itself (icosphere, cube sphere, uv sphere, the diamond wing of test/helpers.jl)
and on this library's NACA wing as a raw quad mesh with a split trailing edge in place of the reference's PAGE wing
(a MeshLibrary file that is not part of the reference).  The wing built through the pipeline must reproduce the
generator's own surface triangle for triangle, chain and corner labels included."""
import numpy as np
import pytest

from vphip import dofmap, geometry as G, mesh
from vphip.mesh import NO_TE, TE_CAP, TE_LOWER, TE_UPPER


def tri_normals(g):
    a, b, c = g.x[g.tris[:, 0]], g.x[g.tris[:, 1]], g.x[g.tris[:, 2]]
    return np.cross(b - a, c - a)


def check_closed_sphere(g, nodes, tris, volume, rtol):
    """test/geometry/test_bodies.jl:1-13"""
    r = g.report
    assert g.nn == nodes and g.nt == tris
    assert len(r.bodies) == 1 and r.bodies[0].closed and r.bodies[0].euler == 2
    assert np.all(r.edge_class == G.EDGE_INTERIOR)
    assert not r.openings
    assert r.counts["kutta_edges"] == 0 and not g.chains
    assert r.bodies[0].volume == pytest.approx(volume, rel=rtol)
    cent = g.x[g.tris].mean(axis=1)
    assert np.all(np.einsum("ij,ij->i", tri_normals(g), cent) > 0)


def diamond_wing(n=4, t=0.1):
    """test/helpers.jl:53-67: rhombus section prism spanning y = +-1, sharp leading and trailing edges, tips open"""
    X = []
    for j in range(n + 1):
        y = -1 + 2 * j / n
        X += [(0.0, y, 0.0), (0.5, y, t / 2), (1.0, y, 0.0), (0.5, y, -t / 2)]
    faces = [[4 * j + k, 4 * (j + 1) + k, 4 * (j + 1) + (k + 1) % 4, 4 * j + (k + 1) % 4]
             for j in range(n) for k in range(4)]
    return np.array(X), faces


def quad_cube(split_face=False):
    """test/helpers.jl:6-22: unit cube of outward quads, the x = 1 face optionally split at z = 1/2 (T-junctions)"""
    X = [(i, j, k) for k in (0, 1) for j in (0, 1) for i in (0, 1)]
    nid = lambda i, j, k: i + 2 * j + 4 * k
    faces = [[nid(0, 0, 0), nid(0, 1, 0), nid(1, 1, 0), nid(1, 0, 0)],
             [nid(0, 0, 1), nid(1, 0, 1), nid(1, 1, 1), nid(0, 1, 1)],
             [nid(0, 0, 0), nid(1, 0, 0), nid(1, 0, 1), nid(0, 0, 1)],
             [nid(0, 1, 0), nid(0, 1, 1), nid(1, 1, 1), nid(1, 1, 0)],
             [nid(0, 0, 0), nid(0, 0, 1), nid(0, 1, 1), nid(0, 1, 0)]]
    if split_face:
        X += [(1.0, 0.0, 0.5), (1.0, 1.0, 0.5)]
        m0, m1 = len(X) - 2, len(X) - 1
        faces += [[nid(1, 0, 0), nid(1, 1, 0), m1, m0], [m0, m1, nid(1, 1, 1), nid(1, 0, 1)]]
    else:
        faces.append([nid(1, 0, 0), nid(1, 1, 0), nid(1, 1, 1), nid(1, 0, 1)])
    return np.array(X, dtype=np.float64), faces


# ---------------------------------------------------------------------------------------------- closed bodies


def test_icosphere():
    ico = mesh.icosphere(3)
    g = G.build_geometry(ico.x, ico.tris)
    check_closed_sphere(g, 642, 1280, 4 * np.pi / 3, 0.02)
    assert g.report.counts["welded_nodes"] == 0 and g.report.counts["flipped_triangles"] == 0
    assert np.array_equal(g.tris, ico.tris)


def test_cube_sphere():
    n = 4
    X, F = mesh.cube_sphere_raw(n)
    assert len(X) == 6 * (n + 1) ** 2
    g = G.build_geometry(X, F)
    check_closed_sphere(g, 6 * n * n + 2, 12 * n * n, 4 * np.pi / 3, 0.1)
    c = g.report.counts
    assert c["welded_nodes"] == 12 * n - 4  # 12(n-1) seam pairs and 8 corner triples
    assert c["split_fold"] + c["split_shortest"] + c["split_tie"] == 6 * n * n


def test_uv_sphere_collapsed_poles():
    nth, nph = 8, 12
    g = G.build_geometry(*mesh.uv_sphere_raw(nth, nph))
    check_closed_sphere(g, (nth - 1) * nph + 2, 2 * nph * (nth - 2) + 2 * nph, 4 * np.pi / 3, 0.1)
    assert g.report.counts["collapsed_elements"] == 2 * nph


def test_ellipsoid():
    ico = mesh.icosphere(3)
    g = G.build_geometry(ico.x * np.array([2.0, 1.0, 0.5]), ico.tris)
    check_closed_sphere(g, 642, 1280, 4 * np.pi / 3, 0.02)


def test_inward_input_is_flipped_outward():
    ico = mesh.icosphere(2)
    gout = G.build_geometry(ico.x, ico.tris)
    gin = G.build_geometry(ico.x, ico.tris[:, ::-1])
    assert gin.report.counts["flipped_triangles"] == gin.nt
    assert np.array_equal(G.canonical(gin.tris), G.canonical(gout.tris))


def test_polygons_are_fanned():
    """A hexagonal prism whose ends are single 6-gons: fanned into 4 triangles each, closed, its exact volume"""
    ang = np.arange(6) * np.pi / 3
    X = np.concatenate([np.stack([np.cos(ang), np.sin(ang), np.zeros(6)], 1),
                        np.stack([np.cos(ang), np.sin(ang), np.ones(6)], 1)])
    faces = [[k, (k + 1) % 6, 6 + (k + 1) % 6, 6 + k] for k in range(6)] + [list(range(5, -1, -1)), list(range(6, 12))]
    g = G.build_geometry(X, faces)
    assert g.nt == 12 + 8 and len(g.report.events("fan")) == 2
    assert g.report.bodies[0].closed and g.report.bodies[0].euler == 2
    assert g.report.bodies[0].volume == pytest.approx(1.5 * np.sqrt(3), rel=1e-12)


def test_degenerate_and_duplicate_faces_dropped():
    ico = mesh.icosphere(1)
    tris = [list(t) for t in ico.tris] + [list(ico.tris[3][::-1]), [0, 0, 5], [1, 2, 1]]
    g = G.build_geometry(ico.x, tris)
    assert g.report.counts["dropped_elements"] == 3
    assert np.array_equal(G.canonical(g.tris), G.canonical(ico.tris))


# ---------------------------------------------------------------------------------------------- units


def test_triangle_quality():
    """test/geometry/test_units.jl:1-13"""
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.5, np.sqrt(3) / 2, 0], [0.5, np.tan(np.radians(1)) / 2, 0]])
    qmin, qmax, asp, skew, area = G.triangle_quality(x, np.array([[0, 1, 2], [0, 1, 3]]))
    assert qmin[0] == pytest.approx(60) and qmax[0] == pytest.approx(60)
    assert asp[0] == pytest.approx(1) and abs(skew[0]) < 1e-12 and area[0] == pytest.approx(np.sqrt(3) / 4)
    assert qmin[1] == pytest.approx(1, abs=1e-9)
    assert qmin[1] < G.GeometryOptions().sliver_angle_deg


def test_quad_diagonal():
    """test/geometry/test_units.jl:15-34 (mirror axis y, 0-based)"""
    opts = G.GeometryOptions()
    x = np.array([[0.0, 0, 0], [2.0, 0, 0], [2.0, 1, 0], [0.0, 3, 0]])
    assert G.choose_diagonal(x, [0, 1, 2, 3], 0, opts) == (1, "shortest")
    assert G.choose_diagonal(x, [1, 2, 3, 0], 0, opts) == (2, "shortest")  # the same diagonal, nodes rotated
    w = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.2, 1.0, 0.4], [0.0, 1, 0]])
    d, reason = G.choose_diagonal(w, [0, 1, 2, 3], 0, opts)
    f13, f24 = G.fold_angle(w[0], w[1], w[2], w[3]), G.fold_angle(w[1], w[2], w[3], w[0])
    assert reason == "fold" and d == (1 if f13 < f24 else 2)
    sq = lambda y0: np.array([[0.0, y0, 0], [1.0, y0, 0], [1.0, y0 + 1, 0], [0.0, y0 + 1, 0]])
    dr, rr = G.choose_diagonal(sq(1.0), [0, 1, 2, 3], 0, opts)
    dl, rl = G.choose_diagonal(sq(-2.0), [0, 1, 2, 3], 0, opts)
    assert rr == "tie_mirror" and rl == "tie_mirror" and dr != dl
    # by hand at y0 = 1: offset 1.5, spread axis x; diagonal 0-2 is (1, 1), 1-3 is (-1, 1): 1.5 > -1.5 picks 0-2
    assert dr == 1
    assert G.choose_diagonal(sq(1.0), [0, 1, 2, 3], 0, G.GeometryOptions(split="d24")) == (2, "fixed")
    # a square straddling the mirror plane: the lowest node id decides
    assert G.choose_diagonal(sq(-0.5), [0, 1, 2, 3], 0, opts) == (1, "tie_id")


@pytest.mark.parametrize("scale,d,merged", [(1.0, 5e-4, True), (1.0, 2e-3, False), (1e-3, 5e-7, True),
                                            (1e3, 2.0, False)])
def test_weld_tolerance_is_relative(scale, d, merged):
    """test/geometry/test_units.jl:36-45"""
    X = scale * np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0], [1.0, 1, 0], [1.0, d / scale, 0]])
    xw, input_node, faces, _ = G.weld(X, [[0, 1, 2], [4, 3, 2]], 1e-3, [])
    assert (len(xw) == 4) == merged
    if merged:
        assert input_node[4] == input_node[1] and np.allclose(xw[input_node[1]], (X[1] + X[4]) / 2)


def test_weld_refuses_to_collapse_an_element():
    """src/geometry/weld.jl:64-75: nodes p and q of one element lie 1.5e-3 apart, beyond their tolerance (1e-3, from
    their shortest non-degenerate edge of length 1), and a node r of another element sits between them, 0.6e-3 from p
    and 0.9e-3 from q, within tolerance of both.  The closest pair p-r merges first; q-r is then refused, since it
    would join p and q."""
    d = 1.5e-3
    X = np.array([(0, 0, 0), (d, 0, 0), (0, 1, 0), (0.4 * d, 0, 0), (0.4 * d, 0, 1), (0.4 * d + 1, 0, 0),
                  (d / 2, -2, 0)], dtype=np.float64)
    log = []
    xw, input_node, faces, _ = G.weld(X, [[0, 1, 2], [3, 4, 5], [1, 0, 6]], 1e-3, log)
    assert input_node[3] == input_node[0] and input_node[1] != input_node[0]
    assert [k for k, _ in log].count("weld_refused") == 1
    assert np.allclose(xw[input_node[0]], (X[0] + X[3]) / 2)
    assert np.allclose(xw[input_node[1]], X[1])
    assert all(len(set(f)) == 3 for f in faces)


def test_ladder_by_hand():
    """src/geometry/cap.jl:28-79 on a hexagon whose chains share their ends: upper (0,0) (1,1) (2,1) (3,0), lower
    (0,0) (0.3,-1) (2.5,-1) (3,0).  After (A0, A1, B1) the diagonals are |A2 - B1| = 2.625 and |A1 - B2| = 2.5, so the
    ladder advances along B; then B is done and A advances; the last triangle closes at the trailing edge."""
    x = np.array([(0, 0, 0), (1, 1, 0), (2, 1, 0), (3, 0, 0), (0.3, -1, 0), (2.5, -1, 0)], dtype=np.float64)
    A, B = [0, 1, 2, 3], [0, 4, 5, 3]
    assert G.ladder(x, A, B, 1e-6) == [[0, 1, 4], [1, 5, 4], [1, 2, 5], [2, 3, 5]]
    # equal diagonals: the chain behind in normalised arclength advances (A reaches 0.631 of its length, B 0.736)
    xs = x.copy()
    xs[4] = (0.5, -1, 0)
    assert G.ladder(xs, A, B, 1e-6) == [[0, 1, 4], [1, 2, 4], [2, 5, 4], [2, 3, 5]]
    # one chain of a single segment: a fan from its far end
    assert G.ladder(x, [0, 3], B, 1e-6) == [[0, 4, 3], [4, 5, 3]]
    loop = [0, 1, 2, 3, 5, 4]  # boundary order; the cap runs the other way on every loop edge
    cap = G.orient_cap(G.ladder(x, A, B, 1e-6), loop)
    along = {(loop[k], loop[(k + 1) % 6]) for k in range(6)}
    assert all((t[k], t[(k + 1) % 3]) not in along for t in cap for k in range(3))


def test_cap_of_a_symmetric_opening():
    """src/geometry/cap.jl:147-181 on the hexagon (0,0) (1,1) (2,1) (3,0) (2,-1) (1,-1) in the x-z plane, trailing
    edge corner (3,0): every ladder step is a tie in length and arclength, and the ladder starts along the upper
    chain (the side facing up_ref), so by hand the diagonals are 1-5, 2-5 and 2-4"""
    x = np.array([(0, 0, 0), (1, 0, 1), (2, 0, 1), (3, 0, 0), (2, 0, -1), (1, 0, -1)], dtype=np.float64)
    op, tris = G.close_opening(x, [0, 1, 2, 3, 4, 5], True, {3}, "cap", G.GeometryOptions())
    assert (op.le, op.te, op.upper, op.lower) == (0, 3, [0, 1, 2, 3], [0, 5, 4, 3])
    edges = {tuple(sorted((t[k], t[(k + 1) % 3]))) for t in tris for k in range(3)}
    assert edges - {(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (0, 5)} == {(1, 5), (2, 5), (2, 4)}
    assert op.planarity < 1e-15


def test_non_manifold_edge():
    """test/geometry/test_units.jl:47-50"""
    X = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0], [0.0, -1, 0], [0.0, 0, 1]])
    with pytest.raises(G.NonManifoldEdgeError) as e:
        G.build_geometry(X, [[0, 1, 2], [1, 0, 3], [0, 1, 4]])
    assert e.value.edge == (0, 1) and e.value.count == 3


def test_non_orientable_surface():
    """A Moebius band: no consistent orientation exists (src/geometry/orient.jl:60-62)"""
    n = 12
    X = []
    for i in range(n):
        u = 2 * np.pi * i / n
        for v in (-0.3, 0.3):
            X.append(((1 + v * np.cos(u / 2)) * np.cos(u), (1 + v * np.cos(u / 2)) * np.sin(u), v * np.sin(u / 2)))
    faces = [[2 * i, 2 * i + 1, 2 * (i + 1) + 1, 2 * (i + 1)] for i in range(n - 1)] + [[2 * n - 2, 2 * n - 1, 0, 1]]
    with pytest.raises(G.GeometryError, match="Non-orientable"):
        G.build_geometry(np.array(X), faces)


def test_t_junctions_are_rejected():
    """The split-face cube of test/geometry/test_pathologies.jl:1-13: its hanging nodes need T-junction constraint
    rows, so the pipeline refuses it; the unsplit cube builds"""
    with pytest.raises(G.GeometryError, match="T-junction"):
        G.build_geometry(*quad_cube(split_face=True))
    g = G.build_geometry(*quad_cube())
    assert g.nt == 12 and g.report.bodies[0].closed and g.report.bodies[0].volume == pytest.approx(1.0, rel=1e-14)


# ---------------------------------------------------------------------------------------------- trailing edges


def test_sharp_leading_edge_is_not_a_trailing_edge():
    """test/geometry/test_pathologies.jl:76-84"""
    X, faces = diamond_wing(n=4)
    g = G.build_geometry(X, faces)
    r = g.report
    kutta = np.nonzero(r.edge_class == G.EDGE_KUTTA)[0]
    assert len(kutta) == 4
    assert np.allclose(g.x[r.edge_nodes[kutta]][..., 0], 1.0)
    assert len(r.openings) == 2 and r.bodies[0].closed
    assert all(g.x[op.te][0] == pytest.approx(1) and g.x[op.le][0] == pytest.approx(0) for op in r.openings)
    # the leading edge becomes a Kutta edge only when forced in
    le = [(4 * j, 4 * (j + 1)) for j in range(4)]
    gi = G.build_geometry(X, faces, G.GeometryOptions(kutta_include=tuple(le), wake_dir=(1.0, 0.0, 0.0)))
    assert gi.report.counts["kutta_edges"] == 8


def _pattern(spec):
    """The diagonal naca_wing takes for every quad (its rule: the shorter diagonal, 0-2 on ties)"""
    X, F = mesh.naca_wing_raw(spec)
    out = []
    for a, b, c, d in F:
        out.append(1 if np.linalg.norm(X[c] - X[a]) <= np.linalg.norm(X[d] - X[b]) * (1 + 1e-12) else 2)
    return tuple(out)


WINGS = {"rect0012": mesh.WingSpec(code="0012", n_around=24, n_span=9, aspect=8.0),
         "swept2412": mesh.WingSpec(code="2412", n_around=24, n_span=12, aspect=10.0, taper=0.5, sweep_c4_deg=25.0)}


@pytest.mark.parametrize("name", list(WINGS))
def test_wing_reproduces_the_generator(name):
    """test/geometry/test_wing.jl:4-46: the raw wing (split trailing edge, open tips) welds the trailing edge copies,
    caps both tips and finds one Kutta chain along +y ending on the caps; with the generator's diagonals its surface
    triangles, Kutta chain and corner labels are the generator's own, and the DOFMap agrees"""
    spec = WINGS[name]
    N, ns = spec.n_around, spec.n_span
    ref = mesh.naca_wing(spec)
    X, F = mesh.naca_wing_raw(spec, split_te=True)
    g = G.build_geometry(X, F, G.GeometryOptions(split="pattern", split_pattern=_pattern(spec)))
    r = g.report
    c = r.counts
    nbody = 2 * N * ns
    assert g.nn == ref.nn and c["welded_nodes"] == ns + 1 and c["flipped_triangles"] == 0
    assert g.nt == nbody + 2 * (N - 2) and c["cap_triangles"] == 2 * (N - 2)
    assert np.array_equal(g.x, ref.x)
    assert np.array_equal(g.tris[:nbody], ref.tris[:nbody])
    assert len(r.bodies) == 1 and r.bodies[0].closed and r.bodies[0].euler == 2
    assert r.bodies[0].volume == pytest.approx(_signed_volume(ref), rel=1e-2)
    assert c["kutta_edges"] == ns and c["te_nodes"] == ns + 1 and len(g.chains) == 1
    ch, cr = g.chains[0], ref.chains[0]
    assert np.array_equal(ch.nodes, cr.nodes) and not ch.closed
    assert np.array_equal(ch.tri_upper, cr.tri_upper) and np.array_equal(ch.tri_lower, cr.tri_lower)
    assert ch.end_type == ("cap", "cap") and min(ch.end_cap_tri) >= nbody
    assert (g.x[ch.nodes[-1]] - g.x[ch.nodes[0]])[1] > 0
    assert np.array_equal(g.corner_side[:nbody], ref.corner_side[:nbody])
    te = set(ch.nodes.tolist())
    for t in range(g.nt):
        for k in range(3):
            v = int(g.tris[t, k])
            if v not in te:
                assert g.corner_side[t, k] == NO_TE
            elif t >= nbody:
                assert g.corner_side[t, k] == TE_CAP
            else:
                assert g.corner_side[t, k] == (TE_UPPER if (t // 2) % N >= N // 2 else TE_LOWER)
    assert np.array_equal(g.tri_cap, np.arange(g.nt) >= nbody)
    # tip caps: flat, outward, the leading edge at the chord's front and the trailing edge at the Kutta chain
    for op in r.openings:
        assert op.mode == "cap" and len(op.loop) == N and op.planarity < 1e-12
        assert op.te in te
        n = tri_normals(g)[list(op.cap_tris)]
        sgn = np.sign(g.x[op.te][1])
        assert np.allclose(n[:, [0, 2]], 0, atol=1e-15) and np.all(n[:, 1] * sgn > 0)
    assert np.sum(r.edge_class == G.EDGE_CAP_SEAM) == 2 * N
    dm, dr = dofmap.build_dofmap(g), dofmap.build_dofmap(ref)
    assert dm.nq == dr.nq and dm.nslot == dr.nslot
    assert np.array_equal(dm.wake_te.node, dr.wake_te.node)
    assert (dm.K != dr.K).nnz == 0


def _signed_volume(s):
    a, b, c = s.x[s.tris[:, 0]], s.x[s.tris[:, 1]], s.x[s.tris[:, 2]]
    return np.sum(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6


def test_wing_triangulation_invariance():
    """test/geometry/test_wing.jl:61-80: the same triangles from the unsplit trailing edge, reversed elements and
    scaled nodes; mirror symmetric about y = 0 except the strip that straddles it"""
    spec = mesh.WingSpec(code="0012", n_around=20, n_span=7, aspect=6.0)
    X, F = mesh.naca_wing_raw(spec, split_te=True)
    g = G.build_geometry(X, F)
    ref = G.canonical(g.tris)
    assert np.array_equal(G.canonical(G.build_geometry(*mesh.naca_wing_raw(spec)).tris), ref)
    assert np.array_equal(G.canonical(G.build_geometry(X, [f[::-1] for f in F]).tris), ref)
    for s in (1e-3, 1e3):
        gs = G.build_geometry(s * X, F)
        assert np.array_equal(G.canonical(gs.tris), ref)
        assert gs.report.bodies[0].volume == pytest.approx(s ** 3 * g.report.bodies[0].volume, rel=1e-10)
    key = lambda p: tuple(np.round(p, 9))
    cent = g.x[g.tris].mean(axis=1)
    have = {key(c) for c in cent}
    y = g.x[g.tris][..., 1]
    straddles = (y.min(axis=1) < 0) & (y.max(axis=1) > 0)
    asym = [t for t in range(g.nt) if key(cent[t] * np.array([1, -1, 1])) not in have]
    assert all(straddles[t] for t in asym)
    assert len(asym) <= 2 * spec.n_around


def test_open_mode():
    """test/geometry/test_wing.jl:82-94, the open mode"""
    spec = mesh.WingSpec(code="0012", n_around=16, n_span=5, aspect=4.0)
    g = G.build_geometry(*mesh.naca_wing_raw(spec, split_te=True), G.GeometryOptions(openings="open"))
    r = g.report
    assert all(op.mode == "open" and len(op.cap_tris) == 0 for op in r.openings)
    assert not r.bodies[0].closed and not g.bodies_closed[0]
    assert np.sum(r.edge_class == G.EDGE_OPEN) == 2 * spec.n_around
    assert g.nt == 2 * spec.n_around * spec.n_span
    assert g.chains[0].end_type == ("open", "open")


def test_kutta_exclude_splits_the_chain():
    spec = mesh.WingSpec(code="0012", n_around=16, n_span=6, aspect=4.0)
    X, F = mesh.naca_wing_raw(spec)
    N = spec.n_around
    g = G.build_geometry(X, F, G.GeometryOptions(kutta_exclude=((3 * N, 4 * N),)))
    assert g.report.counts["kutta_edges"] == spec.n_span - 1 and len(g.chains) == 2
    assert sorted(ch.end_type for ch in g.chains) == [("cap", "dangling"), ("dangling", "cap")]
    # corners at the dangling ends still carry their surface label, and every chain segment separates the surfaces
    for ch in g.chains:
        for tu, tl in zip(ch.tri_upper, ch.tri_lower):
            assert (tu // 2) % N >= N // 2 > (tl // 2) % N


def test_notched_trailing_edge():
    """A section with a V notch in its trailing edge: the two tips are sharp convex wedges pointing downstream (Kutta
    edges, two chains), the notch apex is a sharp wedge whose bisector also points downstream but which is concave, so
    it is not a trailing edge (the convexity test of src/geometry/kutta.jl:22)"""
    sec = [(0.0, 0.0), (0.5, 0.1), (1.0, 0.05), (0.7, 0.0), (1.0, -0.05), (0.5, -0.1)]  # counterclockwise in (x, z)
    n, m = 4, len(sec)
    X = np.array([(px, -1 + 2 * j / n, pz) for j in range(n + 1) for px, pz in sec])
    faces = [[m * j + k, m * j + (k + 1) % m, m * (j + 1) + (k + 1) % m, m * (j + 1) + k]
             for j in range(n) for k in range(m)]
    g = G.build_geometry(X, faces)
    r = g.report
    assert r.counts["kutta_edges"] == 2 * n and len(g.chains) == 2
    ke = r.edge_nodes[r.edge_class == G.EDGE_KUTTA]
    assert np.allclose(g.x[ke][..., 0], 1.0)
    assert sorted(round(float(g.x[ch.nodes[0]][2]), 6) for ch in g.chains) == [-0.05, 0.05]
    assert r.bodies[0].closed and r.bodies[0].euler == 2


def test_wing_down_wake_reverses_tau():
    """With the wake leaving along -x the trailing edge is the leading edge of the section, which is not sharp: no
    Kutta edges; with up_ref along -z, tau runs along -y and the sides swap"""
    spec = mesh.WingSpec(code="0012", n_around=16, n_span=6, aspect=4.0)
    X, F = mesh.naca_wing_raw(spec)
    assert G.build_geometry(X, F, G.GeometryOptions(wake_dir=(-1.0, 0.0, 0.0))).report.counts["kutta_edges"] == 0
    g = G.build_geometry(X, F, G.GeometryOptions(up_ref=(0.0, 0.0, -1.0)))
    ch = g.chains[0]
    assert (g.x[ch.nodes[-1]] - g.x[ch.nodes[0]])[1] < 0
    N = spec.n_around
    assert all((tu // 2) % N < N // 2 for tu in ch.tri_upper)
