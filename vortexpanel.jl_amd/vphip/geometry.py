"""Geometry pipeline: a raw polygon surface into the `Surface` the hot path reads.

A raw surface is node coordinates and faces of three or more nodes, in any orientation, possibly unwelded (every face
on its own nodes) and with open ends.  `build_geometry` welds, triangulates, orients, finds the Kutta (trailing)
edges, caps the openings and labels the corners at the trailing edge, following src/geometry/pipeline.jl:227-275 for
the stages it carries:

- weld (src/geometry/weld.jl:19-129): nodes closer than `weld_rel` times their shortest incident edge merge, closest
  pairs first, refused when the merge would join two nodes of one face that are farther apart than their tolerance;
  faces then collapse repeated nodes, and degenerate and duplicate faces are dropped;
- quad split (src/geometry/quadsplit.jl:29-110): the shorter diagonal, or the one with the smaller fold, with a tie
  rule that is invariant to node order and mirror symmetric; larger polygons are fanned;
- half-edges (src/geometry/halfedge.jl:27-62): an edge of more than two faces raises `NonManifoldEdgeError`;
- orientation (src/geometry/orient.jl:32-117): consistent across shared edges, then outward by signed volume;
- Kutta edges (src/geometry/kutta.jl:18-61): the dihedral wedge test, then the include and exclude lists;
- openings (src/geometry/openings.jl:14-116, src/geometry/cap.jl:1-181): boundary loops capped with a ladder of
  triangles on their own nodes (`openings="cap"`) or left open (`"open"`);
- Kutta chains and corner sides (src/geometry/kutta.jl:90-146, src/geometry/pipeline.jl:20-39 and :150-177).

Not carried: T-junctions, small gaps and non-matching trailing edges (src/geometry/pairs.jl), zipped openings and
input edge tags.  The first three need constraint rows the indexing contract here does not build, so their presence
(detected with the reference's own criteria, src/geometry/pairs.jl:30-60 and :89-102) raises `GeometryError` instead
of giving a surface the solver would treat wrongly.

Indices are 0-based.  Host code: the pipeline runs once per geometry, before any device work.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.spatial import cKDTree

from .mesh import NO_TE, TE_CAP, TE_LOWER, TE_UPPER, KuttaChain, Surface

BODY_TRI, CAP_TRI = 0, 1
EDGE_INTERIOR, EDGE_KUTTA, EDGE_CAP_SEAM, EDGE_OPEN = 0, 1, 2, 3


class GeometryError(ValueError):
    """A surface the pipeline cannot turn into a valid `Surface`"""


class NonManifoldEdgeError(GeometryError):
    """An edge shared by more than two faces (src/geometry/halfedge.jl:41-46)"""

    def __init__(self, edge, points, count):
        self.edge, self.points, self.count = edge, points, count
        super().__init__(f"edge {edge} at {points[0]} - {points[1]} is shared by {count} faces")


@dataclass(frozen=True)
class GeometryOptions:
    """The switches of src/geometry/options.jl:1-59 that this pipeline carries.

    weld_rel: nodes closer than this times their shortest incident edge merge.
    split: "auto" (shorter diagonal, or the smaller fold when either fold exceeds split_fold_deg), "shortest",
        "min_fold", "d13", "d24" or "pattern" (diagonal of face f from split_pattern[f], 1 for nodes 0-2, 2 for 1-3).
    split_tie_rel: diagonals within this relative length (or fold, in radians) are tied; ties are broken by the
        mirror rule about the plane x[split_mirror[0]] = split_mirror[1] (None to skip it), then by node id.
    tj_rel, tj_tmin, gap_rel, gap_angle_deg, gap_overlap: the T-junction and gap criteria, used to reject them.
    openings: "cap" or "open".
    te_angle_deg: a convex edge whose faces meet at an included angle of at most this, pointing along wake_dir, is
        a Kutta edge; kutta_include / kutta_exclude force input node pairs in or out.
    wake_dir, up_ref: the trailing edge direction tau satisfies (wake_dir x tau) . up_ref >= 0, so that "upper" is
        the side facing up_ref.
    sliver_angle_deg, sliver_aspect: quality thresholds of the report.
    """
    weld_rel: float = 1e-3
    split: str = "auto"
    split_pattern: tuple = ()
    split_fold_deg: float = 1.0
    split_tie_rel: float = 1e-6
    split_mirror: tuple | None = (1, 0.0)
    tj_rel: float = 1e-3
    tj_tmin: float = 1e-3
    gap_rel: float = 0.05
    gap_angle_deg: float = 10.0
    gap_overlap: float = 0.25
    openings: str = "cap"
    te_angle_deg: float = 30.0
    kutta_include: tuple = ()
    kutta_exclude: tuple = ()
    wake_dir: tuple = (1.0, 0.0, 0.0)
    up_ref: tuple = (0.0, 0.0, 1.0)
    sliver_angle_deg: float = 5.0
    sliver_aspect: float = 20.0

    def check(self):
        """src/geometry/options.jl:61-66"""
        if self.split not in ("auto", "shortest", "min_fold", "d13", "d24", "pattern"):
            raise ValueError(f"Unknown split rule {self.split}")
        if self.openings not in ("cap", "open"):
            raise ValueError(f"Unknown openings mode {self.openings} (cap or open; zipped openings are not built)")
        return self


@dataclass
class Opening:
    """src/geometry/types.jl:62-76: loop nodes in boundary half-edge order, corners, the two chains from the leading
    to the trailing edge corner, the mode and the cap triangles"""
    loop: list
    le: int
    te: int
    upper: list
    lower: list
    mode: str
    cap_tris: range
    planarity: float


@dataclass
class Body:
    """src/geometry/types.jl:104-111"""
    tris: np.ndarray
    nodes: np.ndarray
    closed: bool
    euler: int
    volume: float


@dataclass
class GeometryReport:
    """What the pipeline did: the welded id of every input node, the input face of every triangle (-1 for caps),
    edges with their classes, openings, bodies, parent quads, counts, quality and the repair log"""
    input_node: np.ndarray
    tri_parent: np.ndarray
    tri_slot: np.ndarray
    tri_kind: np.ndarray
    edge_nodes: np.ndarray
    edge_class: np.ndarray
    openings: list
    bodies: list
    quads: np.ndarray
    quad_parent: np.ndarray
    quad_diag: np.ndarray
    te_nodes: np.ndarray
    counts: dict
    stats: dict
    quality: dict
    log: list = field(default_factory=list)

    def events(self, kind):
        return [m for k, m in self.log if k == kind]


# --------------------------------------------------------------------------------------------------------------------
# small helpers


class _UnionFind:
    """src/geometry/util.jl:66-96: the smaller root wins"""

    def __init__(self, n):
        self.p = list(range(n))

    def find(self, i):
        p = self.p
        while p[i] != i:
            p[i] = p[p[i]]
            i = p[i]
        return i

    def unite(self, i, j):
        ri, rj = self.find(i), self.find(j)
        if ri == rj:
            return ri
        r, s = min(ri, rj), max(ri, rj)
        self.p[s] = r
        return r

    def components(self):
        groups = {}
        for i in range(len(self.p)):
            groups.setdefault(self.find(i), []).append(i)
        return [groups[r] for r in sorted(groups)]


def _norm(v):
    return math.sqrt(float(v[0]) ** 2 + float(v[1]) ** 2 + float(v[2]) ** 2)


def _tri_normals(x, tris):
    n = np.cross(x[tris[:, 1]] - x[tris[:, 0]], x[tris[:, 2]] - x[tris[:, 0]])
    L = np.linalg.norm(n, axis=1)
    return n / np.where(L > 0, L, 1.0)[:, None]


def _point_segment(p, a, b):
    """Distance of p from segment [a, b] and the parameter of the nearest point"""
    d = b - a
    L2 = float(d @ d)
    t = 0.0 if L2 == 0 else min(max(float((p - a) @ d) / L2, 0.0), 1.0)
    return _norm(p - (a + t * d)), t


def _signed_volume(x, tris, ts, centre):
    """src/geometry/orient.jl:8-16"""
    t = tris[ts]
    a, b, c = x[t[:, 0]] - centre, x[t[:, 1]] - centre, x[t[:, 2]] - centre
    return float(np.sum(np.einsum("ij,ij->i", a, np.cross(b, c)))) / 6


# --------------------------------------------------------------------------------------------------------------------
# weld and split


def weld(x, faces, rel, log):
    """src/geometry/weld.jl:19-129.  Returns welded positions, the welded id of every input node, the surviving faces
    in welded ids and their input face ids."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    ea, eb, ef = [], [], []
    for f, face in enumerate(faces):
        m = len(face)
        for k in range(m):
            ea.append(face[k])
            eb.append(face[(k + 1) % m])
            ef.append(f)
    ea, eb, ef = np.array(ea, np.int64), np.array(eb, np.int64), np.array(ef, np.int64)
    node_faces = [set() for _ in range(n)]
    for a, f in zip(ea.tolist(), ef.tolist()):
        node_faces[a].add(f)
    keep = ea != eb
    ka, kb = ea[keep], eb[keep]
    L = np.linalg.norm(x[ka] - x[kb], axis=1)
    Lmax = np.zeros(n)
    np.maximum.at(Lmax, ka, L)
    np.maximum.at(Lmax, kb, L)
    ell = np.full(n, np.inf)
    for v in (ka, kb):
        ok = L > rel * Lmax[v]
        np.minimum.at(ell, v[ok], L[ok])
    finite = np.sort(ell[np.isfinite(ell)])
    ell_default = finite[(len(finite) + 1) // 2 - 1] if len(finite) else 1.0
    tol = rel * np.where(np.isfinite(ell), ell, ell_default)

    pairs = np.zeros((0, 2), np.int64)
    if n > 1:
        pairs = cKDTree(x).query_pairs(r=max(float(tol.max()), np.finfo(float).eps), output_type="ndarray")
    if len(pairs):
        d = np.linalg.norm(x[pairs[:, 0]] - x[pairs[:, 1]], axis=1)
        ok = d <= np.minimum(tol[pairs[:, 0]], tol[pairs[:, 1]])
        pairs, d = pairs[ok], d[ok]
        i, j = pairs.min(axis=1), pairs.max(axis=1)
        order = np.lexsort((j, i, d))
        cand = list(zip(d[order].tolist(), i[order].tolist(), j[order].tolist()))
    else:
        cand = []

    uf = _UnionFind(n)
    members = {}
    for dd, i, j in cand:
        ri, rj = uf.find(i), uf.find(j)
        if ri == rj:
            continue
        mi, mj = members.get(ri, [ri]), members.get(rj, [rj])
        refused = any(node_faces[p] & node_faces[q] and _norm(x[p] - x[q]) > min(tol[p], tol[q])
                      for p in mi for q in mj)
        if refused:
            log.append(("weld_refused", f"nodes {i} and {j} are within tolerance but would merge nodes of one element"))
            continue
        r = uf.unite(ri, rj)
        s = rj if r == ri else ri
        members[r] = mi + mj if r == ri else mj + mi
        members.pop(s, None)

    roots = np.array([uf.find(i) for i in range(n)], np.int64)
    reps = np.nonzero(roots == np.arange(n))[0]
    new_id = np.full(n, -1, np.int64)
    new_id[reps] = np.arange(len(reps))
    input_node = new_id[roots]
    xw = np.zeros((len(reps), 3))
    np.add.at(xw, input_node, x)
    xw /= np.bincount(input_node, minlength=len(reps))[:, None]
    for r, mem in members.items():
        if len(mem) > 1:
            P = x[mem]
            spread = float(np.max(np.linalg.norm(P[:, None] - P[None], axis=2)))
            log.append(("weld", f"merged input nodes {sorted(mem)} (spread {spread:.3g})"))
            if spread > float(np.min(tol[mem])):
                log.append(("weld_wide", f"merged cluster {sorted(mem)} wider than its tolerance"))

    out, parent, seen = [], [], {}
    for f, face in enumerate(faces):
        c = []
        for i in face:
            v = int(input_node[i])
            if not c or c[-1] != v:
                c.append(v)
        if len(c) > 1 and c[0] == c[-1]:
            c.pop()
        if len(c) < 3 or len(set(c)) != len(c):
            log.append(("drop_degenerate", f"input element {f} collapses to nodes {c}"))
            continue
        if len(c) < len(face):
            log.append(("collapse", f"input element {f}: {len(face)} -> {len(c)} nodes"))
        key = tuple(sorted(c))
        if key in seen:
            log.append(("drop_duplicate", f"input element {f} duplicates input element {seen[key]}"))
            continue
        seen[key] = f
        out.append(c)
        parent.append(f)
    return xw, input_node, out, parent


def _sub(u, v):
    return (u[0] - v[0], u[1] - v[1], u[2] - v[2])


def _cross(u, v):
    return (u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0])


def _dot(u, v):
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]


def fold_angle(a, b, c, d):
    """src/geometry/quadsplit.jl:13-20: angle between the normals of (a, b, c) and (a, c, d)"""
    n1 = _cross(_sub(b, a), _sub(c, a))
    n2 = _cross(_sub(c, a), _sub(d, a))
    m = math.sqrt(_dot(n1, n1) * _dot(n2, n2))
    if m <= 0:
        return 0.0
    return math.acos(min(max(_dot(n1, n2) / m, -1.0), 1.0))


def choose_diagonal(x, q, parent, opts: GeometryOptions):
    """src/geometry/quadsplit.jl:29-67: (1 for diagonal q0-q2, 2 for q1-q3, reason)"""
    if opts.split == "d13":
        return 1, "fixed"
    if opts.split == "d24":
        return 2, "fixed"
    if opts.split == "pattern":
        d = int(opts.split_pattern[parent])
        if d not in (1, 2):
            raise ValueError(f"split_pattern[{parent}] = {d} is not 1 or 2")
        return d, "fixed"
    a, b, c, d = (tuple(float(w) for w in x[v]) for v in q)
    L13, L24 = _norm(_sub(c, a)), _norm(_sub(d, b))
    tie = opts.split_tie_rel
    if opts.split in ("min_fold", "auto"):
        f13 = fold_angle(a, b, c, d)
        f24 = fold_angle(b, c, d, a)
        if opts.split == "min_fold" or max(f13, f24) > math.radians(opts.split_fold_deg):
            if abs(f13 - f24) > tie:
                return (1 if f13 < f24 else 2), "fold"
    if abs(L13 - L24) > tie * max(L13, L24):
        return (1 if L13 < L24 else 2), "shortest"
    if opts.split_mirror is not None:
        m, m0 = opts.split_mirror
        size = max(L13, L24)
        offset = (a[m] + b[m] + c[m] + d[m]) / 4 - m0
        if abs(offset) > tie * size:
            ext = [max(a[k], b[k], c[k], d[k]) - min(a[k], b[k], c[k], d[k]) for k in range(3)]
            others = [k for k in range(3) if k != m]
            p = others[0] if ext[others[0]] >= ext[others[1]] else others[1]
            v13, v24 = _sub(c, a), _sub(d, b)
            k13 = v13[m] * v13[p] * offset
            k24 = v24[m] * v24[p] * offset
            if abs(k13 - k24) > tie * size ** 2 * abs(offset):
                return (1 if k13 > k24 else 2), "tie_mirror"
    return (1 if min(q[0], q[2]) < min(q[1], q[3]) else 2), "tie_id"


def split_faces(x, faces, parents, opts, log):
    """src/geometry/quadsplit.jl:69-110"""
    tris, tri_parent, tri_slot = [], [], []
    quads, quad_parent, quad_diag, quad_reason = [], [], [], []
    for face, parent in zip(faces, parents):
        n = len(face)
        if n == 3:
            tris.append(list(face))
            tri_parent.append(parent)
            tri_slot.append(0)
        elif n == 4:
            a, b, c, d = face
            diag, reason = choose_diagonal(x, face, parent, opts)
            halves = ([a, b, c], [a, c, d]) if diag == 1 else ([a, b, d], [b, c, d])
            for s, t in enumerate(halves):
                tris.append(t)
                tri_parent.append(parent)
                tri_slot.append(s + 1)
            quads.append([a, b, c, d])
            quad_parent.append(parent)
            quad_diag.append(diag)
            quad_reason.append(reason)
        else:
            log.append(("fan", f"input element {parent}: {n}-gon fanned into {n - 2} triangles"))
            for k in range(1, n - 1):
                tris.append([face[0], face[k], face[k + 1]])
                tri_parent.append(parent)
                tri_slot.append(0)
    return (np.array(tris, np.int64).reshape(-1, 3), np.array(tri_parent, np.int64), np.array(tri_slot, np.int8),
            np.array(quads, np.int64).reshape(-1, 4), np.array(quad_parent, np.int64), np.array(quad_diag, np.int8),
            quad_reason)


# --------------------------------------------------------------------------------------------------------------------
# half-edges: half-edge h = 3t + k runs from tris[t][k] to tris[t][(k + 1) % 3]


class HalfEdges:
    """src/geometry/halfedge.jl:12-62: twins (-1 on the boundary), the edge of every half-edge, and the unique edges
    (sorted node pairs, in order of first appearance) with their one or two half-edges"""

    def __init__(self, x, tris):
        nt = len(tris)
        a = tris.reshape(-1)
        b = tris[:, [1, 2, 0]].reshape(-1)
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        nn = int(max(len(x), (tris.max() + 1) if nt else 0))
        key = lo * nn + hi
        _, first, inv, counts = np.unique(key, return_index=True, return_inverse=True, return_counts=True)
        order = np.argsort(first, kind="stable")
        rank = np.empty_like(order)
        rank[order] = np.arange(len(order))
        self.he_edge = rank[inv.reshape(-1)]
        counts = counts[order]
        self.edge_nodes = np.stack([lo, hi], axis=1)[first[order]]
        if len(counts) and counts.max() > 2:
            e = int(np.argmax(counts > 2))
            u, v = (int(w) for w in self.edge_nodes[e])
            raise NonManifoldEdgeError((u, v), (x[u], x[v]), int(counts[e]))
        s = np.argsort(self.he_edge, kind="stable")
        start = np.concatenate([[0], np.cumsum(counts)])
        ne = len(counts)
        self.edge_he = np.full((ne, 2), -1, np.int64)
        self.edge_he[:, 0] = s[start[:-1]]
        two = counts == 2
        self.edge_he[two, 1] = s[start[:-1][two] + 1]
        self.twin = np.full(3 * nt, -1, np.int64)
        h1, h2 = self.edge_he[two, 0], self.edge_he[two, 1]
        self.twin[h1] = h2
        self.twin[h2] = h1
        self.tris = tris
        self._edge_of = None

    @property
    def ne(self):
        return len(self.edge_nodes)

    def edge_id(self, a, b):
        """Edge of the node pair in either order, -1 if there is none"""
        if self._edge_of is None:
            self._edge_of = {(int(u), int(v)): e for e, (u, v) in enumerate(self.edge_nodes.tolist())}
        return self._edge_of.get((min(a, b), max(a, b)), -1)

    def origin(self, h):
        return int(self.tris[h // 3, h % 3])

    def dest(self, h):
        return int(self.tris[h // 3, (h % 3 + 1) % 3])


def _he_next(h):
    return 3 * (h // 3) + (h % 3 + 1) % 3


def _he_prev(h):
    return 3 * (h // 3) + (h % 3 + 2) % 3


def vertex_fans(nn, tris, twin):
    """src/geometry/halfedge.jl:71-97: number of edge-connected fans at every node (more than one: a bow-tie)"""
    node_tris = [[] for _ in range(nn)]
    for t, tri in enumerate(tris.tolist()):
        for v in tri:
            node_tris[v].append(t)
    fans = np.zeros(nn, np.int64)
    for v in range(nn):
        ts = node_tris[v]
        if not ts:
            continue
        if len(ts) == 1:
            fans[v] = 1
            continue
        local = {t: i for i, t in enumerate(ts)}
        uf = _UnionFind(len(ts))
        for i, t in enumerate(ts):
            for k in range(3):
                if tris[t, k] != v and tris[t, (k + 1) % 3] != v:
                    continue
                g = twin[3 * t + k]
                if g >= 0 and (g // 3) in local:
                    uf.unite(i, local[g // 3])
        fans[v] = len(uf.components())
    return fans


# --------------------------------------------------------------------------------------------------------------------
# orientation


def orient(x, tris, he: HalfEdges, log):
    """src/geometry/orient.jl:32-117 without the T-junction and gap links: consistent flips within every
    edge-connected patch, then each patch (a body) flipped outward by its signed volume"""
    nt = len(tris)
    flip = np.zeros(nt, bool)
    patch = np.full(nt, -1, np.int64)
    npatch = 0
    for seed in range(nt):
        if patch[seed] >= 0:
            continue
        patch[seed] = npatch
        stack = [seed]
        while stack:
            t = stack.pop()
            for k in range(3):
                h = 3 * t + k
                g = int(he.twin[h])
                if g < 0:
                    continue
                u = g // 3
                want = bool(flip[t]) ^ (he.origin(h) == he.origin(g))
                if patch[u] < 0:
                    patch[u] = npatch
                    flip[u] = want
                    stack.append(u)
                elif flip[u] != want:
                    a, b = he.origin(h), he.dest(h)
                    raise GeometryError(f"Non-orientable surface: cannot orient triangles {t} and {u} consistently "
                                        f"across edge ({a}, {b}) at {x[a]}")
        npatch += 1
    oriented = np.where(flip[:, None], tris[:, [0, 2, 1]], tris)
    volume = np.zeros(npatch)
    for p in range(npatch):
        ts = np.nonzero(patch == p)[0]
        nodes = np.unique(oriented[ts])
        centre = x[nodes].mean(axis=0)
        V = _signed_volume(x, oriented, ts, centre)
        if V < 0:
            flip[ts] = ~flip[ts]
            V = -V
        ext = float(np.max(x[nodes].max(axis=0) - x[nodes].min(axis=0)))
        if V <= math.sqrt(np.finfo(float).eps) * ext ** 3:
            log.append(("flat_body", f"body {p} encloses (almost) no volume; its outward direction is arbitrary"))
        volume[p] = V
    return flip, patch, volume


# --------------------------------------------------------------------------------------------------------------------
# T-junctions and gaps: detected with the reference's criteria and rejected


def _boundary_pathologies(x, he: HalfEdges, opts: GeometryOptions):
    """T-junctions (src/geometry/pairs.jl:30-60) and paired boundary edges (src/geometry/pairs.jl:89-102, the gap
    and non-matching trailing edge candidates), as messages"""
    bed = np.nonzero(he.edge_he[:, 1] < 0)[0]
    if len(bed) == 0:
        return []
    found = []
    en = he.edge_nodes[bed]
    bnodes = np.unique(en)
    tree = cKDTree(x[bnodes])
    L = np.linalg.norm(x[en[:, 1]] - x[en[:, 0]], axis=1)
    mid = (x[en[:, 0]] + x[en[:, 1]]) / 2
    for i in range(len(bed)):
        a, b = int(en[i, 0]), int(en[i, 1])
        tol = opts.tj_rel * L[i]
        for j in tree.query_ball_point(mid[i], L[i] / 2 + tol):
            v = int(bnodes[j])
            if v in (a, b):
                continue
            d, t = _point_segment(x[v], x[a], x[b])
            if d <= tol and opts.tj_tmin < t < 1 - opts.tj_tmin:
                found.append(f"node {v} hangs on boundary edge ({a}, {b}) (a T-junction)")
    if found:
        return found
    cos_gap = math.cos(math.radians(opts.gap_angle_deg))
    Lmax = float(L.max())
    mtree = cKDTree(mid)
    for i in range(len(bed)):
        p, q = int(en[i, 0]), int(en[i, 1])
        for j in mtree.query_ball_point(mid[i], L[i] / 2 + Lmax / 2 + opts.gap_rel * L[i]):
            if j <= i:
                continue
            r, s = int(en[j, 0]), int(en[j, 1])
            if {p, q} & {r, s}:
                continue
            if _edge_pair(x, p, q, r, s, cos_gap, opts):
                found.append(f"boundary edges ({p}, {q}) and ({r}, {s}) face each other across a gap")
    return found


def _overlap_on(xp, xq, xr, xs):
    L = _norm(xq - xp)
    u = (xq - xp) / L
    tr, ts = float((xr - xp) @ u), float((xs - xp) @ u)
    return max(0.0, min(tr, ts)), min(L, max(tr, ts)), u


def _edge_pair(x, p, q, r, s, cos_gap, opts):
    xp, xq, xr, xs = x[p], x[q], x[r], x[s]
    Le, Lf = _norm(xq - xp), _norm(xs - xr)
    if not (Le > 0 and Lf > 0):
        return False
    if abs(float((xq - xp) @ (xs - xr))) / (Le * Lf) < cos_gap:
        return False
    lo, hi, u = _overlap_on(xp, xq, xr, xs)
    if hi - lo < opts.gap_overlap * min(Le, Lf):
        return False
    lo2, hi2, u2 = _overlap_on(xr, xs, xp, xq)
    if hi2 - lo2 < opts.gap_overlap * min(Le, Lf):
        return False
    d = max(_point_segment(xp + lo * u, xr, xs)[0], _point_segment(xp + hi * u, xr, xs)[0],
            _point_segment(xr + lo2 * u2, xp, xq)[0], _point_segment(xr + hi2 * u2, xp, xq)[0])
    return d <= opts.gap_rel * min(Le, Lf)


# --------------------------------------------------------------------------------------------------------------------
# Kutta edges and chains


def classify_kutta_edges(x, tris, he: HalfEdges, input_node, opts: GeometryOptions, log):
    """src/geometry/kutta.jl:28-61 with kutta_source = :dihedral: edges shared by two triangles whose outward faces
    form a sharp (included angle at most te_angle_deg), convex wedge pointing along wake_dir (the wedge test of
    src/geometry/kutta.jl:18-26), then the include and exclude lists (input node pairs)"""
    cte = math.cos(math.radians(opts.te_angle_deg))
    wake = np.asarray(opts.wake_dir, dtype=np.float64)
    inner = np.nonzero(he.edge_he[:, 1] >= 0)[0]
    kutta = set()
    if len(inner):
        nrm = _tri_normals(x, tris)
        t1, t2 = he.edge_he[inner, 0] // 3, he.edge_he[inner, 1] // 3
        n1, n2 = nrm[t1], nrm[t2]
        a, b = he.edge_nodes[inner, 0], he.edge_nodes[inner, 1]
        T2 = tris[t2]
        opp = np.where((T2 != a[:, None]) & (T2 != b[:, None]), T2, -1).max(axis=1)
        sharp = np.einsum("ij,ij->i", n1, n2) <= -cte
        convex = np.einsum("ij,ij->i", x[opp] - x[a], n1) < 0
        along = (n1 + n2) @ wake > 0
        for e in inner[sharp & convex & along]:
            kutta.add((int(he.edge_nodes[e, 0]), int(he.edge_nodes[e, 1])))
    for a, b in opts.kutta_exclude:
        wa, wb = int(input_node[a]), int(input_node[b])
        kutta.discard((min(wa, wb), max(wa, wb)))
    for a, b in opts.kutta_include:
        wa, wb = int(input_node[a]), int(input_node[b])
        e = he.edge_id(wa, wb)
        if e >= 0 and he.edge_he[e, 1] >= 0:
            kutta.add((min(wa, wb), max(wa, wb)))
        else:
            log.append(("kutta_include_ignored", f"kutta_include edge ({a}, {b}) is not shared by two triangles"))
    return kutta


def orient_along_tau(x, nodes, opts: GeometryOptions):
    """src/geometry/kutta.jl:87-95: reverse unless (wake_dir x tau) . up_ref >= 0 for the chord tau of the chain"""
    if len(nodes) < 2:
        return nodes
    tau = x[nodes[-1]] - x[nodes[0]]
    if not _norm(tau) > 0:
        tau = x[nodes[1]] - x[nodes[0]]
    s = float(np.cross(np.asarray(opts.wake_dir, dtype=np.float64), tau) @ np.asarray(opts.up_ref, dtype=np.float64))
    return nodes[::-1] if s < -math.sqrt(np.finfo(float).eps) * _norm(tau) else nodes


def _arclength(x, nodes):
    d = np.linalg.norm(np.diff(x[np.asarray(nodes)], axis=0), axis=1)
    return np.concatenate([[0.0], np.cumsum(d)])


def kutta_chains(x, tris, he: HalfEdges, kutta, opts: GeometryOptions):
    """src/geometry/kutta.jl:104-146: Kutta edges ordered into chains between end and branch nodes (closed when they
    loop), directed along tau; the upper triangle of each segment owns the half-edge along tau"""
    adj = {}
    for p in sorted(kutta):
        for v in p:
            adj.setdefault(v, []).append(p)
    used = set()
    raw = []

    def walk(v0, p0):
        nodes = [v0]
        v, p = v0, p0
        while True:
            used.add(p)
            w = p[1] if p[0] == v else p[0]
            nodes.append(w)
            if w == v0:
                return nodes, True
            if len(adj[w]) != 2:
                return nodes, False
            q = adj[w][1] if adj[w][0] == p else adj[w][0]
            if q in used:
                return nodes, False
            v, p = w, q

    for v in sorted(adj):
        if len(adj[v]) == 2:
            continue
        for p in adj[v]:
            if p not in used:
                raw.append(walk(v, p))
    for p in sorted(kutta):
        if p not in used:
            raw.append(walk(p[0], p))

    chains = []
    for nodes, closed in raw:
        nodes = orient_along_tau(x, nodes, opts)
        edges, tu, tl = [], [], []
        for k in range(len(nodes) - 1):
            a, b = nodes[k], nodes[k + 1]
            e = he.edge_id(a, b)
            h1, h2 = int(he.edge_he[e, 0]), int(he.edge_he[e, 1])
            up, low = (h1, h2) if he.origin(h1) == a else (h2, h1)
            edges.append(e)
            tu.append(up // 3)
            tl.append(low // 3)
        chains.append(KuttaChain(nodes=np.array(nodes, np.int64), tri_upper=np.array(tu, np.int64),
                                 tri_lower=np.array(tl, np.int64), closed=closed, s=_arclength(x, nodes),
                                 edges=np.array(edges, np.int64)))
    return chains


def label_fan(corner_side, tris, he: HalfEdges, tri_kind, stop_edge, t0, v, came, side):
    """src/geometry/pipeline.jl:20-39: label the corners at node v of the triangles reached from t0 by rotating
    around v away from edge `came`, until a stop edge, a boundary, a cap triangle or t0 again"""
    t = t0
    for _ in range(256):
        kv = int(np.nonzero(tris[t] == v)[0][0])
        corner_side[t, kv] = side
        hv = 3 * t + kv
        hp = _he_prev(hv)
        hc = hp if he.he_edge[hv] == came else hv
        ec = int(he.he_edge[hc])
        if stop_edge[ec]:
            return
        g = int(he.twin[hc])
        if g < 0:
            return
        u = g // 3
        if u == t0 or tri_kind[u] == CAP_TRI:
            return
        t, came = u, ec


# --------------------------------------------------------------------------------------------------------------------
# openings and caps


def find_openings(x, tris, he: HalfEdges, opts: GeometryOptions, log):
    """src/geometry/openings.jl:14-116: boundary half-edges traced into loops (the next half-edge leaves the end node
    in the same fan), open chains joined end to start within the gap tolerance. Returns (nodes, closed) pairs."""
    nh = len(he.twin)
    open_he = np.nonzero(he.twin < 0)[0].tolist()
    if not open_he:
        return []
    is_open = he.twin < 0
    leaving = {}
    for h in open_he:
        leaving.setdefault(he.origin(h), []).append(h)
    nxt = np.full(nh, -1, np.int64)
    for h in open_he:
        cands = leaving.get(he.dest(h), [])
        if not cands:
            continue
        if len(cands) == 1:
            nxt[h] = cands[0]
            continue
        g = _he_next(h)
        steps = 0
        while he.twin[g] >= 0 and steps < 1000:
            g = _he_next(int(he.twin[g]))
            steps += 1
        nxt[h] = g if is_open[g] else cands[0]
    prv = np.full(nh, -1, np.int64)
    for h in open_he:
        if nxt[h] >= 0:
            prv[nxt[h]] = h
    visited = np.zeros(nh, bool)
    chains, loops = [], []
    for h0 in open_he:
        if visited[h0] or prv[h0] >= 0:
            continue
        hs, h = [], h0
        while h >= 0 and not visited[h]:
            visited[h] = True
            hs.append(h)
            h = int(nxt[h])
        chains.append(hs)
    for h0 in open_he:
        if visited[h0]:
            continue
        nodes, h = [], h0
        while not visited[h]:
            visited[h] = True
            nodes.append(he.origin(h))
            h = int(nxt[h])
        loops.append((nodes, True))
    if not chains:
        return loops

    def first_node(c):
        return he.origin(c[0])

    def last_node(c):
        return he.dest(c[-1])

    def local_len(c):
        return (_norm(x[he.dest(c[0])] - x[he.origin(c[0])]) + _norm(x[he.dest(c[-1])] - x[he.origin(c[-1])])) / 2

    succ = [-1] * len(chains)
    taken_start = [False] * len(chains)
    for i in range(len(chains)):
        best, bestd = -1, math.inf
        for j in range(len(chains)):
            if taken_start[j]:
                continue
            d = _norm(x[first_node(chains[j])] - x[last_node(chains[i])])
            if d <= opts.gap_rel * min(local_len(chains[i]), local_len(chains[j])) and d < bestd:
                best, bestd = j, d
        if best >= 0:
            succ[i] = best
            taken_start[best] = True
    done = [False] * len(chains)
    for i in range(len(chains)):
        if done[i]:
            continue
        start, seen = i, {i}
        while True:
            p = succ.index(start) if start in succ else -1
            if p < 0 or p in seen:
                break
            seen.add(p)
            start = p
        nodes, j, lastj, closed = [], start, start, False
        while j >= 0 and not done[j]:
            done[j] = True
            nodes += [he.origin(h) for h in chains[j]]
            lastj = j
            j = succ[j]
            if j == start:
                closed = True
        if not closed:
            nodes.append(last_node(chains[lastj]))
            log.append(("opening_unclosed", f"boundary chain of {len(nodes)} nodes from node {nodes[0]} does not "
                                            "close into a loop; left open"))
        loops.append((nodes, closed))
    return loops


def _loop_angle(x, nodes, i):
    m = len(nodes)
    a = x[nodes[(i - 1) % m]] - x[nodes[i]]
    b = x[nodes[(i + 1) % m]] - x[nodes[i]]
    return math.acos(min(max(float(a @ b) / (_norm(a) * _norm(b)), -1.0), 1.0))


def loop_corners(x, nodes, te_candidates):
    """src/geometry/cap.jl:9-26: trailing edge corner (a trailing edge node, the sharpest if several, else the
    sharpest node), leading edge corner (the node farthest from it), and the two chains from the leading to the
    trailing edge corner"""
    m = len(nodes)
    angles = [_loop_angle(x, nodes, i) for i in range(m)]
    cand = [i for i in range(m) if nodes[i] in te_candidates]
    ite = int(np.argmin(angles)) if not cand else cand[int(np.argmin([angles[i] for i in cand]))]
    ile = int(np.argmax([_norm(x[nodes[i]] - x[nodes[ite]]) for i in range(m)]))
    fwd = [nodes[(ile + k) % m] for k in range((ite - ile) % m + 1)]
    bwd = [nodes[(ile - k) % m] for k in range((ile - ite) % m + 1)]
    return nodes[ile], nodes[ite], fwd, bwd


def ladder(x, A, B, tie):
    """src/geometry/cap.jl:28-79: triangles between two chains sharing their first (leading edge) and last
    (trailing edge) nodes; from the leading edge the ladder advances along the chain whose new diagonal is shorter,
    ties going to the chain behind in normalised arclength, then alternating"""
    p, q = len(A) - 1, len(B) - 1
    tris = []
    if p == 1 and q == 1:
        return tris
    if p == 1:
        return [[B[j], B[j + 1], B[q]] for j in range(q - 1)]
    if q == 1:
        return [[A[i], A[i + 1], A[p]] for i in range(p - 1)]
    sA, sB = _arclength(x, A), _arclength(x, B)
    sA, sB = sA / sA[-1], sB / sB[-1]
    tris.append([A[0], A[1], B[1]])
    i = j = 1
    toggle = False
    while not (i == p - 1 and j == q - 1):
        if i == p - 1:
            adv_a = False
        elif j == q - 1:
            adv_a = True
        else:
            dA = _norm(x[A[i + 1]] - x[B[j]])
            dB = _norm(x[A[i]] - x[B[j + 1]])
            if abs(dA - dB) > tie * max(dA, dB):
                adv_a = dA < dB
            elif abs(sA[i + 1] - sB[j + 1]) > tie:
                adv_a = sA[i + 1] < sB[j + 1]
            else:
                toggle = not toggle
                adv_a = toggle
        if adv_a:
            tris.append([A[i], A[i + 1], B[j]])
            i += 1
        else:
            tris.append([A[i], B[j + 1], B[j]])
            j += 1
    tris.append([A[p - 1], A[p], B[q - 1]])
    return tris


def orient_cap(tris, loop):
    """src/geometry/cap.jl:81-89: cap triangles run opposite to the loop's boundary half-edges"""
    m = len(loop)
    along = {(loop[k], loop[(k + 1) % m]) for k in range(m)}
    out = []
    for t in tris:
        directed = ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))
        out.append([t[0], t[2], t[1]] if any(d in along for d in directed) else list(t))
    return out


def _plane_deviation(P):
    c = P.mean(axis=0)
    C = (P - c).T @ (P - c)
    n = np.linalg.eigh(C)[1][:, 0]
    return float(np.max(np.abs((P - c) @ n)))


def close_opening(x, nodes, closed, te_candidates, mode, opts: GeometryOptions):
    """src/geometry/cap.jl:147-181 for modes cap and open"""
    P = x[nodes]
    diam = float(np.max(np.linalg.norm(P[:, None] - P[None], axis=2))) if len(nodes) > 1 else 0.0
    planarity = _plane_deviation(P) / diam if len(nodes) >= 3 and diam > 0 else 0.0
    le, te, A, B = loop_corners(x, nodes, te_candidates)
    up = np.asarray(opts.up_ref, dtype=np.float64)
    mean_up = lambda c: float(np.mean(x[c] @ up))
    upper, lower = (A, B) if mean_up(A) >= mean_up(B) else (B, A)
    tris = []
    if mode == "cap" and closed:
        tris = orient_cap(ladder(x, upper, lower, opts.split_tie_rel), nodes)
    actual = mode if closed else "open"
    return Opening(loop=list(nodes), le=le, te=te, upper=upper, lower=lower, mode=actual, cap_tris=range(0),
                   planarity=planarity), tris


# --------------------------------------------------------------------------------------------------------------------
# quality


def triangle_quality(x, tris):
    """src/geometry/quality.jl:1-27: smallest and largest angle (degrees), aspect ratio R/(2r), equiangle skew, area"""
    a, b, c = x[tris[:, 0]], x[tris[:, 1]], x[tris[:, 2]]
    la, lb, lc = (np.linalg.norm(v, axis=1) for v in (c - b, a - c, b - a))
    area = np.linalg.norm(np.cross(b - a, c - a), axis=1) / 2

    def ang(u, v):
        cs = np.einsum("ij,ij->i", u, v) / (np.linalg.norm(u, axis=1) * np.linalg.norm(v, axis=1))
        return np.degrees(np.arccos(np.clip(cs, -1, 1)))

    th = np.stack([ang(b - a, c - a), ang(a - b, c - b), ang(a - c, b - c)], axis=1)
    qmin, qmax = th.min(axis=1), th.max(axis=1)
    s = (la + lb + lc) / 2
    with np.errstate(divide="ignore", invalid="ignore"):
        aspect = np.where(area > 0, la * lb * lc * s / (8 * area ** 2), np.inf)
    skew = np.maximum((qmax - 60) / 120, (60 - qmin) / 60)
    return qmin, qmax, aspect, skew, area


# --------------------------------------------------------------------------------------------------------------------
# the pipeline


def build_geometry(nodes, faces, opts: GeometryOptions | None = None) -> Surface:
    """src/geometry/pipeline.jl:227-275 (build_geometry) for the stages this module carries.

    `nodes`: (n, 3) coordinates; `faces`: sequences of three or more 0-based node ids.  Returns the `Surface` with
    its `report` (a `GeometryReport`).  Raises `NonManifoldEdgeError` for an edge of more than two faces,
    `GeometryError` for a non-orientable surface or one with T-junctions or gaps."""
    opts = (opts or GeometryOptions()).check()
    log = []
    xin = np.asarray(nodes, dtype=np.float64).reshape(-1, 3)
    faces = [[int(v) for v in f] for f in faces]
    for f, face in enumerate(faces):
        if len(face) < 3 or min(face) < 0 or max(face) >= len(xin):
            raise GeometryError(f"input element {f} has {len(face)} nodes or a node id outside 0..{len(xin) - 1}")

    x, input_node, wfaces, wparent = weld(xin, faces, opts.weld_rel, log)
    tris, tri_parent, tri_slot, quads, quad_parent, quad_diag, quad_reason = split_faces(x, wfaces, wparent, opts, log)
    if len(tris) == 0:
        raise GeometryError("no faces survive welding")

    he1 = HalfEdges(x, tris)
    fans = vertex_fans(len(x), tris, he1.twin)
    for v in np.nonzero(fans > 1)[0]:
        log.append(("bowtie", f"node {v} joins {fans[v]} fans"))
    flip, patch, _ = orient(x, tris, he1, log)
    tris = np.where(flip[:, None], tris[:, [0, 2, 1]], tris)
    he2 = HalfEdges(x, tris)
    bad = _boundary_pathologies(x, he2, opts)
    if bad:
        raise GeometryError(f"{len(bad)} T-junction or gap candidates (first: {bad[0]}); they need constraint rows "
                            "this library does not build")

    kutta = classify_kutta_edges(x, tris, he2, input_node, opts, log)
    loops = find_openings(x, tris, he2, opts, log)
    te_candidates = {v for p in kutta for v in p}
    tri_kind = np.zeros(len(tris), np.int8)
    tri_parent = tri_parent.copy()
    tri_slot = tri_slot.copy()
    openings = []
    all_tris = [tris]
    ntri = len(tris)
    for nodes_l, closed in loops:
        op, captris = close_opening(x, nodes_l, closed, te_candidates, opts.openings, opts)
        op.cap_tris = range(ntri, ntri + len(captris))
        if captris:
            all_tris.append(np.array(captris, np.int64))
        ntri += len(captris)
        openings.append(op)
        log.append((f"opening_{op.mode}", f"opening of {len(op.loop)} nodes ({op.mode}"
                                          f"{f', {len(captris)} triangles' if op.mode == 'cap' else ''})"))
    ncap = ntri - len(tris)
    tris = np.concatenate(all_tris).astype(np.int64)
    tri_kind = np.concatenate([tri_kind, np.full(ncap, CAP_TRI, np.int8)])
    tri_parent = np.concatenate([tri_parent, np.full(ncap, -1, np.int64)])
    tri_slot = np.concatenate([tri_slot, np.zeros(ncap, np.int8)])

    # final connectivity and edge classes
    he = HalfEdges(x, tris)
    ne = he.ne
    cls = np.full(ne, EDGE_INTERIOR, np.int8)
    boundary = he.edge_he[:, 1] < 0
    cls[boundary] = EDGE_OPEN
    inner = ~boundary
    kin = np.array([(int(a), int(b)) in kutta for a, b in he.edge_nodes.tolist()], bool) if ne else np.zeros(0, bool)
    cls[inner & kin] = EDGE_KUTTA
    k1, k2 = tri_kind[he.edge_he[:, 0] // 3], tri_kind[np.maximum(he.edge_he[:, 1], 0) // 3]
    cls[inner & ~kin & (k1 != k2)] = EDGE_CAP_SEAM

    # trailing edge chains and fans
    chains = kutta_chains(x, tris, he, kutta, opts)
    corner_side = np.full((len(tris), 3), NO_TE, np.uint8)
    stop = cls == EDGE_KUTTA
    for ch in chains:
        for k in range(len(ch.edges)):
            for v in (int(ch.nodes[k]), int(ch.nodes[k + 1])):
                label_fan(corner_side, tris, he, tri_kind, stop, int(ch.tri_upper[k]), v, int(ch.edges[k]), TE_UPPER)
                label_fan(corner_side, tris, he, tri_kind, stop, int(ch.tri_lower[k]), v, int(ch.edges[k]), TE_LOWER)
    te_nodes = np.unique(np.concatenate([c.nodes for c in chains])) if chains else np.zeros(0, np.int64)
    te_set = set(te_nodes.tolist())
    cap_of = {}
    for t in np.nonzero(tri_kind == CAP_TRI)[0]:
        for k in range(3):
            v = int(tris[t, k])
            cap_of.setdefault(v, int(t))
            if v in te_set:
                corner_side[t, k] = TE_CAP
    open_nodes = set(he.edge_nodes[cls == EDGE_OPEN].reshape(-1).tolist())
    kdeg = {}
    for p in kutta:
        for v in p:
            kdeg[v] = kdeg.get(v, 0) + 1

    def end_info(v):
        if v in cap_of:
            return "cap", cap_of[v]
        if v in open_nodes:
            return "open", -1
        if kdeg.get(v, 0) > 2:
            return "branch", -1
        return "dangling", -1

    for ch in chains:
        if ch.closed:
            ch.end_type, ch.end_cap_tri = ("closed", "closed"), (-1, -1)
            continue
        (t1, c1), (t2, c2) = end_info(int(ch.nodes[0])), end_info(int(ch.nodes[-1]))
        ch.end_type, ch.end_cap_tri = (t1, t2), (c1, c2)

    # bodies: triangles connected through edges
    uf = _UnionFind(len(tris))
    for h in np.nonzero(he.twin > np.arange(len(he.twin)))[0]:
        uf.unite(int(h) // 3, int(he.twin[h]) // 3)
    open_loop_nodes = {v for op in openings if op.mode != "cap" for v in op.loop}
    bodies = []
    tri_body = np.zeros(len(tris), np.int32)
    for b, ts in enumerate(uf.components()):
        ts = np.array(ts, np.int64)
        tri_body[ts] = b
        bn = np.unique(tris[ts])
        bedges = np.unique(he.he_edge[(3 * ts[:, None] + np.arange(3)).reshape(-1)])
        closed = not np.any(cls[bedges] == EDGE_OPEN) and not (set(bn.tolist()) & open_loop_nodes)
        centre = x[bn].mean(axis=0)
        bodies.append(Body(tris=ts, nodes=bn, closed=bool(closed), euler=int(len(bn) - len(bedges) + len(ts)),
                           volume=_signed_volume(x, tris, ts, centre)))

    # parent quads in the final orientation (src/geometry/pipeline.jl:239-248)
    first_half = {int(tri_parent[t]): t for t in range(len(tri_parent)) if tri_slot[t] == 1}
    fq = []
    for q in range(len(quads)):
        a, b, c, d = (int(v) for v in quads[q])
        half = (a, b, c) if quad_diag[q] == 1 else (a, b, d)
        t = tuple(int(v) for v in tris[first_half[int(quad_parent[q])]])
        same = any(t == (half[k % 3], half[(k + 1) % 3], half[(k + 2) % 3]) for k in range(3))
        fq.append([a, b, c, d] if same else [a, d, c, b])
    fq = np.array(fq, np.int64).reshape(-1, 4)

    qmin, qmax, asp, skew, area = triangle_quality(x, tris)
    sliver = (qmin < opts.sliver_angle_deg) | (asp > opts.sliver_aspect)
    nev = lambda k: sum(1 for kk, _ in log if kk == k)
    counts = dict(nodes=len(x), triangles=len(tris), cap_triangles=int(ncap), edges=ne, bodies=len(bodies),
                  closed_bodies=sum(b.closed for b in bodies), input_nodes=len(xin), input_elements=len(faces),
                  welded_nodes=nev("weld"), dropped_elements=nev("drop_degenerate") + nev("drop_duplicate"),
                  collapsed_elements=nev("collapse"), split_shortest=quad_reason.count("shortest"),
                  split_fold=quad_reason.count("fold"),
                  split_tie=quad_reason.count("tie_mirror") + quad_reason.count("tie_id"),
                  split_fixed=quad_reason.count("fixed"), flipped_triangles=int(flip.sum()),
                  bowtie_vertices=nev("bowtie"), openings_capped=sum(op.mode == "cap" for op in openings),
                  openings_open=sum(op.mode == "open" for op in openings), kutta_edges=len(kutta),
                  kutta_chains=len(chains), te_nodes=len(te_nodes), slivers=int(sliver.sum()))
    stats = dict(min_angle=float(qmin.min()), max_angle=float(qmax.max()), max_aspect=float(asp.max()),
                 max_skew=float(skew.max()),
                 max_cap_planarity=max((op.planarity for op in openings), default=0.0))
    report = GeometryReport(input_node=input_node, tri_parent=tri_parent, tri_slot=tri_slot, tri_kind=tri_kind,
                            edge_nodes=he.edge_nodes, edge_class=cls, openings=openings, bodies=bodies, quads=fq,
                            quad_parent=quad_parent, quad_diag=quad_diag, te_nodes=te_nodes, counts=counts,
                            stats=stats, quality=dict(min_angle=qmin, max_angle=qmax, aspect=asp, skew=skew,
                                                      area=area, sliver=sliver), log=log)
    return Surface(x=np.ascontiguousarray(x), tris=np.ascontiguousarray(tris, dtype=np.int32), tri_body=tri_body,
                   corner_side=corner_side, chains=chains, bodies_closed=[b.closed for b in bodies],
                   tri_cap=tri_kind == CAP_TRI, report=report)


def canonical(tris):
    """Triangles rotated to start at their smallest node and sorted, to compare triangulations"""
    t = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
    k = np.argmin(t, axis=1)
    r = np.stack([t[np.arange(len(t)), (k + j) % 3] for j in range(3)], axis=1)
    return r[np.lexsort(r.T[::-1])]
