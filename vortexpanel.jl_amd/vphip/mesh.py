"""Surface meshes for the Galerkin hot path, as the arrays the reference's `SurfaceGeometry` carries.

The reference builds its `SurfaceGeometry` with a mesh-repair pipeline (src/geometry/pipeline.jl:48-275) from
MeshLibrary meshes; that pipeline is outside this library's scope.  The generators here produce the repaired result
directly (welded nodes, outward triangles, Kutta chains with corner sides, tip caps), following the reference's
construction rules where they exist:

- `icosphere`: src/meshtools/bodies.jl:18-41 (same node and triangle numbering, 1-based there, 0-based here).
- `uv_ellipsoid`: src/meshtools/bodies.jl:82-108 (latitude-longitude sphere, scaled), welded as build_geometry would
  (poles merged, seam merged), quads split along the shorter diagonal (src/geometry/quadsplit.jl:29-67, the
  `:shortest` branch).
- `abuse_jitter`, `abuse_slivers`: src/meshtools/abuse.jl:93-192 (clearance-bounded tangential jitter) and
  :339-360 (slivers), with a numpy PCG64 stream instead of StableRNGs and a radial projection onto the ellipsoid.
- `naca_wing`: a structured NACA 4-digit wing (closed trailing edge, so upper and lower surfaces share the trailing
  edge nodes, which become a Kutta chain), with strip-triangulated tip caps; corner sides as
  src/geometry/pipeline.jl:150-177 labels them.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

NO_TE, TE_UPPER, TE_LOWER, TE_CAP = 0, 1, 2, 3


@dataclass
class KuttaChain:
    """src/geometry/types.jl:85-100: trailing edge nodes along tau, with the upper and lower triangle of each segment"""
    nodes: np.ndarray
    tri_upper: np.ndarray
    tri_lower: np.ndarray
    closed: bool = False
    s: np.ndarray | None = None        # arclength at the nodes
    edges: np.ndarray | None = None    # edge ids of the segments (in the geometry pipeline's edge numbering)
    end_type: tuple | None = None      # "cap", "open", "branch" or "dangling" at the first and last node
    end_cap_tri: tuple | None = None   # a cap triangle at each end node, -1 if none


@dataclass
class Surface:
    """The `SurfaceGeometry` fields the hot path reads (src/geometry/types.jl:174-199), 0-based."""
    x: np.ndarray                      # (nn, 3) float64
    tris: np.ndarray                   # (nt, 3) int32, outward normal along (x2 - x1) x (x3 - x1)
    tri_body: np.ndarray               # (nt,) int32
    corner_side: np.ndarray            # (nt, 3) uint8: NO_TE / TE_UPPER / TE_LOWER / TE_CAP
    chains: list = field(default_factory=list)
    touch_pairs: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    bodies_closed: list = field(default_factory=lambda: [True])
    tri_cap: np.ndarray | None = None  # (nt,) bool, cap triangles
    report: object | None = None       # what the geometry pipeline did, when it built the surface

    @property
    def nn(self):
        return len(self.x)

    @property
    def nt(self):
        return len(self.tris)

    def digest(self) -> str:
        """SHA-256 of the node coordinates and triangles (the fixture hash of a generated mesh)"""
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.x, dtype=np.float64).tobytes())
        h.update(np.ascontiguousarray(self.tris, dtype=np.int32).tobytes())
        return h.hexdigest()


def _normalize(v):
    return v / np.linalg.norm(v)


def _orient_outward(X, f, outward):
    """src/meshtools/bodies.jl:4-8 (a quad adds the normal of its second half)"""
    n = np.cross(X[f[1]] - X[f[0]], X[f[2]] - X[f[0]])
    if len(f) == 4:
        n = n + np.cross(X[f[2]] - X[f[0]], X[f[3]] - X[f[0]])
    return f[::-1] if np.dot(n, outward) < 0 else f


def _closed_surface(X, tris):
    nt = len(tris)
    return Surface(x=np.ascontiguousarray(X, dtype=np.float64), tris=np.ascontiguousarray(tris, dtype=np.int32),
                   tri_body=np.zeros(nt, np.int32), corner_side=np.zeros((nt, 3), np.uint8),
                   tri_cap=np.zeros(nt, bool))


def icosphere(level: int, r: float = 1.0, center=(0.0, 0.0, 0.0)) -> Surface:
    """Icosahedron subdivided `level` times onto the sphere, 10*4^level + 2 nodes (src/meshtools/bodies.jl:18-41)"""
    phi = (1 + np.sqrt(5.0)) / 2
    P = [(-1.0, phi, 0), (1.0, phi, 0), (-1.0, -phi, 0), (1.0, -phi, 0), (0.0, -1, phi), (0.0, 1, phi),
         (0.0, -1, -phi), (0.0, 1, -phi), (phi, 0, -1), (phi, 0, 1), (-phi, 0, -1), (-phi, 0, 1)]
    P = [_normalize(np.array(p, dtype=np.float64)) for p in P]
    F = [(1, 12, 6), (1, 6, 2), (1, 2, 8), (1, 8, 11), (1, 11, 12), (2, 6, 10), (6, 12, 5), (12, 11, 3), (11, 8, 7),
         (8, 2, 9), (4, 10, 5), (4, 5, 3), (4, 3, 7), (4, 7, 9), (4, 9, 10), (5, 10, 6), (3, 5, 12), (7, 3, 11),
         (9, 7, 8), (10, 9, 2)]
    F = [(a - 1, b - 1, c - 1) for a, b, c in F]
    for _ in range(level):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                P.append(_normalize(P[a] + P[b]))
                mid[key] = len(P) - 1
            return mid[key]

        G = []
        for a, b, c in F:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            G += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = G
    c0 = np.asarray(center, dtype=np.float64)
    X = np.array([r * p + c0 for p in P])
    tris = [_orient_outward(X, np.array(f), X[list(f)].sum(0) / 3 - c0) for f in F]
    return _closed_surface(X, np.array(tris))


def combine(parts: list) -> Surface:
    """Several surfaces as one geometry of as many bodies (SurfaceGeometry with tri_body, src/geometry/types.jl:174-199):
    nodes and triangles concatenated in order, body k's triangles labelled k.  Chains, touch pairs and caps are offset
    by the nodes and triangles before them."""
    xs, tris, body, side, chains, touch, closed, caps = [], [], [], [], [], [], [], []
    n0 = t0 = 0
    for k, s in enumerate(parts):
        if len(s.bodies_closed) != 1:
            raise ValueError("combine takes single-body surfaces")
        xs.append(s.x)
        tris.append(s.tris + n0)
        body.append(np.full(s.nt, k, np.int32))
        side.append(s.corner_side)
        for c in s.chains:
            chains.append(KuttaChain(c.nodes + n0, c.tri_upper + t0, c.tri_lower + t0, c.closed))
        touch.append(s.touch_pairs + t0)
        closed += list(s.bodies_closed)
        caps.append(s.tri_cap if s.tri_cap is not None else np.zeros(s.nt, bool))
        n0 += s.nn
        t0 += s.nt
    return Surface(np.vstack(xs), np.vstack(tris).astype(np.int32), np.concatenate(body), np.vstack(side), chains,
                   np.vstack(touch).astype(np.int32), closed, np.concatenate(caps))


def uv_ellipsoid(nth: int, nph: int, axes=(1.0, 1.0, 1.0)) -> Surface:
    """Welded latitude-longitude sphere (src/meshtools/bodies.jl:82-96) scaled by `axes` (:103-108).

    Nodes: north pole, nth-1 rings of nph nodes, south pole: 2 + (nth-1) nph nodes, 2 nph (nth-1) triangles.
    """
    ax = np.asarray(axes, dtype=np.float64)
    X = [np.array([0.0, 0.0, 1.0])]
    for i in range(1, nth):
        th = np.pi * i / nth
        for j in range(nph):
            ph = 2 * np.pi * j / nph
            X.append(np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]))
    X.append(np.array([0.0, 0.0, -1.0]))
    X = np.array(X) * ax
    S = len(X) - 1

    def nid(i, j):  # ring i in 1..nth-1
        return 1 + (i - 1) * nph + (j % nph)

    tris = []
    for j in range(nph):  # north cap
        tris.append(_orient_outward(X, np.array([0, nid(1, j), nid(1, j + 1)]), X[[0, nid(1, j), nid(1, j + 1)]].mean(0)))
    for i in range(1, nth - 1):
        for j in range(nph):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            L13, L24 = np.linalg.norm(X[c] - X[a]), np.linalg.norm(X[d] - X[b])
            out = X[[a, b, c, d]].mean(0)
            if L13 <= L24 * (1 + 1e-12):
                t1, t2 = [a, b, c], [a, c, d]
            else:
                t1, t2 = [a, b, d], [b, c, d]
            tris.append(_orient_outward(X, np.array(t1), out))
            tris.append(_orient_outward(X, np.array(t2), out))
    for j in range(nph):  # south cap
        tris.append(_orient_outward(X, np.array([S, nid(nth - 1, j + 1), nid(nth - 1, j)]),
                                    X[[S, nid(nth - 1, j), nid(nth - 1, j + 1)]].mean(0)))
    return _closed_surface(X, np.array(tris))


def _node_clearance(X, tris):
    """src/meshtools/abuse.jl:98-120: area-weighted normals and the clearance (smallest height over a non-incident edge)"""
    nn = len(X)
    n = np.zeros((nn, 3))
    clr = np.full(nn, np.inf)
    for f in tris:
        a = np.cross(X[f[1]] - X[f[0]], X[f[2]] - X[f[0]])
        for k in range(3):
            v = f[k]
            n[v] += a
            j = (k + 1) % 3  # the edge of f that does not touch v runs from f[k+1] to f[k+2]
            p, q = X[f[j]], X[f[(j + 1) % 3]]
            e = q - p
            t = np.clip(np.dot(X[v] - p, e) / max(np.dot(e, e), np.finfo(float).eps), 0, 1)
            clr[v] = min(clr[v], np.linalg.norm(X[v] - (p + t * e)))
    nrm = np.linalg.norm(n, axis=1)
    n = n / np.where(nrm > 0, nrm, 1)[:, None]
    return n, clr


def ellipsoid_projector(axes):
    ax = np.asarray(axes, dtype=np.float64)
    return lambda p: p / np.sqrt(np.sum((p / ax) ** 2))


def abuse_jitter(s: Surface, delta: float, rng: np.random.Generator, project) -> Surface:
    """Tangential jitter by up to delta x clearance, then back onto the surface (src/meshtools/abuse.jl:174-192)"""
    X = s.x.copy()
    normals, clr = _node_clearance(s.x, s.tris)
    for v in range(len(X)):
        nv = normals[v]
        t1 = _normalize(np.cross(nv, [1.0, 0, 0]) if abs(nv[0]) < 0.9 else np.cross(nv, [0, 1.0, 0]))
        t2 = np.cross(nv, t1)
        th = 2 * np.pi * rng.random()
        p = s.x[v] + delta * clr[v] * rng.random() * (np.cos(th) * t1 + np.sin(th) * t2)
        X[v] = project(p)
    return Surface(x=X, tris=s.tris.copy(), tri_body=s.tri_body.copy(), corner_side=s.corner_side.copy(),
                   chains=s.chains, touch_pairs=s.touch_pairs, bodies_closed=s.bodies_closed, tri_cap=s.tri_cap)


def abuse_slivers(s: Surface, n: int, ratio: float, rng: np.random.Generator, project) -> tuple[Surface, int]:
    """Move n nodes (no two in one triangle) along one of their edges to `ratio` of its length
    (src/meshtools/abuse.jl:339-360).  Returns the mesh and the number of sliver nodes made."""
    X = s.x.copy()
    nn = len(X)
    nel = [[] for _ in range(nn)]
    for e, f in enumerate(s.tris):
        for v in f:
            nel[v].append(e)
    blocked = np.zeros(nn, bool)
    made = 0
    for v in rng.permutation(nn):
        if made == n:
            break
        if blocked[v]:
            continue
        nbrs = sorted({int(u) for e in nel[v] for u in s.tris[e] if u != v})
        if not nbrs:
            continue
        w = nbrs[rng.integers(len(nbrs))]
        X[v] = project(X[w] + ratio * (X[v] - X[w]))
        for e in nel[v]:
            blocked[s.tris[e]] = True
        made += 1
    return Surface(x=X, tris=s.tris.copy(), tri_body=s.tri_body.copy(), corner_side=s.corner_side.copy(),
                   chains=s.chains, touch_pairs=s.touch_pairs, bodies_closed=s.bodies_closed, tri_cap=s.tri_cap), made


def abused_ellipsoid(nth=71, nph=142, axes=(2.0, 1.0, 0.5), jitter=0.25, sliver_frac=0.05, sliver_ratio=0.1,
                     seed=20261203) -> Surface:
    """The abused ellipsoid: uv mesh scaled to `axes`, jittered by `jitter` of each node's clearance and with slivers
    around enough nodes that about `sliver_frac` of the triangles touch one, projected radially onto the surface."""
    rng = np.random.Generator(np.random.PCG64(seed))
    proj = ellipsoid_projector(axes)
    s = uv_ellipsoid(nth, nph, axes)
    s = abuse_jitter(s, jitter, rng, proj)
    s, _ = abuse_slivers(s, int(round(sliver_frac * s.nt / 6)), sliver_ratio, rng, proj)
    return s


# --------------------------------------------------------------------------------------------------------------------
# raw meshes (nodes and faces before repair), the input of the geometry pipeline


def cube_sphere_raw(n: int, r: float = 1.0):
    """src/meshtools/bodies.jl:43-74: each cube face an n x n equiangular grid projected onto the sphere, faces built
    on their own nodes (the cube edges and corners duplicated), outward quads"""
    X, faces = [], []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            a, b = [k for k in range(3) if k != axis]
            base = len(X)
            xi = np.tan(np.pi / 4 * (2 * np.arange(n + 1) / n - 1))
            for j in range(n + 1):
                for i in range(n + 1):
                    p = np.zeros(3)
                    p[axis], p[a], p[b] = sgn, xi[i], xi[j]
                    X.append(r * p / np.linalg.norm(p))
            nid = lambda i, j: base + j * (n + 1) + i
            Xa = np.array(X)
            for j in range(n):
                for i in range(n):
                    f = [nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)]
                    faces.append(list(_orient_outward(Xa, np.array(f), Xa[f].mean(axis=0))))
    return np.array(X), faces


def uv_sphere_raw(nth: int, nph: int, r: float = 1.0):
    """src/meshtools/bodies.jl:76-101: latitude-longitude quads with every coincident node repeated (poles nph + 1
    times, the seam meridian twice), so the pole quads collapse to triangles once welded"""
    X = []
    for i in range(nth + 1):
        for j in range(nph + 1):
            th, ph = np.pi * i / nth, 2 * np.pi * j / nph
            X.append(r * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]))
    X = np.array(X)
    nid = lambda i, j: i * (nph + 1) + j
    faces = []
    for i in range(nth):
        for j in range(nph):
            f = [nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)]
            faces.append(list(_orient_outward(X, np.array(f), X[f].mean(axis=0))))
    return X, faces


def naca4(code: str, xc: np.ndarray):
    """Camber line and half thickness of a NACA 4-digit section, closed trailing edge coefficient -0.1036"""
    m, p, t = int(code[0]) / 100, int(code[1]) / 10, int(code[2:]) / 100
    yt = 5 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3 - 0.1036 * xc ** 4)
    if m == 0:
        return np.zeros_like(xc), np.zeros_like(xc), yt
    yc = np.where(xc < p, m / p ** 2 * (2 * p * xc - xc ** 2), m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
    dy = np.where(xc < p, 2 * m / p ** 2 * (p - xc), 2 * m / (1 - p) ** 2 * (p - xc))
    return yc, np.arctan(dy), yt


@dataclass
class WingSpec:
    code: str = "0012"
    n_around: int = 160       # panels around the section (even)
    n_span: int = 60          # spanwise quads
    aspect: float = 8.0
    taper: float = 1.0
    sweep_c4_deg: float = 0.0
    mean_chord: float = 1.0


def _wing_nodes(spec: WingSpec):
    """Section ring nodes of every station, (n_span + 1) * n_around rows (see naca_wing for the numbering)"""
    N = spec.n_around
    assert N % 2 == 0 and N >= 8
    nh = N // 2
    b = spec.aspect * spec.mean_chord
    cr = 2 * spec.mean_chord / (1 + spec.taper)
    beta = np.pi * np.arange(nh + 1) / nh
    xc = (1 - np.cos(beta)) / 2                     # 0 at the leading edge, 1 at the trailing edge
    yc, th, yt = naca4(spec.code, xc)
    xu, zu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, zl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    # ring: TE, lower from TE to LE, LE, upper from LE to TE (the TE point once)
    sx = np.concatenate([[xc[nh]], xl[nh - 1:0:-1], [xc[0]], xu[1:nh]])
    sz = np.concatenate([[yc[nh]], zl[nh - 1:0:-1], [yc[0]], zu[1:nh]])
    ys = np.linspace(-b / 2, b / 2, spec.n_span + 1)
    tan_s = np.tan(np.radians(spec.sweep_c4_deg))
    X = []
    for y in ys:
        c = cr * (1 - (1 - spec.taper) * abs(y) / (b / 2))
        xle = abs(y) * tan_s + 0.25 * cr - 0.25 * c
        for k in range(N):
            X.append((xle + c * sx[k], y, c * sz[k]))
    return np.array(X, dtype=np.float64)


def naca_wing(spec: WingSpec) -> Surface:
    """Structured wing, tip caps, one Kutta chain along the trailing edge.

    Section ring: index 0 the trailing edge, 1..nh-1 the lower surface toward the leading edge (index nh), then the
    upper surface back toward the trailing edge; cosine spacing.  Stations y from -b/2 to b/2 (uniform); chord
    c(y) = c_r (1 - (1 - taper) |y|/(b/2)), quarter-chord line swept by sweep_c4_deg.
    """
    N = spec.n_around
    nh = N // 2
    X = _wing_nodes(spec)
    nid = lambda j, k: j * N + (k % N)
    tris, side, cap = [], [], []
    for j in range(spec.n_span):
        for k in range(N):
            a, bq, cq, d = nid(j, k), nid(j, k + 1), nid(j + 1, k + 1), nid(j + 1, k)
            # outward: (a, b, c, d) runs counterclockwise about the outward normal
            L13, L24 = np.linalg.norm(X[cq] - X[a]), np.linalg.norm(X[d] - X[bq])
            if L13 <= L24 * (1 + 1e-12):
                t1, t2 = [a, bq, cq], [a, cq, d]
            else:
                t1, t2 = [a, bq, d], [bq, cq, d]
            surf = TE_LOWER if k < nh else TE_UPPER
            for t in (t1, t2):
                tris.append(t)
                side.append([surf if X_is_te(v, N) else NO_TE for v in t])
                cap.append(False)
    # tip caps: strip between lower point k and upper point N-k, fan at both ends
    for j, sgn in ((0, -1.0), (spec.n_span, 1.0)):
        ring = [nid(j, k) for k in range(N)]
        ctris = [[ring[0], ring[1], ring[N - 1]]]
        for k in range(1, nh - 1):
            lo0, lo1, up0, up1 = ring[k], ring[k + 1], ring[N - k], ring[N - k - 1]
            ctris += [[lo0, lo1, up0], [lo1, up1, up0]]
        ctris.append([ring[nh - 1], ring[nh], ring[nh + 1]])
        for t in ctris:
            t = _orient_outward(X, np.array(t), np.array([0.0, sgn, 0.0]))
            tris.append(list(t))
            side.append([TE_CAP if X_is_te(v, N) else NO_TE for v in t])
            cap.append(True)
    tris = np.array(tris, dtype=np.int32)
    # orientation check: outward means positive enclosed volume
    vol = np.sum(np.einsum('ij,ij->i', X[tris[:, 0]], np.cross(X[tris[:, 1]], X[tris[:, 2]]))) / 6
    assert vol > 0, "wing triangles must face outward"
    # the Kutta chain along the trailing edge nodes; tau such that the upper triangle owns the half edge along tau
    te = np.array([nid(j, 0) for j in range(spec.n_span + 1)])
    tri_up, tri_lo = [], []
    edge_owner = {}
    for t, f in enumerate(tris):
        for k in range(3):
            edge_owner[(int(f[k]), int(f[(k + 1) % 3]))] = t
    for j in range(spec.n_span):
        a, c = int(te[j]), int(te[j + 1])
        t_ac, t_ca = edge_owner[(a, c)], edge_owner[(c, a)]
        tri_up.append(t_ac)
        tri_lo.append(t_ca)
    up_is_upper = all(side_of(tris[t], side, t, te) == TE_UPPER for t in tri_up)
    if not up_is_upper:  # reverse tau
        te = te[::-1].copy()
        tri_up, tri_lo = [], []
        for j in range(spec.n_span):
            a, c = int(te[j]), int(te[j + 1])
            tri_up.append(edge_owner[(a, c)])
            tri_lo.append(edge_owner[(c, a)])
    chain = KuttaChain(nodes=te, tri_upper=np.array(tri_up), tri_lower=np.array(tri_lo), closed=False)
    nt = len(tris)
    return Surface(x=X, tris=tris, tri_body=np.zeros(nt, np.int32), corner_side=np.array(side, dtype=np.uint8),
                   chains=[chain], tri_cap=np.array(cap))


def X_is_te(v, N):
    return int(v) % N == 0


def side_of(tri, side, t, te):
    for k in range(3):
        if side[t][k] != NO_TE:
            return side[t][k]
    return NO_TE


def naca_wing_raw(spec: WingSpec, split_te: bool = False):
    """The wing of `naca_wing` as a raw quad mesh with open tips: quad (j, k) joins ring points k and k + 1 of
    stations j and j + 1, counterclockwise about the outward normal.  With `split_te` the upper surface ends on a
    second copy of every trailing edge node (as meshes with a split trailing edge come), which welding merges."""
    N = spec.n_around
    X = _wing_nodes(spec)
    ns = spec.n_span + 1
    if split_te:
        X = np.concatenate([X, X[::N]])
        nid = lambda j, k: j * N + k if k < N else ns * N + j
    else:
        nid = lambda j, k: j * N + (k % N)
    faces = [[nid(j, k), nid(j, k + 1), nid(j + 1, k + 1), nid(j + 1, k)]
             for j in range(spec.n_span) for k in range(N)]
    return X, faces
