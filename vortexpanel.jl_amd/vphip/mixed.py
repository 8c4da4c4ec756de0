"""Mixed meshes: the input's quadrilaterals kept whole as isoparametric quads, every other face a flat triangle.

Host side of src/assembly/mixed.jl (`MixedMesh`, the element order and slots, the element touch pairs) and of
src/post/mixed.jl (`surface_pressure`, `pressure_force`, `spanwise_loading`), plus the isoparametric quad's geometry
(src/elements/iso_quad.jl) that these read.  The assembly itself runs on the device (`vphip.api.assemble_mixed`).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class MixedMesh:
    """src/assembly/mixed.jl:13-23 (0-based): quads first, then triangles.

    quad_nodes, quad_slots (nq, 4): corners in the quad's outward order and their DOF slots; quad_tris (nq, 2): the
    geometry's two triangles of each quad; tri_nodes, tri_slots (nt, 3) and tri_ids (nt,): every other triangle (caps,
    input triangles) and its geometry triangle; element_of (geometry nt,): the element of every geometry triangle."""
    x: np.ndarray
    quad_nodes: np.ndarray
    quad_slots: np.ndarray
    quad_tris: np.ndarray
    tri_nodes: np.ndarray
    tri_slots: np.ndarray
    tri_ids: np.ndarray
    element_of: np.ndarray

    @property
    def nq(self):
        return len(self.quad_nodes)

    @property
    def ntri(self):
        return len(self.tri_nodes)

    @property
    def ne(self):
        return self.nq + self.ntri

    def __repr__(self):
        return f"MixedMesh({self.nq} quads, {self.ntri} triangles)"


def mixed_mesh(g, dm) -> MixedMesh:
    """MixedMesh(g, dm), src/assembly/mixed.jl:26-58.  A welded quad is kept when both its halves survived (the
    geometry pipeline's `tri_slot` 1 and 2 under its `tri_parent`); its slots come from those triangles' corner slots,
    so trailing-edge copies follow theirs.  A surface built without the pipeline (no report) has no quads."""
    tris = np.asarray(g.tris, dtype=np.int64)
    nt = len(tris)
    rep = getattr(g, "report", None)
    halves = {}
    if rep is not None:
        for t in range(nt):
            if rep.tri_slot[t] > 0:
                halves.setdefault(int(rep.tri_parent[t]), []).append(t)
    element_of = np.full(nt, -1, np.int64)
    qn, qs, qt = [], [], []
    if rep is not None:
        for k, q in enumerate(np.asarray(rep.quads, dtype=np.int64).reshape(-1, 4)):
            ts = halves.get(int(rep.quad_parent[k]), [])
            if len(ts) != 2:
                continue
            slots = []
            for v in q:
                for t in ts:
                    i = np.nonzero(tris[t] == v)[0]
                    if len(i):
                        slots.append(int(dm.corner_slot[t][i[0]]))
                        break
                else:
                    raise ValueError(f"node {v} of quad {k} is on neither of its triangles")
            qn.append([int(v) for v in q])
            qs.append(slots)
            qt.append(ts)
            element_of[ts] = len(qn) - 1
    nq = len(qn)
    tn, tsl, tid = [], [], []
    for t in range(nt):
        if element_of[t] >= 0:
            continue
        tn.append(tris[t].tolist())
        tsl.append([int(s) for s in dm.corner_slot[t]])
        tid.append(t)
        element_of[t] = nq + len(tn) - 1
    as2 = lambda a, k: np.asarray(a, dtype=np.int64).reshape(-1, k)
    return MixedMesh(np.asarray(g.x, dtype=np.float64), as2(qn, 4), as2(qs, 4), as2(qt, 2), as2(tn, 3), as2(tsl, 3),
                     np.asarray(tid, np.int64), element_of)


def element_touch_pairs(g, mesh: MixedMesh) -> np.ndarray:
    """The geometry's touch pairs as element pairs (src/assembly/mixed.jl:453-457), both orders, no self pairs"""
    out = set()
    for a, b in np.asarray(g.touch_pairs, dtype=np.int64).reshape(-1, 2):
        ea, eb = int(mesh.element_of[a]), int(mesh.element_of[b])
        if ea != eb:
            out.add((ea, eb))
            out.add((eb, ea))
    return np.array(sorted(out), dtype=np.int64).reshape(-1, 2)


# ------------------------------------------------------------------------------------------------------------------
# the isoparametric quad's geometry (src/elements/iso_quad.jl:40-92), vectorised over parameters


def shape(xi, eta):
    """N1..N4 at (xi, eta): array (..., 4)"""
    xi, eta = np.asarray(xi, dtype=np.float64), np.asarray(eta, dtype=np.float64)
    return np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta], -1)


def shape_derivatives(xi, eta):
    """dN/dxi, dN/deta at (xi, eta): two arrays (..., 4)"""
    xi, eta = np.asarray(xi, dtype=np.float64), np.asarray(eta, dtype=np.float64)
    one = np.ones_like(xi)
    return (np.stack([-(1 - eta), 1 - eta, eta, -eta], -1), np.stack([-(1 - xi), -xi, xi, (1 - xi) * one], -1))


def point(qx, xi, eta):
    """x(xi, eta) = sum N_a x_a for corners qx (4, 3)"""
    return shape(xi, eta) @ np.asarray(qx, dtype=np.float64)


def tangents(qx, xi, eta):
    """a1 = dx/dxi, a2 = dx/deta"""
    qx = np.asarray(qx, dtype=np.float64)
    xi, eta = np.asarray(xi, dtype=np.float64)[..., None], np.asarray(eta, dtype=np.float64)[..., None]
    a1 = (1 - eta) * (qx[1] - qx[0]) + eta * (qx[2] - qx[3])
    a2 = (1 - xi) * (qx[3] - qx[0]) + xi * (qx[2] - qx[1])
    return a1, a2


def area_normal(qx, xi, eta):
    a1, a2 = tangents(qx, xi, eta)
    return np.cross(a1, a2)


def gradient(qx, mu, xi, eta):
    """Surface gradient of the nodal doublet values mu (4,) at (xi, eta): mu_xi a^1 + mu_eta a^2
    (src/elements/iso_quad.jl:84-91)"""
    a1, a2 = tangents(qx, xi, eta)
    nJ = np.cross(a1, a2)
    J2 = np.sum(nJ * nJ, -1)[..., None]
    Nx, Ne = shape_derivatives(xi, eta)
    mx = (Nx @ np.asarray(mu, dtype=np.float64))[..., None]
    me = (Ne @ np.asarray(mu, dtype=np.float64))[..., None]
    return (mx * np.cross(a2, nJ) + me * np.cross(nJ, a1)) / J2


def sheet_strength(qx, mu, xi, eta):
    """gamma = n x grad mu at (xi, eta) (src/elements/iso_quad.jl:76-81)"""
    a1, a2 = tangents(qx, xi, eta)
    J = np.linalg.norm(np.cross(a1, a2), axis=-1)[..., None]
    Nx, Ne = shape_derivatives(xi, eta)
    mu = np.asarray(mu, dtype=np.float64)
    return ((Nx @ mu)[..., None] * a2 - (Ne @ mu)[..., None] * a1) / J


# ------------------------------------------------------------------------------------------------------------------
# post-processing of mixed meshes (src/post/mixed.jl)


def surface_pressure(mesh: MixedMesh, mu, vinf, m: int = 3):
    """src/post/mixed.jl:17-44: the m x m Gauss points of every quad (gradient route: u = V - (V.n)n + grad mu), then
    the centroid of every triangle.  Returns a dict of x, dA (outward area vectors), cp and element."""
    V = np.asarray(vinf, dtype=np.float64)
    V2 = V @ V
    xg, wg = np.polynomial.legendre.leggauss(m)
    xg, wg = (xg + 1) / 2, wg / 2
    XI, ETA = np.meshgrid(xg, xg, indexing="ij")
    W = np.outer(wg, wg).ravel()
    xi, eta = XI.ravel(), ETA.ravel()
    mu = np.asarray(mu, dtype=np.float64)
    xs, dA, cp, el = [], [], [], []
    for e in range(mesh.nq):
        qx = mesh.x[mesh.quad_nodes[e]]
        nJ = area_normal(qx, xi, eta)
        n = nJ / np.linalg.norm(nJ, axis=1)[:, None]
        u = V - (n @ V)[:, None] * n + gradient(qx, mu[mesh.quad_slots[e]], xi, eta)
        xs.append(point(qx, xi, eta))
        dA.append(W[:, None] * nJ)
        cp.append(1 - np.sum(u * u, 1) / V2)
        el.append(np.full(len(xi), e))
    for t in range(mesh.ntri):
        p = mesh.x[mesh.tri_nodes[t]]
        N = np.cross(p[1] - p[0], p[2] - p[0])
        A2 = np.linalg.norm(N)
        n = N / A2
        g = np.zeros(3)  # grad mu = sum mu_b (n x e_b) / 2A, e_b the edge opposite b
        for b in range(3):
            eb = p[(b + 2) % 3] - p[(b + 1) % 3]
            g += mu[mesh.tri_slots[t][b]] * np.cross(n, eb) / A2
        u = V - (V @ n) * n + g
        xs.append(p.mean(0)[None])
        dA.append((A2 / 2 * n)[None])
        cp.append(np.array([1 - u @ u / V2]))
        el.append(np.array([mesh.nq + t]))
    return dict(x=np.concatenate(xs), dA=np.concatenate(dA), cp=np.concatenate(cp), element=np.concatenate(el))


def pressure_force(pts, origin=(0.0, 0.0, 0.0)):
    """src/post/mixed.jl:52-60: F = -sum Cp dA, M = sum (x - origin) x (-Cp dA), in units of q_inf"""
    f = -pts["cp"][:, None] * pts["dA"]
    return f.sum(0), np.cross(pts["x"] - np.asarray(origin, dtype=np.float64), f).sum(0)


def spanwise_loading(pts, wte_x, line_nodes, mu_w, vinf):
    """src/post/mixed.jl:68-91: each point's pressure force to the strip of the trailing-edge line its position falls
    in; returns s0, s1, Gamma, lift_kj, lift_p, drag_p per strip"""
    nodes = np.asarray(line_nodes, dtype=np.int64)
    x0 = wte_x[nodes[0]]
    t = wte_x[nodes[-1]] - x0
    t = t / np.linalg.norm(t)
    V = np.asarray(vinf, dtype=np.float64)
    d = V / np.linalg.norm(V)
    eL = np.cross(d, t)
    eL = eL / np.linalg.norm(eL)
    s = (wte_x[nodes] - x0) @ t
    n = len(nodes) - 1
    lift, drag = np.zeros(n), np.zeros(n)
    for x, a, c in zip(pts["x"], pts["dA"], pts["cp"]):
        k = np.searchsorted(s, (x - x0) @ t, side="right") - 1
        if 0 <= k < n:
            lift[k] -= c * (a @ eL)
            drag[k] -= c * (a @ d)
    G = (mu_w[nodes[:-1]] + mu_w[nodes[1:]]) / 2
    ds = np.diff(s)
    return dict(s0=s[:n], s1=s[1:], Gamma=G, lift_kj=2 * G / np.linalg.norm(V), lift_p=lift / ds, drag_p=drag / ds)
