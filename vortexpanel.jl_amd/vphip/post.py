"""Post-processing of a Galerkin solution: the N-body velocity sums on the GPU (surface_velocity, body_velocity, blob
sums through libvortexhip) and the per-element and force formulas on the host.

Mirrors src/post/Post.jl and src/post/forces.jl: same names, same argument meaning.  The velocity sums have no CPU
fallback (they raise when the library is missing); the host formulas are O(nt) or O(nte^2) array arithmetic.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .api import Context, _check, _surface_c, _wake_c, default_context

PRECISIONS = {"fp64": 0, "mixed": 1}


def _prec(p):
    if p not in PRECISIONS:
        raise ValueError(f"precision must be one of {tuple(PRECISIONS)}, not {p!r}")
    return PRECISIONS[p]


def element_values(dm, mu):
    """Corner doublet values of every element, src/post/Post.jl:19"""
    return np.asarray(mu, dtype=np.float64)[dm.corner_slot]


def sheet_strengths(els, dm, mu):
    """Vortex sheet gamma = n x grad(mu) of every element, Post.jl:22-23 (els: api.Elements)"""
    return np.einsum("tk,tkj->tj", element_values(dm, mu), els.gamma)


def surface_velocity(g, dm, sigma, mu, vinf, wake=None, mu_wake=None, kernel_far=10.0, precision="fp64",
                     ctx: Context | None = None):
    """Velocity at every body triangle's centroid from the singularities, Post.jl:32-64 (GPU N-body sum):
    V∞ + sum_s sigma_s w_s + gamma_s x w_s, the own element on its normal's side, plus the wake's sheets and tip
    filaments from the wake node doublet values mu_wake (`wake_values`)."""
    ctx = ctx or default_context()
    s, ks = _surface_c(g, dm)
    w, kw = _wake_c(wake)
    if wake is not None and mu_wake is None:
        raise ValueError("a wake needs mu_wake")
    sig, m = L.f64(sigma), L.f64(mu)
    if sig.shape != (len(g.tris),) or m.shape != (dm.nslot,):
        raise ValueError("sigma must have one value per triangle and mu one per slot")
    mw = L.f64(mu_wake) if wake is not None else np.zeros(1)
    u = np.zeros((len(g.tris), 3))
    rc = ctx.lib.vp_surface_velocity(ctx.h, C.byref(s), C.byref(w) if w is not None else None, L.ptr(sig), L.ptr(m),
                                     L.ptr(mw), L.ptr(L.f64(vinf)), C.c_double(kernel_far), _prec(precision), L.ptr(u))
    _check(ctx.h, rc)
    return u


def body_velocity(g, sigma, gamma, xs, kernel_far=10.0, precision="fp64", ctx: Context | None = None):
    """Velocity of the body's sources and sheets at off-body points, without the freestream, Post.jl:92-112"""
    ctx = ctx or default_context()
    keep = dict(x=L.f64(g.x), tris=L.i32(g.tris))
    s = L.SurfaceC(len(g.x), L.ptr(keep["x"]), len(g.tris), L.ptr(keep["tris"], C.c_int), None, 0, 0, None)
    X = L.f64(xs).reshape(-1, 3)
    gam = L.f64(gamma).reshape(-1, 3)
    sig = L.f64(sigma)
    if sig.shape != (len(g.tris),) or gam.shape != (len(g.tris), 3):
        raise ValueError("sigma and gamma must have one entry per triangle")
    u = np.zeros_like(X)
    rc = ctx.lib.vp_body_velocity(ctx.h, C.byref(s), L.ptr(sig), L.ptr(gam), len(X), L.ptr(X),
                                  C.c_double(kernel_far), _prec(precision), L.ptr(u))
    _check(ctx.h, rc)
    return u


def blob_velocities(bx, alpha, delta, xs, precision="fp64", ctx: Context | None = None):
    """Velocity of Rosenhead-Moore particles (positions bx, strengths alpha, cores delta) at the points xs by direct
    summation, src/freewake/step.jl:176-187"""
    ctx = ctx or default_context()
    bx, ba = L.f64(bx).reshape(-1, 3), L.f64(alpha).reshape(-1, 3)
    bd = L.f64(delta).reshape(-1)
    if not (len(bx) == len(ba) == len(bd)):
        raise ValueError("blob arrays disagree in length")
    X = L.f64(xs).reshape(-1, 3)
    u = np.zeros_like(X)
    rc = ctx.lib.vp_blob_velocities(ctx.h, len(bd), L.ptr(bx), L.ptr(ba), L.ptr(bd), len(X), L.ptr(X),
                                    _prec(precision), L.ptr(u))
    _check(ctx.h, rc)
    return u


def _tree_stats(st):
    return {k: getattr(st, k) for k, _ in L.TreeStatsC._fields_}


def blob_velocities_tree(bx, alpha, delta, xs, theta=0.5, degree=6, precision="fp64", ctx: Context | None = None):
    """blob_velocities by the cluster-particle treecode (vp_blob_velocities_tree): target clusters well separated from
    a source cluster (box radius <= theta x distance) interpolate its field from the (degree+1)^3 Chebyshev points of
    their box; every other pair is summed directly, each blob with its own core.  precision 'mixed' evaluates the
    kernel in FP32 (interpolation and sums FP64).  Returns (u, stats)."""
    ctx = ctx or default_context()
    bx, ba = L.f64(bx).reshape(-1, 3), L.f64(alpha).reshape(-1, 3)
    bd = L.f64(delta).reshape(-1)
    if not (len(bx) == len(ba) == len(bd)):
        raise ValueError("blob arrays disagree in length")
    X = L.f64(xs).reshape(-1, 3)
    u = np.zeros_like(X)
    o = L.TreeOptionsC(float(theta), int(degree), _prec(precision))
    st = L.TreeStatsC()
    rc = ctx.lib.vp_blob_velocities_tree(ctx.h, len(bd), L.ptr(bx), L.ptr(ba), L.ptr(bd), len(X), L.ptr(X),
                                         C.byref(o), L.ptr(u), C.byref(st))
    _check(ctx.h, rc)
    return u, _tree_stats(st)


def tree_coverage(bx, xs, theta=0.5, degree=6):
    """The treecode's lists built on the host only (no device work): (number of target leaves whose near and far
    source ranges do not cover every source exactly once, stats)"""
    lib = L.load()
    bx, X = L.f64(bx).reshape(-1, 3), L.f64(xs).reshape(-1, 3)
    o = L.TreeOptionsC(float(theta), int(degree), 0)
    bad, st = C.c_longlong(), L.TreeStatsC()
    rc = lib.vp_tree_coverage(len(bx), L.ptr(bx), len(X), L.ptr(X), C.byref(o), C.byref(bad), C.byref(st))
    if rc != 0:
        raise ValueError(f"vp_tree_coverage failed ({rc})")
    return bad.value, _tree_stats(st)


class WakeField:
    """The free-wake step's velocity sums (wake_velocities, src/freewake/step.jl:166-173) on the device as one HIP graph
    (vp_wake_field_*): the body's sources and sheets (as body_velocity) plus the wake's blobs (as blob_velocities)
    at the targets, plus an optional base velocity.  The body's elements are set up once; each call rewrites the
    graph's node arguments, so the target and blob counts may change from call to call."""

    def __init__(self, g, kernel_far=10.0, precision="fp64", ctx: Context | None = None):
        self.ctx = ctx or default_context()
        x, tris = L.f64(g.x), L.i32(g.tris)
        s = L.SurfaceC(len(x), L.ptr(x), len(tris), L.ptr(tris, C.c_int), None, 0, 0, None)
        h = C.c_void_p()
        _check(self.ctx.h, self.ctx.lib.vp_wake_field_create(self.ctx.h, C.byref(s), C.c_double(kernel_far),
                                                              _prec(precision), C.byref(h)))
        self.h = h
        self.nt = len(tris)

    def __call__(self, sigma, gamma, xs, bx=None, alpha=None, delta=None, base=None):
        sig, gam = L.f64(sigma), L.f64(gamma).reshape(-1, 3)
        if sig.shape != (self.nt,) or gam.shape != (self.nt, 3):
            raise ValueError("sigma and gamma must have one entry per triangle")
        X = L.f64(xs).reshape(-1, 3)
        if bx is None:
            bx, ba, bd = np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0)
        else:
            bx, ba, bd = L.f64(bx).reshape(-1, 3), L.f64(alpha).reshape(-1, 3), L.f64(delta).reshape(-1)
        if not (len(bx) == len(ba) == len(bd)):
            raise ValueError("blob arrays disagree in length")
        b = L.f64(base) if base is not None else None
        u = np.zeros_like(X)
        rc = self.ctx.lib.vp_wake_field_eval(self.h, L.ptr(sig), L.ptr(gam), len(X), L.ptr(X), len(bd), L.ptr(bx),
                                             L.ptr(ba), L.ptr(bd), L.ptr(b), L.ptr(u))
        _check(self.ctx.h, rc)
        return u

    def info(self):
        """(graphs built, node updates, graph launches)"""
        c = np.zeros(3, dtype=np.int64)
        _check(self.ctx.h, self.ctx.lib.vp_wake_field_info(self.h, L.ptr(c, C.c_longlong)))
        return tuple(int(v) for v in c)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.vp_wake_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gradient_velocity(els, dm, mu, vinf):
    """V∞ - (V∞.n)n + grad(mu) per element, Post.jl:72-73"""
    V = np.asarray(vinf, dtype=np.float64)
    g = np.einsum("tk,tkj->tj", element_values(dm, mu), els.gradl)
    return V - (els.n @ V)[:, None] * els.n + g


def pressure_coefficient(u, vinf, dmudt=None):
    """1 - |u|^2/|V∞|^2 (- 2 dmu/dt/|V∞|^2), Post.jl:82-84"""
    V = np.asarray(vinf, dtype=np.float64)
    uu = np.einsum("ij,ij->i", u, u)
    return 1 - (uu if dmudt is None else uu + 2 * np.asarray(dmudt)) / V.dot(V)


def pressure_force(els, cp, origin=(0.0, 0.0, 0.0)):
    """F = -sum Cp n A and its moment about origin, in units of q∞, src/post/forces.jl:11-19"""
    f = -(np.asarray(cp) * els.area)[:, None] * els.n
    return f.sum(0), np.cross(els.centroid - np.asarray(origin, dtype=np.float64), f).sum(0)


def te_segments(wte):
    """Segments of the trailing edge lines along tau, forces.jl:22-23"""
    out = []
    for line, closed in zip(wte.lines, wte.closed):
        m = len(line)
        out += [(line[i], line[(i + 1) % m]) for i in range(m if closed else m - 1)]
    return out


def kutta_joukowski_force(wte, mu_w, vinf):
    """F = (2/|V∞|^2) V∞ x integral of mu_w dl along the trailing edge lines, forces.jl:34-39"""
    V = np.asarray(vinf, dtype=np.float64)
    x = np.asarray(wte.x, dtype=np.float64)
    G = np.zeros(3)
    for a, b in te_segments(wte):
        G += (mu_w[a] + mu_w[b]) / 2 * (x[b] - x[a])
    return 2 / V.dot(V) * np.cross(V, G)


def _te_structure(wte):
    """trefftz_drag's index structure of the trailing edge lines, built once for many calls: the segments' ends
    (sa, sb), the vertices (vp), the segments' mean circulation as a matrix on mu_w (S, nseg x nte) and the vertices'
    shed circulation as a matrix on it (D, nv x nseg)"""
    sa, sb, vp, S_rows, D_rows = [], [], [], [], []
    nte = len(wte.x)
    seg0 = 0
    for line, closed in zip(wte.lines, wte.closed):
        line = list(line)
        m = len(line)
        nseg = m if closed else m - 1
        for i in range(nseg):
            row = np.zeros(nte)
            row[line[i]] += 0.5
            row[line[(i + 1) % m]] += 0.5
            S_rows.append(row)
            sa.append(line[i])
            sb.append(line[(i + 1) % m])
        for i in range(m):
            vp.append(line[i])
            D_rows.append((seg0 + i if i < nseg else -1,
                           seg0 + i - 1 if i > 0 else (seg0 + nseg - 1 if closed else -1)))
        seg0 += nseg
    nseg = len(sa)
    D = np.zeros((len(vp), nseg))
    for v, (after, before) in enumerate(D_rows):
        if after >= 0:
            D[v, after] += 1.0
        if before >= 0:
            D[v, before] -= 1.0
    S = np.array(S_rows).reshape(nseg, nte)
    return np.array(sa, dtype=np.int64), np.array(sb, dtype=np.int64), np.array(vp, dtype=np.int64), S, D


def trefftz_drag_batch(wte, MW, V):
    """trefftz_drag for many freestreams at once: MW (nte x na) trailing edge doublets, V (na x 3); the same
    formulas (forces.jl:53-81) vectorised over the angles"""
    V = np.asarray(V, dtype=np.float64)
    sa, sb, vp, S, Dm = _te_structure(wte)
    if len(sa) == 0:
        return np.zeros(len(V))
    x = np.asarray(wte.x, dtype=np.float64)
    G = S @ np.asarray(MW, dtype=np.float64)                      # (nseg, na) mean circulation per segment
    VG = Dm @ G                                                   # (nv, na) shed circulation per vertex
    d = V / np.linalg.norm(V, axis=1)[:, None]                    # (na, 3)
    P = x[None] - np.einsum("ti,ai->at", x, d)[:, :, None] * d[:, None, :]   # (na, nte, 3)
    A, B = P[:, sa], P[:, sb]
    c = (A + B) / 2
    ell = np.linalg.norm(B - A, axis=2)                           # (na, nseg)
    n = np.cross(d[:, None, :], (B - A) / ell[:, :, None])
    r = c[:, :, None, :] - P[:, vp][:, None, :, :]                # (na, nseg, nv, 3)
    w = np.cross(d[:, None, None, :], r) / np.sum(r * r, axis=3)[..., None]
    v = np.einsum("ak,askj->asj", VG.T / (4 * np.pi), w)
    Dd = np.sum(G.T * np.einsum("asj,asj->as", n, v) * ell, axis=1)
    return 2 * Dd / np.einsum("aj,aj->a", V, V)


def trefftz_drag(wte, mu_w, vinf):
    """Induced drag in the Trefftz plane from the strips' mean circulation and the trailing vortices at the trailing
    edge nodes (midpoint normalwash on the lifting line), in units of q∞, forces.jl:53-81.  Vectorised over strips."""
    V = np.asarray(vinf, dtype=np.float64)
    d = V / np.linalg.norm(V)
    x = np.asarray(wte.x, dtype=np.float64)
    P = x - np.outer(x @ d, d)
    mu_w = np.asarray(mu_w, dtype=np.float64)
    vp, vg, sa, sb, sg = [], [], [], [], []
    for line, closed in zip(wte.lines, wte.closed):
        line = list(line)
        m = len(line)
        nseg = m if closed else m - 1
        G = [(mu_w[line[i]] + mu_w[line[(i + 1) % m]]) / 2 for i in range(nseg)]
        for i in range(m):
            before = G[i - 1] if i > 0 else (G[nseg - 1] if closed else 0.0)
            after = G[i] if i < nseg else 0.0
            vp.append(P[line[i]])
            vg.append(after - before)
        for i in range(nseg):
            sa.append(P[line[i]])
            sb.append(P[line[(i + 1) % m]])
            sg.append(G[i])
    if not sg:
        return 0.0
    vp, vg, sa, sb, sg = (np.asarray(v) for v in (vp, vg, sa, sb, sg))
    c = (sa + sb) / 2
    ell = np.linalg.norm(sb - sa, axis=1)
    n = np.cross(d, (sb - sa) / ell[:, None])
    r = c[:, None, :] - vp[None, :, :]
    v = np.einsum("k,skj->sj", vg / (4 * np.pi), np.cross(d, r) / np.sum(r * r, axis=2)[:, :, None])
    D = np.sum(sg * np.einsum("sj,sj->s", n, v) * ell)
    return 2 * D / V.dot(V)
