"""FP64 CPU oracle of the Galerkin hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the checker.  The pair
integrals, assembly, velocity sums and explicit right hand side are the C restatement in vp_oracle.c (built by
oracle/Makefile into oracle/lib/libvporacle.so); the bordered dense solve restates src/assembly/solve.jl:15-115 on
LAPACK (scipy's getrf/getrs/gecon, the routines the reference itself calls through LinearAlgebra); forces restate
src/post/forces.jl:11-81 and the velocity routes src/post/Post.jl:19-84.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libvporacle.so")
_lib = None

TIER_NAMES = ("self", "edge", "vertex", "contact", "near", "far")


class Opts(C.Structure):
    """AssemblyOptions, src/assembly/assemble.jl:26-41 (test 0 hat, 1 onering, 2 dual)"""
    _fields_ = [("test", C.c_int), ("far_ratio", C.c_double), ("far_points", C.c_int), ("near_points", C.c_int),
                ("near_ratio", C.c_double), ("near_depth", C.c_int), ("touch_levels", C.c_int),
                ("touch_ratio", C.c_double), ("touch_order", C.c_int), ("corner_angle", C.c_double),
                ("contact_rel", C.c_double), ("kernel_far", C.c_double)]


def opts(**kw) -> Opts:
    d = dict(test=0, far_ratio=3.0, far_points=3, near_points=7, near_ratio=1.5, near_depth=3, touch_levels=6,
             touch_ratio=0.15, touch_order=3, corner_angle=75.0, contact_rel=0.1, kernel_far=10.0)
    tests = {"hat": 0, "onering": 1, "dual": 2}
    if isinstance(kw.get("test"), str):
        kw["test"] = tests[kw["test"]]
    d.update(kw)
    return Opts(**d)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def source_velocity(a, b, c, x, side=0, far=10.0, plane_tol=1e-12, mode=0):
    """mode 0 dispatch (source_velocity), 1 exact, 2 multipole; src/kernels/source_triangle.jl:75-138"""
    w = np.zeros(3)
    rc = lib().vpo_source_velocity(_p(_f64(a)), _p(_f64(b)), _p(_f64(c)), _p(_f64(x)), C.c_int(side),
                                   C.c_double(far), C.c_double(plane_tol), C.c_int(mode), _p(w))
    if rc:
        raise ValueError("degenerate triangle")
    return w


def inplane_fraction(a, b, c, x, tol=1e-12):
    f = lib().vpo_inplane_fraction
    f.restype = C.c_double
    return f(_p(_f64(a)), _p(_f64(b)), _p(_f64(c)), _p(_f64(x)), C.c_double(tol))


def segment_velocity(a, b, x):
    va, vb = np.zeros(3), np.zeros(3)
    lib().vpo_segment_velocity(_p(_f64(a)), _p(_f64(b)), _p(_f64(x)), _p(va), _p(vb))
    return va, vb


def blob_velocity(x, xp, al, delta):
    u = np.zeros(3)
    lib().vpo_blob_velocity(_p(_f64(x)), _p(_f64(xp)), _p(_f64(al)), C.c_double(delta), _p(u))
    return u


def source_velocity_many(a, b, c, xs, side=0, far=10.0, plane_tol=1e-12, mode=0):
    """source_velocity at the points xs (n, 3) -> (n, 3)"""
    xs = _f64(np.asarray(xs).reshape(-1, 3))
    w = np.zeros_like(xs)
    rc = lib().vpo_source_velocity_many(_p(_f64(a)), _p(_f64(b)), _p(_f64(c)), len(xs), _p(xs), C.c_int(side),
                                        C.c_double(far), C.c_double(plane_tol), C.c_int(mode), _p(w))
    if rc:
        raise ValueError("degenerate triangle")
    return w


def segment_velocity_many(a, b, xs):
    """segment_velocity at the points xs (n, 3) -> va, vb (n, 3) each"""
    xs = _f64(np.asarray(xs).reshape(-1, 3))
    va, vb = np.zeros_like(xs), np.zeros_like(xs)
    lib().vpo_segment_velocity_many(_p(_f64(a)), _p(_f64(b)), len(xs), _p(xs), _p(va), _p(vb))
    return va, vb


RULE_KINDS = dict(classify=-1, far=0, near_base=1, self=2, edge0=3, edge1=4, edge2=5, vertex0=6, vertex1=7, vertex2=8,
                  near=9, vertex_rule0=10, vertex_rule1=11, vertex_rule2=12, contact=13)


def pair_geom(tx, sx, tid=(0, 1, 2), sid=(3, 4, 5), kind="classify", sigma=0.0, self_pair=False, o: Opts | None = None):
    """The rule and pair_block of test triangle tx and source triangle sx (3 corners each), shared corners given by
    equal node ids: kind 'classify' (pair_rule; wake_pair_rule when no id is shared) or a fixed rule (RULE_KINDS).
    Returns (tier or None, lam (n, 3), w (n,), blk (3, 3), rb (3,)), or None when kind 'contact' gives no rule."""
    o = o or opts()
    tier = C.c_int(-1)
    tx, sx = _f64(np.asarray(tx).reshape(3, 3)), _f64(np.asarray(sx).reshape(3, 3))
    ti, si = _i32(tid), _i32(sid)
    args = (_p(tx), _p(sx), _p(ti, C.c_int), _p(si, C.c_int), int(self_pair), RULE_KINDS[kind], C.byref(o),
            C.c_double(sigma), C.byref(tier))
    n = lib().vpo_pair_geom(*args, 0, None, None, None, None)
    if n == -2:
        return None
    if n < 0:
        raise ValueError(f"vpo_pair_geom failed ({n})")
    lam, w, blk, rb = np.zeros((n, 3)), np.zeros(n), np.zeros(9), np.zeros(3)
    lib().vpo_pair_geom(*args, n, _p(lam), _p(w), _p(blk), _p(rb))
    return (TIER_NAMES[tier.value] if tier.value >= 0 else None), lam, w, blk.reshape(3, 3), rb


def segment_geom(tx, fa, fb, o: Opts | None = None):
    """segment_rule and segment_block of test triangle tx and the filament fa -> fb: (tier, npoints, ca, cb)"""
    o = o or opts()
    tier = C.c_int(-1)
    ca, cb = np.zeros(3), np.zeros(3)
    n = lib().vpo_segment_geom(_p(_f64(np.asarray(tx).reshape(3, 3))), _p(_f64(fa)), _p(_f64(fb)), C.byref(o),
                               C.byref(tier), _p(ca), _p(cb))
    if n < 0:
        raise ValueError(f"vpo_segment_geom failed ({n})")
    return TIER_NAMES[tier.value], n, ca, cb


def rule(kind: int, o: Opts | None = None):
    """kind 0 far, 1 near, 2 self, 3..5 edge k, 6..8 vertex k (TierRules, src/elements/rules.jl:161-174)"""
    o = o or opts()
    n = lib().vpo_rule(C.byref(o), kind, None, None)
    lam, w = np.zeros((n, 3)), np.zeros(n)
    lib().vpo_rule(C.byref(o), kind, _p(lam), _p(w))
    return lam, w


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().vpo_gauss_legendre(n, _p(x), _p(w))
    return x, w


def graded_gauss(levels, ratio, order):
    x, w = np.zeros(4096), np.zeros(4096)
    m = lib().vpo_graded_gauss(levels, C.c_double(ratio), order, _p(x), _p(w))
    return x[:m], w[:m]


def _wake_arrays(wake):
    if wake is None:
        return 0, np.zeros((1, 3)), 0, np.zeros((1, 3), np.int32), 0, np.zeros((1, 2), np.int32)
    return (len(wake.x), _f64(wake.x), len(wake.tris), _i32(wake.tris), len(wake.filaments), _i32(wake.filaments))


def assemble(g, dm, sigma, vinf, o: Opts | None = None, wake=None, nthreads=0, tiers=False):
    """src/assembly/assemble.jl:377-447: returns A (nslot x (nslot + nw)), b, pairs and points per tier"""
    o = o or opts()
    nwx, wx, nwt, wt, nwf, wf = _wake_arrays(wake)
    ns = dm.nslot
    A = np.zeros((ns + nwx, ns), dtype=np.float64)  # column-major nslot x ncol == row-major (ncol, nslot)
    b = np.zeros(ns)
    pairs = np.zeros(6, np.int64)
    points = np.zeros(6, np.int64)
    tp = _i32(g.touch_pairs.reshape(-1, 2)) if len(g.touch_pairs) else np.zeros((1, 2), np.int32)
    tr = np.zeros((g.nt, g.nt + nwt), np.uint8) if tiers else None
    rc = lib().vpo_assemble(g.nn, _p(_f64(g.x)), g.nt, _p(_i32(g.tris), C.c_int), _p(_i32(dm.corner_slot), C.c_int),
                            ns, _p(_f64(sigma)), _p(_f64(vinf)), len(g.touch_pairs), _p(tp, C.c_int),
                            nwx, _p(wx), nwt, _p(wt, C.c_int), nwf, _p(wf, C.c_int), C.byref(o), nthreads,
                            _p(A), _p(b), _p(pairs, C.c_longlong), _p(points, C.c_longlong),
                            _p(tr, C.c_ubyte) if tiers else None)
    if rc:
        raise ValueError(f"vpo_assemble failed ({rc})")
    out = (A.T, b, dict(zip(TIER_NAMES, pairs.tolist())), dict(zip(TIER_NAMES, points.tolist())))
    return out + (tr,) if tiers else out


def classify(g, o: Opts | None = None, wake=None, nthreads=0, tiers=True):
    """assemble's classification alone (pair_rule, wake_pair_rule, segment_rule; no pair blocks): the tier of every
    (body test triangle, source triangle) pair (nt, nt + wake triangles) uint8 and the pairs and points per tier"""
    o = o or opts()
    nwx, wx, nwt, wt, nwf, wf = _wake_arrays(wake)
    pairs = np.zeros(6, np.int64)
    points = np.zeros(6, np.int64)
    tp = _i32(g.touch_pairs.reshape(-1, 2)) if len(g.touch_pairs) else np.zeros((1, 2), np.int32)
    tr = np.zeros((g.nt, g.nt + nwt), np.uint8) if tiers else None
    rc = lib().vpo_classify(g.nn, _p(_f64(g.x)), g.nt, _p(_i32(g.tris), C.c_int), len(g.touch_pairs), _p(tp, C.c_int),
                            nwx, _p(wx), nwt, _p(wt, C.c_int), nwf, _p(wf, C.c_int), C.byref(o), nthreads,
                            _p(pairs, C.c_longlong), _p(points, C.c_longlong), _p(tr, C.c_ubyte) if tiers else None)
    if rc:
        raise ValueError(f"vpo_classify failed ({rc})")
    return tr, dict(zip(TIER_NAMES, pairs.tolist())), dict(zip(TIER_NAMES, points.tolist()))


def assemble_rows(g, dm, sigma, vinf, t_begin, t_end, o: Opts | None = None, wake=None, nthreads=0, full=False):
    """assemble restricted to the test triangles [t_begin, t_end) (the bounded sample of the CPU baseline);
    returns the number of pair integrals done, or with full=True also the partial A (nslot x (nslot + nw)) and b,
    whose rows are complete for every slot whose triangles all lie in the range (the geometry's touch pairs included)"""
    o = o or opts()
    nwx, wx, nwt, wt, nwf, wf = _wake_arrays(wake)
    ns = dm.nslot
    A = np.zeros((ns + nwx, ns))
    b = np.zeros(ns)
    pairs, points = np.zeros(6, np.int64), np.zeros(6, np.int64)
    ntp = len(g.touch_pairs) if full else 0
    tp = _i32(g.touch_pairs.reshape(-1, 2)) if ntp else np.zeros((1, 2), np.int32)
    rc = lib().vpo_assemble_rows(g.nn, _p(_f64(g.x)), g.nt, _p(_i32(g.tris), C.c_int),
                                 _p(_i32(dm.corner_slot), C.c_int), ns, _p(_f64(sigma)), _p(_f64(vinf)), ntp,
                                 _p(tp, C.c_int), nwx, _p(wx), nwt, _p(wt, C.c_int), nwf, _p(wf, C.c_int), C.byref(o),
                                 nthreads, _p(A), _p(b), _p(pairs, C.c_longlong), _p(points, C.c_longlong), None,
                                 int(t_begin), int(t_end))
    if rc:
        raise ValueError(f"vpo_assemble_rows failed ({rc})")
    return (int(pairs.sum()), A.T, b) if full else int(pairs.sum())


def pair(g, sigma, t, s, o: Opts | None = None, wake=None, touch=0):
    """One pair block (tier, 3x3 block, RHS, points) of test triangle t and source s (s >= nt: wake triangle)"""
    o = o or opts()
    nwx, wx, nwt, wt, _, _ = _wake_arrays(wake)
    blk, rb, npts = np.zeros(9), np.zeros(3), C.c_int(0)
    tier = lib().vpo_pair(g.nn, _p(_f64(g.x)), g.nt, _p(_i32(g.tris), C.c_int), _p(_f64(sigma)), nwx, _p(wx), nwt,
                          _p(wt, C.c_int), t, s, touch, C.byref(o), _p(blk), _p(rb), C.byref(npts))
    return tier, blk.reshape(3, 3), rb, npts.value


# --------------------------------------------------------------------------- elements and host-side restatements

def elements(x, tris):
    """FlatTriangle data (src/elements/flat_triangle.jl:48-55): n, area, centroid, gradlambda (nt,3,3), gamma (nt,3,3)"""
    x = np.asarray(x, dtype=np.float64)
    a, b, c = x[tris[:, 0]], x[tris[:, 1]], x[tris[:, 2]]
    nA = np.cross(b - a, c - a)
    A2 = np.linalg.norm(nA, axis=1)
    n = nA / A2[:, None]
    e = np.stack([c - b, a - c, b - a], axis=1)
    gl = np.cross(n[:, None, :], e) / A2[:, None, None]
    gam = -e / A2[:, None, None]
    return dict(n=n, area=A2 / 2, centroid=(a + b + c) / 3, gradl=gl, gamma=gam)


def source_strengths(el, vinf):
    """src/assembly/assemble.jl:364"""
    return -(el["n"] @ np.asarray(vinf, dtype=np.float64))


def bordered_system(dm, A, b, W=None):
    """src/assembly/solve.jl:15-42 and src/dofmap/system.jl:26-50"""
    Cf = dm.C if W is None else __import__("scipy.sparse", fromlist=["vstack"]).vstack([dm.C, W]).tocsr()
    if A.shape != (dm.nslot, Cf.shape[0]):
        raise ValueError(f"A is {A.shape}, expected ({dm.nslot}, {Cf.shape[0]})")
    Ar = np.asarray((dm.C.T @ (Cf.T @ A.T).T))
    nq = dm.nq
    rows = []
    for bi in range(len(dm.body_closed)):
        e = np.zeros(nq)
        e[dm.body_q[bi]] = 1.0
        cvec = np.asarray(dm.C.T @ (dm.slot_area * (dm.slot_body == bi)))
        rows.append((e, cvec))
    nb = len(rows)
    M = np.zeros((nq + nb, nq + nb))
    M[:nq, :nq] = Ar
    scale = np.sum(np.abs(np.diag(Ar))) / nq
    se = np.zeros(nb)
    for k, (e, cvec) in enumerate(rows):
        se[k] = scale / np.linalg.norm(e)
        M[:nq, nq + k] = se[k] * e
        M[nq + k, :nq] = cvec * (scale / np.linalg.norm(cvec))
    rhs = np.concatenate([dm.C.T @ b, np.zeros(nb)])
    return M, rhs, se


def solve_galerkin(dm, A, b, W=None):
    """src/assembly/solve.jl:51-59 on LAPACK getrf/getrs/gecon"""
    M, rhs, se = bordered_system(dm, A, b, W)
    lu, piv = sla.lu_factor(M)
    anorm = np.linalg.norm(M, 1)
    rcond, _ = sla.lapack.dgecon(lu, anorm, norm="1")
    x = sla.lu_solve((lu, piv), rhs)
    q = x[:dm.nq]
    return dict(q=q, mu=dm.C @ q, lam=se * x[dm.nq:], cond=1 / rcond,
                residual=np.linalg.norm(M @ x - rhs) / np.linalg.norm(rhs), M=M, rhs=rhs)


def consistency(dm, A, b, W=None):
    """src/assembly/solve.jl:111-115"""
    import scipy.sparse as sp
    Cf = dm.C if W is None else sp.vstack([dm.C, W]).tocsr()
    ACf = np.asarray((Cf.T @ A.T).T)
    nA = np.max(np.abs(ACf))  # Julia's norm(matrix, Inf) is the largest entry
    return dict(rows=np.max(np.abs(ACf.sum(1))) / nA, columns=np.max(np.abs(ACf.sum(0))) / nA,
                rhs=abs(b.sum()) / np.abs(b).sum())


def gradient_velocity(el, dm, mu, vinf):
    """src/post/Post.jl:72-73"""
    V = np.asarray(vinf, dtype=np.float64)
    m = np.asarray(mu)[dm.corner_slot]
    g = np.einsum('tk,tkj->tj', m, el["gradl"])
    return V - (el["n"] @ V)[:, None] * el["n"] + g


def sheet_strengths(el, dm, mu):
    """src/post/Post.jl:22-23"""
    m = np.asarray(mu)[dm.corner_slot]
    return np.einsum('tk,tkj->tj', m, el["gamma"])


def pressure_coefficient(u, vinf, dmudt=None):
    """src/post/Post.jl:82-84"""
    V = np.asarray(vinf, dtype=np.float64)
    uu = np.einsum('ij,ij->i', u, u)
    return 1 - (uu if dmudt is None else uu + 2 * np.asarray(dmudt)) / V.dot(V)


def pressure_force(el, cp, origin=(0.0, 0.0, 0.0)):
    """src/post/forces.jl:11-19"""
    f = -(cp * el["area"])[:, None] * el["n"]
    return f.sum(0), np.cross(el["centroid"] - np.asarray(origin), f).sum(0)


def trefftz_drag(wte, mu_w, vinf):
    """src/post/forces.jl:53-81"""
    d = np.asarray(vinf, dtype=np.float64)
    d = d / np.linalg.norm(d)
    P = lambda x: x - np.dot(x, d) * d
    vort, strips = [], []
    for line, closed in zip(wte.lines, wte.closed):
        m = len(line)
        nseg = m if closed else m - 1
        G = [(mu_w[line[i]] + mu_w[line[(i + 1) % m]]) / 2 for i in range(nseg)]
        for i in range(m):
            before = G[i - 1] if i > 0 else (G[nseg - 1] if closed else 0.0)
            after = G[i] if i < nseg else 0.0
            vort.append((P(wte.x[line[i]]), after - before))
        for i in range(nseg):
            strips.append((P(wte.x[line[i]]), P(wte.x[line[(i + 1) % m]]), G[i]))
    D = 0.0
    for a, b, Gk in strips:
        c, l = (a + b) / 2, np.linalg.norm(b - a)
        n = np.cross(d, (b - a) / l)
        v = sum(dG / (4 * np.pi) * np.cross(d, c - p) / np.sum((c - p) ** 2) for p, dG in vort)
        D += Gk * np.dot(n, v) * l
    return 2 * D / np.dot(vinf, vinf)


def kutta_joukowski_force(wte, mu_w, vinf):
    """src/post/forces.jl:34-39"""
    V = np.asarray(vinf, dtype=np.float64)
    G = np.zeros(3)
    for line, closed in zip(wte.lines, wte.closed):
        m = len(line)
        for i in range(m if closed else m - 1):
            a, b = line[i], line[(i + 1) % m]
            G += (mu_w[a] + mu_w[b]) / 2 * (wte.x[b] - wte.x[a])
    return 2 / V.dot(V) * np.cross(V, G)


def body_velocity(g, sigma, gam, xs, kernel_far=10.0, nthreads=0):
    xs = _f64(xs).reshape(-1, 3)
    u = np.zeros_like(xs)
    rc = lib().vpo_body_velocity(g.nn, _p(_f64(g.x)), g.nt, _p(_i32(g.tris), C.c_int), _p(_f64(sigma)),
                                 _p(_f64(gam)), len(xs), _p(xs), C.c_double(kernel_far), nthreads, _p(u))
    if rc:
        raise ValueError("body_velocity failed")
    return u


def surface_velocity(g, dm, sigma, mu, vinf, wake=None, mu_wake=None, kernel_far=10.0, nthreads=0):
    """src/post/Post.jl:32-64"""
    el = elements(g.x, g.tris)
    gam = sheet_strengths(el, dm, mu)
    nwx, wx, nwt, wt, nwf, wf = _wake_arrays(wake)
    if wake is not None:
        wel = elements(wake.x, wake.tris)
        wm = np.asarray(mu_wake)[wake.tris]
        wgam = np.einsum('tk,tkj->tj', wm, wel["gamma"])
        wmu = _f64(mu_wake)
    else:
        wgam, wmu = np.zeros((1, 3)), np.zeros(1)
    u = np.zeros((g.nt, 3))
    rc = lib().vpo_surface_velocity(g.nn, _p(_f64(g.x)), g.nt, _p(_i32(g.tris), C.c_int), _p(_f64(sigma)),
                                    _p(_f64(gam)), _p(_f64(vinf)), nwx, _p(wx), nwt, _p(wt, C.c_int),
                                    _p(_f64(wgam)), nwf, _p(wf, C.c_int), _p(wmu), C.c_double(kernel_far), nthreads,
                                    _p(u))
    if rc:
        raise ValueError("surface_velocity failed")
    return u


def blob_velocities(bx, ba, bd, xs, nthreads=0):
    """src/freewake/step.jl:176-187"""
    xs = _f64(xs).reshape(-1, 3)
    u = np.zeros_like(xs)
    lib().vpo_blob_velocities(len(bd), _p(_f64(bx)), _p(_f64(ba)), _p(_f64(bd)), len(xs), _p(xs), nthreads, _p(u))
    return u


def explicit_rhs(g, dm, sheets=None, segments=None, blobs=None, o: Opts | None = None, nthreads=0):
    """src/assembly/explicit.jl:31-86.  sheets = (x, tris, corner_mu); segments = (ab (n,6), gam (n,2));
    blobs = (x, alpha, delta)"""
    o = o or opts()
    if sheets is None:
        sx, st, smu = np.zeros((1, 3)), np.zeros((0, 3), np.int32), np.zeros((1, 3))
    else:
        sx, st, smu = _f64(sheets[0]), _i32(sheets[1]), _f64(sheets[2])
    if segments is None:
        sg, sgam = np.zeros((1, 6)), np.zeros((1, 2))
        nseg = 0
    else:
        sg, sgam = _f64(segments[0]).reshape(-1, 6), _f64(segments[1]).reshape(-1, 2)
        nseg = len(sg)
    if blobs is None:
        bx, ba, bd = np.zeros((1, 3)), np.zeros((1, 3)), np.zeros(1)
        nb = 0
    else:
        bx, ba, bd = _f64(blobs[0]), _f64(blobs[1]), _f64(blobs[2])
        nb = len(bd)
    b = np.zeros(dm.nslot)
    rc = lib().vpo_explicit_rhs(g.nn, _p(_f64(g.x)), g.nt, _p(_i32(g.tris), C.c_int),
                                _p(_i32(dm.corner_slot), C.c_int), dm.nslot, len(sx), _p(sx), len(st),
                                _p(st if len(st) else np.zeros((1, 3), np.int32), C.c_int), _p(smu), nseg, _p(sg),
                                _p(sgam), nb, _p(bx), _p(ba), _p(bd), C.byref(o), nthreads, _p(b))
    if rc:
        raise ValueError("explicit_rhs failed")
    return b


# ------------------------------------------------------------------------------------------------------------------
# Isoparametric quads and mixed meshes (vp_oracle_quad.c: src/elements/iso_quad.jl, quad_field.jl, quad_rules.jl,
# src/assembly/mixed.jl)

SQUARE_KINDS = dict(far=0, near=1, self=2, edge0=3, edge1=4, edge2=5, edge3=6, vertex0=7, vertex1=8, vertex2=9,
                    vertex3=10, vertex_base=11)


def quad_field(qx, x, nx, vinf, side=0, polar=1.0, P=None, d=0.0):
    """quad_field(Q, x, nx, V; side, polar) (src/elements/quad_field.jl:40-46), or polar_field about P at distance d:
    returns (c[4], s)"""
    c = np.zeros(4)
    s = np.zeros(1)
    rc = lib().vpo_quad_field(_p(_f64(qx)), _p(_f64(x)), _p(_f64(nx)), _p(_f64(vinf)), C.c_int(side),
                              C.c_double(polar), _p(_f64(P)) if P is not None else None, C.c_double(d), _p(c), _p(s))
    if rc == -1:
        raise ValueError("quad_field: x lies on the quad's boundary, where the sheet's field is singular")
    if rc:
        raise ValueError("degenerate quadrilateral")
    return c, float(s[0])


def quad_project(qx, p):
    """project(Q, p), src/elements/iso_quad.jl:108-146: (P, d)"""
    P = np.zeros(2)
    d = np.zeros(1)
    if lib().vpo_quad_project(_p(_f64(qx)), _p(_f64(p)), _p(P), _p(d), None):
        raise ValueError("degenerate quadrilateral")
    return P, float(d[0])


def quad_info(qx):
    """IsoQuad's derived data (src/elements/iso_quad.jl:24-37): dict of area, diam, reach, centroid, n"""
    out = np.zeros(9)
    if lib().vpo_quad_project(_p(_f64(qx)), None, None, None, _p(out)):
        raise ValueError("degenerate quadrilateral")
    return dict(area=out[0], diam=out[1], reach=out[2], centroid=out[3:6].copy(), n=out[6:9].copy())


def square_rule(kind, o: Opts | None = None, quad_levels=4):
    """A rule of SquareTiers (src/elements/quad_rules.jl:73-87, square_tiers of src/assembly/mixed.jl:70-71):
    points (n, 2) and weights; vertex_base is barycentric (n, 3)"""
    o = o or opts()
    k = SQUARE_KINDS[kind] if isinstance(kind, str) else kind
    n = lib().vpo_square_rule(C.byref(o), quad_levels, k, None, None)
    st = 3 if k == 11 else 2
    p, w = np.zeros((n, st)), np.zeros(n)
    lib().vpo_square_rule(C.byref(o), quad_levels, k, _p(p), _p(w))
    return p, w


QUAD_RULE_KINDS = dict(quad=0, tri=1, wake=2, segment=3, vertex_square=4, contact=5)


def quad_rule(tx, kind, sx=None, tn=(0, 1, 2, 3), sn=(4, 5, 6, 7), self_pair=False, touch=False,
              o: Opts | None = None, quad_levels=4, corner=0):
    """The tier and rule of test quad tx for a source (quad_pair_rule, quad_wake_rule, quad_segment_rule, or
    vertex_square / square_contact_rule alone; src/assembly/mixed.jl:118-369): (tier name, points (n, 2), weights);
    the rule is None for a contact test that finds no contact"""
    o = o or opts()
    k = QUAD_RULE_KINDS[kind]
    tier = C.c_int(0)
    sxa = _f64(sx) if sx is not None else np.zeros(12)
    sel = corner if k == 4 else int(self_pair)
    args = (_p(_f64(tx)), _p(_i32(tn), C.c_int), k, _p(sxa), _p(_i32(sn), C.c_int), sel, int(touch), C.byref(o),
            quad_levels, C.byref(tier))
    n = lib().vpo_quad_rule(*args, None, None, 0)
    if n < 0:
        raise ValueError("degenerate element")
    p, w = np.zeros((n, 2)), np.zeros(n)
    lib().vpo_quad_rule(*args, _p(p), _p(w), n)
    if k == 5 and n == 0:
        return TIER_NAMES[tier.value], None, None
    return TIER_NAMES[tier.value], p, w


def quad_block(tx, sx, p, w, vinf, test=0, side=0):
    """block(T::IsoQuad, S::IsoQuad, rule, V, test, side), src/assembly/mixed.jl:376-388: (4 x 4 block, rb[4])"""
    p, w = _f64(np.asarray(p).reshape(-1, 2)), _f64(w)
    blk, rb = np.zeros((4, 4)), np.zeros(4)
    rc = lib().vpo_quad_block(_p(_f64(tx)), _p(_f64(sx)), len(w), _p(p), _p(w), _p(_f64(vinf)), C.c_int(test),
                              C.c_int(side), _p(blk), _p(rb))
    if rc:
        raise ValueError(f"vpo_quad_block failed ({rc})")
    return blk, rb


def assemble_mixed(x, mesh, nslot, vinf, touch=None, o: Opts | None = None, wake=None, quad_levels=4, nthreads=0,
                   tiers=False, rows=None):
    """assemble(g, dm, mesh::MixedMesh, V∞; opts, wake), src/assembly/mixed.jl:444-505.  mesh: quad_nodes,
    quad_slots (nq, 4), tri_nodes, tri_slots (nt, 3), quads first; touch: element pairs (k, 2).  Returns A
    (nslot x (nslot + nw)), b, pairs and points per tier (and the tiers, ne x (ne + wake triangles)).  rows = (e0, e1)
    assembles the rows of test elements e0..e1-1 only."""
    o = o or opts()
    nwx, wx, nwt, wt, nwf, wf = _wake_arrays(wake)
    qn, qs = _i32(mesh.quad_nodes).reshape(-1, 4), _i32(mesh.quad_slots).reshape(-1, 4)
    tn, ts = _i32(mesh.tri_nodes).reshape(-1, 3), _i32(mesh.tri_slots).reshape(-1, 3)
    nq, nt = len(qn), len(tn)
    A = np.zeros((nslot + nwx, nslot))
    b = np.zeros(nslot)
    pairs, points = np.zeros(6, np.int64), np.zeros(6, np.int64)
    tp = _i32(np.asarray(touch if touch is not None else np.zeros((0, 2)), dtype=np.int64).reshape(-1, 2))
    ntp = len(tp)
    if ntp == 0:
        tp = np.zeros((1, 2), np.int32)
    tr = np.zeros((nq + nt, nq + nt + nwt), np.uint8) if tiers else None
    x = _f64(np.asarray(x).reshape(-1, 3))
    rc = lib().vpo_assemble_mixed(len(x), _p(x), nq, _p(qn if nq else np.zeros((1, 4), np.int32), C.c_int),
                                  _p(qs if nq else np.zeros((1, 4), np.int32), C.c_int), nt,
                                  _p(tn if nt else np.zeros((1, 3), np.int32), C.c_int),
                                  _p(ts if nt else np.zeros((1, 3), np.int32), C.c_int), nslot, _p(_f64(vinf)),
                                  ntp, _p(tp, C.c_int), nwx, _p(wx), nwt, _p(wt, C.c_int), nwf, _p(wf, C.c_int),
                                  C.byref(o), quad_levels, nthreads, _p(A), _p(b), _p(pairs, C.c_longlong),
                                  _p(points, C.c_longlong), _p(tr, C.c_ubyte) if tiers else None,
                                  *(rows if rows is not None else (0, nq + nt)))
    if rc == -4:
        raise ValueError("quad_field: a test point lies on a source quad's boundary")
    if rc:
        raise ValueError(f"vpo_assemble_mixed failed ({rc})")
    out = (A.T, b, dict(zip(TIER_NAMES, pairs.tolist())), dict(zip(TIER_NAMES, points.tolist())))
    return out + (tr,) if tiers else out


def quad_field_many(qx, xs, nxs, vinf, side=0, P=None):
    """quad_field (or, with P (n, 2), polar_field about P at distance 0) at the points xs with directions nxs:
    returns c (n, 4), s (n)"""
    xs = _f64(np.asarray(xs).reshape(-1, 3))
    nxs = _f64(np.broadcast_to(np.asarray(nxs, dtype=np.float64).reshape(-1, 3), xs.shape))
    n = len(xs)
    c, s = np.zeros((n, 4)), np.zeros(n)
    Pa = _f64(np.asarray(P).reshape(-1, 2)) if P is not None else None
    rc = lib().vpo_quad_field_many(_p(_f64(qx)), n, _p(xs), _p(nxs), _p(_f64(vinf)), C.c_int(side),
                                   _p(Pa) if Pa is not None else None, _p(c), _p(s))
    if rc == -1:
        raise ValueError("quad_field: x lies on the quad's boundary, where the sheet's field is singular")
    if rc:
        raise ValueError("degenerate quadrilateral")
    return c, s
