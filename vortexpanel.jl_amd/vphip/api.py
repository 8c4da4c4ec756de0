"""Host-side mirror of the reference's Galerkin hot-path API over the C ABI.

Names and argument meaning follow the reference: `AssemblyOptions` (src/assembly/assemble.jl:26-41),
`build_elements` (src/elements/flat_triangle.jl:53), `source_strengths` (assemble.jl:364),
`assemble(g, dm, els, sigma, Vinf; opts, wake)` (assemble.jl:377-447), with the same errors: a bad option raises
ValueError (the reference's ArgumentError), a size mismatch raises ValueError naming the dimensions (its
DimensionMismatch).  The computation runs on the GPU through libvortexhip; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

TIER_NAMES = ("self", "edge", "vertex", "contact", "near", "far")
TESTS = {"hat": 0, "onering": 1, "dual": 2}


class VortexHipError(RuntimeError):
    pass


class ConvergenceError(VortexHipError):
    """FP64 iterative refinement did not meet its stopping test (VP_E_CONVERGENCE)"""


@dataclass
class AssemblyOptions:
    """src/assembly/assemble.jl:26-41 (`threaded` has no effect on the device); precision 'fp64' or 'mixed'"""
    test: str = "hat"
    far_ratio: float = 3.0
    far_points: int = 3
    near_points: int = 7
    near_ratio: float = 1.5
    near_depth: int = 3
    touch_levels: int = 6
    touch_ratio: float = 0.15
    touch_order: int = 3
    corner_angle: float = 75.0
    contact_rel: float = 0.1
    kernel_far: float = 10.0
    quad_touch_levels: int = 4
    threaded: bool = True
    precision: str = "fp64"

    def to_c(self) -> L.AssemblyOptionsC:
        """check_options, src/assembly/assemble.jl:43-46"""
        if self.test not in TESTS:
            raise ValueError(f"test must be hat, onering or dual, not {self.test}")
        if self.precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be fp64 or mixed, not {self.precision}")
        kf = float(self.kernel_far)
        return L.AssemblyOptionsC(TESTS[self.test], self.far_ratio, self.far_points, self.near_points, self.near_ratio,
                                  self.near_depth, self.touch_levels, self.touch_ratio, self.touch_order,
                                  self.corner_angle, self.contact_rel, kf if np.isfinite(kf) else 1e300,
                                  0 if self.precision == "fp64" else 1, self.quad_touch_levels)


def _check(ctx, rc):
    if rc != 0:
        msg = L.load().vp_last_error(ctx).decode() if ctx else "no context"
        if rc in (L.VP_E_ARGUMENT, L.VP_E_DIMENSION):
            raise ValueError(msg)
        if rc == L.VP_E_CONVERGENCE:
            raise ConvergenceError(f"vortexhip status {rc}: {msg}")
        raise VortexHipError(f"vortexhip status {rc}: {msg}")


class Context:
    """A device and its stream (vp_create / vp_destroy)"""

    def __init__(self, device: int = 0):
        lib = L.load()
        h = C.c_void_p()
        rc = lib.vp_create(device, C.byref(h))
        if rc != 0:
            raise VortexHipError(f"vp_create({device}) failed with status {rc} (no GPU?)")
        self.h = h
        self.lib = lib

    def close(self):
        if self.h:
            self.lib.vp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(self.h, self.lib.vp_sync(self.h))


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class Elements:
    """FlatTriangle data of a geometry (src/elements/flat_triangle.jl:12-31), as arrays"""
    n: np.ndarray
    area: np.ndarray
    centroid: np.ndarray
    gradl: np.ndarray
    gamma: np.ndarray

    def __len__(self):
        return len(self.area)


def build_elements(g) -> Elements:
    """src/elements/flat_triangle.jl:53 (host arrays; the device builds its own records in elem_setup)"""
    x = np.asarray(g.x, dtype=np.float64)
    t = g.tris
    a, b, c = x[t[:, 0]], x[t[:, 1]], x[t[:, 2]]
    nA = np.cross(b - a, c - a)
    A2 = np.linalg.norm(nA, axis=1)
    if np.any(A2 <= 0):
        raise ValueError("degenerate triangle")
    n = nA / A2[:, None]
    e = np.stack([c - b, a - c, b - a], axis=1)
    return Elements(n=n, area=A2 / 2, centroid=(a + b + c) / 3, gradl=np.cross(n[:, None, :], e) / A2[:, None, None],
                    gamma=-e / A2[:, None, None])


def source_strengths(els: Elements, vinf, v_body=None) -> np.ndarray:
    """src/assembly/assemble.jl:364"""
    V = np.asarray(vinf, dtype=np.float64) - (0 if v_body is None else np.asarray(v_body, dtype=np.float64))
    return -(els.n @ V)


def _surface_c(g, dm):
    keep = dict(x=L.f64(g.x), tris=L.i32(g.tris), cs=L.i32(dm.corner_slot),
                touch=L.i32(np.asarray(g.touch_pairs).reshape(-1, 2)) if len(g.touch_pairs) else np.zeros((1, 2), np.int32))
    s = L.SurfaceC(len(g.x), L.ptr(keep["x"]), len(g.tris), L.ptr(keep["tris"], C.c_int), L.ptr(keep["cs"], C.c_int),
                   dm.nslot, len(g.touch_pairs), L.ptr(keep["touch"], C.c_int))
    return s, keep


def _wake_c(w):
    if w is None:
        return None, {}
    keep = dict(x=L.f64(w.x), tris=L.i32(w.tris.reshape(-1, 3)), fil=L.i32(w.filaments.reshape(-1, 2)))
    if len(keep["tris"]) == 0:
        keep["tris"] = np.zeros((1, 3), np.int32)
    if len(keep["fil"]) == 0:
        keep["fil"] = np.zeros((1, 2), np.int32)
    c = L.WakeC(len(w.x), L.ptr(keep["x"]), len(w.tris), L.ptr(keep["tris"], C.c_int), len(w.filaments),
                L.ptr(keep["fil"], C.c_int))
    return c, keep


class AssemblyPlan:
    """vp_plan_*: geometry, element records, tiles and rules resident on the device; A and b stay on the device"""

    def __init__(self, g, dm, wake=None, opts: AssemblyOptions | None = None, ctx: Context | None = None,
                 record_tiers: bool = False, timing: bool = False, wake_only: bool = False):
        self.ctx = ctx or default_context()
        self.opts = opts or AssemblyOptions()
        oc = self.opts.to_c()
        s, self._ks = _surface_c(g, dm)
        w, self._kw = _wake_c(wake)
        lib = self.ctx.lib
        self.nslot, self.nt = dm.nslot, len(g.tris)
        self.nw = 0 if wake is None else len(wake.x)
        self.nwt = 0 if wake is None else len(wake.tris)
        h = C.c_void_p()
        rc = lib.vp_plan_create_ex(self.ctx.h, C.byref(s), C.byref(w) if w is not None else None, C.byref(oc),
                                   (1 if record_tiers else 0) | (2 if timing else 0) | (4 if wake_only else 0),
                                   C.byref(h))
        _check(self.ctx.h, rc)
        self.h = h
        self.record_tiers = record_tiers
        self.pairs = None
        self.points = None

    def assemble(self, sigma, vinf):
        sig = L.f64(sigma)
        if sig.shape != (self.nt,):
            raise ValueError(f"sigma has {sig.shape}, expected ({self.nt},)")
        pairs, points = np.zeros(6, np.int64), np.zeros(6, np.int64)
        rc = self.ctx.lib.vp_plan_assemble(self.h, L.ptr(sig), L.ptr(L.f64(vinf)), L.ptr(pairs, C.c_longlong),
                                           L.ptr(points, C.c_longlong))
        _check(self.ctx.h, rc)
        self.pairs = dict(zip(TIER_NAMES, pairs.tolist()))
        self.points = dict(zip(TIER_NAMES, points.tolist()))
        return self

    def timings(self):
        """Device times (ms) of the last assembly's kernels, from HIP events on the plan's stream"""
        ms = np.zeros(8)
        _check(self.ctx.h, self.ctx.lib.vp_plan_timings(self.h, L.ptr(ms)))
        return dict(far=float(ms[0]), classify=float(ms[1]), near=float(ms[2]), touching=float(ms[3]),
                    freestream=float(ms[4]), near_single_pairs=int(ms[5]), touching_leaves=int(ms[6]),
                    far_exact_points=int(ms[7]))

    def tiers(self):
        """Tier index of every (body test triangle, source triangle) pair of the last assembly"""
        out = np.zeros((self.nt, self.nt + self.nwt), np.uint8)
        _check(self.ctx.h, self.ctx.lib.vp_plan_tiers(self.h, L.ptr(out, C.c_uint8)))
        return out

    def wake_product(self, mu_wake):
        """A_wake mu_wake (nslot): the wake columns of the last assembly times the wake nodes' doublets"""
        out = np.zeros(self.nslot)
        mw = L.f64(mu_wake)
        if mw.shape != (self.nw,):
            raise ValueError(f"mu_wake has {mw.shape}, expected ({self.nw},)")
        _check(self.ctx.h, self.ctx.lib.vp_plan_wake_product(self.h, L.ptr(mw) if self.nw else None, L.ptr(out)))
        return out

    def update_wake(self, x):
        """New coordinates (nw x 3) for the wake's nodes, same triangles and filaments: the plan is reused (assemble
        again before reading)"""
        wx = L.f64(x).reshape(-1, 3)
        if wx.shape != (self.nw, 3):
            raise ValueError(f"wake coordinates have {wx.shape}, expected ({self.nw}, 3)")
        _check(self.ctx.h, self.ctx.lib.vp_plan_update_wake(self.h, L.ptr(wx)))

    def blob_rhs(self, bx, alpha, delta):
        """explicit_rhs of vortex particles on the plan's body (nslot values), src/assembly/explicit.jl:60-79"""
        bx, ba = L.f64(bx).reshape(-1, 3), L.f64(alpha).reshape(-1, 3)
        bd = L.f64(delta).reshape(-1)
        out = np.zeros(self.nslot)
        _check(self.ctx.h, self.ctx.lib.vp_plan_blob_rhs(self.h, len(bd), L.ptr(bx), L.ptr(ba), L.ptr(bd), L.ptr(out)))
        return out

    def read_b(self):
        """b of the last assembly (A stays on the device)"""
        b = np.zeros(self.nslot)
        _check(self.ctx.h, self.ctx.lib.vp_plan_read(self.h, None, L.ptr(b)))
        return b

    def read(self):
        A = np.zeros((self.nslot + self.nw, self.nslot))  # column-major nslot x ncol
        b = np.zeros(self.nslot)
        _check(self.ctx.h, self.ctx.lib.vp_plan_read(self.h, L.ptr(A), L.ptr(b)))
        return A.T, b

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.vp_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def assemble(g, dm, els, sigma, vinf, opts: AssemblyOptions | None = None, external=None, wake=None,
             ctx: Context | None = None):
    """`assemble(g, dm, els, sigma, Vinf; opts, external, wake)` -> (A, b, stats), src/assembly/assemble.jl:377-447.

    A is nslot x (nslot + nwake) (rows: test slots; columns: slots then wake nodes), b has nslot entries, stats holds
    the pairs and points per tier.  `external` (a Julia closure in the reference) is not crossable: pass explicit
    sources to `explicit_rhs` instead.
    """
    if external is not None:
        raise ValueError("external velocity closures are not supported on the device; use explicit_rhs")
    if len(els) != len(g.tris):
        raise ValueError(f"els has {len(els)} elements, the geometry {len(g.tris)} triangles")
    plan = AssemblyPlan(g, dm, wake, opts, ctx)
    import time
    t0 = time.perf_counter()
    plan.assemble(sigma, vinf)
    A, b = plan.read()
    dt = time.perf_counter() - t0
    stats = dict(pairs=plan.pairs, points=plan.points, time=dt)
    plan.close()
    return A, b, stats


def assemble_mixed(g, dm, mesh, vinf, opts: AssemblyOptions | None = None, external=None, wake=None,
                   ctx: Context | None = None, tiers: bool = False):
    """`assemble(g, dm, mesh::MixedMesh, Vinf; opts, external, wake)` -> (A, b, stats), src/assembly/mixed.jl:444-505,
    on the device (vp_assemble_mixed): the quads' inner integrals by quad_field, their test integrals on the
    parameter-square tiers, the triangles' as `assemble`.  `mesh` is a `vphip.mixed.MixedMesh`; the sources are
    -Vinf.n on every element (so no sigma argument).  FP64.  With tiers=True, stats["tiers"] holds the tier of every
    (element, body element or wake triangle) pair."""
    from .mixed import element_touch_pairs

    if external is not None:
        raise ValueError("external velocity closures are not supported on the device; use explicit_rhs")
    o = (opts or AssemblyOptions()).to_c()
    ctx = ctx or default_context()
    lib = L.load()
    keep = dict(x=L.f64(mesh.x), qn=L.i32(mesh.quad_nodes.reshape(-1, 4)), qs=L.i32(mesh.quad_slots.reshape(-1, 4)),
                tn=L.i32(mesh.tri_nodes.reshape(-1, 3)), ts=L.i32(mesh.tri_slots.reshape(-1, 3)),
                tp=L.i32(element_touch_pairs(g, mesh).reshape(-1, 2)))
    for k, w in (("qn", 4), ("qs", 4), ("tn", 3), ("ts", 3), ("tp", 2)):
        if len(keep[k]) == 0:
            keep[k] = np.zeros((1, w), np.int32)
    nq, nt = mesh.nq, mesh.ntri
    ntp = len(element_touch_pairs(g, mesh))
    m = L.MixedMeshC(len(mesh.x), L.ptr(keep["x"]), nq, L.ptr(keep["qn"], C.c_int), L.ptr(keep["qs"], C.c_int), nt,
                     L.ptr(keep["tn"], C.c_int), L.ptr(keep["ts"], C.c_int), dm.nslot, ntp, L.ptr(keep["tp"], C.c_int))
    wc, wkeep = _wake_c(wake)
    nw = len(wake.x) if wake is not None else 0
    nwt = len(wake.tris) if wake is not None else 0
    A = np.zeros((dm.nslot + nw, dm.nslot))  # column-major nslot x ncol
    b = np.zeros(dm.nslot)
    pairs, points = np.zeros(6, np.int64), np.zeros(6, np.int64)
    tr = np.zeros((nq + nt, nq + nt + nwt), np.uint8) if tiers else None
    import time
    t0 = time.perf_counter()
    rc = lib.vp_assemble_mixed(ctx.h, C.byref(m), C.byref(wc) if wc is not None else None, L.ptr(L.f64(vinf)),
                               C.byref(o), L.ptr(A), L.ptr(b), L.ptr(pairs, C.c_longlong), L.ptr(points, C.c_longlong),
                               L.ptr(tr, C.c_uint8) if tiers else None)
    dt = time.perf_counter() - t0
    _check(ctx.h, rc)
    names = ("self", "edge", "vertex", "contact", "near", "far")
    stats = dict(pairs=dict(zip(names, pairs.tolist())), points=dict(zip(names, points.tolist())), time=dt)
    if tiers:
        stats["tiers"] = tr
    return A.T, b, stats
