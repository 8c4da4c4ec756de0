"""Free wake: the implicit trailing edge panel factored once with the body, free panel rows shed every step, their
conversion into vortex particles, forward-Euler advection and the loads of every step.

Mirrors src/freewake/state.jl and src/freewake/step.jl (same names, same argument meaning, MESA's step order).  The
host keeps the wake's bookkeeping (rows, conversions, merging: small and branchy); the device does the work that
scales: the right hand side of the known sheets, filaments and particles (`explicit_rhs`, a wake-only assembly plan and
the particle kernels), the solve with the factored matrix, and the velocities at every free node and particle (the body
N-body sum and the blob sum).  Positions, strengths and doublets stay FP64; with precision 'mixed' the velocity sums use
FP32 kernels with FP64 sums (the circulation bookkeeping stays FP64).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
import warnings
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import post
from .api import (AssemblyOptions, AssemblyPlan, Context, _check, _surface_c, build_elements, default_context,
                  source_strengths)
from .dofmap import wake_rows
from .solve import factorize_galerkin, solve_factored
from .wake import te_panel


# ------------------------------------------------------------------------------------------ options and cores --------

@dataclass
class FreeWakeOptions:
    """src/freewake/state.jl:27-39"""
    dt: float
    te_len: float | None = None
    te_direction: str = "freestream"
    max_rows: int | str = "auto"
    merge_rows: int | str = 1
    core: str | float = "auto"
    overlap: float = 1.5
    core_factor: float = 1.5
    core_floor: float | None = None
    advection: str = "free"
    delete_beyond: float = math.inf

    def check(self):
        """check_options, state.jl:41-51"""
        if not self.dt > 0:
            raise ValueError("dt must be positive")
        if not (self.max_rows == "auto" or (isinstance(self.max_rows, int) and self.max_rows >= 0)):
            raise ValueError("max_rows must be auto or at least 0")
        if not (self.merge_rows == "auto" or (isinstance(self.merge_rows, int) and self.merge_rows >= 1)):
            raise ValueError("merge_rows must be auto or at least 1")
        if self.advection not in ("free", "freestream"):
            raise ValueError(f"advection must be free or freestream, not {self.advection}")
        if self.te_direction not in ("freestream", "bisector"):
            raise ValueError("te_direction must be freestream or bisector")
        if not (self.core in ("auto", "scaled") or (isinstance(self.core, (int, float)) and self.core > 0)):
            raise ValueError(f"core must be auto, scaled or a positive number, not {self.core}")
        if not (self.overlap > 0 and self.core_factor > 0):
            raise ValueError("overlap and core_factor must be positive")
        return self


def _norm(v):
    return math.sqrt(float(np.dot(v, v)))


class CoreRule:
    """Core radius of particles and of the blobs that stand for wake panels and filaments, state.jl:68-92"""

    def __init__(self, o: FreeWakeOptions, vinf):
        self.floor = float(o.core_floor) if o.core_floor is not None else _norm(vinf) * o.dt
        self.overlap, self.factor = float(o.overlap), float(o.core_factor)
        if isinstance(o.core, (int, float)) and not isinstance(o.core, bool):
            self.mode, self.value = "fixed", float(o.core)
        else:
            self.mode, self.value = o.core, 0.0

    def tri(self, p, q, r):
        if self.mode == "fixed":
            return self.value
        if self.mode == "auto":
            return self.overlap * max(_norm(q - p), _norm(r - q), _norm(p - r)) / 2
        return max(self.factor * math.sqrt(_norm(np.cross(q - p, r - p))), self.floor)

    def seg(self, a, b):
        if self.mode == "fixed":
            return self.value
        if self.mode == "auto":
            return self.overlap * _norm(b - a) / 2
        return max(self.factor * _norm(b - a) / 2, self.floor)


@dataclass
class Topology:
    """Strips of the free wake (state.jl:99-121): neighbouring wake trailing edge nodes along tau, and the end nodes of
    open lines (the wake's tips)"""
    nte: int
    segments: list
    first_nodes: list
    last_nodes: list

    @staticmethod
    def of(dm):
        wte = dm.wake_te
        segs, firsts, lasts = [], [], []
        for line, closed in zip(wte.lines, wte.closed):
            line = [int(v) for v in line]
            m = len(line)
            segs += [(line[i], line[i + 1]) for i in range(m - 1)]
            if closed:
                segs.append((line[m - 1], line[0]))
            else:
                firsts.append(line[0])
                lasts.append(line[m - 1])
        return Topology(wte.n, segs, firsts, lasts)


# ------------------------------------------------------------------------------------------ state and elements ------

@dataclass
class Blobs:
    """Vortex particles (src/kernels/blob.jl:7-11): positions, strengths, cores"""
    x: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    alpha: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    delta: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def __len__(self):
        return len(self.delta)

    @staticmethod
    def cat(parts):
        parts = [p for p in parts if len(p)]
        if not parts:
            return Blobs()
        return Blobs(np.vstack([p.x for p in parts]), np.vstack([p.alpha for p in parts]),
                     np.concatenate([p.delta for p in parts]))

    def take(self, keep):
        return Blobs(self.x[keep], self.alpha[keep], self.delta[keep])


@dataclass
class FreeWakeState:
    """state.jl:123-143: node rows newest (R0) to oldest, their frozen doublets, mu_line, particles and merge data"""
    x: list
    mu: list
    mu_line: np.ndarray
    particles: Blobs
    conversions: int
    tag: np.ndarray
    group: np.ndarray
    batch_rows: np.ndarray
    batch_length: np.ndarray
    batch_width: np.ndarray

    @staticmethod
    def empty(nte, nstrips=0):
        return FreeWakeState([], [], np.zeros(nte), Blobs(), 0, np.zeros(0, np.int64), np.zeros(0, np.int8),
                             np.zeros(nstrips, np.int64), np.zeros(nstrips), np.zeros(nstrips))

    def keep_particles(self, keep):
        """keep_particles!, state.jl:146-151"""
        self.particles = self.particles.take(keep)
        self.tag = self.tag[keep]
        self.group = self.group[keep]


@dataclass
class WakeElements:
    """Known vortex elements (state.jl:157-165): triangles with corner doublets (those with an area also as sheets) and
    filament segments with their end circulations"""
    corners: list = field(default_factory=list)
    corner_mu: list = field(default_factory=list)
    sheets: list = field(default_factory=list)
    sheet_mu: list = field(default_factory=list)
    segments: list = field(default_factory=list)
    segment_gamma: list = field(default_factory=list)


def nondegenerate(p, q, r):
    """state.jl:171: |(q - p) x (r - p)| > 1e-12 max squared edge, on Python floats (a 3-vector np.cross costs more
    than the arithmetic)"""
    p0, p1, p2 = p.tolist()
    q0, q1, q2 = q.tolist()
    r0, r1, r2 = r.tolist()
    u0, u1, u2 = q0 - p0, q1 - p1, q2 - p2
    v0, v1, v2 = r0 - p0, r1 - p1, r2 - p2
    w0, w1, w2 = r0 - q0, r1 - q1, r2 - q2
    c0, c1, c2 = u1 * v2 - u2 * v1, u2 * v0 - u0 * v2, u0 * v1 - u1 * v0
    return math.sqrt(c0 * c0 + c1 * c1 + c2 * c2) > 1e-12 * max(u0 * u0 + u1 * u1 + u2 * u2,
                                                                 v0 * v0 + v1 * v1 + v2 * v2,
                                                                 w0 * w0 + w1 * w1 + w2 * w2)


def add_sheet(e, p, q, r, mp, mq, mr):
    """add_sheet!, state.jl:174-183"""
    mu = np.array([mp, mq, mr], dtype=np.float64)
    e.corners.append((p, q, r))
    e.corner_mu.append(mu)
    if nondegenerate(p, q, r):
        e.sheets.append((p, q, r))
        e.sheet_mu.append(mu)


def add_segment(e, a, b, ga, gb):
    e.segments.append((a, b))
    e.segment_gamma.append((float(ga), float(gb)))


def panel_row(e, st, topo, i):
    """panel_row!, state.jl:188-201 (i 0-based: the row between node rows i and i + 1)"""
    x0, x1, m0, m1 = st.x[i], st.x[i + 1], st.mu[i], st.mu[i + 1]
    for A, B in topo.segments:
        add_sheet(e, x0[A], x1[A], x0[B], m0[A], m1[A], m0[B])
        add_sheet(e, x1[A], x1[B], x0[B], m1[A], m1[B], m0[B])
    for k in topo.first_nodes:
        add_segment(e, x1[k], x0[k], m1[k], m0[k])
    for k in topo.last_nodes:
        add_segment(e, x0[k], x1[k], m0[k], m1[k])


def far_end(e, st, topo, i):
    """far_end!, state.jl:204-212"""
    x, mu = st.x[i], st.mu[i]
    for A, B in topo.segments:
        gA, gB = mu[A] - st.mu_line[A], mu[B] - st.mu_line[B]
        if gA == 0 and gB == 0:
            continue
        add_segment(e, x[B], x[A], gB, gA)


def vortex_elements(st, topo):
    """vortex_elements(st, topo), state.jl:222-237"""
    e = WakeElements()
    nr = len(st.x)
    for i in range(nr - 1):
        panel_row(e, st, topo, i)
    if nr >= 1:
        for A, B in topo.segments:
            add_segment(e, st.x[0][A], st.x[0][B], st.mu[0][A], st.mu[0][B])
        far_end(e, st, topo, nr - 1)
    return e


def rigid_elements(w, mu_w):
    """vortex_elements(w::RigidWake, mu_w), state.jl:240-250"""
    e = WakeElements()
    mn = np.asarray(mu_w)[w.te]
    for t in w.tris:
        add_sheet(e, w.x[t[0]], w.x[t[1]], w.x[t[2]], mn[t[0]], mn[t[1]], mn[t[2]])
    for a, b in w.filaments:
        add_segment(e, w.x[a], w.x[b], mn[a], mn[b])
    return e


def sheet_vorticity(p, q, r, mu):
    """state.jl:258: the vorticity of a triangle's sheet without its area"""
    return ((mu[0] + mu[1]) * (q - p) + (mu[1] + mu[2]) * (r - q) + (mu[2] + mu[0]) * (p - r)) / 2


def segment_blobs(a, b, ga, gb, delta):
    """src/kernels/blob.jl:33-38: two 2-point Gauss blobs carrying a filament's strength and first moment"""
    s = 1 / (2 * math.sqrt(3))
    xs, al = [], []
    for xi in (0.5 - s, 0.5 + s):
        xs.append(a + xi * (b - a))
        al.append((((1 - xi) * ga + xi * gb) / 2) * (b - a))
    return xs, al, [delta, delta]


def free_wake_blobs(e, core):
    """state.jl:268-278: a particle at each triangle's centroid with its sheet's vorticity, two per filament"""
    xs, al, dl = [], [], []
    for (p, q, r), mu in zip(e.corners, e.corner_mu):
        xs.append((p + q + r) / 3)
        al.append(sheet_vorticity(p, q, r, mu))
        dl.append(core.tri(p, q, r))
    for (a, b), (ga, gb) in zip(e.segments, e.segment_gamma):
        x2, a2, d2 = segment_blobs(a, b, ga, gb, core.seg(a, b))
        xs += x2
        al += a2
        dl += d2
    if not dl:
        return Blobs()
    return Blobs(np.array(xs, dtype=np.float64).reshape(-1, 3), np.array(al, dtype=np.float64).reshape(-1, 3),
                 np.array(dl, dtype=np.float64))


def convert_oldest(st, topo, core, merge=None):
    """convert_oldest!, state.jl:292-328: the oldest panel row into particles; returns how many"""
    nr = len(st.x)
    if nr < 2:
        return 0
    e = WakeElements()
    panel_row(e, st, topo, nr - 2)
    far_end(e, st, topo, nr - 1)
    blobs = free_wake_blobs(e, core)
    tag, group = np.zeros(len(blobs), np.int64), np.zeros(len(blobs), np.int8)
    if merge is not None:
        up, down = st.x[nr - 2], st.x[nr - 1]
        for s, (A, B) in enumerate(topo.segments):
            k = merge[s]
            if k <= 1:
                continue
            m = k - st.batch_rows[s]
            for h, (u, v) in enumerate(((1 / 3, 1 / 3), (2 / 3, 2 / 3))):
                tag[2 * s + h] = s + 1
                group[2 * s + h] = 1 if u + (m - 1 + v) / k < 1 else 2
            st.batch_rows[s] += 1
            st.batch_length[s] += (_norm(down[A] - up[A]) + _norm(down[B] - up[B])) / 2
            st.batch_width[s] = _norm(up[B] - up[A])
    st.particles = Blobs.cat([st.particles, blobs])
    st.tag = np.concatenate([st.tag, tag])
    st.group = np.concatenate([st.group, group])
    st.mu_line = st.mu[nr - 2].copy()
    st.x.pop()
    st.mu.pop()
    st.conversions += 1
    if merge is not None:
        for s in range(len(topo.segments)):
            if merge[s] > 1 and st.batch_rows[s] >= merge[s]:
                merge_strip(st, s, core)
    return len(blobs)


def merge_strip(st, s, core):
    """merge_strip!, state.jl:337-357 (strip s 0-based; tags hold s + 1, 0 when final)"""
    idx = np.nonzero(st.tag == s + 1)[0]
    o = np.zeros(3)
    d = core.tri(o, o + np.array([st.batch_width[s], 0, 0]), o + np.array([st.batch_width[s], st.batch_length[s], 0]))
    mx, ma = [], []
    for gi in (1, 2):
        sel = [i for i in idx if st.group[i] == gi]
        if not sel:
            continue
        al = np.sum(st.particles.alpha[sel], axis=0)
        w = np.array([_norm(st.particles.alpha[i]) for i in sel])
        if w.sum() > 0:
            x = np.sum(w[:, None] * st.particles.x[sel], axis=0) / w.sum()
        else:
            x = np.sum(st.particles.x[sel], axis=0) / len(sel)
        mx.append(x)
        ma.append(al)
    st.keep_particles(st.tag != s + 1)
    merged = Blobs(np.array(mx).reshape(-1, 3), np.array(ma).reshape(-1, 3), np.full(len(mx), d))
    st.particles = Blobs.cat([st.particles, merged])
    st.tag = np.concatenate([st.tag, np.zeros(len(mx), np.int64)])
    st.group = np.concatenate([st.group, np.zeros(len(mx), np.int8)])
    st.batch_rows[s] = 0
    st.batch_length[s] = 0.0


def vorticity_moments(b):
    """state.jl:365-369: total vorticity and impulse of a set of particles"""
    if not len(b):
        return np.zeros(3), np.zeros(3)
    return b.alpha.sum(0), np.cross(b.x, b.alpha).sum(0) / 2


# ------------------------------------------------------------------------------------------ explicit rhs -------------

class ExplicitRHS:
    """explicit_rhs (src/assembly/explicit.jl:31-86) on one body, through the vp_explicit_* handle: the known sheets and
    filaments take the assembly's tiers (a wake-only plan: exactly their assembled columns times their doublets, one
    wake node per corner and filament end), the particles the far rule or near_rule.  The handle moves its plan of the
    known elements in place while their counts stay (the free rows move every step, keep their triangles) and keeps
    a plan of the body for the particles."""

    def __init__(self, g, dm, opts: AssemblyOptions | None = None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.nslot = dm.nslot
        s, _keep = _surface_c(g, dm)  # the handle copies the arrays
        oc = (opts or AssemblyOptions()).to_c()
        h = C.c_void_p()
        _check(self.ctx.h, self.ctx.lib.vp_explicit_create(self.ctx.h, C.byref(s), C.byref(oc), C.byref(h)))
        self.h = h

    def __call__(self, e: WakeElements | None, blobs: Blobs | None):
        ns = len(e.sheets) if e is not None else 0
        nf = len(e.segments) if e is not None else 0
        sx = L.f64(np.asarray(e.sheets, dtype=np.float64).reshape(ns, 9)) if ns else None
        sm = L.f64(np.asarray(e.sheet_mu, dtype=np.float64).reshape(ns, 3)) if ns else None
        fx = L.f64(np.asarray(e.segments, dtype=np.float64).reshape(nf, 6)) if nf else None
        fg = L.f64(np.asarray(e.segment_gamma, dtype=np.float64).reshape(nf, 2)) if nf else None
        nb = len(blobs) if blobs is not None else 0
        bx = L.f64(blobs.x) if nb else None
        ba = L.f64(blobs.alpha) if nb else None
        bd = L.f64(blobs.delta) if nb else None
        out = np.zeros(self.nslot)
        rc = self.ctx.lib.vp_explicit_rhs(self.h, ns, L.ptr(sx), L.ptr(sm), nf, L.ptr(fx), L.ptr(fg), nb, L.ptr(bx),
                                          L.ptr(ba), L.ptr(bd), L.ptr(out))
        _check(self.ctx.h, rc)
        return out

    def info(self):
        """(plans of known elements built, moved in place, calls)"""
        c = np.zeros(3, dtype=np.int64)
        _check(self.ctx.h, self.ctx.lib.vp_explicit_info(self.h, L.ptr(c, C.c_longlong)))
        return tuple(int(v) for v in c)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.vp_explicit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def explicit_rhs(g, dm, e: WakeElements | None, blobs: Blobs | None, opts: AssemblyOptions | None = None,
                 ctx: Context | None = None, handle: ExplicitRHS | None = None):
    """-int psi_i n.u_src dA for every slot (src/assembly/explicit.jl:31-86) on the device: one call of an
    ExplicitRHS handle (a temporary one unless `handle` is given)"""
    h = handle or ExplicitRHS(g, dm, opts, ctx)
    try:
        return h(e, blobs)
    finally:
        if handle is None:
            h.close()


# ------------------------------------------------------------------------------------------ the simulation ----------

def first_core(core, edge, topo, row):
    """step.jl:65-70"""
    w = max(_norm(edge[B] - edge[A]) for A, B in topo.segments)
    o = np.zeros(3)
    return core.tri(o, o + np.array([w, 0, 0]), o + np.array([w, row, 0]))


def conversion_distance(max_rows, core, edge, topo, te_len, row):
    """step.jl:72-84"""
    d = first_core(core, edge, topo, row)
    return (te_len + (max_rows + 0.5) * row) / d, d


def rows_for_distance(cores, core, edge, topo, te_len, row):
    """step.jl:86-88"""
    return max(0, math.ceil((cores * first_core(core, edge, topo, row) - te_len) / row - 0.5))


def check_conversion_distance(max_rows, core, edge, topo, te_len, row):
    """step.jl:90-98"""
    d, dl = conversion_distance(max_rows, core, edge, topo, te_len, row)
    if d < 2:
        need = rows_for_distance(2, core, edge, topo, te_len, row)
        warnings.warn(f"Particles would be created {d:.2f} cores behind the trailing edge (core {dl:.3g}), which biases "
                      f"the lift; max_rows = {need} keeps them two cores away")
    return d


class FreeWakeHost:
    """FreeWakeSim (step.jl:8-58) with its step (step!, step.jl:113-146) written out on the host, state.jl's
    bookkeeping line by line; the device work is in `_explicit_rhs`, `_solve` and `_velocities`.  It is the readable
    mirror of the reference's step, and the checker's frame: the oracle's simulation overrides those three methods
    with the FP64 CPU oracle.  FreeWakeSim runs the same step in the native engine (vp_freewake_*).

    precision: 'fp64' or 'mixed' (the factorization in FP32 with FP64 refinement, FP32 velocity kernels with FP64 sums)."""

    def __init__(self, g, dm, vinf, opts: FreeWakeOptions, assembly: AssemblyOptions | None = None,
                 precision: str = "fp64", ctx: Context | None = None):
        opts.check()
        self.g, self.dm, self.opts = g, dm, opts
        self.V = np.asarray(vinf, dtype=np.float64)
        self.assembly = assembly or AssemblyOptions()
        self.precision = precision
        self.ctx = ctx
        t0 = time.perf_counter()
        self.els = build_elements(g)
        self.sigma = source_strengths(self.els, self.V)
        te_len = opts.te_len if opts.te_len is not None else 0.3 * _norm(self.V) * opts.dt
        self.te = te_panel(g, dm, self.V, length=te_len, direction=opts.te_direction)
        self._setup()
        self.t_setup = time.perf_counter() - t0
        self.topo = Topology.of(dm)
        nte = self.topo.nte
        self.edge = self.te.x[nte:2 * nte].copy()  # the trailing edge panel's downstream edge (its row 1)
        self.core = CoreRule(opts, self.V)
        row = _norm(self.V) * opts.dt
        if opts.max_rows == "auto":
            self.keep_rows = rows_for_distance(3, self.core, self.edge, self.topo, self.te.rows[1], row)
        else:
            self.keep_rows = int(opts.max_rows)
            check_conversion_distance(self.keep_rows, self.core, self.edge, self.topo, self.te.rows[1], row)
        widths = [_norm(self.edge[B] - self.edge[A]) for A, B in self.topo.segments]
        if opts.merge_rows == "auto":
            self.merge = [max(1, int(round(w / row))) for w in widths]
        else:
            self.merge = None if opts.merge_rows == 1 else [int(opts.merge_rows)] * len(widths)
        self.state = FreeWakeState.empty(nte, len(self.topo.segments))
        self.n, self.t = 0, 0.0
        self.mu = np.zeros(dm.nslot)
        self.mu_w = np.zeros(nte)
        self.cp = np.zeros(len(g.tris))
        self.I_prev = np.zeros(3)
        self.history = []

    # --- device work (overridable) ---
    def _setup(self):
        """assemble the body with the trailing edge panel and factor the bordered matrix once (step.jl:38-46)"""
        ctx = self.ctx or default_context()
        self.ctx = ctx
        o = AssemblyOptions(**{**self.assembly.__dict__, "precision": self.precision})
        plan = AssemblyPlan(self.g, self.dm, self.te, o, ctx=ctx)
        plan.assemble(self.sigma, self.V)
        self.b0 = plan.read_b()
        self.F = factorize_galerkin(self.dm, plan, W=wake_rows(self.dm, self.te.te), mode=self.precision)
        plan.close()
        self.explicit = ExplicitRHS(self.g, self.dm, self.assembly, ctx)  # its plans are moved from step to step
        self.field = post.WakeField(self.g, kernel_far=self.assembly.kernel_far, precision=self.precision, ctx=ctx)

    def _explicit_rhs(self, we, particles):
        return self.explicit(we, particles)

    def _solve(self, rhs):
        out = solve_factored(self.F, rhs)
        return out["q"], out["mu"], out["lam"]

    def _velocities(self, gam, blobs, X):
        """the body's kernels and every wake blob at X: one replay of the wake field's HIP graph"""
        if len(blobs):
            return self.field(self.sigma, gam, X, blobs.x, blobs.alpha, blobs.delta)
        return self.field(self.sigma, gam, X)

    # --- the step ---
    def step(self):
        """step!, step.jl:113-146"""
        st, o, dm = self.state, self.opts, self.dm
        we = vortex_elements(st, self.topo)
        t0 = time.perf_counter()
        rhs = self._explicit_rhs(we, st.particles)
        t_rhs = time.perf_counter() - t0
        q, mu, lam = self._solve(self.b0 + rhs)
        mu_w = dm.K @ q
        loads = self.step_loads(q, mu, mu_w, we)
        if not st.x:
            st.x.append(self.edge.copy())
            st.mu.append(mu_w.copy())
        nodes = np.vstack(st.x)
        X = np.vstack([nodes, st.particles.x])
        t0 = time.perf_counter()
        U = self.wake_velocities(mu, mu_w, we, X)
        t_vel = time.perf_counter() - t0
        nn = len(nodes)
        k = 0
        for i, row in enumerate(st.x):
            st.x[i] = row + o.dt * U[k:k + len(row)]
            k += len(row)
        st.particles.x = st.particles.x + o.dt * U[nn:]
        st.x.insert(0, self.edge.copy())
        st.mu.insert(0, mu_w.copy())
        while len(st.x) - 1 > self.keep_rows:
            convert_oldest(st, self.topo, self.core, self.merge)
        if math.isfinite(o.delete_beyond):
            self.delete_far()
        rec = dict(n=self.n, t=self.t, **loads, rows=len(st.x) - 1, particles=len(st.particles), lam=float(lam[0]),
                   umax=float(np.max(np.linalg.norm(U, axis=1))) if len(U) else 0.0, t_rhs=t_rhs, t_vel=t_vel)
        self.history.append(rec)
        self.mu, self.mu_w = mu, mu_w
        self.n += 1
        self.t += o.dt
        return rec

    def step_loads(self, q, mu, mu_w, we):
        """step.jl:149-163"""
        dm, els, V = self.dm, self.els, self.V
        u = post.gradient_velocity(els, dm, mu, V)
        dmu = np.zeros(dm.nslot) if self.n == 0 else (mu - self.mu) / self.opts.dt
        cp = post.pressure_coefficient(u, V, post.element_values(dm, dmu).sum(1) / 3)
        self.cp = cp
        Fp, Mp = post.pressure_force(els, cp)
        Fkj = post.kutta_joukowski_force(dm.wake_te, mu_w, V)
        gam = post.sheet_strengths(els, dm, mu)
        body = Blobs(els.centroid.copy(), gam * els.area[:, None], np.zeros(len(els)))
        blobs = Blobs.cat([body, free_wake_blobs(rigid_elements(self.te, mu_w), self.core),
                           free_wake_blobs(we, self.core), self.state.particles])
        Om, I = vorticity_moments(blobs)
        Fimp = np.full(3, np.nan) if self.n == 0 else -(2 / V.dot(V)) * (I - self.I_prev) / self.opts.dt
        self.I_prev = I
        return dict(mu_w=mu_w, F_pressure=Fp, M_pressure=Mp, F_kj=Fkj, F_impulse=Fimp, Omega=Om, I=I)

    def wake_velocities(self, mu, mu_w, we, X):
        """step.jl:166-173: V∞ + the body's exact kernels + every wake element as blobs"""
        if self.opts.advection == "freestream":
            return np.tile(self.V, (len(X), 1))
        gam = post.sheet_strengths(self.els, self.dm, mu)
        blobs = Blobs.cat([free_wake_blobs(rigid_elements(self.te, mu_w), self.core), free_wake_blobs(we, self.core),
                           self.state.particles])
        return self.V + self._velocities(gam, blobs, X)

    def delete_far(self):
        """delete_far!, step.jl:190-195"""
        d = self.V / _norm(self.V)
        limit = max(float(x @ d) for x in self.edge) + self.opts.delete_beyond * self.te.span
        self.state.keep_particles(self.state.particles.x @ d <= limit)

    def run(self, nsteps, progress=None, series=None):
        """run!, step.jl:203-216: nsteps steps, progress(sim, record) after each.  With an export.FreeWakeSeries, every
        series.every steps and the last are written as a ParaView frame: the wake as the step's solve saw it (written
        before the step moves it), the body's pressure and the trailing edge panel from that solve, at the step's
        time; the .pvd collections at the end"""
        from . import export
        for k in range(1, nsteps + 1):
            save = series is not None and (self.n % series.every == 0 or k == nsteps)
            stem = os.path.join(series.dir, f"frame_{self.n:05d}") if save else ""
            wake_files = export.write_free_wake_vtk(self, stem) if save else []
            rec = self.step()
            if save:
                export.add_frame(series, rec["t"], export.write_body_vtk(self, stem) + wake_files)
            if progress is not None:
                progress(self, rec)
        if series is not None:
            export.write_collections(series)
        return self.history

    def close(self):
        self.F.close()
        self.explicit.close()
        self.field.close()


# ------------------------------------------------------------------------------------------ the native engine -------

class FreeWakeSim(FreeWakeHost):
    """FreeWakeSim (step.jl:8-58) whose step runs in the native engine (vp_freewake_*): the rows, conversions,
    merging, loads, advection and deletion in C++, the device work through the library's handles (explicit_rhs, the
    kept factorization, the velocity sums as one HIP graph).  The set-up (trailing edge panel, its assembly with the
    body, the bordered factorization, the resolved options) is FreeWakeHost's; `state`, `mu`, `mu_w`, `cp` are read
    back from the engine."""

    def __init__(self, g, dm, vinf, opts: FreeWakeOptions, assembly: AssemblyOptions | None = None,
                 precision: str = "fp64", ctx: Context | None = None, tree: tuple | None = None):
        """tree = (theta, degree): the wake blobs' velocities by the treecode (post.blob_velocities_tree) instead of
        direct sums"""
        self._h = None
        self.tree = tree
        super().__init__(g, dm, vinf, opts, assembly, precision, ctx)
        self._make_engine()

    def _setup(self):
        """the body with the trailing edge panel, assembled and factored once (step.jl:38-46)"""
        ctx = self.ctx or default_context()
        self.ctx = ctx
        o = AssemblyOptions(**{**self.assembly.__dict__, "precision": self.precision})
        plan = AssemblyPlan(self.g, self.dm, self.te, o, ctx=ctx)
        plan.assemble(self.sigma, self.V)
        self.b0 = plan.read_b()
        self.F = factorize_galerkin(self.dm, plan, W=wake_rows(self.dm, self.te.te), mode=self.precision)
        plan.close()

    def _make_engine(self):
        import scipy.sparse as sp
        from .api import _wake_c
        dm, topo, core = self.dm, self.topo, self.core
        i32 = lambda a: L.i32(np.asarray(a, dtype=np.int64).reshape(-1)) if len(a) else np.zeros(1, np.int32)
        Cm, Km = sp.csr_matrix(dm.C), sp.csr_matrix(dm.K)
        from .post import te_segments
        k = dict(V=L.f64(self.V), b0=L.f64(self.b0), Cp=L.i32(Cm.indptr), Ci=L.i32(Cm.indices), Cv=L.f64(Cm.data),
                 Kp=L.i32(Km.indptr), Ki=L.i32(Km.indices), Kv=L.f64(Km.data), tex=L.f64(dm.wake_te.x),
                 teseg=i32(te_segments(dm.wake_te)), edge=L.f64(self.edge), strips=i32(topo.segments),
                 merge=i32(self.merge) if self.merge is not None else None, first=i32(topo.first_nodes),
                 last=i32(topo.last_nodes), pte=L.i32(self.te.te), opts=self.assembly.to_c())
        wc, k["wk"] = _wake_c(self.te)
        k["wc"] = wc
        mode = {"fixed": 0, "auto": 1, "scaled": 2}[core.mode]
        su = L.FreeWakeSetupC(
            L.ptr(k["V"]), self.opts.dt, 1 if self.opts.advection == "freestream" else 0, float(self.opts.delete_beyond),
            float(self.te.span), int(self.keep_rows), mode, core.value, core.overlap, core.factor, core.floor,
            C.pointer(k["opts"]), 1 if self.precision == "mixed" else 0, self.F.h, L.ptr(k["b0"]), dm.nq,
            L.ptr(k["Cp"], C.c_int), L.ptr(k["Ci"], C.c_int), L.ptr(k["Cv"]), topo.nte, L.ptr(k["Kp"], C.c_int),
            L.ptr(k["Ki"], C.c_int), L.ptr(k["Kv"]), L.ptr(k["tex"]), len(te_segments(dm.wake_te)),
            L.ptr(k["teseg"], C.c_int), L.ptr(k["edge"]), len(topo.segments), L.ptr(k["strips"], C.c_int),
            L.ptr(k["merge"], C.c_int) if k["merge"] is not None else None, len(topo.first_nodes),
            L.ptr(k["first"], C.c_int), len(topo.last_nodes), L.ptr(k["last"], C.c_int), C.pointer(wc),
            L.ptr(k["pte"], C.c_int), float(self.tree[0]) if self.tree else 0.0, int(self.tree[1]) if self.tree else 0)
        s, keep = _surface_c(self.g, dm)
        h = C.c_void_p()
        _check(self.ctx.h, self.ctx.lib.vp_freewake_create(self.ctx.h, C.byref(s), C.byref(su), C.byref(h)))
        self._h = h
        self._state_cache = None

    def step(self):
        """step!, step.jl:113-146, in the engine"""
        rec = L.FreeWakeRecordC()
        mw = np.zeros(self.topo.nte)
        _check(self.ctx.h, self.ctx.lib.vp_freewake_step(self._h, C.byref(rec), L.ptr(mw)))
        out = dict(n=rec.n, t=rec.t, mu_w=mw, F_pressure=np.array(rec.F_pressure), M_pressure=np.array(rec.M_pressure),
                   F_kj=np.array(rec.F_kj), F_impulse=np.array(rec.F_impulse), Omega=np.array(rec.Omega),
                   I=np.array(rec.I), rows=rec.rows, particles=rec.particles, lam=rec.lam, umax=rec.umax,
                   t_rhs=rec.t_rhs, t_vel=rec.t_vel, iters=rec.iters)
        self.history.append(out)
        self.n += 1
        self.t += self.opts.dt
        self._state_cache = None
        return out

    def _read(self):
        """the engine's state as a FreeWakeState, with mu, mu_w, cp and I_prev"""
        if self._state_cache is not None:
            return self._state_cache
        c = np.zeros(6, dtype=np.int64)
        _check(self.ctx.h, self.ctx.lib.vp_freewake_counts(self._h, L.ptr(c, C.c_longlong)))
        nr, npt, nte, ns = (int(v) for v in c[:4])
        rx, rm, ml = np.zeros((nr, nte, 3)), np.zeros((nr, nte)), np.zeros(nte)
        px, pa, pd = np.zeros((npt, 3)), np.zeros((npt, 3)), np.zeros(npt)
        tag, grp = np.zeros(npt, np.int64), np.zeros(npt, np.int32)
        br, bl, bw = np.zeros(ns, np.int64), np.zeros(ns), np.zeros(ns)
        mu, mw, cp, Ip, t = np.zeros(self.dm.nslot), np.zeros(nte), np.zeros(len(self.g.tris)), np.zeros(3), C.c_double()
        _check(self.ctx.h, self.ctx.lib.vp_freewake_read(
            self._h, L.ptr(rx), L.ptr(rm), L.ptr(ml), L.ptr(px), L.ptr(pa), L.ptr(pd), L.ptr(tag, C.c_longlong),
            L.ptr(grp, C.c_int), L.ptr(br, C.c_longlong), L.ptr(bl), L.ptr(bw), L.ptr(mu), L.ptr(mw), L.ptr(cp),
            L.ptr(Ip), C.byref(t)))
        st = FreeWakeState([r for r in rx], [r for r in rm], ml, Blobs(px, pa, pd), int(c[5]), tag, grp.astype(np.int8),
                           br, bl, bw)
        self._state_cache = (st, mu, mw, cp, Ip)
        return self._state_cache

    # the engine's state, read back on access (assignments before the engine exists set up the host's empty state)
    @property
    def state(self):
        return self._read()[0] if getattr(self, "_h", None) else self._state0

    @state.setter
    def state(self, v):
        self._state0 = v

    @property
    def mu(self):
        return self._read()[1] if getattr(self, "_h", None) else self._mu0

    @mu.setter
    def mu(self, v):
        self._mu0 = v

    @property
    def mu_w(self):
        return self._read()[2] if getattr(self, "_h", None) else self._muw0

    @mu_w.setter
    def mu_w(self, v):
        self._muw0 = v

    @property
    def cp(self):
        return self._read()[3] if getattr(self, "_h", None) else self._cp0

    @cp.setter
    def cp(self, v):
        self._cp0 = v

    @property
    def I_prev(self):
        return self._read()[4] if getattr(self, "_h", None) else self._I0

    @I_prev.setter
    def I_prev(self, v):
        self._I0 = v

    def close(self):
        if getattr(self, "_h", None):
            self.ctx.lib.vp_freewake_destroy(self._h)
            self._h = None
        self.F.close()
