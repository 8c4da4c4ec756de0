"""Angle-of-attack sweeps of a wing as one multi-right-hand-side solve, with the loads of each angle.

The reference solves a wing per angle (validation/common/wing_case.jl:34-69: assemble, solve_galerkin, pressure and
Trefftz loads), and its wake follows normalize(V∞) (src/wake/rigid.jl:109,128), so its matrix changes with the angle.
With a body-fixed wake (built once along the wing's x axis) the matrix is the same for every angle, and b is linear in
V∞ (the sources sigma = -V∞.n and the freestream term are): b(alpha) = cos(alpha) b(e_x) + sin(alpha) b(e_z).  The
sweep assembles twice, factors once and solves all angles as one batch on the device.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import post
from .api import AssemblyOptions, AssemblyPlan, Context, build_elements, default_context, source_strengths
from .dofmap import wake_rows
from .solve import factorize_galerkin, solve_factored


def wing_freestream(alpha_deg):
    """Unit freestream at angle of attack alpha (degrees) in the wing's frame and the lift direction,
    validation/common/wing_case.jl:7"""
    a = np.radians(alpha_deg)
    return np.array([np.cos(a), 0.0, np.sin(a)]), np.array([-np.sin(a), 0.0, np.cos(a)])


def wing_loads(els, dm, mu, q, vinf, e_lift, area_ref):
    """Pressure force (gradient-route Cp, constant per triangle), Kutta-Joukowski force and Trefftz drag of one
    solution as coefficients on area_ref, validation/common/wing_case.jl:53-67"""
    V = np.asarray(vinf, dtype=np.float64)
    cp = post.pressure_coefficient(post.gradient_velocity(els, dm, mu, V), V)
    Fp, Mp = post.pressure_force(els, cp)
    mu_w = dm.K @ q
    Fkj = post.kutta_joukowski_force(dm.wake_te, mu_w, V)
    return dict(CL=float(Fp @ e_lift) / area_ref, CL_KJ=float(Fkj @ e_lift) / area_ref,
                CDi_p=float(Fp @ V) / area_ref, CDi_T=post.trefftz_drag(dm.wake_te, mu_w, V) / area_ref,
                CY=float(Fp[1]) / area_ref, mu_w=mu_w)


def wing_loads_batch(els, dm, MU, Q, alphas, area_ref):
    """wing_loads for many angles at once (MU nslot x na, Q nq x na).  The pressure route, restated for all angles:
    the surface gradient g = sum_k mu_k gradl_k (Post.jl:72-73) is one sparse (3 nt x nslot) matrix times MU, and with
    u = V - (V.n) n + g, |u|^2 = |V|^2 - 2 (V.n)^2 + (V.n)^2 |n|^2 + 2 (V.g - (V.n)(n.g)) + |g|^2 (Post.jl:82-84) and
    the force -sum_t cp_t A_t n_t (forces.jl:11-19) are dense products over (triangles x angles); the
    Kutta-Joukowski circulation is a linear map of mu_w, and the Trefftz drag is vectorised over the angles."""
    import scipy.sparse as sp
    a = np.radians(np.asarray(alphas, dtype=np.float64))
    V = np.stack([np.cos(a), np.zeros_like(a), np.sin(a)], axis=1)          # (na, 3)
    eL = np.stack([-np.sin(a), np.zeros_like(a), np.cos(a)], axis=1)
    n, nt, na = els.n, len(els.area), len(a)
    rows = (3 * np.arange(nt)[:, None, None] + np.arange(3)[None, None, :]).repeat(3, axis=1)   # (nt, 3 corners, 3)
    cols = np.broadcast_to(np.asarray(dm.corner_slot)[:, :, None], (nt, 3, 3))
    Gm = sp.csr_matrix((np.asarray(els.gradl).reshape(-1), (rows.reshape(-1), cols.reshape(-1))), shape=(3 * nt, dm.nslot))
    G = np.asarray(Gm @ np.asarray(MU, dtype=np.float64)).reshape(nt, 3, na)
    Vn = n @ V.T                                                             # (nt, na)
    VG = np.einsum("aj,tja->ta", V, G)
    nG = np.einsum("tj,tja->ta", n, G)
    VV = np.einsum("aj,aj->a", V, V)
    uu = VV[None] - 2 * Vn ** 2 + Vn ** 2 * np.einsum("tj,tj->t", n, n)[:, None] + 2 * (VG - Vn * nG) + np.einsum("tja,tja->ta", G, G)
    cp = 1 - uu / VV[None]
    F = -(cp.T @ (els.area[:, None] * n))                                    # (na, 3)
    MW = dm.K @ np.asarray(Q)
    x = np.asarray(dm.wake_te.x, dtype=np.float64)
    W = np.zeros((len(x), 3))
    for sa_, sb_ in post.te_segments(dm.wake_te):                            # forces.jl:34-39, linear in mu_w
        W[sa_] += (x[sb_] - x[sa_]) / 2
        W[sb_] += (x[sb_] - x[sa_]) / 2
    Fkj = (2 / VV)[:, None] * np.cross(V, MW.T @ W)
    return dict(CL=np.einsum("aj,aj->a", F, eL) / area_ref, CL_KJ=np.einsum("aj,aj->a", Fkj, eL) / area_ref,
                CDi_T=post.trefftz_drag_batch(dm.wake_te, MW, V) / area_ref,
                CDi_p=np.einsum("aj,aj->a", F, V) / area_ref, CY=F[:, 1] / area_ref)


@dataclass
class AlphaSweep:
    """One wing, one body-fixed wake, one factorization; `solve(alphas)` gives the loads of every angle."""
    g: object
    dm: object
    wake: object
    area_ref: float
    precision: str = "mixed"
    opts: AssemblyOptions | None = None
    ctx: Context | None = None
    timings: dict = field(default_factory=dict)

    def __post_init__(self):
        import time
        self.ctx = self.ctx or default_context()
        self.els = build_elements(self.g)
        o = self.opts or AssemblyOptions(precision=self.precision)
        t0 = time.perf_counter()
        self.plan = AssemblyPlan(self.g, self.dm, self.wake, o, ctx=self.ctx)
        b = []
        for V in (np.array([1.0, 0, 0]), np.array([0, 0, 1.0])):
            self.plan.assemble(source_strengths(self.els, V), V)
            b.append(self.plan.read_b())
        self.b_x, self.b_z = b
        self.ctx.sync()
        t1 = time.perf_counter()
        self.W = wake_rows(self.dm, self.wake.te)
        self.F = factorize_galerkin(self.dm, self.plan, W=self.W, mode=self.precision)
        self.ctx.sync()
        self.timings.update(assemble_s=t1 - t0, factor_s=time.perf_counter() - t1)

    def rhs(self, alphas):
        a = np.radians(np.asarray(alphas, dtype=np.float64))
        return np.outer(self.b_x, np.cos(a)) + np.outer(self.b_z, np.sin(a))

    def solve(self, alphas):
        import time
        t0 = time.perf_counter()
        out = solve_factored(self.F, self.rhs(alphas), cond=False)
        t1 = time.perf_counter()
        al = np.atleast_1d(np.asarray(alphas, dtype=np.float64))
        MU, Q = out["mu"].reshape(self.dm.nslot, -1), out["q"].reshape(self.dm.nq, -1)
        ld = wing_loads_batch(self.els, self.dm, MU, Q, al, self.area_ref)
        self.timings.update(solve_s=t1 - t0, loads_s=time.perf_counter() - t1, refinement_iters=out["iters"])
        return dict(alpha=al, mu=out["mu"], q=out["q"], lam=out["lam"], CL=ld["CL"], CL_KJ=ld["CL_KJ"],
                    CDi_T=ld["CDi_T"], CDi_p=ld["CDi_p"])

    def close(self):
        self.F.close()
        self.plan.close()
