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
    """wing_loads for many angles at once (MU nslot x na, Q nq x na): the per-element pressure route as one batched
    contraction over the angles, the trailing-edge loads per angle"""
    a = np.radians(np.asarray(alphas, dtype=np.float64))
    V = np.stack([np.cos(a), np.zeros_like(a), np.sin(a)], axis=1)          # (na, 3)
    eL = np.stack([-np.sin(a), np.zeros_like(a), np.cos(a)], axis=1)
    n = els.n
    M = np.asarray(MU).T[:, dm.corner_slot]                                  # (na, nt, 3) corner doublets
    u = (V[:, None, :] - np.einsum("aj,tj->at", V, n)[:, :, None] * n[None]  # gradient_velocity, Post.jl:72-73
         + np.einsum("atk,tkj->atj", M, els.gradl))
    cp = 1 - np.einsum("atj,atj->at", u, u) / np.einsum("aj,aj->a", V, V)[:, None]  # pressure_coefficient, :82-84
    F = -np.einsum("at,tj->aj", cp, els.area[:, None] * n)                 # pressure_force per angle, forces.jl:11-19
    MW = dm.K @ np.asarray(Q)
    CL_KJ, CDi_T = [], []
    for k in range(len(a)):
        CL_KJ.append(float(post.kutta_joukowski_force(dm.wake_te, MW[:, k], V[k]) @ eL[k]) / area_ref)
        CDi_T.append(post.trefftz_drag(dm.wake_te, MW[:, k], V[k]) / area_ref)
    return dict(CL=np.einsum("aj,aj->a", F, eL) / area_ref, CL_KJ=np.array(CL_KJ), CDi_T=np.array(CDi_T),
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
