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
        out = solve_factored(self.F, self.rhs(alphas))
        t1 = time.perf_counter()
        loads = []
        for k, al in enumerate(np.atleast_1d(alphas)):
            V, eL = wing_freestream(al)
            loads.append(wing_loads(self.els, self.dm, out["mu"][:, k], out["q"][:, k], V, eL, self.area_ref))
        self.timings.update(solve_s=t1 - t0, loads_s=time.perf_counter() - t1, refinement_iters=out["iters"])
        return dict(alpha=np.atleast_1d(alphas), mu=out["mu"], q=out["q"], lam=out["lam"],
                    CL=np.array([l["CL"] for l in loads]), CL_KJ=np.array([l["CL_KJ"] for l in loads]),
                    CDi_T=np.array([l["CDi_T"] for l in loads]), CDi_p=np.array([l["CDi_p"] for l in loads]))

    def close(self):
        self.F.close()
        self.plan.close()
