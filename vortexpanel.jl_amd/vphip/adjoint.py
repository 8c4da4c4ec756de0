"""The discrete adjoint of a wing solution: dCL/dx and dCDi/dx for every node coordinate, and dCL/dalpha.

No counterpart in the reference (it has no derivatives); the objectives are the reference's own loads
(validation/common/wing_case.jl:53-67: CL from the pressure force of the gradient-route Cp, CDi from the Trefftz plane,
src/post/forces.jl:11-81).  For an objective J(x, q) of the bordered solution M(x) [q; l] = [Cᵀ b(x); 0]:

    dJ/dx = dJ/dx|_q + d/dx [ lamᵀ (b(x) - A(x) [C; W] q) ],   lam = C psi_q,   Mᵀ psi = [dJ/dq; 0]

(the border rows drop out: J does not see a uniform doublet, so psi has no border part).  The adjoint solve runs on
the device (vp_factor_solve_transposed), the bracket is the backward assembly (vp_plan_adjoint, which recomputes
the pair integrals instead of storing dA/dx), and the explicit terms dJ/dx|_q, dJ/dq, dJ/dalpha are per-element
host formulas differentiated by complex steps (exact to rounding).

What is held fixed: the quadrature rules of the forward assembly (tiers and piece positions in barycentric
coordinates), the wake's offsets from its trailing edge nodes (its nodes move with them) and its direction, the DOFMap.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as L
from .api import AssemblyOptions, AssemblyPlan, Context, _check, build_elements, default_context, source_strengths
from .dofmap import wake_rows
from .solve import factorize_galerkin, solve_factored
from .sweep import wing_freestream

H = 1e-30  # complex step


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _dot(a, b):
    return np.sum(a * b, axis=-1)


def cl_terms(X, m, V, eL, area_ref):
    """Per-triangle contributions to CL = F.eL/S with F = -sum Cp A n (pressure_force, forces.jl:11-19) and the
    gradient-route Cp (Post.jl:72-84), as an analytic function of the vertices X (..., nt, 3, 3) and corner doublets m
    (nt, 3): complex arguments give complex-step derivatives, leading axes a batch of them."""
    a, b, c = X[..., 0, :], X[..., 1, :], X[..., 2, :]
    N = _cross(b - a, c - a)
    A2 = np.sqrt(_dot(N, N))
    n = N / A2[..., None]
    e = np.stack([c - b, a - c, b - a], axis=-2)
    gradl = _cross(n[..., None, :], e) / A2[..., None, None]
    u = V - _dot(n, V)[..., None] * n + np.einsum("tk,...tkj->...tj", m, gradl)
    cp = 1 - _dot(u, u) / _dot(V, V)
    return -cp * (A2 / 2) * _dot(n, eL) / area_ref, u, gradl, A2, n


def cl_explicit(g, dm, x, mu, alpha_deg, area_ref):
    """CL and its partial derivatives at fixed doublets: d/dmu (nslot), d/dx (nn, 3), d/dalpha (per degree).

    With N = (b - a) x (c - a) (q = N.N), E = sum_k m_k e_k (e_k the edge opposite corner k) each term is
    f = -(1 - |u|^2/|V|^2) (N.eL)/(2S), u = V - (N.V) N/q + (N x E)/q, so d/dx goes through d/dN and d/dE by hand."""
    V, eL = wing_freestream(alpha_deg)
    tris = g.tris
    X = x[tris]
    m = np.asarray(mu)[dm.corner_slot]
    f, u, gradl, A2, n = cl_terms(X, m, V, eL, area_ref)
    CL = float(np.sum(f))
    VV = np.dot(V, V)
    # d/du of each term, then d/dmu through grad(lambda_k)
    dfdu = (2 * (A2 / 2) * _dot(n, eL) / (area_ref * VV))[:, None] * u
    dmu = np.bincount(dm.corner_slot.ravel(), np.einsum("tj,tkj->tk", dfdu, gradl).ravel(), minlength=dm.nslot)
    a, b, c = X[:, 0], X[:, 1], X[:, 2]
    N = _cross(b - a, c - a)
    q = _dot(N, N)
    E = m[:, 0, None] * (c - b) + m[:, 1, None] * (a - c) + m[:, 2, None] * (b - a)
    gu = dfdu  # df/du
    NV, gN, NE = N @ V, _dot(gu, N), _cross(N, E)
    dN = (-(1 - _dot(u, u) / VV) / (2 * area_ref))[:, None] * eL \
        - (gN / q)[:, None] * V - (NV / q)[:, None] * gu + (2 * NV * gN / q ** 2)[:, None] * N \
        + _cross(E, gu) / q[:, None] - (2 * _dot(gu, NE) / q ** 2)[:, None] * N
    dE = _cross(gu, N) / q[:, None]
    db, dc = _cross(c - a, dN), _cross(dN, b - a)
    grads = [-(db + dc) + (m[:, 1] - m[:, 2])[:, None] * dE, db + (m[:, 2] - m[:, 0])[:, None] * dE,
             dc + (m[:, 0] - m[:, 1])[:, None] * dE]
    dx = np.zeros_like(x)
    for k in range(3):
        for d in range(3):
            dx[:, d] += np.bincount(tris[:, k], grads[k][:, d], minlength=len(x))
    al = np.radians(alpha_deg) + 1j * H
    fa, *_ = cl_terms(X, m, np.array([np.cos(al), 0, np.sin(al)]), np.array([-np.sin(al), 0, np.cos(al)]), area_ref)
    dalpha = float(np.sum(fa).imag / H) * np.pi / 180
    return CL, dmu, dx, dalpha


def _trefftz_topology(lines, closed):
    """The trailing edge lines as index arrays: segment ends (ia, ib) and, per trailing vortex point, its node and the
    segments after and before it along its line (-1: none)"""
    ia, ib, vnode, vafter, vbefore = [], [], [], [], []
    base = 0
    for line, cl in zip(lines, closed):
        line = list(line)
        m = len(line)
        nseg = m if cl else m - 1
        for i in range(nseg):
            ia.append(line[i])
            ib.append(line[(i + 1) % m])
        for i in range(m):
            vnode.append(line[i])
            vbefore.append(base + i - 1 if i > 0 else (base + nseg - 1 if cl else -1))
            vafter.append(base + i if i < nseg else -1)
        base += nseg
    return tuple(np.asarray(v, dtype=np.int64) for v in (ia, ib, vnode, vafter, vbefore))


def _trefftz_batch(P, mu_w, topo, d):
    """trefftz_drag's D (forces.jl:53-81, before the 2/|V∞|^2 factor) for a batch: P (B, nte, 3), mu_w (B, nte)"""
    ia, ib, vnode, vafter, vbefore = topo
    Pp = P - _dot(P, d)[..., None] * d
    G = (mu_w[:, ia] + mu_w[:, ib]) / 2
    Gz = np.concatenate([G, np.zeros((len(G), 1), dtype=G.dtype)], axis=1)  # column -1: no segment
    vg = Gz[:, vafter] - Gz[:, vbefore]
    sa, sb = Pp[:, ia], Pp[:, ib]
    c = (sa + sb) / 2
    ell = np.sqrt(_dot(sb - sa, sb - sa))
    n = _cross(d, (sb - sa) / ell[..., None])
    r = c[:, :, None, :] - Pp[:, vnode][:, None, :, :]
    v = np.einsum("bk,bskj->bsj", vg / (4 * np.pi), _cross(d, r) / _dot(r, r)[..., None])
    return np.sum(G * _dot(n, v) * ell, axis=-1)


def trefftz_of(P, lines, closed, mu_w, d):
    """trefftz_drag (forces.jl:53-81) as an analytic function of the trailing edge node positions P and mu_w (complex
    steps welcome), D in units of q∞ before the 2/|V∞|^2 factor"""
    topo = _trefftz_topology(lines, closed)
    return _trefftz_batch(np.asarray(P)[None], np.asarray(mu_w)[None], topo, np.asarray(d))[0]


def cdi_explicit(dm, x, mu_w, alpha_deg, area_ref):
    """CDi (Trefftz) and its partial derivatives: d/dmu_w (nte), d/dx (nn, 3) through the trailing edge nodes,
    d/dalpha (per degree).

    In the plane normal to d (points projected there) each (segment s, trailing vortex k) term of D is
    G_s g_k (l_s . r_sk) / (4 pi |r_sk|^2), l_s the segment, r_sk from the vortex to the segment's middle, since
    ell n.(d x r) = (d x l).(d x r) = l.r for vectors normal to d; its derivatives are taken by hand."""
    wte = dm.wake_te
    V, _ = wing_freestream(alpha_deg)
    scale = 2 / np.dot(V, V) / area_ref
    d = V / np.sqrt(np.dot(V, V))
    P = np.asarray(x)[wte.node].astype(float)
    mw = np.asarray(mu_w, dtype=float)
    nte = len(mw)
    topo = _trefftz_topology(wte.lines, wte.closed)
    ia, ib, vnode, vafter, vbefore = topo
    Pp = P - (P @ d)[:, None] * d
    G = (mw[ia] + mw[ib]) / 2
    Gz = np.append(G, 0.0)
    vg = Gz[vafter] - Gz[vbefore]
    sa, sb = Pp[ia], Pp[ib]
    ls = sb - sa
    r = ((sa + sb) / 2)[:, None, :] - Pp[vnode][None, :, :]  # (nseg, nv, 3)
    ir2 = 1 / _dot(r, r)
    lr = _dot(ls[:, None, :], r)
    Mk = lr * ir2 / (4 * np.pi)  # (nseg, nv)
    CD = float(G @ Mk @ vg) * scale
    # d/dmu_w: through G_s and g_k
    dG = Mk @ vg
    dvg = G @ Mk
    dGz = np.zeros(len(G) + 1)
    np.add.at(dGz, vafter, dvg)
    np.add.at(dGz, vbefore, -dvg)
    dG = dG + dGz[:-1]
    dmw = (np.bincount(ia, dG / 2, minlength=nte) + np.bincount(ib, dG / 2, minlength=nte)) * scale
    # d/dx: through the segment ends and the vortex points (projected positions), then the projection
    w = (G[:, None] * vg[None, :]) / (4 * np.pi)  # (nseg, nv)
    dT_dl = (w * ir2)[..., None] * r
    dT_dr = (w * ir2)[..., None] * ls[:, None, :] - (2 * w * lr * ir2 ** 2)[..., None] * r
    gl = dT_dl.sum(1)
    gr = dT_dr.sum(1)
    dPp = np.zeros_like(P)
    np.add.at(dPp, ia, -gl + gr / 2)
    np.add.at(dPp, ib, gl + gr / 2)
    np.add.at(dPp, vnode, -dT_dr.sum(0))
    dP = (dPp - (dPp @ d)[:, None] * d) * scale
    dx = np.zeros_like(x)
    np.add.at(dx, wte.node, dP)
    al = np.radians(alpha_deg) + 1j * H
    Vc = np.array([np.cos(al), 0, np.sin(al)])
    da = _trefftz_batch(P[None].astype(complex), mw[None].astype(complex), topo, Vc / np.sqrt(np.sum(Vc * Vc)))[0]
    da = da * 2 / np.sum(Vc * Vc) / area_ref
    return CD, dmw, dx, float(da.imag / H) * np.pi / 180


class WingAdjoint:
    """Forward solve of a wing with a rigid wake, then the adjoint gradients of CL and CDi.

    precision: 'fp64' (FP64 assembly and LU) or 'mixed' (FP32 multipole far field, FP32 LU on MFMA; FP64
    refinement of the forward and the adjoint solves)."""

    def __init__(self, g, dm, wake, alpha_deg, area_ref, precision="fp64", opts: AssemblyOptions | None = None,
                 ctx: Context | None = None):
        self.g, self.dm, self.wake, self.alpha, self.S = g, dm, wake, float(alpha_deg), float(area_ref)
        self.ctx = ctx or default_context()
        self.precision = precision
        self.timings = {}
        V, _ = wing_freestream(alpha_deg)
        self.V = V
        t0 = time.perf_counter()
        els = build_elements(g)
        self.plan = AssemblyPlan(g, dm, wake, opts or AssemblyOptions(precision=precision), ctx=self.ctx)
        t1 = time.perf_counter()
        self.plan.assemble(source_strengths(els, V), V)
        self.b = self.plan.read_b()
        t2 = time.perf_counter()
        self.W = wake_rows(dm, wake.te)
        self.F = factorize_galerkin(dm, self.plan, W=self.W, mode=precision)
        self.ctx.sync()
        t3 = time.perf_counter()
        sol = solve_factored(self.F, self.b, cond=False)
        self.ctx.sync()
        t4 = time.perf_counter()
        self.timings.update(forward_s=t4 - t0, plan_s=t1 - t0, assemble_s=t2 - t1, factor_s=t3 - t2, solve_s=t4 - t3)
        self.q, self.mu = sol["q"], sol["mu"]
        self._bd = None
        self.mu_w = dm.K @ self.q                   # trailing edge doublet jumps
        self.mu_wn = np.asarray(self.W @ self.q)    # wake node values (columns after the slots)

    def loads(self):
        CL, *_ = cl_explicit(self.g, self.dm, self.g.x, self.mu, self.alpha, self.S)
        CD, *_ = cdi_explicit(self.dm, self.g.x, self.mu_w, self.alpha, self.S)
        return dict(CL=CL, CDi=CD)

    def _adjoint_weights(self, dJdq, border=False):
        """lam = C psi_q for M^T psi = [dJ/dq; 0] (dJdq (nq,), or (nq, k) for k objectives solved as one batch); with
        border, also psi's border part (one value per body)"""
        F = self.F
        dJ = np.asarray(dJdq, dtype=np.float64)
        R = np.zeros((F.n,) + dJ.shape[1:])
        R[:F.nq] = dJ
        X = F.solve_bordered_t(R)
        lam = F.C @ X[:F.nq]
        return (lam, X[F.nq:F.nq + F.nb].copy()) if border else lam

    def _border_grad(self, psi_b):
        """-psi_bᵀ d/dx (c~ q) at fixed q, (nn, 3).  The border rows c~_k = c_k scale/|c_k| (solve.jl:30-35, with
        c_k = Cᵀ (slot_area on body k) and scale = se_k |e_k|) depend on the nodes through the slot areas, a third of
        each corner's triangle area (build.jl:199-212); c~_k q = 0 at the solution, so only d(c_k . q)/dx is left, and
        d area/d a = n x (c - b) / 2 for the triangle (a, b, c).  (The border column's se_k moves with A's diagonal,
        but it multiplies the border unknown, zero to rounding for a closed body's consistent right side.)"""
        g, dm = self.g, self.dm
        gx = np.zeros_like(g.x)
        X = g.x[g.tris]
        N = _cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0])
        n = N / np.sqrt(_dot(N, N))[:, None]
        dA = 0.5 * np.stack([_cross(n, X[:, 2] - X[:, 1]), _cross(n, X[:, 0] - X[:, 2]), _cross(n, X[:, 1] - X[:, 0])],
                            axis=1)
        for k, psi in enumerate(np.atleast_1d(psi_b)):
            on = dm.slot_body == k
            c = dm.C.T @ (dm.slot_area * on)
            scale = self.F.se[k] * np.sqrt(len(dm.body_q[k]))
            mt = (self.mu * on)[dm.corner_slot].sum(1) / 3
            np.add.at(gx, g.tris, (-psi * scale / np.linalg.norm(c) * mt)[:, None, None] * dA)
        return gx

    def _dV(self):
        a = np.radians(self.alpha)
        return np.array([-np.sin(a), 0.0, np.cos(a)])  # dV∞/dalpha per radian

    def _backward(self, lams):
        """d/dx [lamᵀ (b - A [mu; mu_wake])] for each row of lams on the device (two per pass of the pair integrals),
        wake node gradients moved to their trailing edge nodes; and lamᵀ db/dalpha (per radian) from the same pass
        (b is linear in V∞, so db/dalpha is b of dV∞/dalpha)"""
        lams = np.atleast_2d(np.asarray(lams, dtype=np.float64))
        nl, nn = len(lams), len(self.g.x)
        mu_full = L.f64(np.concatenate([self.mu, self.mu_wn]))
        G = np.zeros((nl, nn + len(self.wake.x), 3))
        lbd = np.zeros(nl)
        _check(self.ctx.h, self.ctx.lib.vp_plan_adjoint(self.plan.h, nl, L.ptr(L.f64(lams)), L.ptr(mu_full),
                                                        L.ptr(L.f64(self.V)), L.ptr(L.f64(self._dV())), L.ptr(G),
                                                        L.ptr(lbd)))
        gx = G[:, :nn].copy()
        te = self.dm.wake_te.node[self.wake.te]
        for k in range(nl):
            np.add.at(gx[k], te, G[k, nn:])
        return gx, lbd

    def _db_dalpha(self):
        """db/dalpha at fixed geometry and wake (per radian) by a second assembly, made once (a check of the
        backward pass's lamᵀ db/dalpha).  The re-assembly leaves the plan's rule lists as they were (they do not
        depend on V∞), which the backward assembly reads."""
        if self._bd is None:
            Vd = self._dV()
            self.plan.assemble(source_strengths(build_elements(self.g), Vd), Vd)
            self._bd = self.plan.read_b()
        return self._bd

    def _explicit(self, objective):
        if objective == "CL":
            J, dmu, dx, da = cl_explicit(self.g, self.dm, self.g.x, self.mu, self.alpha, self.S)
            return J, self.F.C.T @ dmu, dx, da
        if objective == "CDi":
            J, dmw, dx, da = cdi_explicit(self.dm, self.g.x, self.mu_w, self.alpha, self.S)
            return J, self.dm.K.T @ dmw, dx, da
        raise ValueError("objective must be CL or CDi")

    def gradients(self, objectives=("CL", "CDi")):
        """{objective: (J, dJ/dx (nn, 3), dJ/dalpha per degree)} for J in objectives (CL, CDi): their adjoint systems
        as one batched solve, one backward pass for every two of them"""
        t0 = time.perf_counter()
        ex = [self._explicit(o) for o in objectives]
        t1 = time.perf_counter()
        lams = np.ascontiguousarray(self._adjoint_weights(np.stack([e[1] for e in ex], 1)).T)  # one solve for all
        t2 = time.perf_counter()
        gb, lbd = self._backward(lams)
        t3 = time.perf_counter()
        self.last_lam, self.last_lam_db = lams, lbd
        key = "+".join(objectives)
        self.timings.update({f"{key}_explicit_s": t1 - t0, f"{key}_adjoint_solve_s": t2 - t1,
                             f"{key}_backward_s": t3 - t2})
        return {o: (e[0], e[2] + gb[k], e[3] + lbd[k] * np.pi / 180) for k, (o, e) in enumerate(zip(objectives, ex))}

    def gradient(self, objective="CL"):
        """dJ/dx (nn, 3) and dJ/dalpha (per degree) of J = CL or CDi"""
        return self.gradients((objective,))[objective]

    def pullback_loads(self, dCL, dCDi):
        """The vector-Jacobian product of (CL, CDi) with cotangents (dCL, dCDi): (x̄ (nn, 3), ᾱ per degree) =
        dCL grad CL + dCDi grad CDi, from one adjoint solve and one backward pass (the adjoint weights are linear in
        the objective's)"""
        eCL, eCD = self._explicit("CL"), self._explicit("CDi")
        lam = self._adjoint_weights(dCL * eCL[1] + dCDi * eCD[1])
        gb, lbd = self._backward(lam[None])
        return dCL * eCL[2] + dCDi * eCD[2] + gb[0], dCL * eCL[3] + dCDi * eCD[3] + lbd[0] * np.pi / 180

    def pullback_mu(self, mu_bar):
        """The vector-Jacobian product of the slot doublets mu(x, alpha) with a cotangent mu_bar (nslot): (x̄ (nn, 3),
        ᾱ per degree).  mu = C q, so dJ/dq = Cᵀ mu_bar; the explicit dependence of an objective on x is its own.  A
        cotangent that is not blind to a uniform doublet also moves the zero-mean border rows (_border_grad)"""
        lam, psi_b = self._adjoint_weights(self.F.C.T @ np.asarray(mu_bar, dtype=np.float64), border=True)
        gb, lbd = self._backward(lam[None])
        return gb[0] + self._border_grad(psi_b), lbd[0] * np.pi / 180

    def close(self):
        self.F.close()
        self.plan.close()


# ------------------------------------------------------------------------------------------ rules (rrule form) -----

def wing_loads_rrule(g, dm, wake, alpha_deg, area_ref, precision="fp64", opts: AssemblyOptions | None = None,
                     ctx: Context | None = None):
    """The reverse rule of (CL, CDi) = wing_loads(x, alpha) in ChainRulesCore's rrule form (Julia's analogue of
    torch.autograd.Function): ((CL, CDi), pullback) with pullback((dCL, dCDi)) -> (x̄ (nn, 3), ᾱ per degree).  The
    forward solve happens here; each pullback call is one adjoint solve and one backward pass.  close() on the
    pullback releases the device state."""
    adj = WingAdjoint(g, dm, wake, alpha_deg, area_ref, precision=precision, opts=opts, ctx=ctx)
    loads = adj.loads()

    def pullback(ybar):
        dCL, dCDi = ybar
        return adj.pullback_loads(float(dCL), float(dCDi))

    pullback.close = adj.close
    pullback.adjoint = adj
    return (loads["CL"], loads["CDi"]), pullback


def doublets_rrule(g, dm, wake, alpha_deg, precision="fp64", opts: AssemblyOptions | None = None,
                   ctx: Context | None = None):
    """The reverse rule of the slot doublets mu = galerkin_doublets(x, alpha) (assemble + solve_galerkin): (mu,
    pullback) with pullback(mu_bar) -> (x̄ (nn, 3), ᾱ per degree).  An objective written on mu and x then needs only
    its own explicit derivatives on top, which the caller's differentiation supplies.  close() on the pullback
    releases the device state."""
    adj = WingAdjoint(g, dm, wake, alpha_deg, 1.0, precision=precision, opts=opts, ctx=ctx)

    def pullback(mu_bar):
        return adj.pullback_mu(mu_bar)

    pullback.close = adj.close
    pullback.adjoint = adj
    return adj.mu.copy(), pullback
