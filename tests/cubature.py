"""Adaptive cubature over a triangle, independent of the quadrature rules under test.

The reference pins its pair rules with HCubature over the unit square mapped onto the test triangle
(test/assembly/test_contact.jl:9-17): (u, v) -> lambda = (1 - u, u (1 - v), u v), dA = 2 A u du dv, so the corner
`apex` sits at u = 0 and the edge apex -> apex + 1 at v = 0.  This module does the same with its own rule: tensor
Gauss-Legendre (numpy's nodes) on rectangles, refined where the error is largest, each rectangle split along the axis
whose higher-order estimate disagrees more (an anisotropic refinement, as HCubature's, so that a singular line along an
edge costs rectangles in one direction only).  The integrands evaluate the oracle's kernels on point batches.
"""
from __future__ import annotations

import numpy as np

_GL = {n: np.polynomial.legendre.leggauss(n) for n in (8, 14)}


def _tensor(lo, hi, nu, nv):
    """Points (k, nu*nv, 2) and weights (k, nu*nv) of the nu x nv Gauss rule on rectangles lo..hi (k, 2)"""
    xu, wu = _GL[nu]
    xv, wv = _GL[nv]
    hu, hv = (hi[:, 0] - lo[:, 0]) / 2, (hi[:, 1] - lo[:, 1]) / 2
    cu, cv = (hi[:, 0] + lo[:, 0]) / 2, (hi[:, 1] + lo[:, 1]) / 2
    U = cu[:, None, None] + hu[:, None, None] * xu[None, :, None]
    V = cv[:, None, None] + hv[:, None, None] * xv[None, None, :]
    U, V = np.broadcast_arrays(U, V)
    W = (hu * hv)[:, None, None] * (wu[:, None] * wv[None, :])[None]
    k = len(lo)
    return np.stack([U.reshape(k, -1), V.reshape(k, -1)], -1), W.reshape(k, -1)


def _estimates(f, lo, hi):
    """Per rectangle: the best estimate (14 x 14 points), its error estimate and the error along each axis"""
    k = len(lo)
    parts = []
    for nu, nv in ((8, 8), (14, 8), (8, 14), (14, 14)):
        P, W = _tensor(lo, hi, nu, nv)
        F = f(P.reshape(-1, 2)).reshape(k, P.shape[1], -1)
        parts.append(np.einsum("kp,kpm->km", W, F))
    base, eu, ev, best = parts
    err_u = np.linalg.norm(eu - base, axis=1)
    err_v = np.linalg.norm(ev - base, axis=1)
    return best, np.linalg.norm(best - base, axis=1) + err_u + err_v, err_u, err_v


def square(f, tol, max_rounds=400, max_rects=100000):
    """Integral over [0,1]^2 of f(uv) (f: (n, 2) -> (n, m)) to an estimated absolute error tol (Euclidean over m).
    Every rectangle stays in the pool with its estimate; each round splits those that hold 90 % of the estimated
    error (at least one), each along the axis whose refinement changed its estimate more."""
    lo, hi = np.zeros((1, 2)), np.ones((1, 2))
    val, err, eu, ev = _estimates(f, lo, hi)
    for _ in range(max_rounds):
        total = err.sum()
        if total <= tol:
            return val.sum(0), total
        order = np.argsort(err)[::-1]
        nsplit = int(np.searchsorted(np.cumsum(err[order]), 0.9 * total)) + 1
        split = np.zeros(len(lo), bool)
        split[order[:nsplit]] = True
        lo_s, hi_s = lo[split], hi[split]
        ax = (ev[split] > eu[split]).astype(int)
        r = np.arange(len(ax))
        mid = (lo_s[r, ax] + hi_s[r, ax]) / 2
        hi_a, lo_b = hi_s.copy(), lo_s.copy()
        hi_a[r, ax] = mid
        lo_b[r, ax] = mid
        nlo, nhi = np.concatenate([lo_s, lo_b]), np.concatenate([hi_a, hi_s])
        nval, nerr, neu, nev = _estimates(f, nlo, nhi)
        keep = ~split
        lo, hi = np.concatenate([lo[keep], nlo]), np.concatenate([hi[keep], nhi])
        val, err = np.concatenate([val[keep], nval]), np.concatenate([err[keep], nerr])
        eu, ev = np.concatenate([eu[keep], neu]), np.concatenate([ev[keep], nev])
        if len(lo) > max_rects:
            break
    raise RuntimeError(f"adaptive cubature did not reach {tol:.1e} (estimate {err.sum():.1e})")


def _bary(uv, apex):
    u, v = uv[:, 0], uv[:, 1]
    lam = np.stack([1 - u, u * (1 - v), u * v], 1)
    return np.roll(lam, apex, axis=1)  # column apex holds 1 - u


def triangle(tx, g, tol, apex=0):
    """Integral over triangle tx (3 corners) of g(x, lam) dA (g: points (n, 3), barycentrics (n, 3) -> (n, m))"""
    tx = np.asarray(tx, float)
    area = 0.5 * np.linalg.norm(np.cross(tx[1] - tx[0], tx[2] - tx[0]))

    def f(uv):
        lam = _bary(uv, apex)
        x = lam @ tx
        return g(x, lam) * (2 * area * uv[:, 0])[:, None]

    val, err = square(f, tol)
    return val, err


def block(tx, sx, source_velocity_many, tol, apex=0, sigma=0.0):
    """The pair block int psi_a n_T.(gamma_b x w_S) dA (hat weights) and its right hand side -sigma int psi_a
    (w_S.n_T) dA (src/assembly/assemble.jl:245-258) by cubature: (3, 3) [test corner, source corner] and (3,).
    gamma_b = -e_b / (2 A_S) with e_b the source's edge opposite corner b, counterclockwise
    (src/elements/flat_triangle.jl:23-31)."""
    tx, sx = np.asarray(tx, float), np.asarray(sx, float)
    nT = np.cross(tx[1] - tx[0], tx[2] - tx[0])
    nT /= np.linalg.norm(nT)
    nS = np.cross(sx[1] - sx[0], sx[2] - sx[0])
    A2S = np.linalg.norm(nS)
    e = np.stack([sx[2] - sx[1], sx[0] - sx[2], sx[1] - sx[0]])
    gam = -e / A2S

    def g(x, lam):
        w = source_velocity_many(sx[0], sx[1], sx[2], x)
        v = np.cross(w, nT)
        c = v @ gam.T                                      # (n, 3): c_b = (w x n_T).gamma_b
        out = (lam[:, :, None] * c[:, None, :]).reshape(len(x), 9)
        rhs = -sigma * (w @ nT)[:, None] * lam
        return np.concatenate([out, rhs], 1)

    val, err = triangle(tx, g, tol, apex)
    return val[:9].reshape(3, 3), val[9:], err


def segment_rows(tx, fa, fb, segment_velocity_many, tol, apex=0):
    """int psi_a n_T.v_a dA and int psi_a n_T.v_b dA of a filament fa -> fb (src/assembly/assemble.jl:312-321)"""
    tx = np.asarray(tx, float)
    nT = np.cross(tx[1] - tx[0], tx[2] - tx[0])
    nT /= np.linalg.norm(nT)

    def g(x, lam):
        va, vb = segment_velocity_many(fa, fb, x)
        return np.concatenate([lam * (va @ nT)[:, None], lam * (vb @ nT)[:, None]], 1)

    val, err = triangle(tx, g, tol, apex)
    return val[:3], val[3:], err
