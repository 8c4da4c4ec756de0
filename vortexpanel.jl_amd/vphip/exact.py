"""Exact potential flow about a fixed ellipsoid in uniform flow (Lamb, Hydrodynamics, section 114), the sphere being
a = b = c; restated from validation/common/exact_flows.jl:12-91 for reporting the errors of a solution."""
from __future__ import annotations

import numpy as np


def carlson_rd(x: float, y: float, z: float) -> float:
    """Carlson's R_D(x, y, z) by the duplication algorithm (Carlson 1995), validation/common/exact_flows.jl:12-31"""
    x0, y0, z0 = float(x), float(y), float(z)
    A0 = (x0 + y0 + 3 * z0) / 5
    Q = (np.spacing(A0) / 4) ** (-1 / 6) * max(abs(A0 - x0), abs(A0 - y0), abs(A0 - z0))
    x, y, z, A = x0, y0, z0, A0
    s, f = 0.0, 1.0
    while f * Q >= abs(A):
        lam = np.sqrt(x) * np.sqrt(y) + np.sqrt(x) * np.sqrt(z) + np.sqrt(y) * np.sqrt(z)
        s += f / (np.sqrt(z) * (z + lam))
        f /= 4
        x, y, z, A = (x + lam) / 4, (y + lam) / 4, (z + lam) / 4, (A + lam) / 4
    X, Y = f * (A0 - x0) / A, f * (A0 - y0) / A
    Z = -(X + Y) / 3
    e2, e3 = X * Y - 6 * Z ** 2, (3 * X * Y - 8 * Z ** 2) * Z
    e4, e5 = 3 * (X * Y - Z ** 2) * Z ** 2, X * Y * Z ** 3
    series = 1 - 3 * e2 / 14 + e3 / 6 + 9 * e2 ** 2 / 88 - 3 * e4 / 22 - 9 * e2 * e3 / 52 + 3 * e5 / 26
    return f * A ** (-1.5) * series + 3 * s


class EllipsoidFlow:
    """validation/common/exact_flows.jl:53-91: depolarization factors, the exact doublet k_mu.x and surface Cp"""

    def __init__(self, axes, vinf):
        a, b, c = (float(v) for v in axes)
        self.axes = np.array([a, b, c])
        self.V = np.asarray(vinf, dtype=np.float64)
        k = 2 * a * b * c / 3
        self.factors = np.array([k * carlson_rd(b ** 2, c ** 2, a ** 2), k * carlson_rd(c ** 2, a ** 2, b ** 2),
                                 k * carlson_rd(a ** 2, b ** 2, c ** 2)])
        self.kmu = self.factors / (2 - self.factors) * self.V
        self.k = 2 * self.V / (2 - self.factors)

    def doublet(self, x):
        return np.asarray(x) @ self.kmu

    def doublet_scale(self):
        return float(np.linalg.norm(self.kmu * self.axes))

    def cp(self, p):
        p = np.atleast_2d(p)
        xs = p / np.sqrt(np.sum((p / self.axes) ** 2, axis=1))[:, None]
        n = xs / self.axes ** 2
        n /= np.linalg.norm(n, axis=1)[:, None]
        u = self.k - (n @ self.k)[:, None] * n
        return 1 - np.sum(u * u, axis=1) / (self.V @ self.V)
