"""Golden vectors of the unit source triangle's velocity w(x) = (1/4pi) int_T (x - y)/|x - y|^3 dA.

Independent of the closed form under test: the polar quadrature of the reference's own test reference
(test/kernels/reference.jl:31-61) restated with mpmath -- the radial integrals analytic, the angle integrated
adaptively between the vertex directions at 30 digits.  Shapes and point categories follow
test/kernels/test_kernels.jl:7-44.  Writes kernel_golden.json next to this script.
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 30

ROT_A = np.array([0.3, -0.7, 0.65]) / np.linalg.norm([0.3, -0.7, 0.65])
TH = 1.1
K = np.array([[0, -ROT_A[2], ROT_A[1]], [ROT_A[2], 0, -ROT_A[0]], [-ROT_A[1], ROT_A[0], 0]])
ROT = np.eye(3) + np.sin(TH) * K + (1 - np.cos(TH)) * K @ K
SHIFT = np.array([0.4, -1.2, 2.5])
SHAPES = {
    "equilateral": ([0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]),
    "right": ([0, 0, 0], [1, 0, 0], [0, 1, 0]),
    "sliver (aspect 100)": ([0, 0, 0], [1, 0, 0], [0.5, 0.01, 0]),
    "skewed (175 deg angle)": ([0, 0, 0], [1, 0, 0], [1.9, 0.08, 0]),
}


def place(p):
    return ROT @ np.asarray(p, dtype=float) + SHIFT


def unit(v):
    return v / np.linalg.norm(v)


def tri_data(a, b, c):
    n = unit(np.cross(b - a, c - a))
    lens = [np.linalg.norm(b - a), np.linalg.norm(c - b), np.linalg.norm(a - c)]
    nu = [np.cross(b - a, n) / lens[0], np.cross(c - b, n) / lens[1], np.cross(a - c, n) / lens[2]]
    return n, max(lens), nu, (a + b + c) / 3


def test_points(a, b, c):
    """test/kernels/test_kernels.jl:21-44"""
    n, h, nu, cen = tri_data(a, b, c)
    mid = (a + b) / 2
    bis = unit(nu[0] + nu[1])
    e = unit(b - a)
    return [
        ("far", cen + 20 * h * unit(ROT @ np.array([1.0, 2.0, 0.5])), 0),
        ("near above", cen + 0.1 * h * n, 0),
        ("near below", cen - 0.1 * h * n + 0.2 * h * e, 0),
        ("very near", cen + 1e-6 * h * n, 0),
        ("in-plane inside, side +1", cen, 1),
        ("in-plane inside, side -1", cen, -1),
        ("in-plane inside, principal value", cen + 0.1 * h * e, 0),
        ("in-plane outside", cen + 2 * h * unit(nu[0] + 0.4 * e), 0),
        ("in-plane on an edge's extension", b + 0.5 * (b - a), 0),
        ("edge-adjacent above", mid + 1e-3 * h * n, 0),
        ("edge-adjacent in-plane inside", mid - 1e-3 * h * nu[0], 1),
        ("edge-adjacent in-plane outside", mid + 1e-3 * h * nu[0], 0),
        ("vertex-adjacent above", b + 1e-3 * h * unit(n + bis), 0),
        ("vertex-adjacent in-plane outside", b + 1e-3 * h * bis, 0),
    ]


def ray_interval(q, e):
    """[rho_in, rho_out] of the ray rho e inside the 2D triangle q (test/kernels/reference.jl:12-28)"""
    lo, hi = mp.mpf(0), mp.inf
    for k in range(3):
        a, b = q[k], q[(k + 1) % 3]
        t = (b[0] - a[0], b[1] - a[1])
        nu = (t[1], -t[0])
        c0 = -(nu[0] * a[0] + nu[1] * a[1])
        c1 = nu[0] * e[0] + nu[1] * e[1]
        if abs(c1) < mp.mpf("1e-40"):
            if c0 > 0:
                return None
        elif c1 > 0:
            hi = min(hi, -c0 / c1)
        else:
            lo = max(lo, -c0 / c1)
    return (lo, hi) if lo < hi else None


def w_polar(a, b, c, x, side):
    """test/kernels/reference.jl:31-61 in mpmath"""
    A = [mp.matrix([mp.mpf(float(v)) for v in p]) for p in (a, b, c)]
    X = mp.matrix([mp.mpf(float(v)) for v in x])
    cr = lambda u, v: mp.matrix([u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]])
    dt = lambda u, v: u[0] * v[0] + u[1] * v[1] + u[2] * v[2]
    nm = lambda u: mp.sqrt(dt(u, u))
    n = cr(A[1] - A[0], A[2] - A[0])
    n = n / nm(n)
    e1 = (A[1] - A[0]) / nm(A[1] - A[0])
    e2 = cr(n, e1)
    z = dt(X - A[0], n)
    size = max(nm(A[1] - A[0]), nm(A[2] - A[1]), nm(A[0] - A[2]))
    if abs(z) <= mp.mpf("1e-12") * size:
        z = mp.mpf(0)
    p = X - z * n
    q = [(dt(v - p, e1), dt(v - p, e2)) for v in A]
    Z = abs(z)

    def integrand(th, comp):
        e = (mp.cos(th), mp.sin(th))
        iv = ray_interval(q, e)
        if iv is None:
            return mp.mpf(0)
        r1, r2 = iv
        if Z > 0:
            F = lambda r: mp.asinh(r / Z) - r / mp.sqrt(r * r + Z * Z)
            tang = -(F(r2) - F(r1))
            normal = z * (1 / mp.sqrt(r1 * r1 + Z * Z) - 1 / mp.sqrt(r2 * r2 + Z * Z))
        else:
            tang = -mp.log(r2) if r1 == 0 else -mp.log(r2 / r1)
            normal = mp.mpf(side) if r1 == 0 else mp.mpf(0)
        return (tang * e[0], tang * e[1], normal)[comp]

    br = sorted(set([mp.mpf(0)] + [mp.atan2(v[1], v[0]) % (2 * mp.pi) for v in q] + [2 * mp.pi]))
    I = [sum(mp.quad(lambda t: integrand(t, k), [br[j], br[j + 1]]) for j in range(len(br) - 1)) for k in range(3)]
    w = (I[0] * e1 + I[1] * e2 + I[2] * n) / (4 * mp.pi)
    return [float(w[i]) for i in range(3)]


def main():
    out = []
    for name, (a, b, c) in SHAPES.items():
        a, b, c = place(a), place(b), place(c)
        for label, x, side in test_points(a, b, c):
            w = w_polar(a, b, c, x, side)
            out.append(dict(shape=name, label=label, a=a.tolist(), b=b.tolist(), c=c.tolist(), x=x.tolist(),
                            side=side, w=w))
            print(name, label, w, flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernel_golden.json"), "w") as f:
        json.dump(dict(source="tests/golden/gen_kernel_golden.py (mpmath polar quadrature, 30 digits)", cases=out), f,
                  indent=1)


if __name__ == "__main__":
    main()
