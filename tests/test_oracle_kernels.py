"""The oracle's kernels against independent references (CPU): mpmath polar-quadrature golden vectors
(tests/golden/kernel_golden.json, test/kernels/test_kernels.jl:46-58), closed forms and invariants
(test/kernels/test_kernels.jl:69-158, test_segment.jl, test_blob.jl)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kernel_golden.json")))["cases"]


@pytest.mark.parametrize("case", GOLD, ids=[f"{c['shape']}|{c['label']}" for c in GOLD])
def test_closed_form_against_polar_quadrature(case):
    w = O.source_velocity(case["a"], case["b"], case["c"], case["x"], side=case["side"], far=np.inf)
    ref = np.array(case["w"])
    assert np.allclose(w, ref, rtol=1e-9, atol=1e-13), (w, ref)


def _tri(shape=3):
    c = GOLD[shape * 14]
    return np.array(c["a"]), np.array(c["b"]), np.array(c["c"])


@pytest.mark.parametrize("shape", range(4))
def test_inplane_limits(shape):
    a, b, c = _tri(shape)
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    cen = (a + b + c) / 3
    up, down, pv = (O.source_velocity(a, b, c, cen, side=s) for s in (1, -1, 0))
    assert abs(up @ n - 0.5) < 1e-12 and abs(down @ n + 0.5) < 1e-12 and abs(pv @ n) < 1e-14
    assert np.allclose(up - down, n)
    diam = max(np.linalg.norm(b - a), np.linalg.norm(c - b), np.linalg.norm(a - c))
    assert np.allclose(up, O.source_velocity(a, b, c, cen + 1e-9 * diam * n), rtol=1e-6)
    assert abs(O.inplane_fraction(a, b, c, (a + b) / 2) - 0.5) < 1e-12
    e1, e2 = b - a, c - a
    alpha = np.arccos(e1 @ e2 / np.linalg.norm(e1) / np.linalg.norm(e2))
    assert abs(O.inplane_fraction(a, b, c, a) - alpha / (2 * np.pi)) < 1e-12


@pytest.mark.parametrize("shape", range(4))
def test_far_field_continuity_and_order(shape):
    """test/kernels/test_kernels.jl:88-101: < 5e-6 at 10 diameters, fourth order beyond"""
    a, b, c = _tri(shape)
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    cen = (a + b + c) / 3
    diam = max(np.linalg.norm(b - a), np.linalg.norm(c - b), np.linalg.norm(a - c))
    for u in (np.array([1.0, 0.2, -0.3]), np.array([-0.4, 1.0, 0.9]), n):
        u = u / np.linalg.norm(u)
        for k, bound in ((1, 5e-6), (2, 5e-7)):
            x = cen + k * 10 * diam * u
            ex = O.source_velocity(a, b, c, x, mode=1)
            fa = O.source_velocity(a, b, c, x, mode=2)
            assert np.linalg.norm(fa - ex) / np.linalg.norm(ex) < bound
        x2 = cen + 20 * diam * u
        assert np.array_equal(O.source_velocity(a, b, c, x2), O.source_velocity(a, b, c, x2, mode=2))


def test_gauss_flux_equals_area():
    """test/kernels/test_kernels.jl:113-123: the flux out of a cube around the triangle is its area"""
    a, b, c = _tri(3)
    cen = (a + b + c) / 3
    diam = max(np.linalg.norm(b - a), np.linalg.norm(c - b), np.linalg.norm(a - c))
    area = np.linalg.norm(np.cross(b - a, c - a)) / 2
    L = 1.5 * diam
    xg, wg = np.polynomial.legendre.leggauss(24)
    xg, wg = L * xg, L * wg
    flux = 0.0
    for axis in range(3):
        i, j = [k for k in range(3) if k != axis]
        for s in (-1.0, 1.0):
            for xi, wi in zip(xg, wg):
                for xj, wj in zip(xg, wg):
                    p = np.zeros(3)
                    p[axis], p[i], p[j] = s * L, xi, xj
                    flux += wi * wj * s * O.source_velocity(a, b, c, cen + p)[axis]
    assert abs(flux / area - 1) < 1e-10


def test_orientation_flip():
    """test/kernels/test_kernels.jl:103-111"""
    a, b, c = _tri(3)
    for case in GOLD[42:56]:
        if case["label"] == "far":
            continue
        x, side = case["x"], case["side"]
        w1 = O.source_velocity(a, c, b, x, side=-side, far=np.inf)
        w2 = O.source_velocity(a, b, c, x, side=side, far=np.inf)
        assert np.allclose(w1, w2, rtol=1e-12, atol=1e-14)


def test_segment_infinite_line_and_quadrature():
    """test/kernels/test_segment.jl:13-36: Biot-Savart by quadrature; long segments tend to Gamma/(2 pi d)"""
    a, b = np.array([0.1, -0.2, 0.3]), np.array([1.2, 0.4, -0.1])
    x = np.array([0.4, 0.9, 0.5])
    va, vb = O.segment_velocity(a, b, x)
    s, w = np.polynomial.legendre.leggauss(60)
    s, w = (s + 1) / 2, w / 2
    ua, ub = np.zeros(3), np.zeros(3)
    dl = b - a
    for si, wi in zip(s, w):
        y = a + si * dl
        r = x - y
        k = np.cross(dl, r) / (4 * np.pi * np.linalg.norm(r) ** 3) * wi
        ua += (1 - si) * k
        ub += si * k
    assert np.allclose(va, ua, rtol=1e-12) and np.allclose(vb, ub, rtol=1e-12)
    L = 1e6
    va, vb = O.segment_velocity([0, 0, -L], [0, 0, L], [0.7, 0, 0])
    assert abs(np.linalg.norm(va + vb) - 1 / (2 * np.pi * 0.7)) < 1e-9
    for p in (a, b):  # zero at its ends
        z = O.segment_velocity(a, b, p)
        assert np.all(z[0] == 0) and np.all(z[1] == 0)
    z = O.segment_velocity([0, 0, 0], [0, 0, 2.0], [0, 0, 1.0])  # and on the segment
    assert np.all(z[0] == 0) and np.all(z[1] == 0)


def test_blob_point_vortex_limit_and_solenoidal():
    """test/kernels/test_blob.jl:7-38"""
    xp, al = np.array([0.2, -0.1, 0.3]), np.array([0.3, 1.0, -0.4])
    x = xp + np.array([30.0, -20.0, 10.0])
    r = x - xp
    u = O.blob_velocity(x, xp, al, 0.1)
    pv = np.cross(al, r) / (4 * np.pi * np.linalg.norm(r) ** 3)
    r2 = r @ r
    assert np.linalg.norm(u - pv * (r2 / (r2 + 0.01)) ** 1.5) / np.linalg.norm(pv) < 1e-13  # the RM core factor
    assert np.linalg.norm(u - pv) / np.linalg.norm(pv) < 2e-5  # point-vortex limit
    assert np.all(O.blob_velocity(xp, xp, al, 0.1) == 0)
    h = 1e-5
    x0 = xp + np.array([0.05, 0.02, -0.03])
    div = sum((O.blob_velocity(x0 + h * e, xp, al, 0.1) - O.blob_velocity(x0 - h * e, xp, al, 0.1))[k] / (2 * h)
              for k, e in enumerate(np.eye(3)))
    assert abs(div) < 1e-6
