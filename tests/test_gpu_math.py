"""The device's own FP64 log and atan2 (the closed-form kernel's transcendentals) against the host C library's, through
the C ABI's probe: within 2 ulp over the whole range, the special values and signed zeros as the library gives them;
and its 1/sqrt of the multipole's distances (normal positive arguments) within 2 ulp, and the square root of the closed
form's squared distances within 3 ulp."""
import ctypes as C

import numpy as np
import pytest

from vphip import _lib as L, api

pytestmark = pytest.mark.gpu


def _probe(fn, a, b=None):
    ctx = api.default_context()
    a = L.f64(a)
    b = L.f64(b) if b is not None else None
    out = np.zeros_like(a)
    rc = ctx.lib.vp_math_probe(ctx.h, fn, len(a), L.ptr(a), L.ptr(b) if b is not None else None, L.ptr(out))
    assert rc == 0
    return out


def _ulps(out, ref):
    return np.abs(out - ref) / np.spacing(np.abs(ref))


def test_log():
    print("device log", flush=True)
    rng = np.random.default_rng(3)
    x = np.concatenate([10.0 ** rng.uniform(-307, 307, 20000), 1 + rng.uniform(-1e-3, 1e-3, 5000),
                        rng.uniform(0.5, 2.0, 5000), 1 + rng.uniform(-1e-12, 1e-12, 1000), [1.0, 2.0, 0.5, np.e],
                        [5e-324, 2.2e-308, 1.7e308, 0.7071067811865476, 1.4142135623730951]])
    out, ref = _probe(0, x), np.log(x)
    nz = ref != 0
    assert np.max(_ulps(out[nz], ref[nz])) <= 2.0
    assert out[~nz].max() == 0.0
    sp = np.array([0.0, -0.0, np.inf, -1.0, np.nan])
    o = _probe(0, sp)
    assert o[0] == -np.inf and o[1] == -np.inf and o[2] == np.inf and np.isnan(o[3]) and np.isnan(o[4])


def test_atan2():
    print("device atan2", flush=True)
    rng = np.random.default_rng(4)
    n = 40000
    y = rng.standard_normal(n) * 10.0 ** rng.uniform(-20, 20, n)
    x = rng.standard_normal(n) * 10.0 ** rng.uniform(-20, 20, n)
    t = np.tan(np.linspace(-np.pi / 2 + 1e-9, np.pi / 2 - 1e-9, 2001))  # every octant boundary region
    y = np.concatenate([y, t, -t, np.ones_like(t)])
    x = np.concatenate([x, np.ones_like(t), -np.ones_like(t), t])
    out, ref = _probe(1, y, x), np.arctan2(y, x)
    nz = ref != 0
    assert np.max(_ulps(out[nz], ref[nz])) <= 2.0
    zs = np.array([0.0, -0.0])
    for yy in zs:
        for xx in np.concatenate([zs, [1.0, -1.0]]):
            o = _probe(1, np.array([yy]), np.array([xx]))[0]
            r = np.arctan2(yy, xx)
            assert o == r and np.signbit(o) == np.signbit(r), (yy, xx, o, r)
    o = _probe(1, np.array([1.0, -2.0]), np.array([0.0, -0.0]))
    assert np.allclose(o, [np.pi / 2, -np.pi / 2], rtol=0, atol=1e-16)


def test_rsqrt():
    print("device rsqrt", flush=True)
    rng = np.random.default_rng(5)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 20000), rng.uniform(0.25, 4.0, 20000), [1.0, 4.0, 2.0, 0.5]])
    out, ref = _probe(2, x), 1.0 / np.sqrt(x)
    assert np.max(_ulps(out, ref)) <= 2.0


def test_distance_sqrt():
    """The closed form's vertex distances: sqrt of zero or normal positive squared distances within 3 ulp"""
    print("device distance sqrt", flush=True)
    rng = np.random.default_rng(6)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 20000), rng.uniform(0.25, 4.0, 20000), [1.0, 4.0, 2.0, 0.5],
                        [2.2250738585072014e-308, 1.7e308]])
    out, ref = _probe(3, x), np.sqrt(x)
    assert np.max(_ulps(out, ref)) <= 3.0
    assert _probe(3, np.array([0.0]))[0] == 0.0


def test_fp32_atan2_log1p():
    """The mixed path's FP32 atan2 (a degree-8 polynomial fitted here, with octant reduction) and log(1 + u) (odd
    series below 0.5, the hardware log above) against numpy's FP64 values of the same FP32 inputs: within 5e-7 of
    atan2 (relative, and 2e-7 absolute near 0) in every octant and at the axes, within 5e-7 relative of log1p (a few
    FP32 ulp: the hardware reciprocal and logarithm are within about 1 ulp)"""
    print("device fp32 atan2 and log1p", flush=True)
    rng = np.random.default_rng(8)
    y = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.uniform(-6, 6, 20000), [0.0, 1.0, -1.0, 0.0, 3.0]])
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.uniform(-6, 6, 20000), [1.0, 0.0, 0.0, -2.0, 3.0]])
    y32, x32 = y.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64)
    out, ref = _probe(4, y32, x32), np.arctan2(y32, x32)
    assert np.all(np.abs(out - ref) <= 5e-7 * np.abs(ref) + 2e-7), np.max(np.abs(out - ref) / (np.abs(ref) + 1e-300))
    u = np.concatenate([10.0 ** rng.uniform(-12, 6, 20000), [0.0, 0.4999999, 0.5, 1.0]]).astype(np.float32)
    u = u.astype(np.float64)
    out, ref = _probe(5, u), np.log1p(u)
    assert np.all(np.abs(out - ref) <= 5e-7 * ref + 1e-300), np.max(np.abs(out - ref) / np.maximum(ref, 1e-300))
