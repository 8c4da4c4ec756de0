"""The exact ellipsoid flow used to measure solutions (vphip/exact.py, validation/common/exact_flows.jl) against the
reference's checks of it (test/assembly/test_ellipsoid.jl:29-41): Carlson's R_D at Carlson's own test values, against
scipy's independent R_D and its defining integral; the depolarization factors summing to 2; the prolate spheroid's
closed form; and the sphere, where the doublet is V.x/2 and Cp = 1 - (9/4) sin^2(theta)."""
import math

import numpy as np
import scipy.integrate as si
import scipy.special as ss

from vphip.exact import EllipsoidFlow, carlson_rd

VE = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])


def test_carlson_rd():
    assert abs(carlson_rd(0, 2, 1) / 1.7972103521034 - 1) < 1e-12    # Carlson (1995)'s test values
    assert abs(carlson_rd(2, 3, 4) / 0.16510527294261 - 1) < 1e-12
    assert abs(carlson_rd(1, 1, 1) - 1) < 1e-14
    x, y, z = 0.3, 2.0, 0.7
    integral = si.quad(lambda t: 1 / ((t + z) * math.sqrt((t + x) * (t + y) * (t + z))), 0, np.inf, epsabs=1e-300,
                       epsrel=1e-14, limit=200)[0]
    assert abs(carlson_rd(x, y, z) / (1.5 * integral) - 1) < 1e-12
    rng = np.random.default_rng(4)
    for x, y, z in rng.uniform(0.01, 10, (20, 3)):
        assert abs(carlson_rd(x, y, z) / ss.elliprd(x, y, z) - 1) < 1e-13


def test_depolarization_factors():
    for axes in ((2.0, 1.0, 0.5), (3.0, 0.5, 0.5), (1.0, 1.0, 0.1)):
        assert abs(EllipsoidFlow(axes, VE).factors.sum() - 2) < 1e-13, axes
    e = math.sqrt(1 - (0.5 / 3) ** 2)  # the prolate spheroid's eccentricity
    f = EllipsoidFlow((3.0, 0.5, 0.5), VE).factors
    assert abs(f[0] / (2 * (1 - e ** 2) / e ** 3 * (math.atanh(e) - e)) - 1) < 1e-12
    assert abs(f[1] - f[2]) < 1e-14


def test_sphere_limit():
    sphere = EllipsoidFlow((1.0, 1.0, 1.0), VE)
    assert np.abs(sphere.factors - 2 / 3).max() < 1e-14
    x = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    assert abs(sphere.doublet(x) - VE @ x / 2) < 1e-14
    assert abs(sphere.cp(x)[0] - (1 - 9 / 4 * (1 - (VE @ x) ** 2))) < 1e-13
