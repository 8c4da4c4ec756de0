"""Quadrature rules of the oracle (test/elements/test_elements.jl:8-47): Gauss-Legendre moments, triangle-rule
monomial exactness, graded rules' log integrals and convergence."""
from math import factorial

import numpy as np
import pytest

from oracle import oracle as O


def mono(a, b, c):
    return 2 * factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 2)


@pytest.mark.parametrize("n", [1, 2, 5, 9])
def test_gauss_legendre(n):
    x, w = O.gauss_legendre(n)
    for k in range(2 * n):
        assert abs(np.sum(w * x ** k) - 1 / (k + 1)) < 1e-14


def test_graded_gauss():
    x, w = O.graded_gauss(6, 0.15, 3)
    assert abs(w.sum() - 1) < 1e-14 and abs(np.sum(w * np.log(x)) + 1) < 1e-6


@pytest.mark.parametrize("npts,degree", [(3, 2), (7, 5)])
def test_triangle_rules(npts, degree):
    kind = 0 if npts == 3 else 1
    lam, w = O.rule(kind, O.opts(far_points=npts) if kind == 0 else O.opts(near_points=npts))
    assert len(w) == npts and abs(w.sum() - 1) < 1e-14
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            for c in range(degree + 1 - a - b):
                assert abs(np.sum(w * lam[:, 0] ** a * lam[:, 1] ** b * lam[:, 2] ** c) - mono(a, b, c)) < 1e-14


def test_graded_rules():
    e, v, s = O.rule(3), O.rule(6), O.rule(2)
    assert (len(e[1]), len(v[1]), len(s[1])) == (1512, 378, 4536)
    for lam, w in (e, v, s):
        assert abs(w.sum() - 1) < 1e-13 and np.all(lam > 0) and np.allclose(lam.sum(1), 1)
        assert abs(np.sum(w * lam[:, 0] ** 2 * lam[:, 1] * lam[:, 2]) / mono(2, 1, 1) - 1) < 1e-12
    assert abs(np.sum(e[1] * np.log(e[0][:, 2])) / -1.5 - 1) < 1e-6
    assert abs(np.sum(v[1] * np.log(1 - v[0][:, 0])) / -0.5 - 1) < 1e-6
    assert abs(np.sum(s[1] * np.log(s[0]).sum(1)) / -4.5 - 1) < 1e-6
    for k in range(3):
        lam, w = O.rule(3 + k)
        assert abs(np.sum(w * np.log(lam[:, (k + 2) % 3])) / -1.5 - 1) < 1e-6
        lam, w = O.rule(6 + k)
        assert abs(np.sum(w * np.log(1 - lam[:, k])) / -0.5 - 1) < 1e-6
    errs = [abs(np.sum(w * np.log(l[:, 2])) + 1.5) for l, w in (O.rule(3, O.opts(touch_levels=L)) for L in (2, 4, 6))]
    assert errs[2] < errs[1] < errs[0]
