"""The blob-sum treecode's lists, built on the host without a device (vp_tree_coverage): for every target leaf, its
near source ranges and the source ranges of its far ancestors must cover every source exactly once, whatever the
point sets' shapes (clouds, flat sheets, lines, duplicated points, disjoint target and source sets, one source), the
separation parameter and the interpolation degree.  A list that dropped or repeated a source would show here."""
import numpy as np
import pytest

from vphip import post

RNG = np.random.default_rng(0)
CASES = {
    "cloud": (RNG.standard_normal((20000, 3)) * [3.0, 3.0, 0.3], None),
    "sheet": (np.c_[RNG.uniform(0, 4, 8000), RNG.uniform(-2, 2, 8000), np.zeros(8000)], None),
    "line": (np.c_[np.linspace(0, 1, 5000), np.zeros(5000), np.zeros(5000)], None),
    "duplicates": (np.repeat(RNG.standard_normal((40, 3)), 150, axis=0), None),
    "disjoint": (RNG.standard_normal((3000, 3)), RNG.standard_normal((5000, 3)) * 2 + [4.0, 0, 0]),
    "one_source": (np.zeros((1, 3)), RNG.standard_normal((700, 3))),
    "no_source": (np.zeros((0, 3)), RNG.standard_normal((700, 3))),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("theta,degree", [(0.5, 5), (0.3, 4), (0.8, 2), (0.5, 1), (0.6, 7)])
def test_every_pair_covered_once(name, theta, degree):
    bx, xs = CASES[name]
    xs = bx if xs is None else xs
    bad, st = post.tree_coverage(bx, xs, theta=theta, degree=degree)
    assert bad == 0
    P = (degree + 1) ** 3
    assert st["proxy_evals"] == P * (st["proxy_evals"] // P)
    assert st["near_evals"] <= len(bx) * len(xs)


def test_far_field_is_used():
    """On the 20,000-point cloud, at theta 0.5 and degree 5, most of the work goes through the proxies and the
    kernel evaluations fall well below the direct sum's"""
    bx = CASES["cloud"][0]
    bad, st = post.tree_coverage(bx, bx, theta=0.5, degree=5)
    assert bad == 0 and st["far_clusters"] > 0
    assert st["proxy_evals"] + st["near_evals"] < 0.7 * len(bx) ** 2


@pytest.mark.parametrize("theta,degree", [(0.0, 5), (1.0, 5), (-0.5, 5), (0.5, 0), (0.5, 8)])
def test_bad_options_raise(theta, degree):
    with pytest.raises(ValueError):
        post.tree_coverage(np.zeros((3, 3)), np.zeros((3, 3)), theta=theta, degree=degree)


def test_non_finite_points_raise():
    x = np.zeros((5, 3))
    x[2, 1] = np.nan
    with pytest.raises(ValueError):
        post.tree_coverage(x, np.zeros((4, 3)))
