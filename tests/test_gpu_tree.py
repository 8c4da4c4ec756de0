"""The blob-sum treecode (vp_blob_velocities_tree) on the device against the oracle's direct FP64 sums
(blob_velocities, src/freewake/step.jl:176-187).  The treecode approximates: its error is measured relative to the
largest velocity, and the bounds below are the measured errors with a margin of about five.  It must converge with the
interpolation degree (about fivefold per degree at theta 0.5), reproduce the direct sum to rounding when the
separation parameter takes nothing far, keep each blob's own core, and not depend on the order of its inputs."""
import numpy as np
import pytest

from oracle import oracle as O
from vphip import post

pytestmark = pytest.mark.gpu


def _cloud(n, seed=0, cores=(0.005, 0.02)):
    rng = np.random.default_rng(seed)
    bx = rng.standard_normal((n, 3)) * np.array([3.0, 3.0, 0.3])
    ba = rng.standard_normal((n, 3)) * 1e-3
    bd = rng.uniform(*cores, n)
    return bx, ba, bd


def _err(u, ref):
    return np.abs(u - ref).max() / np.abs(ref).max()


def test_nothing_far_is_the_direct_sum():
    bx, ba, bd = _cloud(3000)
    print("tree theta 0.05, 3000 blobs", flush=True)
    u, st = post.blob_velocities_tree(bx, ba, bd, bx, theta=0.05, degree=4)
    assert st["far_pairs"] == 0
    assert _err(u, O.blob_velocities(bx, ba, bd, bx)) < 1e-13


def test_converges_with_degree():
    """20,000 blobs with cores from 0.005 to 0.02, every blob a target: error against the oracle for degrees 3..7"""
    bx, ba, bd = _cloud(20000, seed=1)
    ref = O.blob_velocities(bx, ba, bd, bx)
    errs = []
    for degree in (3, 4, 5, 6, 7):
        print(f"tree theta 0.5 degree {degree}, 20000 blobs", flush=True)
        u, st = post.blob_velocities_tree(bx, ba, bd, bx, theta=0.5, degree=degree)
        errs.append(_err(u, ref))
        print(f"  error {errs[-1]:.2e} proxy {st['proxy_evals']:.2e} near {st['near_evals']:.2e}", flush=True)
        assert st["far_pairs"] > 0
    assert all(b < a / 2 for a, b in zip(errs, errs[1:])), errs
    assert errs[2] < 5e-6 and errs[4] < 1e-7, errs


@pytest.mark.parametrize("theta,degree,bound", [(0.5, 5, 5e-6), (0.7, 5, 5e-5), (0.4, 6, 1e-7)])
def test_full_size_against_the_direct_sum(theta, degree, bound):
    """The free wake's end size, 84,000 blobs at all 84,000 blob positions, against the device's direct FP64 sums
    (themselves within 1e-12 of the oracle at this size) and on 300 targets against the oracle"""
    bx, ba, bd = _cloud(84000)
    print(f"tree theta {theta} degree {degree}, 84000 blobs", flush=True)
    u, st = post.blob_velocities_tree(bx, ba, bd, bx, theta=theta, degree=degree)
    ud = post.blob_velocities(bx, ba, bd, bx)
    e = _err(u, ud)
    pick = np.arange(0, 84000, 280)
    eo = np.abs(u[pick] - O.blob_velocities(bx, ba, bd, bx[pick])).max() / np.abs(ud).max()
    print(f"  error {e:.2e} (oracle sample {eo:.2e}) work/direct {(st['proxy_evals'] + st['near_evals']) / 84000**2:.3f} "
          f"build {st['build_ms']:.1f} ms device {st['device_ms']:.2f} ms", flush=True)
    assert e < bound and eo < bound


def test_mixed_precision():
    """FP32 kernels relative to a nearby origin (interpolation and sums FP64): 84,000 blobs within 5e-5 of the
    oracle's FP64 sums, the direct FP32 path's tolerance, and within 1e-5 of the FP64 treecode"""
    bx, ba, bd = _cloud(84000)
    print("tree mixed theta 0.5 degree 5, 84000 blobs", flush=True)
    u, st = post.blob_velocities_tree(bx, ba, bd, bx, theta=0.5, degree=5, precision="mixed")
    u64, _ = post.blob_velocities_tree(bx, ba, bd, bx, theta=0.5, degree=5)
    pick = np.arange(0, 84000, 280)
    eo = np.abs(u[pick] - O.blob_velocities(bx, ba, bd, bx[pick])).max() / np.abs(u64).max()
    e64 = _err(u, u64)
    print(f"  oracle sample {eo:.2e}, against the FP64 treecode {e64:.2e}, device {st['device_ms']:.2f} ms", flush=True)
    assert eo < 5e-5 and e64 < 1e-5


def test_separate_targets_planar_and_duplicates():
    """Targets apart from the blobs (a plane of targets beside the cloud), a flat sheet of blobs (boxes of zero
    thickness), and blobs stacked 150 deep on 40 points (clusters of identical points are never interpolated; the
    stacks' strong sources at the separation limit make this the least accurate case, 3e-6 measured at degree 6)"""
    rng = np.random.default_rng(3)
    bx, ba, bd = _cloud(6000, seed=3)
    xs = np.c_[rng.uniform(-6, 6, 5000), rng.uniform(-6, 6, 5000), np.full(5000, 0.7)]
    u, _ = post.blob_velocities_tree(bx, ba, bd, xs, theta=0.5, degree=6)
    e1 = _err(u, O.blob_velocities(bx, ba, bd, xs))
    sheet = np.c_[rng.uniform(0, 4, 8000), rng.uniform(-2, 2, 8000), np.zeros(8000)]
    sa, sd = rng.standard_normal((8000, 3)) * 1e-3, np.full(8000, 0.02)
    u, st = post.blob_velocities_tree(sheet, sa, sd, sheet, theta=0.5, degree=6)
    assert st["far_pairs"] > 0
    e2 = _err(u, O.blob_velocities(sheet, sa, sd, sheet))
    dup = np.repeat(rng.standard_normal((40, 3)), 150, axis=0)
    da, dd = rng.standard_normal((6000, 3)) * 1e-3, np.full(6000, 0.05)
    u, _ = post.blob_velocities_tree(dup, da, dd, dup, theta=0.5, degree=6)
    e3 = _err(u, O.blob_velocities(dup, da, dd, dup))
    print(f"  separate {e1:.2e} sheet {e2:.2e} duplicates {e3:.2e}", flush=True)
    assert e1 < 1e-6 and e2 < 1e-6 and e3 < 1.5e-5


def test_each_blob_keeps_its_core():
    """Cores spread over two decades: the treecode's error stays at the uniform-core level (sources are never merged,
    so no blob's core is averaged away: an averaged core would be off by about 1.5 (δ² - δ̄²) / r², some 1e-2 here)"""
    bx, ba, bd = _cloud(12000, seed=4, cores=(0.001, 0.1))
    u, _ = post.blob_velocities_tree(bx, ba, bd, bx, theta=0.5, degree=6)
    assert _err(u, O.blob_velocities(bx, ba, bd, bx)) < 1e-6


def test_input_order_does_not_matter():
    bx, ba, bd = _cloud(10000, seed=5)
    xs = bx[::3] + 0.01
    u, _ = post.blob_velocities_tree(bx, ba, bd, xs, theta=0.5, degree=5)
    p = np.random.default_rng(6).permutation(len(bx))
    q = np.random.default_rng(7).permutation(len(xs))
    u2, _ = post.blob_velocities_tree(bx[p], ba[p], bd[p], xs[q], theta=0.5, degree=5)
    assert np.abs(u2 - u[q]).max() < 1e-13 * np.abs(u).max()


def test_empty_and_single():
    u, _ = post.blob_velocities_tree(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.ones((5, 3)))
    assert u.shape == (5, 3) and not u.any()
    u, _ = post.blob_velocities_tree(np.ones((2, 3)), np.ones((2, 3)), np.ones(2), np.zeros((0, 3)))
    assert u.shape == (0, 3)
    bx, ba, bd = np.array([[0.3, -0.2, 0.1]]), np.array([[0.0, 0.0, 1.0]]), np.array([0.05])
    xs = np.random.default_rng(8).standard_normal((500, 3))
    u, _ = post.blob_velocities_tree(bx, ba, bd, xs, theta=0.5, degree=5)
    assert _err(u, O.blob_velocities(bx, ba, bd, xs)) < 1e-14


def test_bad_arguments_raise():
    x = np.zeros((4, 3))
    with pytest.raises(ValueError):
        post.blob_velocities_tree(x, x, np.ones(4), x, theta=1.2)
    with pytest.raises(ValueError):
        post.blob_velocities_tree(x, x, np.ones(4), x, degree=9)
    with pytest.raises(ValueError):
        post.blob_velocities_tree(x, x, np.ones(4), x, precision="fp16")
    with pytest.raises(ValueError):
        post.blob_velocities_tree(x, x, np.ones(3), x)
