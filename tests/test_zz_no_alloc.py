"""The no-allocation entry point vp_blob_velocities_device (device arrays in and out, partial sums in the library's
fixed per-GPU scratch array) against the allocating vp_blob_velocities on the same inputs: two calls in a row, the
second smaller (the scratch array keeps the first call's contents), in FP64 and mixed precision; and a size beyond the
scratch array rejected on the host."""
import ctypes as C

import numpy as np
import pytest

from vphip import api

pytestmark = pytest.mark.gpu

D = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(D)


class DeviceArray:
    """A device copy of a float64 host array through the library's own allocation and copy helpers"""

    def __init__(self, ctx, a):
        self.ctx, self.a = ctx, np.ascontiguousarray(a, dtype=np.float64)
        self.p = C.c_void_p()
        assert ctx.lib.vp_device_alloc(ctx.h, C.c_longlong(max(self.a.nbytes, 8)), C.byref(self.p)) == 0
        assert ctx.lib.vp_copy_to_device(ctx.h, self.p, _ptr(self.a), C.c_longlong(self.a.nbytes)) == 0

    @property
    def dptr(self):
        return C.cast(self.p, D)

    def read(self):
        out = np.empty_like(self.a)
        assert self.ctx.lib.vp_copy_from_device(self.ctx.h, _ptr(out), self.p, C.c_longlong(out.nbytes)) == 0
        return out

    def free(self):
        assert self.ctx.lib.vp_device_free(self.ctx.h, self.p) == 0


def _case(rng, nb, npts):
    bx = rng.uniform(-1, 1, (nb, 3))
    al = rng.standard_normal((nb, 3)) * 0.01
    dl = rng.uniform(0.02, 0.08, nb)
    xs = rng.uniform(-1.2, 1.2, (npts, 3))
    return bx, al, dl, xs


@pytest.mark.parametrize("precision", [0, 1])
def test_blob_velocities_no_alloc(precision):
    ctx = api.default_context()
    lib = ctx.lib
    rng = np.random.default_rng(17 + precision)
    for nb, npts in ((6000, 5000), (1500, 900)):   # the second call smaller than the first
        print(f"no-alloc blob sums precision={precision} nb={nb} np={npts}", flush=True)
        bx, al, dl, xs = _case(rng, nb, npts)
        ref = np.zeros((npts, 3))
        assert lib.vp_blob_velocities(ctx.h, nb, _ptr(bx), _ptr(al), _ptr(dl), npts, _ptr(xs), precision, _ptr(ref)) == 0
        dev = [DeviceArray(ctx, a) for a in (bx, al, dl, xs, np.zeros((npts, 3)))]
        rc = lib.vp_blob_velocities_device(ctx.h, nb, dev[0].dptr, dev[1].dptr, dev[2].dptr, npts, dev[3].dptr,
                                           precision, dev[4].dptr)
        assert rc == 0, lib.vp_last_error(ctx.h)
        assert lib.vp_sync(ctx.h) == 0
        u = dev[4].read()
        for d in dev:
            d.free()
        err = np.abs(u - ref).max() / np.abs(ref).max()
        assert err <= (1e-14 if precision == 0 else 1e-6), err   # the same chunks, added in the same order


def test_blob_velocities_no_alloc_rejects_oversize():
    ctx = api.default_context()
    lib = ctx.lib
    x = DeviceArray(ctx, np.zeros((4, 3)))
    too_many = (1 << 22) // 3 + 1   # one target more than the scratch array holds; rejected before any launch
    rc = lib.vp_blob_velocities_device(ctx.h, 1, x.dptr, x.dptr, x.dptr, too_many, x.dptr, 0, x.dptr)
    x.free()
    assert rc == -2 and b"scratch" in lib.vp_last_error(ctx.h)
