"""The free wake's host logic (vphip.freewake, src/freewake/step.jl:113-173) with its device work done by the FP64 CPU
oracle instead (assemble + LAPACK LU, explicit_rhs, body_velocity, blob sums) -- TEST INFRASTRUCTURE ONLY: the
reference against which the device path is checked step by step (tests/), and the free-wake step of bench.py's CPU
baseline legs.  nthreads: the oracle's OpenMP threads (0: its default)."""
import numpy as np
import scipy.linalg as sla

from . import oracle as O
from vphip import dofmap, freewake as FW


def oracle_opts(a):
    return O.opts(**{k: getattr(a, k) for k in ("test", "far_ratio", "near_ratio", "near_depth", "touch_levels",
                                                "touch_ratio", "touch_order", "corner_angle", "contact_rel", "kernel_far")})


def oracle_explicit_rhs(g, dm, we, particles, a, nthreads=0):
    sheets = segments = blobs = None
    if we is not None and we.sheets:
        sx = np.array([c for tri in we.sheets for c in tri])
        sheets = (sx, np.arange(len(sx), dtype=np.int32).reshape(-1, 3), np.array(we.sheet_mu))
    if we is not None and we.segments:
        segments = (np.array([np.concatenate([p, q]) for p, q in we.segments]), np.array(we.segment_gamma))
    if particles is not None and len(particles):
        blobs = (particles.x, particles.alpha, particles.delta)
    return O.explicit_rhs(g, dm, sheets, segments, blobs, oracle_opts(a), nthreads=nthreads)


class OracleFreeWakeSim(FW.FreeWakeHost):
    nthreads = 0

    def _setup(self):
        A, b, _, _ = O.assemble(self.g, self.dm, self.sigma, self.V, oracle_opts(self.assembly), wake=self.te,
                                nthreads=self.nthreads)
        self.b0 = b
        M, _, se = O.bordered_system(self.dm, A, b, dofmap.wake_rows(self.dm, self.te.te))
        self._lu, self._se = sla.lu_factor(M), se

    def _explicit_rhs(self, we, particles):
        return oracle_explicit_rhs(self.g, self.dm, we, particles, self.assembly, self.nthreads)

    def _solve(self, rhs):
        nq = self.dm.nq
        R = np.concatenate([self.dm.C.T @ rhs, np.zeros(len(self._se))])
        x = sla.lu_solve(self._lu, R)
        return x[:nq], self.dm.C @ x[:nq], self._se * x[nq:]

    def _velocities(self, gam, blobs, X):
        u = O.body_velocity(self.g, self.sigma, gam, X, self.assembly.kernel_far, nthreads=self.nthreads)
        if len(blobs):
            u = u + O.blob_velocities(blobs.x, blobs.alpha, blobs.delta, X, nthreads=self.nthreads)
        return u

    def close(self):
        pass
