"""The bordered dense solve on the device: `factorize_galerkin`, `solve_galerkin` (src/assembly/solve.jl:15-99).

The bordered matrix [Cᵀ A [C; W]  se e; sc cᵀ 0] is formed on the GPU from A (host array or an AssemblyPlan's device
matrix), factored there (randomized LU on MFMA, FP64 or FP32) and solved with FP64 iterative refinement.  The
host computes only the sparse products with C (Cᵀ b, C q) and the border vectors.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .api import AssemblyPlan, Context, _check, default_context
from .dofmap import DOFMapping, bordered_rows

MODES = {"fp64": 0, "mixed": 1}


class ConstraintsC(C.Structure):
    _fields_ = [("nslot", C.c_int), ("nq", C.c_int), ("nw", C.c_int), ("nb", C.c_int),
                ("C_rowptr", C.POINTER(C.c_int)), ("C_col", C.POINTER(C.c_int)), ("C_val", C.POINTER(C.c_double)),
                ("W_rowptr", C.POINTER(C.c_int)), ("W_col", C.POINTER(C.c_int)), ("W_val", C.POINTER(C.c_double)),
                ("e", C.POINTER(C.c_double)), ("c", C.POINTER(C.c_double))]


def _constraints(dm: DOFMapping, W):
    Cc = dm.C.tocsr()
    keep = dict(cp=L.i32(Cc.indptr), ci=L.i32(Cc.indices), cv=L.f64(Cc.data))
    nw = 0
    if W is not None:
        Wc = sp.csr_matrix(W)
        nw = Wc.shape[0]
        keep.update(wp=L.i32(Wc.indptr), wi=L.i32(Wc.indices) if Wc.nnz else np.zeros(1, np.int32),
                    wv=L.f64(Wc.data) if Wc.nnz else np.zeros(1))
    rows = bordered_rows(dm)
    keep["e"] = L.f64(np.array([r[1] for r in rows]).reshape(len(rows), dm.nq))
    keep["c"] = L.f64(np.array([r[2] for r in rows]).reshape(len(rows), dm.nq))
    k = ConstraintsC(dm.nslot, dm.nq, nw, len(rows), L.ptr(keep["cp"], C.c_int), L.ptr(keep["ci"], C.c_int),
                     L.ptr(keep["cv"]), L.ptr(keep.get("wp"), C.c_int) if nw else None,
                     L.ptr(keep.get("wi"), C.c_int) if nw else None, L.ptr(keep.get("wv")) if nw else None,
                     L.ptr(keep["e"]), L.ptr(keep["c"]))
    return k, keep, len(rows)


@dataclass
class GalerkinFactorization:
    """src/assembly/solve.jl:65-72: the device factorization handle plus se, C, nq, nb, and cond (estimated on its
    first use and kept).  last_resid: the reference's relative residual |R - M X|_2 / |R|_2 of the last solve (its
    worst right hand side)."""
    h: object
    ctx: Context
    se: np.ndarray
    C: sp.csr_matrix
    nq: int
    nb: int
    n: int
    mode: str
    last_iters: int = 0
    last_resid: float = 0.0
    _cond: float | None = None

    @property
    def cond(self) -> float:
        """1-norm condition number estimate of the bordered matrix (1/rcond of gecon, solve.jl:85)"""
        if self._cond is None:
            c = C.c_double()
            _check(self.ctx.h, self.ctx.lib.vp_factor_cond(self.h, C.byref(c)))
            self._cond = c.value
        return self._cond

    @property
    def precision_info(self) -> tuple[str, int]:
        """(the factors' current precision, FP64 fallbacks taken): mixed factors whose refinement does not converge
        are replaced by an FP64 factorization of the same matrix (LAPACK dsgesv's fallback)"""
        m, fb = C.c_int(), C.c_int()
        _check(self.ctx.h, self.ctx.lib.vp_factor_mode(self.h, C.byref(m), C.byref(fb)))
        return ("fp64" if m.value == 0 else "mixed"), fb.value

    @property
    def gmres_info(self) -> tuple[int, int]:
        """(solves finished by GMRES-IR, GMRES iterations): mixed factors whose plain refinement stops short are first
        used as the preconditioner of restarted GMRES on M (in FP64) before an FP64 refactorization"""
        sv, it = C.c_longlong(), C.c_longlong()
        _check(self.ctx.h, self.ctx.lib.vp_factor_gmres(self.h, C.byref(sv), C.byref(it)))
        return sv.value, it.value

    @property
    def inverse_info(self) -> tuple[bool, int]:
        """(explicit inverse of the factors in use, narrow applications counted before it was made): narrow solves
        apply the inverse by one matrix-vector product once a factorization has served 24 of them"""
        u, a = C.c_int(), C.c_longlong()
        _check(self.ctx.h, self.ctx.lib.vp_factor_inverse(self.h, C.byref(u), C.byref(a)))
        return bool(u.value), a.value

    def solve_bordered(self, R):
        """X = M^-1 R for R (n,) or (n, nrhs)"""
        R = np.asarray(R, dtype=np.float64)
        one = R.ndim == 1
        R2 = L.f64(R.reshape(self.n, -1))
        X = np.zeros_like(R2)
        it, res = C.c_int(), C.c_double()
        _check(self.ctx.h, self.ctx.lib.vp_factor_solve(self.h, L.ptr(R2), R2.shape[1], L.ptr(X), C.byref(it),
                                                        C.byref(res)))
        self.last_iters, self.last_resid = it.value, res.value
        return X[:, 0] if one else X

    def solve_bordered_t(self, R):
        """X = M^-T R for R (n,) or (n, nrhs), through the same factors read transposed"""
        R = np.asarray(R, dtype=np.float64)
        one = R.ndim == 1
        R2 = L.f64(R.reshape(self.n, -1))
        X = np.zeros_like(R2)
        it, res = C.c_int(), C.c_double()
        _check(self.ctx.h, self.ctx.lib.vp_factor_solve_transposed(self.h, L.ptr(R2), R2.shape[1], L.ptr(X),
                                                                   C.byref(it), C.byref(res)))
        self.last_iters, self.last_resid = it.value, res.value
        return X[:, 0] if one else X

    def close(self):
        if self.h:
            self.ctx.lib.vp_factor_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def factorize_galerkin(dm: DOFMapping, A, W=None, mode: str = "fp64", ctx: Context | None = None) -> GalerkinFactorization:
    """`factorize_galerkin(dm, A; W)`, src/assembly/solve.jl:81-87.  A: host array nslot x (nslot + nw) or an
    AssemblyPlan (its device matrix is used in place)."""
    if mode not in MODES:
        raise ValueError(f"mode must be fp64 or mixed, not {mode}")
    nw = 0 if W is None else W.shape[0]
    k, keep, nb = _constraints(dm, W)
    if isinstance(A, AssemblyPlan):
        ctx = A.ctx
        if A.nslot != dm.nslot or A.nw != nw:
            raise ValueError(f"A is ({A.nslot}, {A.nslot + A.nw}), expected ({dm.nslot}, {dm.nslot + nw})")
        dA, db, lda = C.c_void_p(), C.c_void_p(), C.c_longlong()
        _check(ctx.h, ctx.lib.vp_plan_device_A(A.h, C.byref(dA), C.byref(db), C.byref(lda)))
        Aptr, ondev, ld = C.cast(dA, C.POINTER(C.c_double)), 1, lda.value
    else:
        ctx = ctx or default_context()
        A = np.asarray(A, dtype=np.float64)
        if A.shape != (dm.nslot, dm.nslot + nw):
            raise ValueError(f"A is {A.shape}, expected ({dm.nslot}, {dm.nslot + nw})")
        keep["A"] = np.asfortranarray(A)
        Aptr, ondev, ld = L.ptr(keep["A"]), 0, dm.nslot
    h = C.c_void_p()
    _check(ctx.h, ctx.lib.vp_factorize(ctx.h, Aptr, ld, ondev, C.byref(k), MODES[mode], C.byref(h)))
    n = C.c_int()
    se = np.zeros(max(nb, 1))
    ctx.lib.vp_factor_info(h, C.byref(n), L.ptr(se))
    return GalerkinFactorization(h=h, ctx=ctx, se=se[:nb], C=dm.C.tocsr(), nq=dm.nq, nb=nb, n=n.value, mode=mode)


def solve_factored(F: GalerkinFactorization, b, cond: bool = True):
    """`solve_galerkin(F, b)`, src/assembly/solve.jl:95-99; b (nslot,) or (nslot, nrhs) for a batch.  Returns q, mu,
    lam and cond as the reference does (cond is estimated on the factorization's first solve; cond=False leaves it
    out, for callers that do not return it), plus the refinement's iters and relative residual resid."""
    b = np.asarray(b, dtype=np.float64)
    one = b.ndim == 1
    B = b.reshape(b.shape[0], -1)
    R = np.zeros((F.n, B.shape[1]))
    R[:F.nq] = F.C.T @ B
    X = F.solve_bordered(R)
    q = X[:F.nq]
    out = dict(q=q[:, 0] if one else q, mu=(F.C @ q)[:, 0] if one else F.C @ q,
               lam=(F.se[:, None] * X[F.nq:F.nq + F.nb])[:, 0] if one else F.se[:, None] * X[F.nq:F.nq + F.nb],
               cond=F.cond if cond else None, iters=F.last_iters, resid=F.last_resid)
    return out


def solve_galerkin(dm_or_F, A=None, b=None, W=None, mode: str = "fp64", ctx: Context | None = None, cond: bool = True):
    """`solve_galerkin(dm, A, b; W)` (src/assembly/solve.jl:51-59) -> q, mu, lambda, cond, residual (the relative
    residual |M x - rhs|_2 / |rhs|_2 of the bordered system); or `solve_galerkin(F, b)` when the first argument is a
    GalerkinFactorization.  cond=False skips the condition estimate (cond is then None)."""
    if isinstance(dm_or_F, GalerkinFactorization):
        return solve_factored(dm_or_F, A if b is None else b)
    F = factorize_galerkin(dm_or_F, A, W=W, mode=mode, ctx=ctx)
    out = solve_factored(F, b, cond=cond)
    out["residual"] = F.last_resid
    F.close()
    return out


def consistency(dm: DOFMapping, A, b, W=None, ctx: Context | None = None) -> dict:
    """`consistency(dm, A, b; W)`, src/assembly/solve.jl:111-115, on the device: the largest absolute row sum and
    column sum of A [C; W] over its largest absolute entry (rows, columns) and |sum b| / sum |b| (rhs).  A: host array
    nslot x (nslot + nw) or an AssemblyPlan (its device matrix is read in place)."""
    nw = 0 if W is None else W.shape[0]
    k, keep, _ = _constraints(dm, W)
    b = L.f64(b)
    if b.shape != (dm.nslot,):
        raise ValueError(f"b has shape {b.shape}, expected ({dm.nslot},)")
    if isinstance(A, AssemblyPlan):
        ctx = A.ctx
        if A.nslot != dm.nslot or A.nw != nw:
            raise ValueError(f"A is ({A.nslot}, {A.nslot + A.nw}), expected ({dm.nslot}, {dm.nslot + nw})")
        dA, db, lda = C.c_void_p(), C.c_void_p(), C.c_longlong()
        _check(ctx.h, ctx.lib.vp_plan_device_A(A.h, C.byref(dA), C.byref(db), C.byref(lda)))
        Aptr, ondev, ld = C.cast(dA, C.POINTER(C.c_double)), 1, lda.value
    else:
        ctx = ctx or default_context()
        A = np.asarray(A, dtype=np.float64)
        if A.shape != (dm.nslot, dm.nslot + nw):
            raise ValueError(f"A is {A.shape}, expected ({dm.nslot}, {dm.nslot + nw})")
        keep["A"] = np.asfortranarray(A)
        Aptr, ondev, ld = L.ptr(keep["A"]), 0, dm.nslot
    out = np.zeros(3)
    _check(ctx.h, ctx.lib.vp_consistency(ctx.h, Aptr, ld, ondev, L.ptr(b), C.byref(k), L.ptr(out)))
    return dict(rows=float(out[0]), columns=float(out[1]), rhs=float(out[2]))
