"""GPU bordered solve (randomized LU on MFMA + FP64 refinement) against the oracle's LAPACK solve
(src/assembly/solve.jl:15-99), and the MFMA GEMM against numpy."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from vphip import _lib as L, api, dofmap, geometry, mesh, post, solve, wake as wk

pytestmark = pytest.mark.gpu

VG = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("shape", [(16, 16, 4), (64, 64, 64), (130, 70, 33), (300, 1, 257), (17, 200, 64),
                                   (260, 300, 64)])
def test_gemm(prec, shape):
    """Every autotuner candidate (the eight tile shapes and the library GEMM, index 8, which takes the trailing
    update's form, subtract without transposed A, and falls back to the default tile otherwise), plain and
    transposed A, set and subtract, against numpy."""
    M, N, K = shape
    print(f"gemm prec={prec} M={M} N={N} K={K}", flush=True)
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A, B, Cm = rng.standard_normal((M, K)), rng.standard_normal((K, N)), rng.standard_normal((M, N))
    if prec == 1:
        A, B, Cm = (x.astype(np.float32).astype(np.float64) for x in (A, B, Cm))
    At = np.ascontiguousarray(A.T)
    ctx = api.default_context()
    for tile in range(9):
        for ta in (0, 1):
            for sub in (0, 1):
                out = np.ascontiguousarray(Cm.copy())
                src = At if ta else A
                rc = ctx.lib.vp_gemm(ctx.h, prec, sub | (ta << 1) | (tile << 4), M, N, K,
                                     src.ctypes.data_as(C.POINTER(C.c_double)), B.ctypes.data_as(C.POINTER(C.c_double)),
                                     out.ctypes.data_as(C.POINTER(C.c_double)))
                assert rc == 0
                ref = (Cm - A @ B) if sub else A @ B
                err = np.abs(out - ref).max() / np.abs(ref).max()
                assert err < (1e-13 if prec == 0 else 2e-6), (tile, ta, sub, err)


@pytest.mark.parametrize("prec,nb", [(0, 64), (1, 64), (0, 37), (1, 37)])
def test_diag_lu_block(prec, nb):
    """The LU's diagonal-block kernel against numpy on a diagonally weighted random block (no pivoting needed): the
    packed factors reproduce the block (FP64 1e-14, FP32 2e-6 of its largest entry), and the inverses of L and U are
    inverses (the identity beyond nb)"""
    print(f"diag_lu prec={prec} nb={nb}", flush=True)
    rng = np.random.default_rng(nb + 100 * prec)
    A = rng.standard_normal((64, 64)) + 8 * np.eye(64)
    if prec == 1:
        A = A.astype(np.float32).astype(np.float64)
    LU, Li, Ui = np.zeros((64, 64)), np.zeros((64, 64)), np.zeros((64, 64))
    ctx = api.default_context()
    rc = ctx.lib.vp_diag_lu_probe(ctx.h, prec, nb, *(x.ctypes.data_as(C.POINTER(C.c_double)) for x in (A, LU, Li, Ui)))
    assert rc == 0
    L = np.tril(LU[:nb, :nb], -1) + np.eye(nb)
    U = np.triu(LU[:nb, :nb])
    tol = 1e-14 if prec == 0 else 2e-6
    assert np.abs(L @ U - A[:nb, :nb]).max() <= tol * np.abs(A).max()
    Lf, Uf = np.eye(64), np.eye(64)
    Lf[:nb, :nb], Uf[:nb, :nb] = L, U
    assert np.abs(Li @ Lf - np.eye(64)).max() <= 50 * tol and np.abs(Ui @ Uf - np.eye(64)).max() <= 50 * tol


def test_gemm_rejects_bad_flags():
    ctx = api.default_context()
    A = np.zeros(4)
    for flags in (4, 9 << 4, 1 << 8):
        rc = ctx.lib.vp_gemm(ctx.h, 0, flags, 2, 2, 2, *(A.ctypes.data_as(C.POINTER(C.c_double)) for _ in range(3)))
        assert rc == L.VP_E_ARGUMENT


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n", [5, 64, 300, 700])
def test_dense_solve(mode, n):
    print(f"dense solve mode={mode} n={n}", flush=True)
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    r = rng.standard_normal(n)
    x = np.zeros(n)
    it = C.c_int()
    ctx = api.default_context()
    rc = ctx.lib.vp_dense_solve(ctx.h, n, np.ascontiguousarray(M).ctypes.data_as(C.POINTER(C.c_double)), mode,
                                r.ctypes.data_as(C.POINTER(C.c_double)), x.ctypes.data_as(C.POINTER(C.c_double)),
                                C.byref(it))
    assert rc == 0
    xs = np.linalg.solve(M, r)
    assert np.abs(x - xs).max() / np.abs(xs).max() < 1e-12, (np.abs(x - xs).max(), it.value)


def _case(g, dm, V, w=None):
    els = api.build_elements(g)
    sig = api.source_strengths(els, V)
    A, b, _ = api.assemble(g, dm, els, sig, V, wake=w)
    W = None if w is None else dofmap.wake_rows(dm, w.te)
    return A, b, W


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_sphere_solve(mode):
    g = mesh.icosphere(3)
    dm = dofmap.build_dofmap(g)
    A, b, _ = _case(g, dm, VG)
    print(f"sphere solve mode={mode}", flush=True)
    sol = solve.solve_galerkin(dm, A, b, mode=mode)
    ref = O.solve_galerkin(dm, A, b)
    e = np.max(np.abs(sol["mu"] - ref["mu"])) / np.max(np.abs(ref["mu"]))
    assert e < 1e-11, e
    assert abs(sol["lam"][0] - ref["lam"][0]) < 1e-14
    assert abs(sol["cond"] / ref["cond"] - 1) < 0.05, (sol["cond"], ref["cond"])
    assert sol["residual"] < 1e-12, sol["residual"]  # test/assembly/test_sphere.jl:52
    # lambda takes up an incompatible right hand side (test/assembly/test_sphere.jl:50-51)
    sh = solve.solve_galerkin(dm, A, b + 1e-3, mode=mode, cond=False)
    assert abs(sh["lam"][0] - 1e-3) < 1e-9


def test_wing_solve_and_batch():
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=24, n_span=8, aspect=6.0))
    dm = dofmap.build_dofmap(g)
    V = np.array([1.0, 0.0, 0.0])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
    A, b, W = _case(g, dm, V, w)
    print("wing solve", flush=True)
    for mode in ("fp64", "mixed"):
        F = solve.factorize_galerkin(dm, A, W=W, mode=mode)
        # batch of angles: b is linear in Vinf with a fixed (body-fixed) wake
        els = api.build_elements(g)
        angs = np.radians(np.linspace(-4, 12, 9))
        Bs = []
        for a in angs:
            Va = np.array([np.cos(a), 0, np.sin(a)])
            _, ba, _ = api.assemble(g, dm, els, api.source_strengths(els, Va), Va, wake=w)
            Bs.append(ba)
        Bm = np.stack(Bs, axis=1)
        out = solve.solve_galerkin(F, Bm)
        for j in range(len(angs)):
            ref = O.solve_galerkin(dm, A, Bm[:, j], W)
            e = np.max(np.abs(out["mu"][:, j] - ref["mu"])) / np.max(np.abs(ref["mu"]))
            assert e < 1e-10, (mode, j, e)
        F.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_pipeline_wing_solve(mode):
    """The wing of the generator rebuilt by the geometry pipeline from its raw quads with the generator's diagonals:
    the same surface triangles, chain and labels, the pipeline's own ladder caps.  The solve agrees with the oracle's
    on it, and its lift with the generator's wing to the caps' share (the caps differ, the lifting surface does not)."""
    spec = mesh.WingSpec(code="0012", n_around=24, n_span=8, aspect=6.0)
    X, F = mesh.naca_wing_raw(spec, split_te=True)
    pattern = tuple(1 if np.linalg.norm(X[c] - X[a]) <= np.linalg.norm(X[d] - X[b]) * (1 + 1e-12) else 2
                    for a, b, c, d in F)
    V = np.array([np.cos(np.radians(6)), 0, np.sin(np.radians(6))])
    cl = []
    for g in (geometry.build_geometry(X, F, geometry.GeometryOptions(split="pattern", split_pattern=pattern)),
              mesh.naca_wing(spec)):
        dm = dofmap.build_dofmap(g)
        w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
        A, b, W = _case(g, dm, V, w)
        print(f"pipeline wing solve mode={mode} nt={g.nt}", flush=True)
        F_ = solve.factorize_galerkin(dm, A, W=W, mode=mode)
        sol = solve.solve_galerkin(F_, b)
        F_.close()
        ref = O.solve_galerkin(dm, A, b, W)
        e = np.max(np.abs(sol["mu"] - ref["mu"])) / np.max(np.abs(ref["mu"]))
        assert e < 1e-10, (mode, e)
        eL = np.array([-np.sin(np.radians(6)), 0, np.cos(np.radians(6))])
        cl.append(post.kutta_joukowski_force(dm.wake_te, dm.K @ sol["q"], V) @ eL)
    assert abs(cl[0] - cl[1]) < 0.01 * abs(cl[1]), cl


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_wing_transposed_solve(mode):
    """Mᵀ X = R through the forward factors (the adjoint system) against numpy's solve of the oracle's bordered
    matrix (src/assembly/solve.jl:15-42), and the duality zᵀ (M^-1 r) = (M^-T z)ᵀ r."""
    g = mesh.naca_wing(mesh.WingSpec(code="2412", n_around=20, n_span=6, aspect=5.0))
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(0.07), 0.0, np.sin(0.07)])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
    A, b, W = _case(g, dm, V, w)
    print(f"wing transposed solve mode={mode}", flush=True)
    M, _, _ = O.bordered_system(dm, A, b, W)
    F = solve.factorize_galerkin(dm, A, W=W, mode=mode)
    rng = np.random.default_rng(5)
    R = rng.standard_normal((F.n, 3))
    X = F.solve_bordered_t(R)
    ref = np.linalg.solve(M.T, R)
    assert np.abs(X - ref).max() / np.abs(ref).max() < 1e-11
    assert 0 <= F.last_resid < 1e-12  # |R - Mᵀ X|_2 / |R|_2, the reference's residual bound (test_sphere.jl:52)
    assert abs(F.last_resid - np.linalg.norm(R - M.T @ X) / np.linalg.norm(R)) < 1e-13
    r, z = rng.standard_normal(F.n), rng.standard_normal(F.n)
    lhs, rhs = z @ F.solve_bordered(r), F.solve_bordered_t(z) @ r
    assert abs(lhs - rhs) < 1e-11 * (abs(lhs) + 1)
    ref_cond = np.linalg.cond(M, 1)
    # a lower bound (Hager-Higham); FP32 factors give it to about 1e-4
    assert 0.3 < F.cond / ref_cond <= 1.0 + (1e-6 if mode == "fp64" else 1e-3), (F.cond, ref_cond)
    F.close()


@pytest.mark.parametrize("mode,graphs", [("fp64", "1"), ("mixed", "1"), ("mixed", "0")])
def test_repeated_solves_replay(mode, graphs, monkeypatch):
    """The factors' inverse is applied eagerly on its first call for a direction and number of right hand sides,
    captured as a HIP graph on the second and replayed after that: every solve of a sequence that changes the
    direction and the batch width (forcing new captures), with eager work between them, against numpy's solve of
    the oracle's bordered matrix; the residual GEMV (1-4 right hand sides) and the MFMA GEMM (more) both take part.
    With VP_GRAPHS=0 every application is eager."""
    monkeypatch.setenv("VP_GRAPHS", graphs)
    g = mesh.naca_wing(mesh.WingSpec(code="2412", n_around=20, n_span=6, aspect=5.0))
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(0.07), 0.0, np.sin(0.07)])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
    A, b, W = _case(g, dm, V, w)
    print(f"repeated solves mode={mode}", flush=True)
    M, _, _ = O.bordered_system(dm, A, b, W)
    F = solve.factorize_galerkin(dm, A, W=W, mode=mode)
    rng = np.random.default_rng(11)
    for k, (nrhs, trans) in enumerate([(1, False)] * 4 + [(1, True)] * 3 + [(3, False)] * 3 + [(1, False)] * 3
                                      + [(6, True)] * 3 + [(2, True)] * 3):
        R = rng.standard_normal((F.n, nrhs))
        X = F.solve_bordered_t(R) if trans else F.solve_bordered(R)
        ref = np.linalg.solve(M.T if trans else M, R)
        assert np.abs(X - ref).max() / np.abs(ref).max() < 1e-11, (k, nrhs, trans)
        assert F.last_resid < 1e-12, (k, F.last_resid)
        np.linalg.norm(M @ ref - R)   # host work between the device calls
    F.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_two_bodies(mode):
    """Two spheres of uneven size and resolution in one geometry: one set of work lists over both bodies' pairs (tier
    counts and entries against the oracle), and the bordered solve with one zero-mean row per body (mu and both border
    multipliers against the oracle's LAPACK solve)."""
    print(f"two bodies mode={mode}", flush=True)
    g = mesh.combine([mesh.icosphere(3), mesh.icosphere(1, r=0.4, center=(2.2, 0.3, -0.1))])
    dm = dofmap.build_dofmap(g)
    els = api.build_elements(g)
    sig = api.source_strengths(els, VG)
    A, b, stats = api.assemble(g, dm, els, sig, VG)
    Ao, bo, po, _ = O.assemble(g, dm, sig, VG)
    assert stats["pairs"] == po
    assert np.abs(A - Ao).max() < 1e-12 * np.abs(Ao).max()
    assert np.abs(b - bo).max() < 1e-12 * np.abs(bo).max()
    sol = solve.solve_galerkin(dm, A, b, mode=mode)
    ref = O.solve_galerkin(dm, Ao, bo)
    assert np.abs(sol["mu"] - ref["mu"]).max() < 1e-10 * np.abs(ref["mu"]).max()
    assert sol["lam"].shape == (2,) and np.abs(sol["lam"] - ref["lam"]).max() < 1e-12
    sh = solve.solve_galerkin(dm, A, b + 1e-3, mode=mode, cond=False)   # the multipliers take up the inconsistency
    rsh = O.solve_galerkin(dm, Ao, bo + 1e-3)
    assert np.abs(sh["lam"] - rsh["lam"]).max() < 1e-12, (sh["lam"], rsh["lam"])
    assert np.abs(sh["lam"] - 1e-3).max() < 1e-5, sh["lam"]


def _dense_solve(M, r, mode):
    ctx = api.default_context()
    n = M.shape[0]
    Mc, rc = np.ascontiguousarray(M), np.ascontiguousarray(r)
    x, it = np.zeros(n), C.c_int()
    st = ctx.lib.vp_dense_solve(ctx.h, n, Mc.ctypes.data_as(C.POINTER(C.c_double)), mode,
                                rc.ctypes.data_as(C.POINTER(C.c_double)), x.ctypes.data_as(C.POINTER(C.c_double)),
                                C.byref(it))
    return st, x, it.value


def test_mixed_refinement_falls_back_to_fp64():
    """A matrix whose condition (1e10) is beyond FP32 refinement: the FP32 factors' refinement stops reducing the
    residual, so the solve replaces them by an FP64 factorization of the same matrix (dsgesv's fallback) and returns
    the FP64 solution, whose residual (1e-13 relative) FP32 factors cannot reach at that condition; a well conditioned
    one converges (5 to 7 steps measured on this matrix: the unpivoted FP32 factors of a dense orthogonal-type matrix
    lose more than a diagonally dominant Galerkin matrix's, which converges in two)."""
    rng = np.random.default_rng(3)
    n = 300
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    x0 = rng.standard_normal(n)
    for lo, fallback in ((1e-1, False), (1e-10, True)):
        M = (Q1 * np.logspace(0, np.log10(lo), n)) @ Q2.T
        r = M @ x0
        print(f"dense solve mode=mixed cond={1 / lo:.0e}", flush=True)
        st, x, it = _dense_solve(M, r, 1)
        assert st == 0, api.default_context().lib.vp_last_error(api.default_context().h)
        ref = np.linalg.solve(M, r)
        assert fallback or it <= 10, it  # 5 to 7 steps here, depending on the factors' rounding
        assert np.linalg.norm(M @ x - r) / np.linalg.norm(r) < 1e-13
        assert np.abs(x - ref).max() / np.abs(ref).max() < 1e-4


def test_consistency_matches_oracle():
    """`consistency` (src/assembly/solve.jl:111-115) on the device, from a host A and from an assembly plan's device A,
    against the oracle's: a closed sphere (row and rhs defects at rounding level, test/assembly/test_sphere.jl:44-47)
    and a wing whose wake rows W join A's wake columns."""
    g = mesh.icosphere(3)
    dm = dofmap.build_dofmap(g)
    A, b, _ = _case(g, dm, VG)
    print("consistency icosphere(3)", flush=True)
    got, ref = solve.consistency(dm, A, b), O.consistency(dm, A, b)
    for key in ("rows", "columns", "rhs"):
        assert abs(got[key] - ref[key]) < 1e-14 + 1e-9 * ref[key], (key, got[key], ref[key])
    assert got["rows"] < 1e-12 and got["rhs"] < 1e-12  # test/assembly/test_sphere.jl:44-47
    gw = mesh.naca_wing(mesh.WingSpec(code="2412", n_around=20, n_span=6, aspect=5.0))
    dmw = dofmap.build_dofmap(gw)
    V = np.array([np.cos(0.07), 0.0, np.sin(0.07)])
    w = wk.build_wake(gw, dmw, V, wk.WakeOptions(length=2.0))
    print("consistency wing with wake", flush=True)
    plan = api.AssemblyPlan(gw, dmw, w)
    plan.assemble(api.source_strengths(api.build_elements(gw), V), V)
    Aw, bw = plan.read()
    W = dofmap.wake_rows(dmw, w.te)
    ref = O.consistency(dmw, Aw, bw, W)
    for src in (Aw, plan):
        got = solve.consistency(dmw, src, bw, W)
        for key in ("rows", "columns", "rhs"):
            assert abs(got[key] - ref[key]) < 1e-14 + 1e-9 * ref[key], (key, got[key], ref[key])


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_sphere_exact_flow_configured(mode):
    """The icosphere(3) sphere at its configured freestream V = (1, 0, 0) on the device path (assembly, bordered solve,
    gradient-route Cp) against the exact flow mu = V.x/2 and Cp = 1 - (9/4) sin^2(theta), at the reference's published
    accuracy (README.md:210-217: mu L2 5.8e-4, max 1.9e-3 of the doublet scale 1/2, Cp L2 9.8e-3, cond 640-650,
    lambda ~ 0), and against the oracle's solution"""
    V = np.array([1.0, 0.0, 0.0])
    g = mesh.icosphere(3)
    dm = dofmap.build_dofmap(g)
    els = api.build_elements(g)
    sig = api.source_strengths(els, V)
    print(f"sphere exact flow V=(1,0,0) mode={mode}", flush=True)
    A, b, _ = api.assemble(g, dm, els, sig, V, api.AssemblyOptions(precision=mode))
    sol = solve.solve_galerkin(dm, A, b, mode=mode)
    err = sol["mu"] - g.x[dm.slot_node] @ V / 2
    err -= np.sum(dm.slot_area * err) / np.sum(dm.slot_area)
    muL2 = np.sqrt(np.sum(dm.slot_area * err ** 2) / np.sum(dm.slot_area)) / 0.5
    cp = post.pressure_coefficient(post.gradient_velocity(els, dm, sol["mu"], V), V)
    c = els.centroid / np.linalg.norm(els.centroid, axis=1)[:, None]
    cpL2 = np.sqrt(np.sum(els.area * (cp - (1 - 9 / 4 * (1 - (c @ V) ** 2))) ** 2) / np.sum(els.area))
    assert abs(muL2 - 5.8e-4) < 0.05e-4 and np.abs(err).max() / 0.5 < 1.9e-3, (muL2, np.abs(err).max() / 0.5)
    assert abs(cpL2 - 9.8e-3) < 0.3e-3, cpL2
    assert 640 < sol["cond"] < 650 and abs(sol["lam"][0]) < 1e-14 and sol["residual"] < 1e-12
    Ao, bo, _, _ = O.assemble(g, dm, O.source_strengths(O.elements(g.x, g.tris), V), V)
    ref = O.solve_galerkin(dm, Ao, bo)
    assert np.abs(sol["mu"] - ref["mu"]).max() < (1e-11 if mode == "fp64" else 1e-7) * np.abs(ref["mu"]).max()


def _conditioned(n, kappa, seed=7):
    """A dense n x n matrix with singular values spaced logarithmically from 1 to 1/kappa"""
    rng = np.random.default_rng(seed)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q1 * np.logspace(0, -np.log10(kappa), n)) @ Q2.T, rng.standard_normal(n)


@pytest.mark.parametrize("kappa", [1e2, 1e3])
def test_gmres_ir_where_refinement_stops_short(kappa):
    """A dense 642 x 642 nonsymmetric matrix of condition 1e2 and 1e3 in the bordered system: the unpivoted FP32
    factors are poor there (their growth, not the condition), so plain refinement stops short (with VP_GMRES=0 the
    solve takes the FP64 refactorization), and GMRES on M in FP64 preconditioned by the FP32 factors (GMRES-IR)
    finishes it in mixed precision, in about 100 of its 240 iterations, its solution the FP64 factorization's to
    1e-10."""
    import os
    g = mesh.icosphere(3)
    dm = dofmap.build_dofmap(g)
    A, b = _conditioned(dm.nslot, kappa)
    ref = solve.solve_galerkin(dm, A, b, mode="fp64")
    scale = np.abs(ref["mu"]).max()
    print(f"gmres-ir kappa {kappa:.0e}: mixed factors", flush=True)
    F = solve.factorize_galerkin(dm, A, mode="mixed")
    out = solve.solve_factored(F, b, cond=False)
    nsolve, nit = F.gmres_info
    prec, fb = F.precision_info
    print(f"  gmres solves {nsolve} iterations {nit} fallbacks {fb} resid {F.last_resid:.2e}", flush=True)
    assert nsolve == 1 and 0 < nit <= 240 and fb == 0 and prec == "mixed"
    assert np.abs(out["mu"] - ref["mu"]).max() < 1e-10 * scale
    F.close()
    os.environ["VP_GMRES"] = "0"
    try:
        F = solve.factorize_galerkin(dm, A, mode="mixed")
        out = solve.solve_factored(F, b, cond=False)
        assert F.gmres_info == (0, 0) and F.precision_info == ("fp64", 1)
        assert np.abs(out["mu"] - ref["mu"]).max() < 1e-10 * scale
        F.close()
    finally:
        os.environ.pop("VP_GMRES", None)


def test_gmres_ir_alone_with_good_factors():
    """VP_GMRES=2 runs GMRES-IR right after the first application of the mixed factors: where the factors are good
    (the sphere's Galerkin matrix, a symmetric positive definite matrix of condition 1e4) it reaches the FP64 solution
    in a few iterations, as a right-preconditioned GMRES must (a wrong preconditioner application would take
    hundreds)"""
    import os
    g = mesh.icosphere(3)
    dm = dofmap.build_dofmap(g)
    V = np.array([1.0, 0.2, 0.1])
    els = api.build_elements(g)
    Ag, bg, _ = api.assemble(g, dm, els, api.source_strengths(els, V), V)
    rng = np.random.default_rng(7)
    Q, _ = np.linalg.qr(rng.standard_normal((dm.nslot, dm.nslot)))
    As = (Q * np.logspace(0, -4, dm.nslot)) @ Q.T
    os.environ["VP_GMRES"] = "2"
    try:
        for name, A, b in (("sphere", Ag, bg), ("spd 1e4", As, rng.standard_normal(dm.nslot))):
            ref = solve.solve_galerkin(dm, A, b, mode="fp64")
            F = solve.factorize_galerkin(dm, A, mode="mixed")
            out = solve.solve_factored(F, b, cond=False)
            print(f"gmres-ir alone {name}: {F.gmres_info} {F.precision_info}", flush=True)
            nsolve, nit = F.gmres_info
            assert nsolve == 1 and 0 < nit <= 30 and F.precision_info == ("mixed", 0)
            assert np.abs(out["mu"] - ref["mu"]).max() < 1e-10 * np.abs(ref["mu"]).max()
            F.close()
    finally:
        os.environ.pop("VP_GMRES", None)


def test_gmres_ir_gives_up_to_the_fp64_refactorization():
    """Condition 1e8: GMRES-IR runs its 8 restarts of 30 steps without meeting the stopping test, and the solve
    falls back to the FP64 refactorization, whose solution is returned"""
    g = mesh.icosphere(3)
    dm = dofmap.build_dofmap(g)
    A, b = _conditioned(dm.nslot, 1e8)
    ref = solve.solve_galerkin(dm, A, b, mode="fp64")
    F = solve.factorize_galerkin(dm, A, mode="mixed")
    out = solve.solve_factored(F, b, cond=False)
    print(f"gmres-ir kappa 1e8: {F.gmres_info} {F.precision_info}", flush=True)
    assert F.gmres_info == (0, 240) and F.precision_info == ("fp64", 1)
    assert np.abs(out["mu"] - ref["mu"]).max() < 1e-10 * np.abs(ref["mu"]).max()
    F.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_explicit_inverse_of_the_factors(mode):
    """Narrow solves through the explicit inverse of the factors (VP_LU_INVERSE=1: formed at the first narrow
    application) give the solves through the triangular sweeps (VP_LU_INVERSE=0) to the refinement's accuracy, for M
    and Mᵀ, one to four right-hand sides, on the wing with its wake; by default the inverse appears at the first solve
    after 24 narrow applications"""
    import os
    g = mesh.naca_wing(mesh.WingSpec(code="0012", n_around=24, n_span=8, aspect=6.0))
    dm = dofmap.build_dofmap(g)
    V = np.array([np.cos(np.radians(5.0)), 0.0, np.sin(np.radians(5.0))])
    w = wk.build_wake(g, dm, V, wk.WakeOptions(length=2.0))
    els = api.build_elements(g)
    A, b, _ = api.assemble(g, dm, els, api.source_strengths(els, V), V, wake=w)
    W = dofmap.wake_rows(dm, w.te)
    rng = np.random.default_rng(5)
    out = {}
    try:
        for pol in ("0", "1"):
            os.environ["VP_LU_INVERSE"] = pol
            F = solve.factorize_galerkin(dm, A, W=W, mode=mode)
            R = rng.standard_normal((F.n, 3)) if pol == "0" else out["R"]
            out["R"] = R
            out[pol] = [F.solve_bordered(R[:, 0]), F.solve_bordered(R), F.solve_bordered_t(R[:, 1])]
            print(f"VP_LU_INVERSE={pol} {mode}: inverse {F.inverse_info}", flush=True)
            assert F.inverse_info[0] == (pol == "1")
            F.close()
        for a, c in zip(out["1"], out["0"]):
            assert np.abs(a - c).max() < 1e-11 * np.abs(c).max()
        os.environ.pop("VP_LU_INVERSE")
        F = solve.factorize_galerkin(dm, A, W=W, mode=mode)
        k = 0
        while F.inverse_info[1] < 24:
            F.solve_bordered(out["R"][:, 1 + k % 2])
            k += 1
        assert not F.inverse_info[0] and k <= 24
        x = F.solve_bordered(out["R"][:, 0])  # the first solve after 24 applications forms it
        used, applies = F.inverse_info
        print(f"  default: inverse {used} after {applies} narrow applications", flush=True)
        assert used and 24 <= applies <= 40
        assert np.abs(x - out["0"][0]).max() < 1e-11 * np.abs(out["0"][0]).max()
        F.close()
    finally:
        os.environ.pop("VP_LU_INVERSE", None)
