"""Angle-of-attack batch on a lifting wing with a body-fixed rigid wake (one matrix, one factorization, one
multi-RHS solve) against the FP64 oracle solved angle by angle (assemble + LAPACK + loads,
validation/common/wing_case.jl:34-67).  FP64: |dCL|, |dCDi| <= 1e-10; mixed (FP32 far field and LU, FP64
refinement): |dCL| <= 1e-6 and |dCDi| <= 1e-7."""
import numpy as np
import pytest

from oracle import oracle as O
from vphip import dofmap, mesh, sweep, wake as wk

pytestmark = pytest.mark.gpu

ALPHAS = np.linspace(-4.0, 12.0, 8)


@pytest.fixture(scope="module")
def wing():
    spec = mesh.WingSpec(code="0012", n_around=32, n_span=10, aspect=8.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    w = wk.build_wake(g, dm, np.array([1.0, 0.0, 0.0]), wk.WakeOptions(length=2.5))
    S = spec.aspect * spec.mean_chord ** 2
    el = O.elements(g.x, g.tris)
    W = dofmap.wake_rows(dm, w.te)
    ref = []
    for a in ALPHAS:
        V, eL = sweep.wing_freestream(a)
        A, b, _, _ = O.assemble(g, dm, O.source_strengths(el, V), V, O.opts(), wake=w)
        sol = O.solve_galerkin(dm, A, b, W)
        cp = O.pressure_coefficient(O.gradient_velocity(el, dm, sol["mu"], V), V)
        Fp, _ = O.pressure_force(el, cp)
        mu_w = dm.K @ sol["q"]
        ref.append((Fp @ eL / S, O.trefftz_drag(dm.wake_te, mu_w, V) / S, b))
    return g, dm, w, S, ref


@pytest.mark.parametrize("precision,tl,td", [("fp64", 1e-10, 1e-10), ("mixed", 1e-6, 1e-7)])
def test_alpha_batch_loads(wing, precision, tl, td):
    print(f"alpha sweep {precision}", flush=True)
    g, dm, w, S, ref = wing
    sw = sweep.AlphaSweep(g, dm, w, S, precision=precision)
    out = sw.solve(ALPHAS)
    CL = np.array([r[0] for r in ref])
    CD = np.array([r[1] for r in ref])
    assert np.abs(out["CL"] - CL).max() <= tl, np.abs(out["CL"] - CL).max()
    assert np.abs(out["CDi_T"] - CD).max() <= td, np.abs(out["CDi_T"] - CD).max()
    # lift grows with the angle and the two force routes agree to discretisation accuracy
    assert np.all(np.diff(out["CL"]) > 0) and np.abs(out["CL"] - out["CL_KJ"]).max() < 0.05 * np.abs(CL).max()
    # the batch's right hand sides are the oracle's assembled ones (b is linear in V∞); the mixed path's FP32 far
    # field and split near pairs move b by about 1e-7 of its largest entry
    Bn = sw.rhs(ALPHAS)
    tb = 1e-12 if precision == "fp64" else 1e-6
    for k in (0, 5):
        assert np.abs(Bn[:, k] - ref[k][2]).max() <= tb * np.abs(ref[k][2]).max()
    sw.close()


def test_alpha_batch_full_size():
    """The benchmark's batch at full size (rectangular NACA 0012 wing, AR 8, 160 x 60 quads: 19,516 triangles, wake to
    20 chords; 64 angles from -4 to 12 deg plus 0 and +-4): mixed within 1e-6 (CL) and 1e-7 (CDi) of FP64 at every
    angle; the symmetric section gives |CL(0)| and |CL(4) + CL(-4)| below 1e-4; the lift slope lies in the lifting-line
    range for AR 8 (0.080-0.092 per degree) and the Trefftz drag's span efficiency CL²/(π AR CDi) in (0.9, 1.05)"""
    spec = mesh.WingSpec(code="0012", n_around=160, n_span=60, aspect=8.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    w = wk.build_wake(g, dm, np.array([1.0, 0.0, 0.0]), wk.WakeOptions(length=2.5))
    print(f"case alpha batch full size nt={g.nt}", flush=True)
    alphas = np.concatenate([np.linspace(-4.0, 12.0, 64), [0.0, 4.0]])
    out = {}
    for p in ("fp64", "mixed"):
        sw = sweep.AlphaSweep(g, dm, w, spec.aspect * spec.mean_chord ** 2, precision=p)
        out[p] = sw.solve(alphas)
        sw.close()
    a, m = out["fp64"], out["mixed"]
    dcl, dcd = np.abs(a["CL"] - m["CL"]).max(), np.abs(a["CDi_T"] - m["CDi_T"]).max()
    print(f"  mixed vs fp64: |dCL| {dcl:.2e} |dCDi| {dcd:.2e}", flush=True)
    assert dcl < 1e-6 and dcd < 1e-7, (dcl, dcd)
    cl0, cl4, clm4 = a["CL"][64], a["CL"][65], a["CL"][0]
    slope = np.polyfit(alphas[16:40], a["CL"][16:40], 1)[0]
    e = cl4 ** 2 / (np.pi * spec.aspect * a["CDi_T"][65])
    print(f"  CL(0) {cl0:.2e} CL(4)+CL(-4) {cl4 + clm4:.2e} slope {slope:.5f}/deg e {e:.4f}", flush=True)
    assert abs(cl0) < 1e-4 and abs(cl4 + clm4) < 1e-4
    assert 0.080 < slope < 0.092
    assert 0.9 < e < 1.05
