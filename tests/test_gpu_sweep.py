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
    # the batch's right hand sides are the oracle's assembled ones (b is linear in V∞); the mixed path's FP32
    # multipole far field moves b by about 1e-10 of its largest entry
    Bn = sw.rhs(ALPHAS)
    tb = 1e-12 if precision == "fp64" else 1e-8
    for k in (0, 5):
        assert np.abs(Bn[:, k] - ref[k][2]).max() <= tb * np.abs(ref[k][2]).max()
    sw.close()
