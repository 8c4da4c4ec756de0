"""Oracle assembly + LAPACK solve on the icosphere against the exact sphere flow and the reference's published
accuracy (README.md:210-217; test/assembly/test_sphere.jl:34-55)."""
import numpy as np
import pytest

from oracle import oracle as O
from vphip import dofmap, mesh

VG = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])


def solve_sphere(level, V=VG, **kw):
    g = mesh.icosphere(level)
    dm = dofmap.build_dofmap(g)
    el = O.elements(g.x, g.tris)
    sig = O.source_strengths(el, V)
    A, b, pairs, points = O.assemble(g, dm, sig, V, O.opts(**kw))
    sol = O.solve_galerkin(dm, A, b)
    mu_ex = g.x[dm.slot_node] @ V / 2
    err = sol["mu"] - mu_ex
    err -= np.sum(dm.slot_area * err) / np.sum(dm.slot_area)
    u = O.gradient_velocity(el, dm, sol["mu"], V)
    cp = O.pressure_coefficient(u, V)
    cpex = 1 - 9 / 4 * (1 - (el["centroid"] / np.linalg.norm(el["centroid"], axis=1)[:, None] @ V) ** 2)
    area = el["area"]
    return dict(g=g, dm=dm, A=A, b=b, sol=sol, pairs=pairs, points=points,
                muL2=np.sqrt(np.sum(dm.slot_area * err ** 2) / np.sum(dm.slot_area)) / 0.5,
                mumax=np.max(np.abs(err)) / 0.5, cpgL2=np.sqrt(np.sum(area * (cp - cpex) ** 2) / np.sum(area)))


@pytest.fixture(scope="module")
def S():
    return {lev: solve_sphere(lev) for lev in (1, 2, 3)}


def test_published_icosphere3_accuracy(S):
    """README.md:210: mu L2 5.8e-4, mu max 1.9e-3, cp L2 (grad route) 9.8e-3, lambda ~1e-19; README.md:217 cond 640-650"""
    r = S[3]
    assert abs(r["muL2"] - 5.8e-4) < 0.05e-4 and abs(r["mumax"] - 1.9e-3) < 0.05e-3
    assert abs(r["cpgL2"] - 9.8e-3) < 0.3e-3
    assert 640 < r["sol"]["cond"] < 650 and abs(r["sol"]["lam"][0]) < 1e-17
    assert abs(O.consistency(r["dm"], r["A"], r["b"])["columns"] - 3.5e-6) < 0.1e-6


def test_sphere_orders_and_diagnostics(S):
    r = S[2]
    assert r["muL2"] < 3e-3 and r["mumax"] < 1e-2
    assert S[1]["muL2"] / S[2]["muL2"] > 3 and S[2]["muL2"] / S[3]["muL2"] > 3.5
    assert S[2]["cpgL2"] / S[3]["cpgL2"] > 1.7
    c = O.consistency(r["dm"], r["A"], r["b"])
    assert c["rows"] < 1e-12 and c["columns"] < 1e-5 and c["rhs"] < 1e-12
    assert abs(r["sol"]["lam"][0]) < 1e-14 and r["sol"]["residual"] < 1e-12 and r["sol"]["cond"] < 1e3
    nt = r["g"].nt
    assert r["pairs"]["self"] == nt and sum(r["pairs"].values()) == nt ** 2


def test_incompatible_rhs_lambda(S):
    r = S[2]
    sh = O.solve_galerkin(r["dm"], r["A"], r["b"] + 1e-3)
    assert abs(sh["lam"][0] - 1e-3) < 1e-9


def test_quadrature_convergence():
    """test/assembly/test_sphere.jl:57-59: finer rules move mu by far less than the discretization error"""
    a = solve_sphere(2)
    f = solve_sphere(2, far_ratio=8.0, near_ratio=4.0, near_depth=5, touch_levels=10, touch_order=4, kernel_far=1e300)
    assert np.max(np.abs(a["sol"]["mu"] - f["sol"]["mu"])) / np.max(np.abs(f["sol"]["mu"])) < 5e-5
