"""Two bodies of uneven size in one geometry (SurfaceGeometry's tri_body, src/geometry/types.jl:174-199; bordered_rows
per body, src/dofmap/system.jl:41-50): the DOFMap's per-body unknowns, one zero-mean border row per body
(src/assembly/solve.jl:30-35), and, far apart, each sphere's doublets against its own exact flow."""
import numpy as np

from oracle import oracle as O
from vphip import dofmap, mesh

V = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])


def two_spheres(d):
    a = mesh.icosphere(2)
    b = mesh.icosphere(2, r=0.5, center=(d, 0.0, 0.0))
    return mesh.combine([a, b]), (np.zeros(3), np.array([d, 0.0, 0.0]))


def test_dofmap_two_bodies():
    g, _ = two_spheres(3.0)
    dm = dofmap.build_dofmap(g)
    assert g.bodies_closed == [True, True] and g.nt == 640
    assert [len(q) for q in dm.body_q] == [162, 162]
    assert np.all(dm.slot_body[:162] == 0) and np.all(dm.slot_body[162:] == 1)
    assert np.abs(np.asarray(dm.C.sum(1)).ravel() - 1).max() == 0
    e, c = zip(*[(r[1], r[2]) for r in dofmap.bordered_rows(dm)])
    assert len(e) == 2 and np.all(np.asarray(e).sum(0) == 1)   # every unknown on exactly one body


def test_two_bodies_solve():
    """Each body's doublets have zero area-weighted mean and a vanishing border multiplier; 40 radii apart, each
    sphere's doublets match V.(x - c)/2 about its own centre to the accuracy of the mesh alone."""
    for d, tol in ((3.0, None), (40.0, 1.2)):
        g, centres = two_spheres(d)
        dm = dofmap.build_dofmap(g)
        el = O.elements(g.x, g.tris)
        A, b, pairs, points = O.assemble(g, dm, O.source_strengths(el, V), V, O.opts())
        assert sum(pairs.values()) == g.nt ** 2 and pairs["self"] == g.nt
        sol = O.solve_galerkin(dm, A, b)
        mu = sol["mu"]
        assert len(sol["lam"]) == 2 and np.abs(sol["lam"]).max() < 1e-10
        assert sol["residual"] < 1e-12
        for k in range(2):
            on = dm.slot_body == k
            assert abs(np.sum(dm.slot_area[on] * mu[on])) < 1e-12 * np.sum(dm.slot_area[on]) * np.abs(mu).max()
        if tol is None:
            continue
        single = mesh.icosphere(2)
        dm1 = dofmap.build_dofmap(single)
        e1 = O.elements(single.x, single.tris)
        A1, b1, _, _ = O.assemble(single, dm1, O.source_strengths(e1, V), V, O.opts())
        mu1 = O.solve_galerkin(dm1, A1, b1)["mu"]
        err1 = np.abs(mu1 - single.x[dm1.slot_node] @ V / 2).max()
        for k, (c, r) in enumerate(zip(centres, (1.0, 0.5))):
            on = dm.slot_body == k
            ex = (g.x[dm.slot_node[on]] - c) @ V / 2
            assert np.abs(mu[on] - ex).max() < tol * err1 * r, (k, np.abs(mu[on] - ex).max(), err1)
