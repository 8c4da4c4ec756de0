"""The α sweep's batched loads (all angles in one pass over the elements) against the per-angle wing_loads
(validation/common/wing_case.jl:53-67): pressure CL, CDi and CY, Kutta-Joukowski CL and Trefftz CDi."""
import numpy as np

from vphip import api, dofmap, mesh, sweep


def test_batched_loads_match_per_angle():
    spec = mesh.WingSpec(code="0012", n_around=40, n_span=10, aspect=8.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    els = api.build_elements(g)
    rng = np.random.default_rng(0)
    al = np.linspace(-4.0, 12.0, 16)
    MU = rng.standard_normal((dm.nslot, len(al)))
    Q = rng.standard_normal((dm.nq, len(al)))
    b = sweep.wing_loads_batch(els, dm, MU, Q, al, 8.0)
    ref = [sweep.wing_loads(els, dm, MU[:, k], Q[:, k], *sweep.wing_freestream(a), 8.0) for k, a in enumerate(al)]
    for key in ("CL", "CL_KJ", "CDi_T", "CDi_p", "CY"):
        r = np.array([x[key] for x in ref])
        assert np.abs(b[key] - r).max() <= 1e-13 * np.abs(r).max(), key
