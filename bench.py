#!/usr/bin/env python3
"""Benchmark of the Galerkin hot path on one MI355X: pair integrals per second of `assemble` on the abused
ellipsoid (semi-axes 2, 1, 0.5; uv mesh 71 x 142: 9,942 nodes, 19,880 triangles; clearance-bounded jitter 0.25 and
5 % sliver triangles; Vinf at 5 degrees), plain FP64.  A step is one whole assembly of A (9,942 x 9,942) and b with
the geometry resident in device memory.  Also reported: the mixed path (FP32 far field and split near pairs), the
swept wing's solve + adjoint in both precisions (the metric's second part, three repeats), the bordered
factorization + solve in both modes with their errors against the exact ellipsoid flow, the far-tier kernel's roofline
(algorithmic FLOP/s from HIP events on the assembly stream), and the FP64 CPU oracle timed on the host on a bounded
sample of the same assembly (rows of test triangles).

python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import copy
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "vortexpanel.jl_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

# algorithmic FP64 flops per far-tier evaluation (add/sub/mul = 1, FMA = 2, sqrt/div/log/atan2 = 1): the closed form
# (src/kernels/source_triangle.jl:75-89) 132, the octupole multipole (:100-119) 258, and the pair_block work per point
# (src/assembly/assemble.jl:248-255) 73
FLOPS_EXACT, FLOPS_MULTIPOLE, FLOPS_BLOCK = 132, 258, 73
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload():
    from vphip import dofmap, mesh
    g = mesh.abused_ellipsoid(nth=71, nph=142, axes=(2.0, 1.0, 0.5), jitter=0.25, sliver_frac=0.05, sliver_ratio=0.1,
                              seed=20261203)
    dm = dofmap.build_dofmap(g)
    a = np.radians(5.0)
    V = np.array([np.cos(a), 0.0, np.sin(a)])
    return g, dm, V


def cpu_threads():
    """Host threads of the CPU legs: the process's CPU share (OMP_NUM_THREADS where it is set, else the CPUs this
    process may run on), or VP_CPU_THREADS"""
    n = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return int(os.environ.get("VP_CPU_THREADS", n))


def cpu_baseline(g, dm, V, threads, target_s=12.0):
    """The FP64 oracle (C, OpenMP) on the host's cores over the same assembly: every test-triangle row when a pilot
    puts it within target_s, else a bounded sample of rows"""
    from oracle import oracle as O
    el = O.elements(g.x, g.tris)
    sig = O.source_strengths(el, V)
    pilot = min(g.nt, max(threads * 32, 512))
    t0 = time.perf_counter()
    O.assemble_rows(g, dm, sig, V, 0, pilot, nthreads=threads)
    tp = time.perf_counter() - t0
    rows = int(min(g.nt, max(pilot, pilot * target_s / max(tp, 1e-3))))
    t0 = time.perf_counter()
    npairs = O.assemble_rows(g, dm, sig, V, 0, rows, nthreads=threads)
    dt = time.perf_counter() - t0
    return dict(value=npairs / dt, unit="pair-integrals/s", cores=threads, kind="port",
                sample=f"FP64 C oracle assemble over test triangles 0..{rows - 1} of {g.nt} ({npairs:.3e} pair "
                       f"integrals, {dt:.1f} s)")


def cpu_lu_leg(dm, A, b, threads, gpu):
    """LU + solve of the ellipsoid's bordered system on the host: the oracle's solve_galerkin (LAPACK getrf, gecon,
    getrs, the routines the reference calls) on `threads` BLAS threads, beside the device's factorization + solve"""
    from threadpoolctl import threadpool_limits
    from oracle import oracle as O
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        O.solve_galerkin(dm, A, b)
        dt = time.perf_counter() - t0
    n = dm.nq + 1
    return dict(cpu_s=dt, gpu_fp64_s=gpu["fp64"], gpu_mixed_s=gpu["mixed"], cores=threads, kind="port",
                sample=f"bordered LU (n = {n}) + condition estimate + one solve, FP64 LAPACK")


def cpu_alpha_leg(threads):
    """The 64-angle batch on the host: the FP64 oracle's assembly of the wing with its body-fixed wake (timed once,
    counted twice: b at V = e_x and at e_z span every angle's right hand side), LAPACK LU of the bordered matrix and
    the 64 solves.  (Row samples of this assembly are not used: each call clears the whole 1 GB A, which skews them.)"""
    from threadpoolctl import threadpool_limits
    from oracle import oracle as O
    from vphip import dofmap, mesh, wake as wk
    spec = mesh.WingSpec(code="0012", n_around=160, n_span=60, aspect=8.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    w = wk.build_wake(g, dm, np.array([1.0, 0.0, 0.0]), wk.WakeOptions(length=2.5))
    el = O.elements(g.x, g.tris)
    ex = np.array([1.0, 0.0, 0.0])
    t0 = time.perf_counter()
    A, b, pairs, _ = O.assemble(g, dm, O.source_strengths(el, ex), ex, wake=w, nthreads=threads)
    ta = time.perf_counter() - t0
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        M, rhs, _ = O.bordered_system(dm, A, b, dofmap.wake_rows(dm, w.te))
        lu = O.sla.lu_factor(M)
        O.sla.lu_solve(lu, np.repeat(rhs[:, None], 64, axis=1))
        tl = time.perf_counter() - t0
    return dict(cpu_s=2 * ta + tl, cpu_assemble_s=ta, cpu_lu_solves_s=tl, cores=threads, kind="port",
                sample=f"FP64 oracle assemble of the {g.nt}-triangle wing with its wake ({sum(pairs.values()):.3e} "
                       f"pair integrals, {ta:.1f} s, counted twice), bordered LU + 64 solves ({tl:.1f} s)")


def cpu_free_wake_leg(fw, threads):
    """One free-wake step on the host at the particle counts of the device run's snapshots: the same host logic over
    the FP64 oracle (explicit_rhs, LAPACK solve with the factors kept, body and blob velocity sums), from the device
    run's state before that step"""
    import copy
    from oracle.freewake_oracle import OracleFreeWakeSim
    from vphip import freewake as FW
    OracleFreeWakeSim.nthreads = threads
    g, dm, V, opts, snaps = fw
    t0 = time.perf_counter()
    sim = OracleFreeWakeSim(g, dm, V, opts)
    tsetup = time.perf_counter() - t0
    out = []
    for snap in snaps:
        for k, v in snap["state"].items():
            setattr(sim, k, copy.deepcopy(v))
        t0 = time.perf_counter()
        r = sim.step()
        dt = time.perf_counter() - t0
        out.append(dict(step=snap["step"], particles=snap["particles"], cpu_s=dt, gpu_s=snap["gpu_s"],
                        cpu_t_rhs=r["t_rhs"], cpu_t_vel=r["t_vel"]))
        log(f"cpu free-wake step {out[-1]}")
    return dict(steps=out, cpu_setup_s=tsetup, cores=threads, kind="port",
                sample="one step each from the device run's state at about 10k, 40k and 84k particles (FP64 oracle; "
                       "the device run is the mixed path)")


def wing_adjoint_bench(repeats=3):
    """Swept tapered NACA 2412 wing (taper 0.5, 25 deg quarter-chord sweep, AR 10, 240 x 120 quads, rigid wake),
    alpha = 4 deg: forward solve plus the adjoint gradients dCL/dx, dCDi/dx (all node coordinates) and dCL/dalpha,
    plain FP64 against mixed (FP32 far and split near tiers forward and backward, FP32 LU on MFMA, FP64 refinement),
    alternated `repeats` times; the medians are reported, with every repeat's phases"""
    from vphip import adjoint as AD, dofmap, mesh, sweep, wake as wk
    spec = mesh.WingSpec(code="2412", n_around=240, n_span=120, aspect=10.0, taper=0.5, sweep_c4_deg=25.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    V, _ = sweep.wing_freestream(4.0)
    w = wk.build_wake(g, dm, V)
    S = spec.aspect * spec.mean_chord ** 2
    out = {"nt": g.nt, "nslot": dm.nslot, "wake_nodes": len(w.x), "repeats": repeats}
    grads, runs = {}, {"fp64": [], "mixed": []}
    for rep in range(repeats):
        for p in ("fp64", "mixed"):
            t0 = time.perf_counter()
            adj = AD.WingAdjoint(g, dm, w, 4.0, S, precision=p)
            t1 = time.perf_counter()
            r = adj.gradients(("CL", "CDi"))  # one backward pass for both
            (CL, gcl, dcl), (CD, gcd, _) = r["CL"], r["CDi"]
            t2 = time.perf_counter()
            grads[p] = (CL, CD, gcl, gcd, dcl)
            runs[p].append(dict(total_s=t2 - t0, forward_s=t1 - t0, adjoint_s=t2 - t1, CL=CL, CDi=CD,
                                dCL_dalpha_per_deg=dcl, pairs=adj.plan.pairs,
                                **{k: v for k, v in adj.timings.items() if k != "forward_s"}))
            adj.close()
            log(f"wing adjoint {p} repeat {rep}: {runs[p][-1]}")
    for p in runs:
        tot = [r_["total_s"] for r_ in runs[p]]
        out[p] = dict(runs[p][int(np.argsort(tot)[len(tot) // 2])], total_s_median=float(np.median(tot)),
                      total_s_all=tot)
    a, b = grads["fp64"], grads["mixed"]
    out["abs_dCL_mixed_vs_fp64"] = abs(a[0] - b[0])
    out["abs_dCDi_mixed_vs_fp64"] = abs(a[1] - b[1])
    out["max_rel_dCLdx_mixed_vs_fp64"] = float(np.abs(a[2] - b[2]).max() / np.abs(a[2]).max())
    out["speedup_mixed_vs_fp64"] = out["fp64"]["total_s_median"] / out["mixed"]["total_s_median"]
    return out


def alpha_sweep_bench():
    """Rectangular NACA 0012 wing, AR 8, 160 x 60 quads, body-fixed rigid wake to 20 chords, 64 angles from -4 to
    12 deg as one multi-RHS solve (mixed precision), CL by pressure and CDi by Trefftz per angle"""
    from vphip import dofmap, mesh, sweep, wake as wk
    spec = mesh.WingSpec(code="0012", n_around=160, n_span=60, aspect=8.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    w = wk.build_wake(g, dm, np.array([1.0, 0.0, 0.0]), wk.WakeOptions(length=2.5))
    t0 = time.perf_counter()
    sw = sweep.AlphaSweep(g, dm, w, spec.aspect * spec.mean_chord ** 2, precision="mixed")
    alphas = np.linspace(-4.0, 12.0, 64)
    r = sw.solve(alphas)
    t1 = time.perf_counter()
    out = dict(nt=g.nt, nslot=dm.nslot, n_alpha=64, total_s=t1 - t0, **sw.timings,
               CL_at_0_4_8_12=[float(np.interp(a, alphas, r["CL"])) for a in (0.0, 4.0, 8.0, 12.0)],
               CDi_at_4=float(np.interp(4.0, alphas, r["CDi_T"])),
               dCL_dalpha_per_deg=float(np.polyfit(alphas[16:40], r["CL"][16:40], 1)[0]))
    sw.close()
    log(f"alpha sweep: {out}")
    return out


def free_wake_bench(nsteps=1000):
    """Impulsively started rectangular NACA 0012 wing, AR 6, 120 x 40 quads, alpha 8 deg, dt = 0.02 chord/V∞, fixed
    particle core 0.1 x the mean panel edge, mixed precision (FP32 velocity kernels with FP64 sums and bookkeeping):
    the whole 1000-step run to ~84,000 particles, with the step time and particle count every 100 steps"""
    from vphip import dofmap, freewake as FW, mesh, sweep
    spec = mesh.WingSpec(code="0012", n_around=120, n_span=40, aspect=6.0)
    g = mesh.naca_wing(spec)
    dm = dofmap.build_dofmap(g)
    V, L = sweep.wing_freestream(8.0)
    edges = np.concatenate([np.linalg.norm(g.x[g.tris[:, (k + 1) % 3]] - g.x[g.tris[:, k]], axis=1) for k in range(3)])
    t0 = time.perf_counter()
    sim = FW.FreeWakeSim(g, dm, V, FW.FreeWakeOptions(dt=0.02, core=float(0.1 * edges.mean())), precision="mixed")
    t1 = time.perf_counter()
    steps, every100, snaps, marks = [], [], [], [10_000, 40_000, 80_000]
    for n in range(1, nsteps + 1):
        snap = None
        if marks and len(sim.state.particles) >= marks[0]:  # the state before this step, for the CPU leg
            marks.pop(0)
            snap = dict(step=n, particles=len(sim.state.particles),
                        state={k: copy.deepcopy(getattr(sim, k)) for k in ("state", "n", "t", "mu", "mu_w", "I_prev")})
        ts = time.perf_counter()
        r = sim.step()
        steps.append(time.perf_counter() - ts)
        if snap is not None:
            snap["gpu_s"] = steps[-1]
            snaps.append(snap)
        if n % 100 == 0:
            every100.append(dict(step=n, particles=r["particles"], mean_step_ms=1e3 * float(np.mean(steps[-100:])),
                                 CL_kj=float(r["F_kj"] @ L / spec.aspect)))
    sim.close()
    out = dict(nt=g.nt, nslot=dm.nslot, setup_s=t1 - t0, steps=nsteps, total_s=float(np.sum(steps)),
               mean_step_s=float(np.mean(steps)), last10_step_s=float(np.mean(steps[-10:])), particles=r["particles"],
               rows=r["rows"], CL_kj=float(r["F_kj"] @ L / spec.aspect),
               CL_pressure=float(r["F_pressure"] @ L / spec.aspect), t_rhs_last=r["t_rhs"], t_vel_last=r["t_vel"],
               every_100=every100)
    log(f"free wake: {out}")
    return out, (g, dm, V, FW.FreeWakeOptions(dt=0.02, core=float(0.1 * edges.mean())), snaps)


def uneven_bodies_bench(steps=5):
    """Bodies of uneven size in one geometry, one assembly: the abused ellipsoid (19,880 triangles), a sphere of
    5,120 triangles and one of 80, apart; FP64 pair integrals per second of the combined assembly against the
    ellipsoid's own, with the far tier as a grid of tiles and as one persistent work queue (VP_FAR_QUEUE)"""
    from vphip import api, dofmap, mesh
    g0, _, V = workload()
    parts = [g0, mesh.icosphere(4, r=0.8, center=(4.0, 0.5, 0.0)), mesh.icosphere(1, r=0.2, center=(-3.0, 0.0, 0.5))]
    out = dict(bodies=[p_.nt for p_ in parts])
    for name, geo in (("ellipsoid", g0), ("three_bodies", mesh.combine(parts))):
        dm = dofmap.build_dofmap(geo)
        els = api.build_elements(geo)
        sig = api.source_strengths(els, V)
        for form in ("grid", "queue"):
            os.environ["VP_FAR_QUEUE"] = "1" if form == "queue" else "0"
            plan = api.AssemblyPlan(geo, dm, None, api.AssemblyOptions(), timing=True)
            for _ in range(2):
                plan.assemble(sig, V)
            plan.ctx.sync()
            t0 = time.perf_counter()
            far = []
            for _ in range(steps):
                plan.assemble(sig, V)
                far.append(plan.timings()["far"])
            plan.ctx.sync()
            dt = (time.perf_counter() - t0) / steps
            npairs = sum(plan.pairs.values())
            out[f"{name}_{form}"] = dict(nt=geo.nt, pairs=npairs, assemble_ms=1e3 * dt, pair_integrals_per_s=npairs / dt,
                                         far_ms=float(np.median(far)), far_pairs_per_ms=plan.pairs["far"] / float(np.median(far)))
            plan.close()
    os.environ.pop("VP_FAR_QUEUE", None)
    for form in ("grid", "queue"):
        out[f"rate_three_bodies_over_ellipsoid_{form}"] = (out[f"three_bodies_{form}"]["pair_integrals_per_s"] /
                                                           out[f"ellipsoid_{form}"]["pair_integrals_per_s"])
    log(f"uneven bodies: {out}")
    return out


def mixed_wing():
    """The 64-angle batch's wing (NACA 0012, AR 8, 160 x 60 quads) from its raw quads through the geometry pipeline,
    its quads kept whole (isoparametric) and its tip caps flat triangles, with the body-fixed rigid wake"""
    from vphip import dofmap, geometry, mesh, mixed, wake as wk
    X, F = mesh.naca_wing_raw(mesh.WingSpec(code="0012", n_around=160, n_span=60, aspect=8.0))
    g = geometry.build_geometry(X, F)
    dm = dofmap.build_dofmap(g)
    w = wk.build_wake(g, dm, np.array([1.0, 0.0, 0.0]), wk.WakeOptions(length=2.5))
    return g, dm, w, mixed.mixed_mesh(g, dm)


def mixed_mesh_bench(repeats=3):
    """Mixed-mesh assembly (src/assembly/mixed.jl:444-598) of mixed_wing() at alpha 4 deg on the device, FP64, the
    median of `repeats`; beside it the same wing split into triangles through the triangle path"""
    from vphip import api, sweep
    g, dm, w, m = mixed_wing()
    V, _ = sweep.wing_freestream(4.0)
    ts, st = [], None
    for _ in range(repeats + 1):
        t0 = time.perf_counter()
        A, b, st = api.assemble_mixed(g, dm, m, V, wake=w)
        ts.append(time.perf_counter() - t0)
    ts = ts[1:]
    els = api.build_elements(g)
    sig = api.source_strengths(els, V)
    plan = api.AssemblyPlan(g, dm, w, api.AssemblyOptions())
    for _ in range(2):
        plan.assemble(sig, V)
    plan.ctx.sync()
    t0 = time.perf_counter()
    plan.assemble(sig, V)
    plan.ctx.sync()
    tt = time.perf_counter() - t0
    tri_pairs, tri_points = sum(plan.pairs.values()), sum(plan.points.values())
    plan.close()
    npairs, npts = sum(st["pairs"].values()), sum(st["points"].values())
    out = dict(quads=m.nq, triangles=m.ntri, nslot=dm.nslot, wake_nodes=len(w.x), assemble_s=float(np.median(ts)),
               assemble_s_all=ts, pairs=st["pairs"], points=st["points"], pair_integrals_per_s=npairs / np.median(ts),
               points_per_s=npts / np.median(ts), split_triangles_assemble_s=tt, split_triangle_pairs=tri_pairs,
               split_triangle_points=tri_points, quads_over_triangles_time=float(np.median(ts)) / tt)
    log(f"mixed mesh: {out}")
    return out


def cpu_mixed_leg(threads, gpu, target_s=12.0):
    """The mixed-mesh assembly on the host: the FP64 oracle (C, OpenMP) over a bounded sample of test elements of
    mixed_wing(), scaled to pair integrals per second"""
    from oracle import oracle as O
    from vphip import mixed, sweep
    g, dm, w, m = mixed_wing()
    V, _ = sweep.wing_freestream(4.0)
    touch = mixed.element_touch_pairs(g, m)
    pilot = max(threads * 4, 64)
    t0 = time.perf_counter()
    O.assemble_mixed(g.x, m, dm.nslot, V, touch=touch, wake=w, nthreads=threads, rows=(0, pilot))
    tp = time.perf_counter() - t0
    rows = int(min(m.ne, max(pilot, pilot * target_s / max(tp, 1e-3))))
    t0 = time.perf_counter()
    _, _, pairs, points = O.assemble_mixed(g.x, m, dm.nslot, V, touch=touch, wake=w, nthreads=threads, rows=(0, rows))
    dt = time.perf_counter() - t0
    npairs = sum(pairs.values())
    return dict(value=npairs / dt, unit="pair-integrals/s", cpu_points_per_s=sum(points.values()) / dt,
                gpu_pair_integrals_per_s=gpu["pair_integrals_per_s"], cores=threads, kind="port",
                sample=f"FP64 C oracle mixed assemble over test elements 0..{rows - 1} of {m.ne} ({npairs:.3e} pair "
                       f"integrals, {dt:.1f} s)")


def blob_tree_bench(sizes=(84_000, 500_000, 1_000_000), theta=0.5, degree=5):
    """The blob sums by the treecode against direct sums, in FP64 and in mixed precision (FP32 kernels): the benchmark's
    blob cloud (spread 3 x 3 x 0.3, strengths 1e-3, cores 0.005 to 0.02) summed at every blob, the wall time of each
    (host arrays in and out), the treecode's host build and device times, its kernel evaluations against the direct
    sum's n^2, and its largest error against the FP64 direct sum relative to the largest velocity"""
    from vphip import post
    rng = np.random.default_rng(0)
    post.blob_velocities_tree(rng.standard_normal((3000, 3)), rng.standard_normal((3000, 3)), np.full(3000, 0.01),
                              rng.standard_normal((3000, 3)), theta=theta, degree=degree)
    out = dict(theta=theta, degree=degree, runs=[])
    for n in sizes:
        bx = rng.standard_normal((n, 3)) * np.array([3.0, 3.0, 0.3])
        ba = rng.standard_normal((n, 3)) * 1e-3
        bd = rng.uniform(0.005, 0.02, n)
        t0 = time.perf_counter()
        ut, st = post.blob_velocities_tree(bx, ba, bd, bx, theta=theta, degree=degree)
        tt = time.perf_counter() - t0
        t0 = time.perf_counter()
        um, sm = post.blob_velocities_tree(bx, ba, bd, bx, theta=theta, degree=degree, precision="mixed")
        tm = time.perf_counter() - t0
        t0 = time.perf_counter()
        ud = post.blob_velocities(bx, ba, bd, bx)
        td = time.perf_counter() - t0
        t0 = time.perf_counter()
        post.blob_velocities(bx, ba, bd, bx, precision="mixed")
        tdm = time.perf_counter() - t0
        scale = np.abs(ud).max()
        out["runs"].append(dict(n=n, direct_s=td, tree_s=tt, speedup=td / tt, build_ms=st["build_ms"],
                                device_ms=st["device_ms"], evals_vs_direct=(st["proxy_evals"] + st["near_evals"]) / n ** 2,
                                far_clusters=st["far_clusters"], error=float(np.abs(ut - ud).max() / scale),
                                direct_mixed_s=tdm, tree_mixed_s=tm, speedup_mixed=tdm / tm,
                                device_mixed_ms=sm["device_ms"], error_mixed=float(np.abs(um - ud).max() / scale)))
        log(f"blob tree n={n}: {out['runs'][-1]}")
    return out


def nbody_bench(g, dm, V, mu):
    """surface_velocity on the ellipsoid solution and blob sums at the free wake's end size (84,000 particles)"""
    from vphip import api, post
    out = {}
    sig = api.source_strengths(api.build_elements(g), V)
    for p in ("fp64", "mixed"):
        post.surface_velocity(g, dm, sig, mu, V, precision=p)
        t0 = time.perf_counter()
        post.surface_velocity(g, dm, sig, mu, V, precision=p)
        dt = time.perf_counter() - t0
        out[f"surface_velocity_{p}_s"] = dt
        out[f"surface_velocity_{p}_evals_per_s"] = g.nt * g.nt / dt
    rng = np.random.default_rng(0)
    nb = 84000
    bx = rng.standard_normal((nb, 3)) * np.array([3.0, 3.0, 0.3])
    ba = rng.standard_normal((nb, 3)) * 1e-3
    bd = np.full(nb, 0.01)
    for p in ("fp64", "mixed"):
        post.blob_velocities(bx[:1000], ba[:1000], bd[:1000], bx[:1000], precision=p)
        t0 = time.perf_counter()
        post.blob_velocities(bx, ba, bd, bx, precision=p)
        dt = time.perf_counter() - t0
        out[f"blob_{p}_s"] = dt
        out[f"blob_{p}_interactions_per_s"] = nb * nb / dt
        out[f"blob_{p}_tflops"] = nb * nb * 28 / dt / 1e12
    log(f"nbody: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)  # the launch autotuner times its candidates on the first two
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the assembly metric (no wing, sweep, N-body)")
    ap.add_argument("--extras", default="nbody,alpha_sweep,wing_adjoint,free_wake,mixed_mesh,uneven_bodies,blob_tree",
                    help="which of the extra workloads to run (comma-separated)")
    ap.add_argument("--no-solve", action="store_true", help="no factorizations or solves (with --no-extras: the "
                    "assemblies alone, the program the far kernel's counter passes profile)")
    ap.add_argument("--copy-calibration", action="store_true", help="only copy 1 GiB device to device three times "
                    "(the known-size reference for the HBM byte counters); loads no library")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 3 << 30}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X, and the path runs on one device "
              "(how it would shard is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from vphip import api, dofmap, solve
    from vphip.exact import EllipsoidFlow

    g, dm, V = workload()
    log(f"workload: nt={g.nt} nn={g.nn} nslot={dm.nslot} mesh sha256 {g.digest()[:16]}")
    ctx = api.default_context()
    els = api.build_elements(g)
    sig = api.source_strengths(els, V)
    plans = {m: api.AssemblyPlan(g, dm, None, api.AssemblyOptions(precision=m), ctx=ctx, timing=True)
             for m in ("fp64", "mixed")}
    for m, p in plans.items():
        for _ in range(max(args.warmup, 1)):
            p.assemble(sig, V)
    ctx.sync()
    # timed region: K plain-FP64 assemblies
    p = plans["fp64"]
    far_ms = []
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p.assemble(sig, V)
        far_ms.append(p.timings()["far"])
    ctx.sync()
    dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    pairs = sum(p.pairs.values())
    tm = p.timings()
    value = pairs / (dt / args.steps)
    log(f"fp64 assemble {ms_step:.2f} ms/step, {value:.3e} pair-integrals/s, kernels {tm}")
    # roofline of the dominant kernel (the far tier)
    far_pairs_n = p.pairs["far"]
    ex = tm["far_exact_points"]
    flops = ex * (FLOPS_EXACT + FLOPS_BLOCK) + (3 * far_pairs_n - ex) * (FLOPS_MULTIPOLE + FLOPS_BLOCK)
    far_avg = float(np.mean(far_ms)) / 1e3
    achieved = flops / far_avg
    traffic = None  # HBM bytes per far-kernel launch, measured by the counter passes of the profile summary
    tpath = os.path.join(ROOT, "profiles", "far_pairs_traffic.json")
    if os.path.exists(tpath):
        traffic = float(json.load(open(tpath))["traffic_bytes_per_launch"])
    # mixed path
    pm = plans["mixed"]
    tms = []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        pm.assemble(sig, V)
        ctx.sync()
        tms.append(time.perf_counter() - t1)
    mixed_ms = 1e3 * float(np.median(tms))
    log(f"mixed assemble {mixed_ms:.2f} ms/step, kernels {pm.timings()}")
    A64, b64 = p.read()
    Amx, bmx = pm.read()
    entry_err = float(np.max(np.abs(Amx - A64)) / np.max(np.abs(A64)))
    # bordered factorization + solve in both modes, errors against the exact flow
    if args.no_solve:
        print(json.dumps({"metric": "pair-integrals/s", "value": value, "unit": "pair-integrals/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
                          "higher_is_better": True, "note": "assemblies only (--no-solve)"}))
        return
    ef = EllipsoidFlow((2.0, 1.0, 0.5), V)
    mu_ex = ef.doublet(g.x[dm.slot_node])
    scale = ef.doublet_scale()
    solves = {}
    for m, plan in (("fp64", p), ("mixed", pm)):
        ctx.sync()
        t1 = time.perf_counter()
        F = solve.factorize_galerkin(dm, plan, mode=m)
        F.cond  # the condition estimate, which the reference's factorization computes (gecon, solve.jl:85)
        ctx.sync()
        tf = time.perf_counter() - t1
        t1 = time.perf_counter()
        out = solve.solve_galerkin(F, b64 if m == "fp64" else bmx)
        ts = time.perf_counter() - t1
        err = out["mu"] - mu_ex
        err -= np.sum(dm.slot_area * err) / np.sum(dm.slot_area)
        solves[m] = dict(factor_s=tf, solve_s=ts, refinement_iters=out["iters"],
                         mu_L2=float(np.sqrt(np.sum(dm.slot_area * err ** 2) / np.sum(dm.slot_area)) / scale),
                         mu_max=float(np.max(np.abs(err)) / scale), mu=out["mu"])
        F.close()
    dmu = float(np.max(np.abs(solves["mixed"]["mu"] - solves["fp64"]["mu"])) / np.max(np.abs(solves["fp64"]["mu"])))
    extras = {}
    want = set() if args.no_extras else set(args.extras.split(","))
    if "nbody" in want:
        extras["nbody"] = nbody_bench(g, dm, V, solves["fp64"]["mu"])
    for m in solves:
        solves[m].pop("mu")
        log(f"{m} solve {solves[m]}")
    for p_ in plans.values():
        p_.close()
    fw = None
    for name, fn in (("alpha_sweep", alpha_sweep_bench), ("wing_adjoint", wing_adjoint_bench),
                     ("free_wake", free_wake_bench), ("mixed_mesh", mixed_mesh_bench),
                     ("uneven_bodies", uneven_bodies_bench), ("blob_tree", blob_tree_bench)):
        if name in want:
            extras[name] = fn()
            if name == "free_wake":
                extras[name], fw = extras[name]
    # CPU legs: the FP64 oracle on the host's CPU share, beside the device figures above
    cpu, legs = None, None
    if not args.no_cpu:
        threads = cpu_threads()
        cpu = cpu_baseline(g, dm, V, threads)
        log(f"cpu baseline {cpu}")
        legs = {"lu_solve": cpu_lu_leg(dm, A64, b64, threads, {m: solves[m]["factor_s"] + solves[m]["solve_s"]
                                                               for m in solves})}
        log(f"cpu lu leg {legs['lu_solve']}")
        if "alpha_sweep" in extras:
            legs["alpha_batch"] = dict(cpu_alpha_leg(threads), gpu_s=extras["alpha_sweep"]["total_s"])
            log(f"cpu alpha leg {legs['alpha_batch']}")
        if fw is not None:
            legs["free_wake_step"] = cpu_free_wake_leg(fw, threads)
        if "mixed_mesh" in extras:
            legs["mixed_mesh"] = cpu_mixed_leg(threads, extras["mixed_mesh"])
            log(f"cpu mixed leg {legs['mixed_mesh']}")
        legs["adjoint"] = "no CPU leg: the reference computes no derivatives (its only route, finite differences, " \
                          "takes two solves per node coordinate: 174,240 for the swept wing)"
    line = {
        "metric": "pair-integrals/s",
        "value": value,
        "unit": "pair-integrals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded abused ellipsoid mesh generated in-process)",
        "config": {"workload": "assemble, abused ellipsoid 2:1:0.5 uv 71x142 (9,942 nodes, 19,880 triangles, jitter "
                               "0.25, 5% slivers), Vinf 5 deg, default TierRules, plain FP64",
                   "nt": g.nt, "nslot": dm.nslot, "pairs_per_step": pairs, "tier_pairs": p.pairs,
                   "tier_points": p.points},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "bound_note": "FP64 vector arithmetic (the kernel issues no MFMA); MI355X's FP64 vector peak, "
                                   "78.6 TFLOP/s, equals its FP64 matrix peak",
                     "frac": achieved / PEAK_FP64, "traffic": traffic,
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE passes, copy-calibrated, "
                                       "profiles/far_pairs_abused_ellipsoid_pmc.md" if traffic else None,
                     "kernel": "far_pairs_queue<double>", "kernel_ms": far_avg * 1e3,
                     "algorithmic_flops": flops, "far_exact_points": ex, "far_points": 3 * far_pairs_n},
        "cpu_baseline": cpu,
        "cpu_legs": legs,
        "mixed": {"assemble_ms": mixed_ms, "pair_integrals_per_s": pairs / (mixed_ms / 1e3),
                  "max_entry_rel_diff_vs_fp64": entry_err, "max_mu_rel_diff_vs_fp64": dmu},
        "solve": solves,
        "gemm_tuning": json.loads(ctx.lib.vp_gemm_tuning().decode()),
        "assembly_tuning": json.loads(ctx.lib.vp_assembly_tuning().decode()),
        "kernels_ms": tm,
        **extras,
    }
    wa = extras.get("wing_adjoint")
    if wa:  # the metric's second part: the swept wing's solve + adjoint, mixed against plain FP64 (target >= 3x)
        line["wing_solve_adjoint"] = {
            "workload": "swept tapered NACA 2412 wing, 58,076 triangles, rigid wake, alpha 4 deg: assemble, factor, "
                        "solve, CL and CDi with dCL/dx, dCDi/dx, dCL/dalpha (adjoint)",
            "fp64_s": wa["fp64"]["total_s_median"], "mixed_s": wa["mixed"]["total_s_median"],
            "speedup_mixed_vs_fp64": wa["speedup_mixed_vs_fp64"], "target_speedup": 3.0,
            "abs_dCL_mixed_vs_fp64": wa["abs_dCL_mixed_vs_fp64"], "dCL_bound": 1e-6,
            "repeats": wa["repeats"]}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
