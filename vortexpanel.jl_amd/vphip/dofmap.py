"""DOFMap: slots, constraint matrix C, Kutta rows K, border rows -- the indexing contract of the hot path.

Follows src/dofmap/build.jl:19-48 (assign_slots), :112-127 (tip_constraints!), :152-217 (resolve, build_dofmap),
src/dofmap/kutta.jl:41-109 (wake_trailing_edge for matching Kutta chains) and src/dofmap/system.jl:17-50
(wake_rows, reduce_system, bordered_rows).  Indices are 0-based.  T-junction, gap and zip constraints and
non-matching trailing edges come from the reference's geometry pipeline, which is outside this library; the
surfaces this library generates have none, and `build_dofmap` rejects inputs that would need them.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from .mesh import NO_TE, TE_CAP, TE_LOWER, TE_UPPER, Surface


@dataclass
class WakeTE:
    """src/dofmap/types.jl:31-38"""
    node: np.ndarray
    x: np.ndarray
    tangent: np.ndarray
    bisector: np.ndarray
    lines: list
    closed: list

    @property
    def n(self):
        return len(self.node)


@dataclass
class DOFMapping:
    """src/dofmap/types.jl:73-93 (0-based; constraint kinds 'free' or 'tip')"""
    nslot: int
    nq: int
    slot_node: np.ndarray
    slot_side: np.ndarray
    lower_slot: np.ndarray      # -1 where the node has no lower copy
    corner_slot: np.ndarray     # (nt, 3) int32
    constraint: list
    free_slot: np.ndarray
    q_of_slot: np.ndarray       # -1 for constrained slots
    C: sp.csr_matrix            # nslot x nq
    wake_te: WakeTE
    K: sp.csr_matrix            # nte x nq
    slot_area: np.ndarray
    slot_body: np.ndarray
    body_q: list
    body_closed: list
    log: list = field(default_factory=list)


def _safe_normalize(v):
    n = np.linalg.norm(v)
    return v / n if n > 0 else np.zeros_like(v)


def _tri_normal(x, tri):
    return _safe_normalize(np.cross(x[tri[1]] - x[tri[0]], x[tri[2]] - x[tri[0]]))


def build_dofmap(g: Surface, cap_te_copy: str = "upper", tip_kutta: str = "zero") -> DOFMapping:
    """build_dofmap, src/dofmap/build.jl:174-217"""
    if cap_te_copy not in ("upper", "lower"):
        raise ValueError(f"cap_te_copy must be upper or lower, not {cap_te_copy}")
    if tip_kutta not in ("free", "zero"):
        raise ValueError(f"tip_kutta must be free or zero, not {tip_kutta}")
    nn, nt = g.nn, g.nt
    log = []
    # assign_slots, build.jl:19-48
    has_up = np.zeros(nn, bool)
    has_low = np.zeros(nn, bool)
    for t in range(nt):
        for k in range(3):
            if g.corner_side[t, k] == TE_UPPER:
                has_up[g.tris[t, k]] = True
            elif g.corner_side[t, k] == TE_LOWER:
                has_low[g.tris[t, k]] = True
    slot_node = list(range(nn))
    slot_side = [TE_UPPER if has_up[v] else TE_LOWER if has_low[v] else NO_TE for v in range(nn)]
    lower_slot = np.full(nn, -1, np.int64)
    for v in range(nn):
        if has_up[v] and has_low[v]:
            slot_node.append(v)
            slot_side.append(TE_LOWER)
            lower_slot[v] = len(slot_node) - 1
    corner_slot = g.tris.astype(np.int64).copy()
    for t in range(nt):
        for k in range(3):
            v = g.tris[t, k]
            if lower_slot[v] < 0:
                continue
            s = g.corner_side[t, k]
            if s == TE_LOWER or (s == TE_CAP and cap_te_copy == "lower"):
                corner_slot[t, k] = lower_slot[v]
            elif s == NO_TE:
                log.append(f"corner {k} of triangle {t} at trailing edge node {v} has no side label; it takes the upper copy")
    nslot = len(slot_node)

    def cslot(t, v):
        k = int(np.nonzero(g.tris[t] == v)[0][0])
        return int(corner_slot[t, k])

    # wake_trailing_edge, kutta.jl:41-109 (matching chains)
    nodes, index, up_terms, low_terms, wedge = [], {}, [], [], []
    lines, closed = [], []
    for ch in g.chains:
        vs = list(ch.nodes[:-1]) if ch.closed and ch.nodes[-1] == ch.nodes[0] else list(ch.nodes)
        nseg = len(ch.tri_upper)
        line = []
        for i, v in enumerate(vs):
            if ch.closed:
                segs = sorted({(i - 1) % nseg, i % nseg}, key=lambda s: [(i - 1) % nseg, i % nseg].index(s))
            else:
                segs = [k for k in (i - 1, i) if 0 <= k < nseg]
            k0 = segs[0]
            bsum = sum(_tri_normal(g.x, g.tris[ch.tri_upper[s]]) + _tri_normal(g.x, g.tris[ch.tri_lower[s]]) for s in segs)
            v = int(v)
            if v in index:
                kk = index[v]
                wedge[kk] = wedge[kk] + bsum
            else:
                nodes.append(v)
                up_terms.append([(cslot(ch.tri_upper[k0], v), 1.0)])
                low_terms.append([(cslot(ch.tri_lower[k0], v), 1.0)])
                wedge.append(bsum)
                kk = index[v] = len(nodes) - 1
            line.append(kk)
        lines.append(line)
        closed.append(bool(ch.closed))
    nte = len(nodes)
    xte = g.x[nodes] if nte else np.zeros((0, 3))
    tangent = np.zeros((nte, 3))
    for line, c in zip(lines, closed):
        m = len(line)
        for i, k in enumerate(line):
            prev = line[i - 1] if i > 0 else (line[-1] if c else k)
            nxt = line[i + 1] if i < m - 1 else (line[0] if c else k)
            tangent[k] = _safe_normalize(g.x[nodes[nxt]] - g.x[nodes[prev]])
    bisector = np.array([_safe_normalize(b) for b in wedge]) if nte else np.zeros((0, 3))
    wake_te = WakeTE(node=np.array(nodes, np.int64), x=xte, tangent=tangent, bisector=bisector, lines=lines, closed=closed)

    # constraints: tip_constraints!, build.jl:112-127 (the generated surfaces carry no T-junctions, gaps or zips)
    cons, kind = {}, {}
    if tip_kutta == "zero":
        for line, c in zip(lines, closed):
            if c:
                continue
            for k in dict.fromkeys((line[0], line[-1])):
                v = nodes[k]
                if lower_slot[v] < 0:
                    log.append(f"tip trailing edge node {v} has no lower copy; its mu_w is left free")
                    continue
                if lower_slot[v] in cons:
                    continue
                cons[int(lower_slot[v])] = [(v, 1.0)]
                kind[int(lower_slot[v])] = "tip"
    # resolve_constraints, build.jl:152-167 (one level: tip constraints point at free slots)
    constraint = [kind.get(s, "free") for s in range(nslot)]
    free_slot = np.array([s for s in range(nslot) if constraint[s] == "free"], np.int64)
    q_of_slot = np.full(nslot, -1, np.int64)
    q_of_slot[free_slot] = np.arange(len(free_slot))
    nq = len(free_slot)
    I, J, V = [], [], []
    for s in range(nslot):
        terms = cons.get(s, [(s, 1.0)])
        for f, w in terms:
            if f in cons:
                raise ValueError("chained constraints are not produced by the generated surfaces")
            I.append(s)
            J.append(q_of_slot[f])
            V.append(w)
    C = sp.csr_matrix((V, (I, J)), shape=(nslot, nq))

    def selection(terms):
        r, c, w = [], [], []
        for k, ts in enumerate(terms):
            for s, wt in ts:
                r.append(k)
                c.append(s)
                w.append(wt)
        return sp.csr_matrix((w, (r, c)), shape=(len(terms), nslot))

    U, L = selection(up_terms), selection(low_terms)
    K = ((U - L) @ C).tocsr()
    K.eliminate_zeros()

    # slot areas and bodies, build.jl:199-212
    slot_area = np.zeros(nslot)
    slot_body = np.zeros(nslot, np.int64)
    for t in range(nt):
        tri = g.tris[t]
        a = np.linalg.norm(np.cross(g.x[tri[1]] - g.x[tri[0]], g.x[tri[2]] - g.x[tri[0]])) / 6
        for s in corner_slot[t]:
            slot_area[s] += a
            slot_body[s] = g.tri_body[t]
    nbody = len(g.bodies_closed)
    body_q = [[] for _ in range(nbody)]
    for q, s in enumerate(free_slot):
        body_q[slot_body[s]].append(q)
    return DOFMapping(nslot=nslot, nq=nq, slot_node=np.array(slot_node, np.int64), slot_side=np.array(slot_side),
                      lower_slot=lower_slot, corner_slot=corner_slot.astype(np.int32), constraint=constraint,
                      free_slot=free_slot, q_of_slot=q_of_slot, C=C, wake_te=wake_te, K=K, slot_area=slot_area,
                      slot_body=slot_body, body_q=[np.array(b, np.int64) for b in body_q],
                      body_closed=list(g.bodies_closed), log=log)


def wake_rows(dm: DOFMapping, te_of_wake_node) -> sp.csr_matrix:
    """src/dofmap/system.jl:17: W = K[te_of_wake_node, :]"""
    return dm.K[np.asarray(te_of_wake_node), :].tocsr()


def bordered_rows(dm: DOFMapping):
    """src/dofmap/system.jl:41-50: per body, e (ones on its unknowns) and c = C^T (slot_area on the body)"""
    rows = []
    for b in range(len(dm.body_closed)):
        e = np.zeros(dm.nq)
        e[dm.body_q[b]] = 1.0
        c = np.asarray(dm.C.T @ (dm.slot_area * (dm.slot_body == b)))
        rows.append((b, e, c))
    return rows


def reproduction_error(dm: DOFMapping, g: Surface, f) -> float:
    """src/dofmap/system.jl:62-66: max |(C q)[s] - f(x_s)| for q = f at the free slots"""
    x = g.x[dm.slot_node]
    fx = np.array([f(p) for p in x])
    q = fx[dm.free_slot]
    return float(np.max(np.abs(dm.C @ q - fx)))
