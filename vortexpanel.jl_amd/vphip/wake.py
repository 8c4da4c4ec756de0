"""Rigid wake strips and the free wake's trailing edge panel (src/wake/rigid.jl:18-169), 0-based."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .dofmap import DOFMapping
from .mesh import Surface


@dataclass
class WakeOptions:
    """src/wake/rigid.jl:18-26"""
    direction: str = "freestream"
    first_row: float | None = None
    growth: float = 1.2
    length: float = 20.0
    ending: str = "truncate"
    far_factor: float = 1e3
    rows: list | None = None

    def check(self):
        """check_options, src/wake/rigid.jl:28-36"""
        if self.direction not in ("freestream", "bisector"):
            raise ValueError(f"direction must be freestream or bisector, not {self.direction}")
        if self.ending not in ("truncate", "semi_infinite", "closed"):
            raise ValueError(f"ending must be truncate, semi_infinite or closed, not {self.ending}")
        if not (self.growth >= 1 and self.length > 0 and self.far_factor > 1):
            raise ValueError("growth must be >= 1, length > 0 and far_factor > 1")
        if self.rows is not None and not (len(self.rows) >= 2 and self.rows[0] == 0 and np.all(np.diff(self.rows) > 0)):
            raise ValueError("rows must start at 0 and increase")
        return self


@dataclass
class RigidWake:
    """src/wake/rigid.jl:53-62"""
    x: np.ndarray          # (nw, 3)
    te: np.ndarray         # (nw,) wake trailing edge node of every wake node
    row: np.ndarray
    tris: np.ndarray       # (nwt, 3) int32
    filaments: np.ndarray  # (nf, 2) int32
    rows: np.ndarray
    span: float
    options: WakeOptions

    @property
    def nw(self):
        return len(self.x)


def te_panel_size(g: Surface) -> float:
    """src/wake/rigid.jl:75-94: median height of the trailing edge triangles over their trailing edge"""
    h = []

    def height(t, a, b):
        c = [v for v in g.tris[t] if v != a and v != b][0]
        return np.linalg.norm(np.cross(g.x[b] - g.x[a], g.x[c] - g.x[a])) / np.linalg.norm(g.x[b] - g.x[a])

    for ch in g.chains:
        for s in range(len(ch.tri_upper)):
            a, b = int(ch.nodes[s]), int(ch.nodes[(s + 1) % len(ch.nodes)])
            h += [height(ch.tri_upper[s], a, b), height(ch.tri_lower[s], a, b)]
    if not h:
        raise ValueError("no trailing edge panels")
    h.sort()
    return h[(len(h) + 1) // 2 - 1]


def build_wake(g: Surface, dm: DOFMapping, vinf, opts: WakeOptions | None = None) -> RigidWake:
    """build_wake, src/wake/rigid.jl:104-151"""
    opts = (opts or WakeOptions()).check()
    wte = dm.wake_te
    n = wte.n
    if n == 0:
        raise ValueError("the geometry has no trailing edge to shed a wake from")
    d = np.asarray(vinf, dtype=np.float64)
    d = d / np.linalg.norm(d)
    span = max(np.linalg.norm(wte.x[i] - wte.x[j]) for i in range(n) for j in range(n))
    if opts.rows is None:
        l1 = te_panel_size(g) if opts.first_row is None else float(opts.first_row)
        rows = [0.0, l1]
        while rows[-1] < opts.length * span:
            rows.append(rows[-1] + (rows[-1] - rows[-2]) * opts.growth)
    else:
        rows = [float(r) for r in opts.rows]
    if opts.ending == "semi_infinite":
        rows.append(rows[-1] * opts.far_factor)
    nrow = len(rows) - 1
    first_dir = wte.bisector if opts.direction == "bisector" else np.tile(d, (n, 1))
    if np.any(np.all(first_dir == 0, axis=1)):
        raise ValueError("a trailing edge node has no bisector; use direction = freestream")

    def nid(k, j):  # wake node of trailing edge node k in row j
        return j * n + k

    x = np.zeros((n * (nrow + 1), 3))
    for j in range(nrow + 1):
        for k in range(n):
            x[nid(k, j)] = wte.x[k] if j == 0 else wte.x[k] + rows[1] * first_dir[k] + (rows[j] - rows[1]) * d
    te = np.tile(np.arange(n), nrow + 1)
    row = np.repeat(np.arange(nrow + 1), n)
    tris, fil = [], []
    for line, closed in zip(wte.lines, wte.closed):
        m = len(line)
        segs = [(line[i], line[i + 1]) for i in range(m - 1)]
        if closed:
            segs.append((line[m - 1], line[0]))
        for A, B in segs:
            for j in range(nrow):
                tris.append((nid(A, j), nid(A, j + 1), nid(B, j)))
                tris.append((nid(A, j + 1), nid(B, j + 1), nid(B, j)))
        if not closed:
            fil += [(nid(line[0], j + 1), nid(line[0], j)) for j in range(nrow)]
            fil += [(nid(line[m - 1], j), nid(line[m - 1], j + 1)) for j in range(nrow)]
        if opts.ending == "closed":
            fil += [(nid(B, nrow), nid(A, nrow)) for A, B in segs]
    return RigidWake(x=x, te=te, row=row, tris=np.array(tris, np.int32).reshape(-1, 3),
                     filaments=np.array(fil, np.int32).reshape(-1, 2), rows=np.array(rows), span=float(span),
                     options=opts)


def te_panel(g: Surface, dm: DOFMapping, vinf, length: float | None = None, direction: str = "freestream") -> RigidWake:
    """src/wake/rigid.jl:162-163: one closed row of wake strips (the free wake's implicit trailing edge panel)"""
    L = te_panel_size(g) if length is None else float(length)
    return build_wake(g, dm, vinf, WakeOptions(rows=[0.0, L], ending="closed", direction=direction))


def wake_values(w: RigidWake, mu_w):
    """src/wake/rigid.jl:169"""
    return np.asarray(mu_w)[w.te]
