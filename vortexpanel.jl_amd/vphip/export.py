"""ParaView output of a free wake run: the free rows, the particles, the body with its pressure and the trailing edge
panel as VTK unstructured grids (.vtu, XML, ASCII), and the .pvd collections that tie the frames to their times.

Mirrors src/freewake/export.jl:7-97 (write_free_wake_vtk, write_body_vtk, FreeWakeSeries, add_frame!,
write_collections): the same files, names, arrays and frame rule.  The reference writes through WriteVTK; here a
small writer of the same XML format, host side (files only, no device work).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

VTK_VERTEX, VTK_TRIANGLE = 1, 5


def _ascii(a, fmt):
    return " ".join(fmt % v for v in np.asarray(a).ravel())


def _data_array(name, values):
    v = np.asarray(values)
    if v.dtype.kind in "iub":
        kind, fmt = "Int64", "%d"
        v = v.astype(np.int64)
    else:
        kind, fmt = "Float64", "%.17g"
    ncomp = 1 if v.ndim == 1 else v.shape[1]
    return (f'        <DataArray type="{kind}" Name="{name}" NumberOfComponents="{ncomp}" format="ascii">\n'
            f"          {_ascii(v, fmt)}\n        </DataArray>\n")


def write_vtu(path, points, cells, cell_type, point_data=None, cell_data=None):
    """An unstructured grid of one cell type (cells: (ncell, nvert) 0-based) with point and cell arrays"""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    cl = np.asarray(cells, dtype=np.int64).reshape(len(cells), -1) if len(cells) else np.zeros((0, 1), np.int64)
    offsets = np.arange(1, len(cl) + 1) * cl.shape[1]
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian">\n')
        f.write(f'  <UnstructuredGrid>\n    <Piece NumberOfPoints="{len(pts)}" NumberOfCells="{len(cl)}">\n')
        f.write("      <Points>\n" + _data_array("Points", pts) + "      </Points>\n")
        f.write("      <Cells>\n" + _data_array("connectivity", cl.ravel()) + _data_array("offsets", offsets)
                + _data_array("types", np.full(len(cl), cell_type, np.int64)) + "      </Cells>\n")
        for tag, data in (("PointData", point_data), ("CellData", cell_data)):
            if data:
                f.write(f"      <{tag}>\n" + "".join(_data_array(k, v) for k, v in data.items()) + f"      </{tag}>\n")
        f.write("    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n")
    return path


def write_free_wake_vtk(sim, prefix):
    """prefix_free.vtu: the free panel rows (their triangles, each node's frozen doublet and row, 0 the newest); and
    prefix_particles.vtu: the particles (strength vectors, magnitudes, cores, pending merge); export.jl:7-32"""
    st, topo = sim.state, sim.topo
    files = []
    nr = len(st.x)
    if nr >= 2:
        pts = np.vstack(st.x)
        cells = []
        for i in range(nr - 1):
            for A, B in topo.segments:
                a0, a1 = i * topo.nte + A, (i + 1) * topo.nte + A
                b0, b1 = i * topo.nte + B, (i + 1) * topo.nte + B
                cells += [(a0, a1, b0), (a1, b1, b0)]
        files.append(write_vtu(prefix + "_free.vtu", pts, cells, VTK_TRIANGLE,
                               point_data={"mu": np.concatenate(st.mu),
                                           "row": np.repeat(np.arange(nr), topo.nte)}))
    p = st.particles
    if len(p):
        files.append(write_vtu(prefix + "_particles.vtu", p.x, np.arange(len(p))[:, None], VTK_VERTEX,
                               point_data={"alpha": p.alpha, "strength": np.linalg.norm(p.alpha, axis=1),
                                           "core": p.delta, "pending": (st.tag > 0).astype(np.int64)}))
    return files


def nodal_average(g, els, values):
    """Area-weighted average of per-triangle values at the nodes, src/post/Post.jl:115-123"""
    acc = np.zeros(len(g.x))
    wsum = np.zeros(len(g.x))
    for k in range(3):
        np.add.at(acc, g.tris[:, k], els.area * values)
        np.add.at(wsum, g.tris[:, k], els.area)
    return acc / np.maximum(wsum, np.finfo(float).eps)


def write_body_vtk(sim, prefix):
    """prefix_body.vtu: the body of the last solved step (Cp per triangle, area-averaged to the nodes, and the doublet
    at the nodes, the upper copy at the trailing edge); prefix_te_panel.vtu: the trailing edge panel with its doublets
    (mu_w); export.jl:34-51"""
    g = sim.g
    nn = len(g.x)
    body = write_vtu(prefix + "_body.vtu", g.x, g.tris, VTK_TRIANGLE,
                     point_data={"Cp_nodes": nodal_average(g, sim.els, sim.cp), "mu": sim.mu[:nn]},
                     cell_data={"Cp": sim.cp})
    te = write_vtu(prefix + "_te_panel.vtu", sim.te.x, sim.te.tris, VTK_TRIANGLE,
                   point_data={"mu": np.asarray(sim.mu_w)[sim.te.te]})
    return [body, te]


_KINDS = ("te_panel", "body", "free", "particles")


@dataclass
class FreeWakeSeries:
    """A ParaView time series of a free wake run (export.jl:53-68): written to dir by FreeWakeSim.run(n, series=...)
    every `every` steps and the last"""
    dir: str
    every: int = 1
    frames: dict = field(default_factory=dict)

    def __post_init__(self):
        os.makedirs(self.dir, exist_ok=True)


def add_frame(series: FreeWakeSeries, t, files):
    """Add files to the series at time t, each to the collection its suffix names (export.jl:70-78)"""
    for f in files:
        kind = next((k for k in _KINDS if f.endswith(f"_{k}.vtu")), None)
        key = kind or os.path.splitext(os.path.basename(f))[0]
        series.frames.setdefault(key, []).append((float(t), f))
    return series


def write_collections(series: FreeWakeSeries):
    """The series' .pvd collections (ParaView reads the frames' times from them); export.jl:80-97"""
    out = []
    for name, entries in series.frames.items():
        path = os.path.join(series.dir, name + ".pvd")
        with open(path, "w") as f:
            f.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n')
            f.write("  <Collection>\n")
            for t, file in entries:
                f.write(f'    <DataSet timestep="{t!r}" group="" part="0" file="{os.path.basename(file)}"/>\n')
            f.write("  </Collection>\n</VTKFile>\n")
        out.append(path)
    return out
