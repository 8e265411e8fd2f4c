"""VTK output of interval fields and dense level arrays (drop-in for
``subsetix_cupy.export_vtk``; the on-disk formats of export_vtk.py:37-487).

* ``write_unstructured_quads_vtu``: one UnstructuredGrid (.vtu) in which every
  listed cell becomes a quad (VTK type 9) with cell arrays ``level``,
  ``ghost_mask`` and ``u``; binary arrays appended raw, each behind a UInt64
  byte count (header_type="UInt64").  Cells come from interval fields or
  from dense masks.
* ``write_rectilinear_grid_vtr``: a dense cell-centred field on a
  RectilinearGrid (.vtr, ASCII), with optional extra cell arrays.
* ``write_pvd``: a ParaView collection of time steps.
* ``save_amr3_step_vtr``: the three dense level files of one step.

Interval fields are read back with ONE device -> host copy per array (the
CSR, the values); the cell coordinates are expanded on the host with
vectorised NumPy (every cell is (row, begin + k) of its interval), so the
exporter adds no kernel and no per-cell device work.
"""

from __future__ import annotations

import os
import struct
from typing import Dict, List, Sequence, Tuple

import numpy as np

QUAD = 9  # VTK_QUAD


def _host(a) -> np.ndarray:
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def interval_cells(field) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(rows, cols, values) of every cell of an interval field, in storage
    order (row by row, interval by interval)."""
    s = field.interval_set
    off = _host(s.row_offsets).astype(np.int64)
    rows = _host(s.rows).astype(np.int64)
    beg = _host(s.begin).astype(np.int64)
    end = _host(s.end).astype(np.int64)
    vals = _host(field.values)
    lens = np.maximum(end - beg, 0)
    n = int(lens.sum())
    if n == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32)
    interval_row = np.repeat(rows, np.diff(off))
    start = np.concatenate([[0], np.cumsum(lens)[:-1]])
    cols = np.arange(n, dtype=np.int64) + np.repeat(beg - start, lens)
    return np.repeat(interval_row, lens).astype(np.int32), cols.astype(np.int32), vals[:n].astype(np.float32)


def _quads(rows, cols, dx, dy, x0, y0):
    """Corner points (4 per cell, counter-clockwise from lower-left) in
    float32, as the reference writer lays them out."""
    dx32, dy32 = np.float32(dx), np.float32(dy)
    xl = np.float32(x0) + cols.astype(np.float32) * dx32
    yb = np.float32(y0) + rows.astype(np.float32) * dy32
    xr, yt = xl + dx32, yb + dy32
    pts = np.zeros((rows.size, 4, 3), np.float32)
    pts[:, :, 0] = np.stack([xl, xr, xr, xl], axis=1)
    pts[:, :, 1] = np.stack([yb, yb, yt, yt], axis=1)
    return pts.reshape(-1, 3)


def _entry_cells(entry):
    """(rows, cols, values, level, dx, dy, x0, y0, ghost flags) of one entry:
    (IntervalField, level, dx, dy, x0, y0, ghost_flag) or
    (mask, u, level, dx, dy, x0, y0[, ghost])."""
    first = entry[0]
    if hasattr(first, "interval_set"):
        if len(entry) < 7:
            raise ValueError("an IntervalField entry is (field, level, dx, dy, x0, y0, ghost_flag)")
        field, lvl, dx, dy, x0, y0, gflag = entry[:7]
        r, c, v = interval_cells(field)
        return r, c, v, int(lvl), float(dx), float(dy), float(x0), float(y0), np.full(r.size, int(gflag), np.int32)
    if len(entry) not in (7, 8):
        raise ValueError("a mask entry is (mask, u, level, dx, dy, x0, y0[, ghost])")
    mask, u, lvl, dx, dy, x0, y0 = entry[:7]
    m = _host(mask).astype(bool)
    r, c = np.nonzero(m)
    v = _host(u).astype(np.float32)[r, c]
    if len(entry) == 8 and entry[7] is not None:
        g = _host(entry[7]).astype(np.int32)[r, c]
    else:
        g = np.zeros(r.size, np.int32)
    return r.astype(np.int32), c.astype(np.int32), v, int(lvl), float(dx), float(dy), float(x0), float(y0), g


def write_unstructured_quads_vtu(path: str, cells: Sequence[tuple], origin=(0.0, 0.0)) -> None:
    """Write the cells of every entry (see ``_entry_cells``) as quads of one
    UnstructuredGrid with cell arrays level / ghost_mask / u."""
    parts = [_entry_cells(e) for e in cells if e]
    parts = [p for p in parts if p[0].size]
    ncell = sum(p[0].size for p in parts)
    level = np.concatenate([np.full(p[0].size, p[3], np.int32) for p in parts]) if parts else np.zeros(0, np.int32)
    ghost = np.concatenate([p[8] for p in parts]) if parts else np.zeros(0, np.int32)
    u = np.concatenate([p[2] for p in parts]).astype(np.float32) if parts else np.zeros(0, np.float32)
    pts = (np.concatenate([_quads(p[0], p[1], p[4], p[5], p[6], p[7]) for p in parts]) if parts
           else np.zeros((0, 3), np.float32))
    conn = np.arange(4 * ncell, dtype=np.int32)
    offs = 4 * np.arange(1, ncell + 1, dtype=np.int32)
    types = np.full(ncell, QUAD, np.uint8)
    arrays = [("CellData", "Int32", "level", 1, level), ("CellData", "Int32", "ghost_mask", 1, ghost),
              ("CellData", "Float32", "u", 1, u), ("Points", "Float32", None, 3, pts),
              ("Cells", "Int32", "connectivity", 1, conn), ("Cells", "Int32", "offsets", 1, offs),
              ("Cells", "UInt8", "types", 1, types)]
    offsets, pos = [], 0
    for a in arrays:
        offsets.append(pos)
        pos += 8 + a[4].nbytes

    def tag(a, off):
        name = f' Name="{a[2]}"' if a[2] else ""
        comps = f' NumberOfComponents="{a[3]}"' if a[3] != 1 else ""
        return f'        <DataArray type="{a[1]}"{name}{comps} format="appended" offset="{off}"/>\n'

    head = ['<?xml version="1.0"?>\n',
            '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt64">\n',
            "  <UnstructuredGrid>\n",
            f'    <Piece NumberOfPoints="{pts.shape[0]}" NumberOfCells="{ncell}">\n']
    for section, open_tag in (("CellData", '      <CellData Scalars="u">\n'), ("Points", "      <Points>\n"),
                              ("Cells", "      <Cells>\n")):
        head.append(open_tag)
        head += [tag(a, off) for a, off in zip(arrays, offsets) if a[0] == section]
        head.append(f"      </{section}>\n")
    head += ["    </Piece>\n", "  </UnstructuredGrid>\n", '  <AppendedData encoding="raw">\n']
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "wb") as fh:
        fh.write("".join(head).encode())
        fh.write(b"_")
        for a in arrays:
            data = np.ascontiguousarray(a[4])
            fh.write(struct.pack("<Q", data.nbytes))
            fh.write(data.tobytes())
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")


def _ascii_array(name: str, arr: np.ndarray, vtk_type: str, per_line: int = 4096) -> str:
    flat = arr.reshape(-1)
    lines = [" ".join(str(x) for x in flat[i:i + per_line]) for i in range(0, flat.size, per_line)]
    body = "".join(f"          {ln}\n" for ln in lines)
    return f'        <DataArray type="{vtk_type}" Name="{name}" format="ascii">\n{body}        </DataArray>\n'


def write_rectilinear_grid_vtr(path: str, u_cell, dx: float, dy: float, *, origin=(0.0, 0.0, 0.0),
                               cell_arrays: Dict[str, object] | None = None) -> None:
    """Dense (H, W) cell-centred field as a RectilinearGrid (.vtr, ASCII);
    boolean / integer extra arrays become Int32, others Float32."""
    u = _host(u_cell)
    h, w = u.shape
    x0, y0, z0 = origin
    parts = [_ascii_array("u", u.astype(np.float32), "Float32")]
    for name, arr in (cell_arrays or {}).items():
        a = _host(arr)
        if a.dtype == np.bool_ or np.issubdtype(a.dtype, np.integer):
            parts.append(_ascii_array(name, a.astype(np.int32), "Int32"))
        else:
            parts.append(_ascii_array(name, a.astype(np.float32), "Float32"))
    xs = np.float32(x0) + np.float32(dx) * np.arange(w + 1, dtype=np.float32)
    ys = np.float32(y0) + np.float32(dy) * np.arange(h + 1, dtype=np.float32)

    def coord(name, v):
        return (f'        <DataArray type="Float32" Name="{name}" NumberOfComponents="1" format="ascii">\n'
                f"          {' '.join(str(x) for x in v)}\n        </DataArray>\n")

    text = ('<?xml version="1.0"?>\n'
            '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian">\n'
            f'  <RectilinearGrid WholeExtent="0 {w} 0 {h} 0 0">\n'
            f'    <Piece Extent="0 {w} 0 {h} 0 0">\n'
            '      <CellData Scalars="u">\n' + "".join(parts) + "      </CellData>\n"
            "      <Coordinates>\n" + coord("X_COORDINATES", xs) + coord("Y_COORDINATES", ys)
            + coord("Z_COORDINATES", np.array([z0], np.float32)) + "      </Coordinates>\n"
            "    </Piece>\n  </RectilinearGrid>\n</VTKFile>\n")
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w", encoding="utf-8") as fh:
        fh.write(text)


def write_pvd(path: str, datasets: List[Tuple[float, List[str]]]) -> None:
    """ParaView collection: one DataSet per (time, file) with part = the
    file's position in its step's list."""
    rows = "".join(f'    <DataSet timestep="{t}" group="" part="{k}" file="{f}"/>\n'
                   for t, files in datasets for k, f in enumerate(files))
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w", encoding="utf-8") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
                 f"  <Collection>\n{rows}  </Collection>\n</VTKFile>\n")


def save_amr3_step_vtr(out_dir: str, base: str, step: int, time_value: float, *, u0, u1, u2, refine0, L1_mask,
                       refine1_mid, L2_mask, dx0: float, dy0: float, dx1: float, dy1: float, dx2: float,
                       dy2: float):
    """The three dense level files of one step (L0 with coarse_only, L1 with
    mid_active / mid_only, L2 with fine_active) and the step's PVD entry."""
    names = [f"{base}_step{step:04d}_L{k}.vtr" for k in range(3)]
    l1 = _host(L1_mask).astype(bool)
    write_rectilinear_grid_vtr(os.path.join(out_dir, names[0]), u0, dx0, dy0,
                               cell_arrays={"coarse_only": ~_host(refine0).astype(bool)})
    write_rectilinear_grid_vtr(os.path.join(out_dir, names[1]), u1, dx1, dy1,
                               cell_arrays={"mid_active": l1, "mid_only": l1 & ~_host(refine1_mid).astype(bool)})
    write_rectilinear_grid_vtr(os.path.join(out_dir, names[2]), u2, dx2, dy2,
                               cell_arrays={"fine_active": _host(L2_mask).astype(bool)})
    return names, (time_value, names)


def read_vtu(path: str) -> Dict[str, np.ndarray]:
    """Arrays of a .vtu written by ``write_unstructured_quads_vtu`` (the
    appended raw block decoded by the XML offsets); for checks and tools."""
    import re

    raw = open(path, "rb").read()
    start = raw.index(b"<AppendedData")
    data0 = raw.index(b"_", start) + 1
    out = {}
    types = {"Int32": np.int32, "Float32": np.float32, "UInt8": np.uint8}
    for m in re.finditer(rb'<DataArray type="(\w+)"(?: Name="(\w+)")?(?: NumberOfComponents="(\d)")? '
                         rb'format="appended" offset="(\d+)"/>', raw[:start]):
        vt, name, comps, off = m.group(1).decode(), m.group(2), m.group(3), int(m.group(4))
        nbytes = struct.unpack("<Q", raw[data0 + off:data0 + off + 8])[0]
        arr = np.frombuffer(raw[data0 + off + 8:data0 + off + 8 + nbytes], dtype=types[vt])
        key = name.decode() if name else "points"
        out[key] = arr.reshape(-1, int(comps)) if comps else arr
    return out


__all__ = ["write_unstructured_quads_vtu", "write_rectilinear_grid_vtr", "write_pvd", "save_amr3_step_vtr",
           "interval_cells", "read_vtu"]
