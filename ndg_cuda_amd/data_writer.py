"""VTK XML output of the solution with a ParaView ".series" time index, without the VTK library.

SEM::Helpers::DataWriter_t (src/helpers/DataWriter_t.cpp:18-546, include/helpers/DataWriter_t.h)
writes, through VTK's parallel unstructured-grid writer, one file per output time named
``<stem>_t<time with 9 decimals, '.' -> '_'>s<ext>`` in the directory of the path it was given,
and keeps ``<path>.series`` (and ``<dir>/<stem>_element<ext>.series``) listing
``{"name", "time"}`` of every file written. This module writes the same files directly:

- ``.vtu``: one UnstructuredGrid file; ``.pvtu``: a PUnstructuredGrid file naming one piece,
  ``<file stem>_0.vtu`` beside it (the single piece of a one-process run).
- Points: the n x n uniform output points of every element (element e's point (i, j) is point
  e * n * n + i * n + j, z = 0); cells: VTK_QUAD (i + 1, j), (i + 1, j + 1), (i, j + 1), (i, j) for
  every output sub-cell (DataWriter_t.cpp:58-71). Point data "Pressure" and "Velocity" (2
  components).
- write_complete_data adds the point data "PressureDerivative", "VelocityDerivative",
  "PressureAnalyticalError" and "VelocityAnalyticalError", and writes a second grid (the
  ``_element`` file) with one quad per element over its output corners and the cell data "N",
  "Index", "Status", "Rotation", "Rank", "PressureError", "VelocityError", "PressureSigma",
  "VelocitySigma", "Refine", "Coarsen", "SplitLevel" (DataWriter_t.cpp:135-503).

Differences: arrays are written uncompressed, inline (``format="binary"``: base64 of a UInt64 byte
count followed by the little-endian values, or ``format="ascii"``), and points as Float64 (VTK's
default points type is Float32).
"""
import base64
import json
import os
from pathlib import Path

import numpy as np

VTK_QUAD = 9
_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int32): "Int32",
              np.dtype(np.int64): "Int64", np.dtype(np.uint64): "UInt64", np.dtype(np.uint8): "UInt8"}


def time_string(time: float) -> str:
    """``_t<time>s`` with 9 decimals and '.' replaced by '_' (DataWriter_t.cpp:113-116)."""
    return ("_t%.9fs" % float(time)).replace(".", "_")


def quad_connectivity(n: int, K: int) -> np.ndarray:
    """(K * (n - 1)^2, 4) point ids of the output sub-cells, element after element, i then j
    (DataWriter_t.cpp:58-71)."""
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    local = np.stack([(i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1, i * n + j], axis=1)
    offsets = (np.arange(K, dtype=np.int64) * n * n)[:, None, None]
    return (local[None, :, :] + offsets).reshape(-1, 4).astype(np.int64)


def element_corners(n: int, x, y) -> np.ndarray:
    """(4 K, 3) corners of every element's output grid: points (0, 0), (n - 1, 0), (n - 1, n - 1),
    (0, n - 1) (DataWriter_t.cpp:175-184)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, n * n)
    y = np.asarray(y, dtype=np.float64).reshape(-1, n * n)
    sel = [0, (n - 1) * n, (n - 1) * n + n - 1, n - 1]
    pts = np.zeros((x.shape[0] * 4, 3))
    pts[:, 0] = x[:, sel].ravel()
    pts[:, 1] = y[:, sel].ravel()
    return pts


class _Array:
    def __init__(self, name, values, components=1):
        self.name = name
        self.values = np.ascontiguousarray(values)
        self.components = components

    @property
    def vtk_type(self):
        return _VTK_TYPES[self.values.dtype]

    def attrs(self):
        a = f'type="{self.vtk_type}" Name="{self.name}"'
        if self.components != 1:
            a += f' NumberOfComponents="{self.components}"'
        return a


def _encode(values: np.ndarray, fmt: str) -> str:
    v = np.ascontiguousarray(values).astype(values.dtype.newbyteorder("<"), copy=False)
    if fmt == "ascii":
        if v.dtype.kind == "f":
            return " ".join(repr(float(a)) for a in v.ravel())
        return " ".join(str(int(a)) for a in v.ravel())
    raw = v.tobytes()
    return base64.b64encode(np.uint64(len(raw)).astype("<u8").tobytes() + raw).decode("ascii")


def _data_array(a: _Array, fmt: str, indent: str) -> str:
    return f'{indent}<DataArray {a.attrs()} format="{fmt}">\n{indent}  {_encode(a.values, fmt)}\n{indent}</DataArray>\n'


def _vtu(points, cells, point_data, cell_data, fmt) -> str:
    n_cells = cells.shape[0]
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
           '  <UnstructuredGrid>\n',
           f'    <Piece NumberOfPoints="{points.shape[0]}" NumberOfCells="{n_cells}">\n',
           '      <PointData>\n']
    out += [_data_array(a, fmt, "        ") for a in point_data]
    out.append('      </PointData>\n      <CellData>\n')
    out += [_data_array(a, fmt, "        ") for a in cell_data]
    out.append('      </CellData>\n      <Points>\n')
    out.append(_data_array(_Array("Points", points.astype(np.float64), 3), fmt, "        "))
    out.append('      </Points>\n      <Cells>\n')
    out.append(_data_array(_Array("connectivity", cells.reshape(-1).astype(np.int64)), fmt, "        "))
    out.append(_data_array(_Array("offsets", 4 * np.arange(1, n_cells + 1, dtype=np.int64)), fmt, "        "))
    out.append(_data_array(_Array("types", np.full(n_cells, VTK_QUAD, dtype=np.uint8)), fmt, "        "))
    out.append('      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n')
    return "".join(out)


def _pvtu(piece: str, point_data, cell_data) -> str:
    def p(a):
        return f'      <PDataArray {a.attrs()}/>\n'
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="PUnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
           '  <PUnstructuredGrid GhostLevel="0">\n', '    <PPointData>\n']
    out += ["  " + p(a) for a in point_data]
    out.append('    </PPointData>\n    <PCellData>\n')
    out += ["  " + p(a) for a in cell_data]
    out.append('    </PCellData>\n    <PPoints>\n')
    out.append('      <PDataArray type="Float64" Name="Points" NumberOfComponents="3"/>\n')
    out.append(f'    </PPoints>\n    <Piece Source="{piece}"/>\n  </PUnstructuredGrid>\n</VTKFile>\n')
    return "".join(out)


def _interleave(a, b):
    return np.stack([np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()], axis=1)


class DataWriter_t:
    """SEM::Helpers::DataWriter_t(output_filename) (DataWriter_t.cpp:18-34): creates (or empties)
    the series files on construction. ``fmt`` is "binary" (default) or "ascii"."""

    def __init__(self, output_filename, fmt: str = "binary"):
        path = Path(output_filename)
        if fmt not in ("binary", "ascii"):
            raise ValueError("DataWriter_t: fmt is 'binary' or 'ascii'")
        self.fmt = fmt
        self.directory_ = path.parent
        self.filename_ = path.stem
        self.extension_ = path.suffix
        if self.extension_ not in (".vtu", ".pvtu"):
            raise ValueError(f"DataWriter_t: the output extension must be .vtu or .pvtu, got '{self.extension_}'")
        self.series_filename_ = Path(str(path) + ".series")
        self.series_element_filename_ = self.directory_ / (self.filename_ + "_element" + self.extension_ + ".series")
        os.makedirs(self.directory_ if str(self.directory_) else ".", exist_ok=True)
        self.create_time_series_file()

    # --- series files (DataWriter_t.cpp:505-546) -------------------------------------------
    def create_time_series_file(self):
        for f in (self.series_filename_, self.series_element_filename_):
            self._write_series(f, {"file-series-version": "1.0", "files": []})

    @staticmethod
    def _write_series(path, doc):
        with open(path, "w") as f:
            f.write(json.dumps(doc, indent=2) + "\n")

    def add_time_series_to_file(self, filename, time, filename_element=None):
        targets = [(self.series_filename_, filename)]
        if filename_element is not None:
            targets.append((self.series_element_filename_, filename_element))
        for path, name in targets:
            with open(path) as f:
                doc = json.load(f)
            doc["files"].append({"name": name, "time": float(time)})
            self._write_series(path, doc)

    # --- grids -------------------------------------------------------------------------------
    def _write_grid(self, name: str, points, cells, point_data, cell_data):
        target = self.directory_ / name
        if self.extension_ == ".pvtu":
            piece = Path(name).stem + "_0.vtu"
            (self.directory_ / piece).write_text(_vtu(points, cells, point_data, cell_data, self.fmt))
            target.write_text(_pvtu(piece, point_data, cell_data))
        else:
            target.write_text(_vtu(points, cells, point_data, cell_data, self.fmt))
        return target

    @staticmethod
    def _points(x, y):
        x = np.asarray(x, dtype=np.float64).ravel()
        pts = np.zeros((x.size, 3))
        pts[:, 0] = x
        pts[:, 1] = np.asarray(y, dtype=np.float64).ravel()
        return pts

    def write_data(self, n_interpolation_points, n_elements, time, x, y, p, u, v):
        """DataWriter_t::write_data (DataWriter_t.cpp:36-133): the output points with "Pressure"
        and "Velocity"; returns the path written."""
        n, K = int(n_interpolation_points), int(n_elements)
        name = self.filename_ + time_string(time) + self.extension_
        point_data = [_Array("Pressure", np.asarray(p, dtype=np.float64).ravel()),
                      _Array("Velocity", _interleave(u, v), 2)]
        out = self._write_grid(name, self._points(x, y), quad_connectivity(n, K), point_data, [])
        self.add_time_series_to_file(name, time)
        return out

    def write_complete_data(self, n_interpolation_points, n_elements, time, rank, x, y, p, u, v, N, dp_dt, du_dt,
                            dv_dt, p_error, u_error, v_error, p_sigma, u_sigma, v_sigma, refine, coarsen,
                            split_level, p_analytical_error, u_analytical_error, v_analytical_error, status,
                            rotation):
        """DataWriter_t::write_complete_data (DataWriter_t.cpp:135-503): the point grid with the
        state, its time derivative and the analytical error, and the element grid with the
        per-element fields; returns (point file, element file)."""
        n, K = int(n_interpolation_points), int(n_elements)
        ts = time_string(time)
        name = self.filename_ + ts + self.extension_
        name_el = self.filename_ + "_element" + ts + self.extension_
        f64 = lambda a: np.asarray(a, dtype=np.float64).ravel()
        i32 = lambda a: np.asarray(a).astype(np.int32).ravel()
        point_data = [_Array("Pressure", f64(p)), _Array("Velocity", _interleave(u, v), 2),
                      _Array("PressureDerivative", f64(dp_dt)),
                      _Array("VelocityDerivative", _interleave(du_dt, dv_dt), 2),
                      _Array("PressureAnalyticalError", f64(p_analytical_error)),
                      _Array("VelocityAnalyticalError", _interleave(u_analytical_error, v_analytical_error), 2)]
        cell_data = [_Array("N", i32(N)), _Array("Index", np.arange(K, dtype=np.uint64)),
                     _Array("Status", i32(status)), _Array("Rotation", i32(rotation)),
                     _Array("Rank", np.full(K, int(rank), dtype=np.int32)),
                     _Array("PressureError", f64(p_error)), _Array("VelocityError", _interleave(u_error, v_error), 2),
                     _Array("PressureSigma", f64(p_sigma)), _Array("VelocitySigma", _interleave(u_sigma, v_sigma), 2),
                     _Array("Refine", i32(refine)), _Array("Coarsen", i32(coarsen)),
                     _Array("SplitLevel", i32(split_level))]
        out = self._write_grid(name, self._points(x, y), quad_connectivity(n, K), point_data, [])
        el_cells = np.arange(4 * K, dtype=np.int64).reshape(K, 4)
        out_el = self._write_grid(name_el, element_corners(n, x, y), el_cells, [], cell_data)
        self.add_time_series_to_file(name, time, name_el)
        return out, out_el


def read_vtu(path):
    """Reads a file written by DataWriter_t back (test and inspection helper): returns
    {"points": (P, 3), "connectivity": (C, 4), "types": (C,), "point_data": {name: array},
    "cell_data": {name: array}}, multi-component arrays shaped (n, components). A .pvtu is
    followed to its piece."""
    import xml.etree.ElementTree as ET
    path = Path(path)
    root = ET.parse(path).getroot()
    if root.get("type") == "PUnstructuredGrid":
        piece = root.find("PUnstructuredGrid").find("Piece").get("Source")
        return read_vtu(path.parent / piece)
    types = {v: k for k, v in _VTK_TYPES.items()}

    def decode(el):
        dt = types[el.get("type")].newbyteorder("<")
        comps = int(el.get("NumberOfComponents", "1"))
        text = (el.text or "").strip()
        if el.get("format") == "ascii":
            vals = np.array(text.split(), dtype=np.float64 if dt.kind == "f" else np.int64).astype(dt)
        else:
            raw = base64.b64decode(text)
            nbytes = int(np.frombuffer(raw[:8], dtype="<u8")[0])
            if nbytes != len(raw) - 8:
                raise ValueError(f"{el.get('Name')}: header says {nbytes} bytes, found {len(raw) - 8}")
            vals = np.frombuffer(raw[8:], dtype=dt)
        return vals.reshape(-1, comps) if comps > 1 else vals

    piece = root.find("UnstructuredGrid").find("Piece")
    out = {"point_data": {}, "cell_data": {}}
    for el in piece.find("PointData"):
        out["point_data"][el.get("Name")] = decode(el)
    for el in piece.find("CellData"):
        out["cell_data"][el.get("Name")] = decode(el)
    out["points"] = decode(piece.find("Points").find("DataArray"))
    cells = {el.get("Name"): decode(el) for el in piece.find("Cells")}
    out["connectivity"] = cells["connectivity"].reshape(-1, 4)
    out["offsets"] = cells["offsets"]
    out["types"] = cells["types"]
    if out["points"].shape[0] != int(piece.get("NumberOfPoints")):
        raise ValueError("point count mismatch")
    if out["connectivity"].shape[0] != int(piece.get("NumberOfCells")):
        raise ValueError("cell count mismatch")
    return out
