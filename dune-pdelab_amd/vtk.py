"""VTK output of a DG solution, the role of SubsamplingVTKWriter with refinementIntervals(L)
(R:dune/pdelab/common/vtk.hh:655-690, as the reference tests use it: each cell is refined into
L^3 sub-hexahedra with their own vertices, so the discontinuous field is written as it is).

The values at the refined vertices come from the device (pdelab_gpu_sample); the file is
written on the host as a VTK XML UnstructuredGrid (.vtu).
"""
import base64
import struct

import numpy as np

VTK_HEXAHEDRON = 12


def _refined_grid(cfg, intervals):
    """Vertex coordinates (ncells (L+1)^3, 3) and hexahedron connectivity (ncells L^3, 8)."""
    n = np.array([cfg.nx, cfg.ny, cfg.nz])
    h = np.array([cfg.lx, cfg.ly, cfg.lz]) / n
    L1 = intervals + 1
    c = np.arange(cfg.nx * cfg.ny * cfg.nz)
    co = np.stack([c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1])], axis=1)
    t = np.arange(L1) / intervals
    lz, ly, lx = np.meshgrid(t, t, t, indexing="ij")  # vertex (i, j, l) at [l, j, i]
    local = np.stack([lx.ravel(), ly.ravel(), lz.ravel()], axis=1)  # x fastest
    pts = ((co[:, None, :] + local[None, :, :]) * h).reshape(-1, 3)
    # sub-hexahedra of one cell, VTK vertex order (bottom face counter-clockwise, then top)
    i, j, l = np.meshgrid(np.arange(intervals), np.arange(intervals), np.arange(intervals), indexing="ij")
    i, j, l = i.transpose(2, 1, 0).ravel(), j.transpose(2, 1, 0).ravel(), l.transpose(2, 1, 0).ravel()

    def vid(a, b, d):
        return a + L1 * (b + L1 * d)

    hexes = np.stack([vid(i, j, l), vid(i + 1, j, l), vid(i + 1, j + 1, l), vid(i, j + 1, l),
                      vid(i, j, l + 1), vid(i + 1, j, l + 1), vid(i + 1, j + 1, l + 1), vid(i, j + 1, l + 1)], axis=1)
    conn = (hexes[None, :, :] + (c * L1 ** 3)[:, None, None]).reshape(-1, 8)
    return pts, conn


def _data_array(name, arr, vtype, ncomp, fmt):
    arr = np.ascontiguousarray(arr)
    head = f'<DataArray type="{vtype}" Name="{name}" NumberOfComponents="{ncomp}" format="{fmt}">'
    if fmt == "ascii":
        body = " ".join(repr(float(v)) if arr.dtype.kind == "f" else str(int(v)) for v in arr.ravel())
    else:
        raw = arr.tobytes()
        body = base64.b64encode(struct.pack("<I", len(raw)) + raw).decode()
    return head + body + "</DataArray>\n"


def write_vtu(path, go, fields, intervals=1, fmt="binary"):
    """Writes the refined grid of `go` with one point field per (name, device vector) item of
    `fields` to `path` (.vtu)."""
    cfg = go.cfg
    pts, conn = _refined_grid(cfg, intervals)
    ncells = conn.shape[0]
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">\n',
           "<UnstructuredGrid>\n", f'<Piece NumberOfPoints="{pts.shape[0]}" NumberOfCells="{ncells}">\n',
           "<PointData>\n"]
    for name, x in fields:
        vals = go.sample(x, intervals).reshape(-1)
        out.append(_data_array(name, vals.astype(np.float64), "Float64", 1, fmt))
    out += ["</PointData>\n", "<Points>\n", _data_array("Coordinates", pts.astype(np.float64), "Float64", 3, fmt),
            "</Points>\n", "<Cells>\n",
            _data_array("connectivity", conn.astype(np.int64).ravel(), "Int64", 1, fmt),
            _data_array("offsets", (8 * np.arange(1, ncells + 1)).astype(np.int64), "Int64", 1, fmt),
            _data_array("types", np.full(ncells, VTK_HEXAHEDRON, np.uint8), "UInt8", 1, fmt),
            "</Cells>\n", "</Piece>\n", "</UnstructuredGrid>\n", "</VTKFile>\n"]
    with open(path, "w") as fh:
        fh.write("".join(out))
    return pts.shape[0], ncells
