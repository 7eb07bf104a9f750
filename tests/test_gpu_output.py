"""GPU parity of the output path: L2 error integration and subsampled evaluation.

* pdelab_gpu_l2_error_squared against the oracle's integrateGridFunction restatement
  (order 10): within 1e-12 relative for a random coefficient vector, within 1e-15 of the
  field's scale for the solution (whose error cancels to ~1e-5 of u); zero to roundoff for an
  interpolated polynomial of the basis degree (quadrature and evaluation are then exact);
* pdelab_gpu_sample: for an interpolated polynomial of the basis degree, the sampled values
  equal the polynomial at the sample points (1e-13); at L = k intervals they are the DoFs;
* the VTK writer: the file parses, with the refined grid's sizes and the sampled values.
"""
import importlib.util
import os
import xml.etree.ElementTree as ET
import base64
import struct

import numpy as np
import pytest

from test_gpu_parity import rel, say

pytestmark = pytest.mark.gpu

QUAD = (0.3, -0.2, 0.5, 0.1, 0.7, -0.4, 0.25)


def _poly(x, y, z):
    q = QUAD
    return q[0] + q[1] * x + q[2] * y + q[3] * z + q[4] * x * x + q[5] * y * y + q[6] * z * z


@pytest.mark.parametrize("k", [1, 2, 3])
def test_l2_error_vs_oracle(pkg, of, k):
    say(f"[gpu] l2 error k={k}")
    cfg = pkg.make_config(nx=4, ny=3, nz=5, degree=k, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    xd = go.vector()
    go.linear_problem_solve(xd, reduction=1e-12)
    x = xd.download()
    # e = u_h - u cancels to ~1e-5 of u (~1) for the solution, so e^2 agrees to ~1e-16 |u| / |e|
    # relative: bound the difference by 1e-15 times the integral of u_h^2 (the field's scale)
    scale = o.l2_error_sq(np.zeros(go.N))
    assert abs(go.l2_error_squared(xd, 10) - o.l2_error_sq(x)) <= 1e-15 * scale
    xr = np.random.default_rng(k).normal(size=go.N)
    assert rel(np.array([go.l2_error_squared(go.vector(xr), 10)]), np.array([o.l2_error_sq(xr)])) <= 1e-12


def test_l2_error_and_sample_of_polynomial(pkg, of):
    say("[gpu] polynomial reproduction of l2 error and sample")
    cfg = pkg.make_config(nx=3, ny=4, nz=2, lx=1.0, ly=2.0, lz=0.5, degree=2, alpha=3.0,
                          problem=pkg.PROBLEM_QUADRATIC, quad=QUAD)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    x = o.interpolate()
    xd = go.vector(x)
    assert go.l2_error_squared(xd, 8) <= 1e-26
    # u_h(0) = 0 gives the integral of u^2, against a numpy tensor Gauss rule
    g, w = np.polynomial.legendre.leggauss(6)
    g, w = 0.5 * (g + 1), 0.5 * w
    X, Y, Z = np.meshgrid(g * 1.0, g * 2.0, g * 0.5, indexing="ij")
    W = np.einsum("i,j,k->ijk", w * 1.0, w * 2.0, w * 0.5)
    assert abs(go.l2_error_squared(go.vector(), 10) - np.sum(W * _poly(X, Y, Z) ** 2)) <= 1e-12
    L = 5
    s = go.sample(xd, L)
    h = np.array([1.0 / 3, 2.0 / 4, 0.5 / 2])
    t = np.arange(L + 1) / L
    for c in range(go.ncells):
        co = np.array([c % 3, (c // 3) % 4, c // 12])
        zz, yy, xx = np.meshgrid(t, t, t, indexing="ij")
        want = _poly((co[0] + xx) * h[0], (co[1] + yy) * h[1], (co[2] + zz) * h[2])
        assert np.abs(s[c] - want).max() <= 1e-13
    # at L = k the refined vertices are the Lagrange nodes
    assert np.abs(go.sample(xd, 2).reshape(-1) - x).max() <= 1e-14


def _read_array(elem):
    raw = base64.b64decode(elem.text)
    n = struct.unpack("<I", raw[:4])[0]
    dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[elem.get("type")]
    return np.frombuffer(raw[4:4 + n], dtype=dt)


def test_vtk_writer(pkg, of, tmp_path):
    say("[gpu] vtk writer")
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("pdelab_vtk", os.path.join(here, "dune-pdelab_amd", "vtk.py"))
    vtk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(vtk)
    cfg = pkg.make_config(nx=2, ny=3, nz=2, degree=2, alpha=3.0, problem=pkg.PROBLEM_QUADRATIC, quad=QUAD)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    xd = go.vector(o.interpolate())
    path = str(tmp_path / "u.vtu")
    npts, nc = vtk.write_vtu(path, go, [("x_h", xd)], intervals=2)
    assert (npts, nc) == (12 * 27, 12 * 8)
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == npts and int(piece.get("NumberOfCells")) == nc
    vals = _read_array(piece.find("PointData/DataArray"))
    pts = _read_array(piece.find("Points/DataArray")).reshape(-1, 3)
    assert np.abs(vals - _poly(pts[:, 0], pts[:, 1], pts[:, 2])).max() <= 1e-13
    conn = _read_array(piece.find("Cells/DataArray[@Name='connectivity']")).reshape(-1, 8)
    # every sub-hexahedron has positive volume in VTK's vertex order
    p = pts[conn]
    vol = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 3] - p[:, 0]), p[:, 4] - p[:, 0])
    assert (vol > 0).all()
