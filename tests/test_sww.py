"""SWW output and gauge helpers on the CPU.

* the NetCDF-3 (64-bit offset) writer: files read back by scipy's independent
  reader with every dimension, attribute and variable, including appended
  and rewritten records;
* the node structure of General_mesh.build_inverted_triangle_structure
  (general_mesh.py:790-855) on the reference's own worked example
  (its docstring, :760-775);
* the frame restatement used by the GPU tests (oracle side: node averages in
  the reference's C loop order, quantity.c:872-914, and the storable mask of
  sww.py:318-356) against closed forms;
* gauge CSV quantities (_quantities2csv, gauge.py:30-81) and calc_bearing.
"""

import math

import numpy as np
import pytest

from anuga_core_amd.domain import _bearing, _gauge_row
from anuga_core_amd.mesh import Mesh, rectangular_cross
from anuga_core_amd.sww import NC3Writer, node_structure, read_sww
from tests.sww_frames import frame_restatement, node_average_restatement


def test_nc3_writer_round_trip(tmp_path):
    path = str(tmp_path / "t.sww")
    nc = NC3Writer(path)
    nc.attrs.update(dict(institution="x", order=2, starttime=1.5, smoothing="Yes"))
    nc.create_dimension("number_of_volumes", 3)
    nc.create_dimension("number_of_points", 5)
    nc.create_dimension("numbers_in_range", 2)
    nc.create_dimension("number_of_timesteps", None)
    nc.create_variable("x", np.float32, ("number_of_points",))
    nc.create_variable("volumes", np.int32, ("number_of_volumes", "numbers_in_range"))
    nc.create_variable("stage", np.float32, ("number_of_timesteps", "number_of_points"))
    nc.create_variable("stage_range", np.float32, ("numbers_in_range",))
    nc.create_variable("stage_c", np.float32, ("number_of_timesteps", "number_of_volumes"))
    nc.create_variable("time", np.float64, ("number_of_timesteps",))
    nc.end_define()
    x = np.arange(5, dtype=np.float32) * 0.5
    nc.write("x", x)
    nc.write("volumes", np.arange(6, dtype=np.int32).reshape(3, 2))
    rng = np.random.default_rng(0)
    frames = [(t, rng.random(5).astype(np.float32), rng.random(3).astype(np.float32))
              for t in (0.0, 0.5, 1.0)]
    for t, s, c in frames:
        nc.append_record({"time": [t], "stage": s, "stage_c": c})
    s2 = rng.random(5).astype(np.float32)
    nc.append_record({"time": [0.5], "stage": s2, "stage_c": frames[1][2]}, index=1)
    nc.set_attribute_array("stage_range", np.array([-1.0, 2.0], np.float32))
    nc.close()
    v, a, d = read_sww(path)
    assert d["number_of_timesteps"] is None and d["number_of_points"] == 5
    assert a["institution"] == "x" and a["order"] == 2 and a["starttime"] == 1.5
    assert np.array_equal(v["x"], x)
    assert np.array_equal(v["volumes"], np.arange(6).reshape(3, 2))
    assert np.array_equal(v["time"], [0.0, 0.5, 1.0])
    assert np.array_equal(v["stage"][0], frames[0][1])
    assert np.array_equal(v["stage"][1], s2)
    assert np.array_equal(v["stage_c"][2], frames[2][2])
    assert np.array_equal(v["stage_range"], [-1.0, 2.0])


def test_node_structure_reference_example():
    """general_mesh.py:760-775: triangles [[1,0,2],[1,2,4],[4,2,5],[3,1,4]]."""
    tri = np.array([[1, 0, 2], [1, 2, 4], [4, 2, 5], [3, 1, 4]])
    counts, vvi = node_structure(tri, 6)
    assert counts.tolist() == [1, 3, 3, 1, 3, 1]
    # the docstring's order: node 0: 1; node 1: 0, 3, 10; node 2: 2, 4, 7;
    # node 3: 9; node 4: 5, 6, 11; node 5: 8 (any order within a node is a
    # valid grouping; numpy's argsort fixes it)
    start = np.concatenate([[0], np.cumsum(counts)])
    groups = [sorted(vvi[start[j]:start[j + 1]].tolist()) for j in range(6)]
    assert groups == [[1], [0, 3, 10], [2, 4, 7], [9], [5, 6, 11], [8]]


def test_frame_restatement_closed_forms():
    p, v, b = rectangular_cross(3, 2, 3.0, 2.0)
    mesh = Mesh(p, v, b)
    N = mesh.number_of_triangles
    x, y = mesh.centroid_coordinates.T
    # linear fields are not preserved by centroid averaging in general, but a
    # constant is, and a node average is the mean of its triangles' values
    counts, vvi = node_structure(mesh.triangles, mesh.number_of_nodes)
    const = np.full(N, 2.5)
    assert np.array_equal(node_average_restatement(const, counts, vvi), np.full(len(counts), 2.5))
    vals = np.arange(N, dtype=np.float64)
    avg = node_average_restatement(vals, counts, vvi)
    for j in range(mesh.number_of_nodes):
        tris = np.nonzero((mesh.triangles == j).any(axis=1))[0]
        assert avg[j] == pytest.approx(vals[tris].mean(), rel=1e-15)
    # storable mask: depth below the threshold -> stage = bed, momenta 0
    z = -0.1 * x
    w = np.where(x < 1.5, z + 0.5, z + 1e-4)
    uh = np.full(N, 0.2)
    vh = np.full(N, -0.1)
    nv, cv = frame_restatement(mesh, w, uh, vh, z, 1e-3)
    zn = node_average_restatement(z, counts, vvi)
    wn = node_average_restatement(w, counts, vvi)
    keep = wn - zn >= 1e-3
    assert keep.any() and (~keep).any()
    assert np.array_equal(nv[0], np.where(keep, wn, zn).astype(np.float32))
    assert np.array_equal(nv[1], np.where(keep, np.float32(0.2), np.float32(0)))
    assert np.array_equal(cv[0], w.astype(np.float32))


def test_gauge_quantities_and_bearing():
    row = _gauge_row(["stage", "elevation", "xmomentum", "ymomentum", "depth", "momentum",
                      "speed", "bearing"], 2.0, 0.5, 0.3, -0.4)
    assert row[:5] == [2.0, 0.5, 0.3, -0.4, 1.5]
    assert row[5] == pytest.approx(0.5)
    assert row[6] == pytest.approx(0.5 / 1.5)
    # calc_bearing: the angle from (0, -1), anticlockwise, in degrees
    assert _bearing(0.0, -1.0) == pytest.approx(0.0)
    assert _bearing(1.0, 0.0) == pytest.approx(90.0)
    assert _bearing(0.0, 1.0) == pytest.approx(180.0)
    assert _bearing(-1.0, 0.0) == pytest.approx(270.0)
    assert math.isnan(_bearing(0.0, 0.0))
    # shallow gauges report zero speed
    assert _gauge_row(["speed"], 0.5005, 0.5, 1.0, 1.0) == [0.0]


def test_find_triangles():
    p, v, b = rectangular_cross(4, 4, 4.0, 4.0)
    mesh = Mesh(p, v, b)
    cen = mesh.centroid_coordinates
    assert np.array_equal(mesh.find_triangles(cen), np.arange(mesh.number_of_triangles))
    assert mesh.find_triangles([[5.0, 1.0]])[0] == -1
