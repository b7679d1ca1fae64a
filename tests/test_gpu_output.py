"""Output on the device (MI355X): SWW frames prepared by the device and gauge
rows recorded on the device, against the oracle's run and the frame
restatement of tests/sww_frames.py (node averages in the reference's C loop
order, storable mask, float32), bit for bit; the file read back with scipy's
independent NetCDF reader.
"""

import csv
import os

import numpy as np
import pytest

from anuga_core_amd.domain import Domain
from anuga_core_amd.sww import read_sww
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S
from tests.sww_frames import frame_restatement, static_restatement

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alg,device_timestep", [("DE0", True), ("DE1", True),
                                                 ("DE1", False)])
def test_sww_file_frames_vs_oracle(tmp_path, alg, device_timestep):
    print("case sww %s device_timestep=%s" % (alg, device_timestep), flush=True)
    mesh, _, q, bc = S.dam_break_de0(m=12, n=10)
    d = S.setup(Domain(mesh=mesh), alg, q, bc)
    d.device_timestep = device_timestep
    d.set_store(True)
    d.set_datadir(str(tmp_path))
    d.set_name("dam.sww")
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=True)), alg, q, bc)
    expect = []
    for t in o.evolve(yieldstep=0.5, finaltime=2.0):
        expect.append((o.relative_time,) + frame_restatement(
            mesh, o.stage_c, o.xmom_c, o.ymom_c, o.elev_c, 1e-3))
    times = [t for t in d.evolve(yieldstep=0.5, finaltime=2.0)]
    d.writer.close()
    v, a, dims = read_sww(os.path.join(str(tmp_path), "dam.sww"))
    assert len(times) == len(expect) == 5
    assert np.array_equal(v["time"], [e[0] for e in expect])
    assert dims["number_of_points"] == mesh.number_of_nodes
    assert dims["number_of_volumes"] == mesh.number_of_triangles
    assert a["smoothing"] == "Yes" and a["order"] == 2 and a["starttime"] == 0.0
    assert np.array_equal(v["x"], mesh.nodes[:, 0].astype(np.float32))
    assert np.array_equal(v["volumes"], mesh.triangles)
    for i, (_, nodes, cents) in enumerate(expect):
        for j, name in enumerate(("stage", "xmomentum", "ymomentum")):
            assert np.array_equal(v[name][i], nodes[j]), (i, name)
            assert np.array_equal(v[name + "_c"][i], cents[j]), (i, name)
    ez, ezc = static_restatement(mesh, q["elevation"])
    ef, efc = static_restatement(mesh, q["friction"])
    assert np.array_equal(v["elevation"], ez) and np.array_equal(v["elevation_c"], ezc)
    assert np.array_equal(v["friction"], ef) and np.array_equal(v["friction_c"], efc)
    assert np.array_equal(v["elevation_range"], [ez.min(), ez.max()])
    st = np.stack([e[1][0] for e in expect])
    assert np.array_equal(v["stage_range"], [st.min(), st.max()])


def _oracle_step_rows(o, tri):
    rows = []
    for name in ("evolve_one_euler_step", "evolve_one_rk2_step"):
        orig = getattr(o, name)

        def wrapped(*a, _orig=orig):
            _orig(*a)
            rows.append(np.stack([o.stage_c[tri], o.xmom_c[tri], o.ymom_c[tri],
                                  o.elev_c[tri]], axis=1).copy())
        setattr(o, name, wrapped)
    return rows


@pytest.mark.parametrize("device_timestep", [True, False])
def test_gauges_every_step_vs_oracle(tmp_path, device_timestep):
    print("case gauges device_timestep=%s" % device_timestep, flush=True)
    mesh, alg, q, bc = S.hills_de1(m=14, n=14, seed=5)
    d = S.setup(Domain(mesh=mesh), alg, q, bc)
    d.device_timestep = device_timestep
    pts = [[1000.0, 1000.0], [300.0, 1700.0], [1850.0, 420.0]]
    d.set_gauges(points=pts, names=["centre", "nw", "se"], capacity=7)   # forces drains
    tri = d._gauges["triangles"]
    assert np.array_equal(tri, mesh.find_triangles(pts))
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=True)), alg, q, bc)
    rows = _oracle_step_rows(o, tri)
    for _ in d.evolve(yieldstep=5.0, finaltime=20.0):
        pass
    for _ in o.evolve(yieldstep=5.0, finaltime=20.0):
        pass
    ts = d.get_gauge_timeseries()
    assert d.dt_history == o.dt_history
    n = len(o.dt_history)
    assert len(ts["time"]) == n == len(rows)
    assert np.array_equal(ts["time"], np.cumsum(o.dt_history))
    R = np.stack(rows)
    for j, name in enumerate(("stage", "xmomentum", "ymomentum", "elevation")):
        assert np.array_equal(ts[name], R[:, :, j]), name
    paths = d.write_gauge_csv(str(tmp_path), quantities=["stage", "depth", "speed"])
    assert [os.path.basename(p) for p in paths] == ["gauge_centre.csv", "gauge_nw.csv",
                                                    "gauge_se.csv"]
    with open(paths[0]) as f:
        table = list(csv.reader(f))
    assert table[0] == ["time", "hours", "stage", "depth", "speed"]
    assert len(table) == n + 1
    assert float(table[1][2]) == ts["stage"][0, 0]
    # removing the gauges
    d.set_gauges(triangles=[])
    for _ in d.evolve(yieldstep=5.0, finaltime=25.0):
        pass
