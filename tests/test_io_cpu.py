"""Output and restart files (libhclpp_io.so) on the CPU, and the output schedule of the oracle's
driver.

* The file layout is that of the reference's decl_hdf5 declaration (io/config.yaml.cpp:53-239): it
  is read back by h5dump -- the HDF5 tools' own reader, independent of the writer -- for dataset
  names, types, shapes and values, and by hclpp_io_read for the restart path (io.cpp:326-431).
* The output schedule (main.cpp:362-385) is pinned by what it must produce: time outputs land
  exactly on time_first_output + k * frequency, iteration outputs every N steps, the final step
  always outputs, and a first output time not after t is the reference's "negative time step".
"""

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from heraclespp_amd import io as hio
from heraclespp_amd import setups
from heraclespp_amd.hydro import Grid
from tests.helpers import make_oracle

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def _state(g, nfx, seed=0):
    rng = np.random.default_rng(seed)
    s3 = g.shape3
    return (rng.random(s3), rng.standard_normal((g.ndim,) + s3), rng.random(s3), rng.random(s3), rng.random(s3),
            rng.random((nfx,) + s3) if nfx else None)


def _interior(g):
    return tuple(slice(g.Ng, g.Ng + g.Nx_glob_ng[d]) if d < g.ndim else slice(0, 1) for d in (2, 1, 0))


def _h5dump(*args):
    return subprocess.run([H5DUMP, *args], capture_output=True, text=True, check=True).stdout


def test_output_filename():
    # io.cpp:62-74
    assert hio.output_filename("result", 0) == "result_00000000.h5"
    assert hio.output_filename("kh", 1234) == "kh_00001234.h5"


@pytest.mark.parametrize("nd,n,nfx", [(1, (7, 1, 1), 0), (2, (6, 5, 1), 2), (3, (5, 4, 3), 1)])
def test_write_then_read_back(tmp_path, nd, n, nfx):
    g = Grid(nd, n, 2, "Cartesian", (0.0, -1.0, 2.0), (1.0, 0.5, 3.0))
    rho, u, P, E, T, fx = _state(g, nfx)
    hdr = hio.Header(output_id=4, iter_output_id=2, time_output_id=1, iter=321, current_time=0.375, gamma=1.4)
    path = hio.write_state(g, None, str(tmp_path), "result", hdr, rho, u, P, E, fx, T=T)
    assert os.path.basename(path) == "result_00000004.h5"
    got, r2, u2, P2, fx2, xs = hio.read_state(g, path, nfx)
    assert got == hdr
    I = _interior(g)
    assert np.array_equal(r2[I], rho[I]) and np.array_equal(P2[I], P[I])
    for d in range(nd):
        assert np.array_equal(u2[d][I], u[d][I])
    for f in range(nfx):
        assert np.array_equal(fx2[f][I], fx[f][I])
    # ghost entries are left as they were (zero here): only the interior is selected
    mask = np.ones(g.shape3, bool)
    mask[I] = False
    assert not np.any(r2[mask]) and not np.any(P2[mask])
    for d in range(3):
        assert np.array_equal(xs[d], g.x[d])


def test_layout_seen_by_h5dump(tmp_path):
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not installed")
    g = Grid(3, (5, 4, 3), 2, "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    rho, u, P, E, T, fx = _state(g, 2, seed=1)
    hdr = hio.Header(output_id=7, iter_output_id=3, time_output_id=6, iter=90, current_time=0.5, gamma=5.0 / 3)
    path = hio.write_state(g, None, str(tmp_path), "run", hdr, rho, u, P, E, fx, T=T)
    head = _h5dump("-H", path)
    want3 = "SIMPLE { ( 3, 4, 5 ) / ( 3, 4, 5 ) }"  # [nz, ny, nx] interior (config.yaml.cpp:90-97)
    dims = dict(re.findall(r'DATASET "(\w+)" \{\s*DATATYPE\s+\S+\s*DATASPACE\s+(SCALAR|SIMPLE \{[^}]*\})', head))
    for name in ("rho", "ux0", "ux1", "ux2", "P", "E", "T", "fx0", "fx1"):
        assert dims[name] == want3, (name, dims[name])
    for name, size in (("x0", 10), ("x1", 9), ("x2", 8), ("x0_ng", 6), ("x1_ng", 5), ("x2_ng", 4)):
        assert dims[name] == f"SIMPLE {{ ( {size} ) / ( {size} ) }}", (name, dims[name])
    for name in ("output_id", "iter_output_id", "time_output_id", "iter", "current_time", "gamma"):
        assert dims[name] == "SCALAR", name
    assert "fx2" not in dims
    types = dict(re.findall(r'DATASET "(\w+)" \{\s*DATATYPE\s+(\S+)', head))
    assert types["iter"] == "H5T_STD_I32LE" and types["current_time"] == "H5T_IEEE_F64LE" and types["rho"] == "H5T_IEEE_F64LE"
    for attr in ("git_build_string", "git_branch", "compile_date", "compile_time"):
        assert f'ATTRIBUTE "{attr}"' in head

    def value(name):
        out = _h5dump("-d", f"/{name}", "-m", "%.17g", path)
        return re.search(r"DATA \{\s*\(0\): ([^\s}]+)", out).group(1)

    assert int(value("iter")) == 90 and int(value("output_id")) == 7 and int(value("time_output_id")) == 6
    assert float(value("current_time")) == 0.5 and float(value("gamma")) == 5.0 / 3
    # x0_ng is the interior node range of x0 (memory_selection start n_ghost)
    out = _h5dump("-d", "/x0_ng", "-m", "%.17g", path)
    block = re.sub(r"\(\d+\):", " ", out.split("DATA {")[1].split("}")[0])
    vals = [float(v) for v in block.replace(",", " ").split()]
    assert np.array_equal(np.array(vals), g.x[0][2:2 + 6])
    # one interior value of rho, element (k, j, i) = (1, 2, 3) of the [nz, ny, nx] dataset
    out = _h5dump("-d", "/rho", "-s", "1,2,3", "-c", "1,1,1", "-m", "%.17g", path)
    v = float(re.search(r"\(1,2,3\): ([^\s,}]+)", out).group(1))
    assert v == rho[2 + 1, 2 + 2, 2 + 3]


def test_read_rejects_another_grid(tmp_path):
    g = Grid(2, (6, 5, 1), 2, "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    rho, u, P, E, T, _ = _state(g, 0)
    path = hio.write_state(g, None, str(tmp_path), "r", hio.Header(0, 0, -1, 0, 0.0, 1.4), rho, u, P, E, T=T)
    other = Grid(2, (7, 5, 1), 2, "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    with pytest.raises(hio.HdfIoError, match="extent of x0 does not match"):
        hio.read_state(other, path)
    with pytest.raises(hio.HdfIoError, match="dataset fx0 not found"):
        hio.read_state(g, path, nfx=1)


# ---- the XDMF side file (io.cpp:433-566)
def _h5_dims(path):
    head = _h5dump("-H", path)
    out = {}
    for name, space in re.findall(r'DATASET "(\w+)" \{\s*DATATYPE\s+\S+\s*DATASPACE\s+(SCALAR|SIMPLE \{[^}]*\})', head):
        m = re.match(r"SIMPLE \{ \( ([^)]*) \)", space)
        out[name] = m.group(1).replace(",", "") if m else "scalar"
    return out


@pytest.mark.parametrize("nd,n,nfx", [(1, (7, 1, 1), 0), (2, (6, 5, 1), 2), (3, (5, 4, 3), 1)])
def test_xdmf_indexes_the_output_files(tmp_path, nd, n, nfx):
    """Every data item of the XDMF file names a dataset of its output file whose shape is the item's
    Dimensions (read by h5dump); grids are the recorded outputs, in order, with their times."""
    import xml.etree.ElementTree as ET

    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not installed")
    g = Grid(nd, n, 2, "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    rho, u, P, E, T, fx = _state(g, nfx)
    # a restarted run: outputs 3..5 of a run, the record holding only those (main.cpp:210)
    record = [(40, 0.1 + 0.2), (80, 1e-7), (120, 2.5)]
    for k, (it, t) in enumerate(record):
        hio.write_state(g, None, str(tmp_path), "run", hio.Header(3 + k, k, k, it, t, 1.4), rho, u, P, E, fx, T=T)
    path = hio.write_xdmf(str(tmp_path), "run", g.Nx_glob_ng, nd, nfx, 5, record)
    assert os.path.basename(path) == "run.xmf"
    text = open(path).read()
    lines = text.splitlines()
    assert lines[:3] == ['<?xml version="1.0"?>', '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>', '<Xdmf Version="2.0">']
    assert lines[3] == "  <Domain>" and lines[-1] == "</Xdmf>"
    root = ET.fromstring("\n".join(lines[2:]))  # the DOCTYPE line is not parsed
    series = root.find("Domain/Grid")
    assert series.attrib == {"Name": "TimeSeries", "GridType": "Collection", "CollectionType": "Temporal"}
    grids = series.findall("Grid")
    assert len(grids) == 3
    nx = [int(v) for v in (list(n) + [1, 1])[:3]]
    want_vars = ["rho", "P", "E", "T"] + [f"fx{f}" for f in range(nfx)] + [f"ux{d}" for d in range(nd)]
    for k, grid in enumerate(grids):
        fname = f"run_{3 + k:08d}.h5"
        dims = _h5_dims(str(tmp_path / fname))
        assert grid.find("Time").attrib["Value"] == ("0.3", "1e-07", "2.5")[k]  # C++ ostream default: 6 digits
        topo = grid.find("Topology").attrib
        assert topo == {"TopologyType": "3DRectMesh", "Dimensions": f"{nx[2] + 1} {nx[1] + 1} {nx[0] + 1}"}
        geo = grid.find("Geometry")
        assert geo.attrib == {"GeometryType": "VXVYVZ"}
        for d, item in enumerate(geo.findall("DataItem")):
            ref = item.text.strip()
            assert ref == f"{fname}:/x{d}_ng"
            assert item.attrib["Dimensions"] == dims[f"x{d}_ng"] == str(nx[d] + 1)
            assert (item.attrib["NumberType"], item.attrib["Precision"], item.attrib["Format"]) == ("Float", "8", "HDF")
        attrs = grid.findall("Attribute")
        assert [a.attrib["Name"] for a in attrs] == want_vars
        for a in attrs:
            assert a.attrib["Center"] == "Cell" and a.attrib["AttributeType"] == "Scalar"
            item = a.find("DataItem")
            name = a.attrib["Name"]
            assert item.text.strip() == f"{fname}:/{name}"
            assert item.attrib["Dimensions"] == dims[name] == f"{nx[2]} {nx[1]} {nx[0]}"
    # rewritten whole at each output: a shorter record leaves a shorter file
    hio.write_xdmf(str(tmp_path), "run", g.Nx_glob_ng, nd, nfx, 3, record[:1])
    assert len(ET.fromstring("\n".join(open(path).read().splitlines()[2:])).findall("Domain/Grid/Grid")) == 1


# ---- the oracle's output schedule (main.cpp:362-385)
def _sod_oracle(n=64):
    s = setups.sod(n)
    o = make_oracle(s)
    o.init_from_prim(s.rho, s.u, s.P)
    return s, o


def test_oracle_time_outputs_land_on_their_times():
    s, o = _sod_oracle()
    freq = 0.013
    o.set_outputs(time_first_output=freq, time_output_frequency=freq)
    times = []
    while True:
        _, ex, out = o.step_outputs(s.cfl, 0.1)
        if out:
            times.append(o.t)
        if ex:
            break
    # every time output at first + k freq (dt clipped to it: t + (next - t), within an ulp of next),
    # then the final output at t_end
    want = [freq + k * freq for k in range(7)]
    assert len(times) == 8 and np.allclose(times[:-1], want, rtol=2e-16, atol=0) and times[-1] == 0.1
    assert o.output_ids() == (-1, 6)


def test_oracle_iteration_outputs_and_final_step():
    s, o = _sod_oracle()
    o.set_outputs(iter_output_frequency=5, iter_output_id=0)  # after the initial output (main.cpp:313-316)
    steps = []
    while True:
        _, ex, out = o.step_outputs(s.cfl, 1.0, max_iter=23)
        if out:
            steps.append(o.iter)
        if ex:
            break
    assert steps == [5, 10, 15, 20, 23] and o.output_ids() == (4, -1)


def test_oracle_output_time_not_after_t_is_an_error():
    s, o = _sod_oracle()
    o.step(s.cfl, 1.0)
    o.set_outputs(time_first_output=0.0, time_output_frequency=0.0001, time_output_id=-1)
    with pytest.raises(RuntimeError, match="the time step is negative"):
        o.step_outputs(s.cfl, 1.0)
