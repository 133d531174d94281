"""Outputs, restart and the loop's stopping rules on the GPU, against the oracle's driver.

* simulate() with a time-output schedule (main.cpp:313-345, :362-385, :532-557), fused and operator
  paths: the same number of files as the oracle's loop, each with the oracle's header (output ids,
  iter, current_time) and, read back from the file, the oracle's primitive state, total energy and
  temperature at that output, bit for bit.
* restart (main.cpp:223-238, then :305-311): a run restarted from an intermediate output file
  equals the oracle restarted from the same file's primitives, bit for bit, with the same later
  outputs; it stays within round-off of the uninterrupted run (E is rebuilt from P).
* the internal-energy check (main.cpp:397-425) runs on the state each iteration starts from: a bad
  initial state stops at iteration 0 (after an output), and a state that turns bad in the last step
  allowed by max_iter ends the run normally, as in the reference.
"""

import os

import numpy as np
import pytest

from heraclespp_amd import io as hio
from heraclespp_amd import setups
from tests.helpers import make_oracle, make_solver, rel_l1

pytestmark = pytest.mark.gpu

FREQ, T_END = 0.004, 0.021


def _kh():
    return setups.kelvin_helmholtz(48, seed=3)


def _oracle_outputs(s, schedule, restart=None):
    """The oracle's loop with outputs: [(header fields, rho, u, P, E)] per output."""
    o = make_oracle(s)
    out = []
    if restart is None:
        o.init_from_prim(s.rho, s.u, s.P)
        ids = [0, 0, -1]  # output_id, iter_output_id, time_output_id after the initial output
        out.append((tuple(ids) + (0, 0.0), o.state("rho").copy(), o.state("u").copy(), o.state("P").copy(),
                    o.state("E").copy()))
    else:
        hdr, rho, u, P, _, _ = restart
        o.init_from_prim(rho, u, P)
        o.set_time(hdr.current_time, hdr.iter)
        ids = [hdr.output_id, hdr.iter_output_id, hdr.time_output_id]
    o.set_outputs(schedule.iter_frequency, schedule.first_output(0.0), schedule.time_frequency, ids[1], ids[2])
    while True:
        _, ex, make = o.step_outputs(s.cfl, T_END, 1 << 30)
        if make:
            ids[0] += 1
            ids[1], ids[2] = o.output_ids()
            out.append((tuple(ids) + (o.iter, o.t), o.state("rho").copy(), o.state("u").copy(), o.state("P").copy(),
                        o.state("E").copy()))
        if ex:
            return out, o


def _check_file(path, grid, want, s):
    hdr, rho, u, P, _, _ = hio.read_state(grid, path)
    (oid, iid, tid, it, t), wrho, wu, wP, wE = want
    assert (hdr.output_id, hdr.iter_output_id, hdr.time_output_id, hdr.iter) == (oid, iid, tid, it), path
    assert hdr.current_time == t and hdr.gamma == s.gamma
    I = (Ellipsis,) + s.interior()
    assert np.array_equal(rho[I], wrho[I]) and np.array_equal(u[I], wu[I]) and np.array_equal(P[I], wP[I]), path
    E = hio.read_field(grid, path, "E")
    T = hio.read_field(grid, path, "T")
    assert np.array_equal(E[I], wE[I]), path
    # analysis/temperature.hpp with perfect_gas.hpp:69-72 (mmw = 1): T = P mmw mp / (rho kb)
    assert np.array_equal(T[I], (wP[I] * 1.0 * 1.672649e-27) / (wrho[I] * 1.380649e-23)), path


def _check_xdmf(xmf, files, grid):
    """The run's XDMF file lists exactly its output files, in order, at their times (printed with
    6 significant digits, as the reference's ostream does)."""
    import xml.etree.ElementTree as ET

    grids = ET.fromstring("\n".join(open(xmf).read().splitlines()[2:])).findall("Domain/Grid/Grid")
    assert len(grids) == len(files)
    for g, path in zip(grids, files):
        name = os.path.basename(path)
        assert g.find("Geometry/DataItem").text.strip() == f"{name}:/x0_ng"
        assert g.find("Attribute/DataItem").text.strip() == f"{name}:/rho"
        t = hio.read_state(grid, path)[0].current_time
        assert g.find("Time").attrib["Value"] == f"{t:g}"


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "operators"])
def test_simulate_time_outputs_match_oracle(cuda, tmp_path, fused):
    s = _kh()
    sched = hio.OutputSchedule(directory=str(tmp_path), prefix="kh", time_frequency=FREQ)
    print(f"[io] KH 48^2 simulate with time outputs every {FREQ}, fused={fused}", flush=True)
    sol = make_solver(s, fused=fused)
    sol.init_from_prim(s.rho, s.u, s.P)
    files = sol.simulate(T_END, 1 << 30, sched)
    want, o = _oracle_outputs(s, sched)
    assert len(files) == len(want) == 7  # initial, t = 0.004 k (k = 1..5), t_end
    for path, w in zip(files, want):
        _check_file(path, sol.grid, w, s)
    _check_xdmf(tmp_path / "kh.xmf", files, sol.grid)
    assert sol.t == o.t and sol.iter == o.iter
    assert np.array_equal(sol.E.cpu().numpy(), o.state("E"))


def test_fp32_storage_outputs_land_on_output_times(cuda, tmp_path):
    s = _kh()
    sched = hio.OutputSchedule(directory=str(tmp_path), prefix="kh32", time_frequency=FREQ)
    sol = make_solver(s, precision="fp32")
    sol.init_from_prim(s.rho, s.u, s.P)
    files = sol.simulate(T_END, 1 << 30, sched)
    assert len(files) == 7
    times = [hio.read_state(sol.grid, f)[0].current_time for f in files]
    want = [0.0] + [FREQ + k * FREQ for k in range(5)] + [T_END]
    assert np.allclose(times, want, rtol=2e-16, atol=0)


def test_restart_matches_oracle_restart(cuda, tmp_path):
    s = _kh()
    sched = hio.OutputSchedule(directory=str(tmp_path), prefix="kh", time_frequency=FREQ)
    full = make_solver(s)
    full.init_from_prim(s.rho, s.u, s.P)
    files = full.simulate(T_END, 1 << 30, sched)
    restart_file = files[3]  # t = 0.012
    print(f"[io] restart from {restart_file}", flush=True)
    sched2 = hio.OutputSchedule(directory=str(tmp_path / "restarted"), prefix="kh", time_frequency=FREQ)
    sol = make_solver(s)
    hdr = sol.restart(restart_file)
    assert hdr.output_id == 3 and sol.iter == hdr.iter and sol.t == hdr.current_time
    files2 = sol.simulate(T_END, 1 << 30, sched2)
    want, o = _oracle_outputs(s, sched2, restart=hio.read_state(sol.grid, restart_file))
    assert len(files2) == len(want) == 3  # no initial output on restart: t = 0.016, 0.020, t_end
    for path, w in zip(files2, want):
        _check_file(path, sol.grid, w, s)
    _check_xdmf(tmp_path / "restarted" / "kh.xmf", files2, sol.grid)  # the restarted run indexes its own outputs
    for name, g in (("rho", sol.rho), ("rhou", sol.rhou), ("E", sol.E)):
        got = g.cpu().numpy()
        assert np.array_equal(got, o.state(name)), name
        # E was rebuilt from P at the restart: the continued run stays at round-off from the full one
        assert rel_l1(got, {"rho": full.rho, "rhou": full.rhou, "E": full.E}[name].cpu().numpy()) <= 1e-13, name


def _static_gravity_column():
    s = setups.sod(32)
    s.bc = ["Periodic", "Periodic"]
    s.gravity = (-1.0, 0.0, 0.0)
    s.rho[...] = 1.0
    s.u[...] = 0.0
    s.P[...] = 1e-10  # the first step's kinetic energy (g dt)^2 / 2 exceeds the internal energy
    return s


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "operators"])
def test_internal_energy_check_runs_on_the_state_each_iteration_starts_from(cuda, tmp_path, fused):
    s = _static_gravity_column()
    # the state turns bad in the step max_iter = 1 allows: the run ends normally (main.cpp:397 checks
    # the state an iteration starts from, and the loop exits after the last step)
    sol = make_solver(s, fused=fused)
    sol.init_from_prim(s.rho, s.u, s.P)
    assert sol.run(1e30, 1) == 1 and sol.iter == 1
    # one more iteration allowed: it starts from the bad state and stops at iteration 1
    sol = make_solver(s, fused=fused)
    sol.init_from_prim(s.rho, s.u, s.P)
    with pytest.raises(RuntimeError, match="iteration = 1: detected invalid volumic internal energy"):
        sol.run(1e30, 2)
    assert sol.iter == 1
    # bad from the start: iteration 0, after an output of that state (main.cpp:399-424)
    s.P[...] = -1.0
    sol = make_solver(s, fused=fused)
    sol.init_from_prim(s.rho, s.u, s.P)
    sched = hio.OutputSchedule(directory=str(tmp_path), prefix="bad")
    with pytest.raises(RuntimeError, match="iteration = 0: detected invalid volumic internal energy"):
        sol.simulate(1e30, 10, sched)
    hdr = hio.read_state(sol.grid, str(tmp_path / "bad_00000000.h5"))[0]
    assert hdr.iter == 0 and hdr.output_id == 0 and sol.iter == 0 and sol.t == 0.0
    assert not (tmp_path / "bad.xmf").exists()  # the output before the error joins no XDMF file
