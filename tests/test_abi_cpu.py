"""CPU-side checks of the product: the C-ABI library loads and exports every
symbol include/*.h declares, argument validation happens before any device
work, and the host mirror / setups / bench front-end behave (no GPU needed)."""

import ctypes
import glob
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import heraclespp_amd as H
from heraclespp_amd import _lib, setups
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="hclpp_hip.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"^\s*(?:int|const char\s*\*|void)\s+(hclpp_\w+)\s*\(", src, re.M))


def test_every_header_is_checked():
    assert {os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h"))} == {"hclpp_hip.h", "hclpp_io.h"}


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == declared


def test_io_library_exports_every_declared_symbol():
    from heraclespp_amd import io as hio

    L = hio.io_lib()
    declared = _declared("hclpp_io.h")
    assert len(declared) == 5
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_config_matches_header_layout():
    # hclpp_config: 7 ints, n[3], bc[6], bc_order[6], 2 doubles, g[3], 1 double, x[3] pointers, 1 int
    assert ctypes.sizeof(_lib.Config) == 4 * 22 + 8 * 6 + 8 * 3 + 8  # trailing int padded to 8
    assert ctypes.sizeof(_lib.Range) == 24


def _cfg(**kw):
    c = _lib.Config()
    c.ndim = kw.get("ndim", 1)
    c.gamma = kw.get("gamma", 1.4)
    c.mmw = kw.get("mmw", 1.0)
    c.ng = 2
    c.n[0], c.n[1], c.n[2] = kw.get("n", (8, 1, 1))
    c.limiter = kw.get("limiter", 1)
    c.riemann = kw.get("riemann", 0)
    c.precision = kw.get("precision", 0)
    for f in range(6):
        c.bc_order[f] = kw.get("bc_order", list(range(6)))[f]
    x = [np.zeros(20) for _ in range(3)]
    c._keep = x
    for d in range(3):
        c.x[d] = x[d].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return c


@pytest.mark.parametrize("kw,msg", [
    (dict(gamma=1.0), "Invalid gamma"),
    (dict(gamma=float("nan")), "Invalid gamma"),
    (dict(mmw=0.0), "Invalid mmw"),
    (dict(ndim=4), "ndim"),
    (dict(limiter=7), "limiter"),
    (dict(riemann=5), "Riemann"),
    (dict(n=(1, 1, 1)), "ghost width"),
    (dict(bc_order=[1, 1, 0, 0, 0, 0]), "priority"),
])
def test_ctx_create_validates_before_device_work(kw, msg):
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.hclpp_ctx_create(ctypes.byref(_cfg(**kw)), ctypes.byref(h))
    assert rc == _lib.HCLPP_ERR_ARG
    assert msg in L.hclpp_last_error().decode()


def test_name_tables_follow_reference_factories():
    L = _lib.lib()
    for name, v in _lib.LIMITER.items():
        assert L.hclpp_limiter_from_name(name.encode()) == v
    for name, v in _lib.RIEMANN.items():
        assert L.hclpp_riemann_from_name(name.encode()) == v
    for name, v in _lib.BC.items():
        assert L.hclpp_bc_from_name(name.encode()) == v
    assert L.hclpp_limiter_from_name(b"Superbee") == -1
    assert L.hclpp_riemann_from_name(b"Roe") == -1
    with pytest.raises(RuntimeError, match="Unknown face reconstruction algorithm: X."):
        H.factory_face_reconstruction("X")
    with pytest.raises(RuntimeError, match="Invalid riemann solver: X."):
        H.factory_godunov_scheme("X", H.PerfectGas(1.4, 1))


@pytest.mark.parametrize("nd,geom", [(1, "Cartesian"), (3, "Cartesian"), (3, "Spherical")])
def test_grid_matches_oracle_geometry(nd, geom):
    lo, hi = ((0.1, 0.0, 0.0), (1.0, np.pi, 2 * np.pi)) if geom == "Spherical" else ((0, 0, 0), (1, 2, 3))
    g = H.Grid(nd, [12, 7, 5], 2, geom, lo, hi)
    o = orc.Oracle(nd, [12, 7, 5], geom=geom, xmin=lo, xmax=hi)
    for d in range(3):
        assert np.array_equal(g.dx[d], o.geometry(d))
        assert np.array_equal(g.x_center[d], o.geometry(3 + d))
    assert g.shape3 == o.shape3


def test_setups_initial_conditions():
    s = setups.kelvin_helmholtz(32)
    inner = s.interior()
    r = s.rho[inner]
    assert set(np.unique(r)) == {1.0, 2.0}
    assert np.all(s.P[inner] == 2.5)
    assert np.max(np.abs(s.u[1][inner])) <= 0.01
    assert np.all(s.rho[0, :2, :] == 0)  # ghosts left for the boundary fill
    s = setups.rayleigh_taylor(16)
    zc = (np.arange(20) - 2 + 0.5) / 16
    P = s.P[2:-2, 5, 5]
    rho = s.rho[2:-2, 5, 5]
    # hydrostatic: dP/dz = rho g between cells of the same fluid
    dP = np.diff(P) / np.diff(zc[2:-2])
    same = np.diff(rho) == 0
    assert np.allclose(dP[same], -0.1 * rho[:-1][same])
    s2 = setups.rayleigh_taylor(16)
    assert np.array_equal(s.u, s2.u)  # seeded perturbation is reproducible
    s = setups.sedov(16, 8, 8)
    assert np.all(s.P[2:-2, 2:-2, 2:4] > 1.0) and np.all(s.P[2:-2, 2:-2, 4:-2] == 1e-5)
    s = setups.smooth_euler(32)
    assert abs(np.mean(s.rho[s.interior()]) - 1) < 0.05


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
