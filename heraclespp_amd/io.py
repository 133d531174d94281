"""Output and restart files of the hydro state in the reference's HDF5 layout.

ctypes binding of libhclpp_io.so (include/hclpp_io.h): write_pdi (src/io/io.cpp:185-324) and
read_pdi (:326-431) with the decl_hdf5 layout of src/io/config.yaml.cpp:53-239 -- one file per
output, <directory>/<prefix>_<output_id:08d>.h5, with the header scalars (output_id,
iter_output_id, time_output_id, iter, current_time, gamma), the node coordinates (x0, x1, x2 with
ghosts; x0_ng, x1_ng, x2_ng interior) and the interior of rho, ux0..ux<ndim-1>, P, E, T and
fx0..fx<nfx-1>, shape [nz, ny, nx].

The device state is copied to the host before writing, as write_pdi does through host views; the
temperature is computed on the GPU first (analysis/temperature.hpp, hclpp_temperature).

write_xdmf writes the XDMF side file <directory>/<prefix>.xmf that indexes the outputs of a run
(XmlWriter, io.cpp:433-566), for viewers that read XDMF.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(HERE, "libhclpp_io.so")


class IoGrid(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("nfx", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("ng", ctypes.c_int),
        ("x", ctypes.POINTER(ctypes.c_double) * 3),
    ]


class IoHeader(ctypes.Structure):
    _fields_ = [
        ("output_id", ctypes.c_int),
        ("iter_output_id", ctypes.c_int),
        ("time_output_id", ctypes.c_int),
        ("iter", ctypes.c_int),
        ("current_time", ctypes.c_double),
        ("gamma", ctypes.c_double),
    ]


_io = None


def io_lib() -> ctypes.CDLL:
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IO_LIB_PATH):
        raise ImportError(f"{IO_LIB_PATH} is not built: run `python -m heraclespp_amd.build --io` (needs the HDF5 C library)")
    L = ctypes.CDLL(IO_LIB_PATH)
    vp = ctypes.c_void_p
    L.hclpp_io_last_error.restype = ctypes.c_char_p
    L.hclpp_io_last_error.argtypes = []
    L.hclpp_io_output_filename.restype = ctypes.c_int
    L.hclpp_io_output_filename.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.hclpp_io_write.restype = ctypes.c_int
    L.hclpp_io_write.argtypes = [ctypes.POINTER(IoGrid), ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(IoHeader)] + [vp] * 6
    L.hclpp_io_read_field.restype = ctypes.c_int
    L.hclpp_io_read_field.argtypes = [ctypes.POINTER(IoGrid), ctypes.c_char_p, ctypes.c_char_p, vp]
    L.hclpp_io_read.restype = ctypes.c_int
    L.hclpp_io_read.argtypes = [ctypes.POINTER(IoGrid), ctypes.c_char_p, ctypes.POINTER(IoHeader)] + [vp] * 7
    _io = L
    return L


class HdfIoError(RuntimeError):
    pass


def _check(rc: int):
    if rc != 0:
        raise HdfIoError(io_lib().hclpp_io_last_error().decode())


def output_filename(prefix: str, output_id: int) -> str:
    """io.cpp:62-74: prefix + '_' + output_id on 8 digits + '.h5'."""
    buf = ctypes.create_string_buffer(4096)
    _check(io_lib().hclpp_io_output_filename(prefix.encode(), int(output_id), buf, len(buf)))
    return buf.value.decode()


def _io_grid(grid, nfx: int, keep: list) -> IoGrid:
    g = IoGrid()
    g.ndim = grid.ndim
    g.nfx = nfx
    g.ng = grid.Ng
    for d in range(3):
        g.n[d] = grid.Nx_glob_ng[d]
        x = np.ascontiguousarray(grid.x[d], dtype=np.float64)
        keep.append(x)
        g.x[d] = x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return g


def _host(t) -> np.ndarray:
    if isinstance(t, torch.Tensor):
        return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())
    return np.ascontiguousarray(t, dtype=np.float64)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


@dataclass
class Header:
    output_id: int
    iter_output_id: int
    time_output_id: int
    iter: int
    current_time: float
    gamma: float


def write_state(grid, eos, directory: str, prefix: str, header: Header, rho, u, P, E, fx=None, T=None) -> str:
    """write_pdi (io.cpp:185-324) of a state given as device tensors or host arrays over the
    ghost-inclusive extents.  T is computed from rho and P over all ghosts (main.cpp:316, :534) when
    not given.  Returns the file path."""
    nfx = 0 if fx is None else int(fx.shape[0])
    if T is None:
        T = temperature(grid, eos, rho, P)
    keep: list = []
    g = _io_grid(grid, nfx, keep)
    h = IoHeader(header.output_id, header.iter_output_id, header.time_output_id, header.iter, header.current_time, header.gamma)
    arrs = [_host(a) for a in (rho, u, P, E, T)]
    fxa = _host(fx) if nfx else None
    os.makedirs(directory, exist_ok=True)
    _check(io_lib().hclpp_io_write(ctypes.byref(g), directory.encode(), prefix.encode(), ctypes.byref(h),
                                   *[_ptr(a) for a in arrs], _ptr(fxa) if nfx else None))
    return os.path.join(directory, output_filename(prefix, header.output_id))


def read_state(grid, filename: str, nfx: int = 0):
    """read_pdi (io.cpp:326-431): the header and host arrays (ghost-inclusive extents, ghosts zero)
    of rho, u, P and fx, plus the node coordinates x0, x1, x2 the file holds (their extents must
    match the grid's, io.cpp:364-366)."""
    keep: list = []
    g = _io_grid(grid, nfx, keep)
    shape3 = grid.shape3
    rho = np.zeros(shape3)
    u = np.zeros((grid.ndim,) + shape3)
    P = np.zeros(shape3)
    fx = np.zeros((nfx,) + shape3) if nfx else None
    xs = [np.zeros(grid.Nx_local_wg[d] + 1) for d in range(3)]
    h = IoHeader()
    _check(io_lib().hclpp_io_read(ctypes.byref(g), filename.encode(), ctypes.byref(h), _ptr(rho), _ptr(u), _ptr(P),
                                  _ptr(fx) if nfx else None, *[_ptr(x) for x in xs]))
    hdr = Header(h.output_id, h.iter_output_id, h.time_output_id, h.iter, h.current_time, h.gamma)
    return hdr, rho, u, P, fx, xs


def read_field(grid, filename: str, dataset: str) -> np.ndarray:
    """One interior dataset of an output file (rho, ux0, P, E, T, fx0, ...) as a host array over the
    ghost-inclusive extents (ghosts zero)."""
    keep: list = []
    g = _io_grid(grid, 0, keep)
    a = np.zeros(grid.shape3)
    _check(io_lib().hclpp_io_read_field(ctypes.byref(g), filename.encode(), dataset.encode(), _ptr(a)))
    return a


def temperature(grid, eos, rho, P):
    """analysis/temperature.hpp:18-26 over all ghosts (on the GPU; returns a device tensor)."""
    ctx = grid.context(eos)
    rho_d = rho if isinstance(rho, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(rho)).cuda()
    P_d = P if isinstance(P, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(P)).cuda()
    T = torch.zeros_like(rho_d)
    r = grid.range.all_ghosts()._c()
    _lib.check(_lib.lib().hclpp_temperature(ctx.handle, ctypes.byref(r), ctx._chk(rho_d, (), "rho"), ctx._chk(P_d, (), "P"),
                                            ctx._chk(T, (), "T")))
    return T


def xdmf_variables(ndim: int, nfx: int) -> list:
    """Cell variables the XDMF file lists, in the reference's order (io.cpp:433-445): rho, P, E, T,
    the passive scalars, then the velocity components."""
    return ["rho", "P", "E", "T"] + [f"fx{f}" for f in range(nfx)] + [f"ux{d}" for d in range(ndim)]


def write_xdmf(directory: str, prefix: str, n, ndim: int, nfx: int, output_id: int, outputs_record) -> str:
    """The XDMF side file <directory>/<prefix>.xmf of XmlWriter (io.cpp:447-566): a temporal collection
    with one rectilinear grid per recorded output, whose node coordinates are the x0_ng, x1_ng and
    x2_ng datasets and whose cell attributes are the datasets of its HDF5 file.  outputs_record holds
    (iter, t) of the outputs written since the start of the run (restart included), the last one being
    output_id; the file is rewritten whole at every output, as the reference does (main.cpp:313-322,
    :533-538).  n: interior cells per dim (3 values)."""
    n = [int(v) for v in (list(n) + [1, 1, 1])[:3]]
    lines = []
    level = 0

    def put(text):
        lines.append("  " * level + text)

    def item(dims, ref):
        nonlocal level
        put(f'<DataItem NumberType="Float" Precision="8" Dimensions="{dims}" Format="HDF">')
        level += 1
        put(ref)
        level -= 1
        put("</DataItem>")

    put('<?xml version="1.0"?>')
    put('<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>')
    put('<Xdmf Version="2.0">')
    level = 1
    put("<Domain>")
    level = 2
    put('<Grid Name="TimeSeries" GridType="Collection" CollectionType="Temporal">')
    first = output_id + 1 - len(outputs_record)
    for i, (_, t) in enumerate(outputs_record):
        fname = output_filename(prefix, first + i)
        level = 3
        put('<Grid Name="output" GridType="Uniform">')
        level = 4
        put(f'<Time Value="{t:g}"/>')
        put(f'<Topology TopologyType="3DRectMesh" Dimensions="{n[2] + 1} {n[1] + 1} {n[0] + 1}"/>')
        put('<Geometry GeometryType="VXVYVZ">')
        for d in range(3):
            level = 5
            item(n[d] + 1, f"{fname}:/x{d}_ng")
        level = 4
        put("</Geometry>")
        for v in xdmf_variables(ndim, nfx):
            put(f'<Attribute Center="Cell" Name="{v}" AttributeType="Scalar">')
            level = 5
            item(f"{n[2]} {n[1]} {n[0]}", f"{fname}:/{v}")
            level = 4
            put("</Attribute>")
        level = 3
        put("</Grid>")
    level = 2
    put("</Grid>")
    level = 1
    put("</Domain>")
    level = 0
    put("</Xdmf>")
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, prefix + ".xmf")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


@dataclass
class OutputSchedule:
    """The [Output] section of the reference's input file (utils/param.cpp:44-49): iteration
    outputs every iter_frequency steps, time outputs every time_frequency from time_first_output
    (default t_ini + time_frequency), files <directory>/<prefix>_<id>.h5."""

    directory: str = "."
    prefix: str = "result"
    iter_frequency: int = 0
    time_frequency: float = 0.0
    time_first_output: float | None = None

    def first_output(self, t_ini: float) -> float:
        return t_ini + self.time_frequency if self.time_first_output is None else float(self.time_first_output)

    @property
    def enabled(self) -> bool:
        return self.iter_frequency > 0 or self.time_frequency > 0
