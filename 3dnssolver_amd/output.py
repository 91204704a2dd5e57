"""Output in the reference's formats, and checkpoints to resume from.

Raw binary, per rank (BinaryWriter, io/io_binary.cpp:3-75):

  results/mesh_p_<rank>.dat           x, y, z of every interior cell of the rank's block (cell
                                      centres as Grid computes them, grid.cpp:24-57), i slowest
  results/fields_<iter>_p_<rank>.dat  rho, p, u, v, w of every interior cell, same order

native-endian float64.

HDF5 (Hdf5Writer, io/hdf5_writer.cpp:6-299), one file for the whole domain:

  /mesh/{x,y,z}_coords                  [nx][ny][nz] float64, the global grid
  /checkpoint_<N>/{rho,p,u,v,w}         [nx][ny][nz] float64, the primitive state
  /checkpoint_<N>/time                  [1] float64, the flow time of checkpoint N

written through the HDF5 C library (serial; the reference writes the same datasets with MPI-IO
hyperslabs, one per rank).  The library is loaded from $NSGPU_HDF5_LIB, the system search path
or the conda tree of this image; HDF5 calls raise Hdf5Error when it is absent.

Checkpoints: the reference has no reader, but its state between steps is the primitive array
alone (TimeScheme::step re-derives U and dt from W, time_scheme.cpp:16-27), so W and the time
resume a run bitwise.  write_checkpoint stores the interior W (raw float64) with a JSON record
of time, step, last dt and the grid; the HDF5 checkpoints above resume the same way.
"""
import ctypes
import ctypes.util
import json
import os

import numpy as np

NGC = 3


def _coords(case, start=(0, 0, 0), count=None):
    """Cell centres of a block: (x0 + dx * 0.5) + i * dx, x0 = dx * start (grid.cpp:24-57)."""
    count = count or (case["nx"], case["ny"], case["nz"])
    out = []
    for n, L, s, c in zip((case["nx"], case["ny"], case["nz"]), case["L"], start, count):
        d = L / n
        out.append((d * s + d * 0.5) + np.arange(c) * d)
    return np.meshgrid(*out, indexing="ij")


def write_mesh(case, directory, rank=0, start=(0, 0, 0), count=None):
    path = os.path.join(directory, "results", f"mesh_p_{rank}.dat")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.stack(_coords(case, start, count), axis=-1).astype(np.float64).tofile(path)
    return path


def write_fields(W, directory, it, rank=0):
    """W: primitive state of the rank's block, padded (nx+6, ny+6, nz+6, 5)."""
    path = os.path.join(directory, "results", f"fields_{it}_p_{rank}.dat")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.ascontiguousarray(W[NGC:-NGC, NGC:-NGC, NGC:-NGC, :], dtype=np.float64).tofile(path)
    return path


def read_fields(path, case, count=None):
    """The interior primitive state (nx, ny, nz, 5) of a fields file."""
    count = count or (case["nx"], case["ny"], case["nz"])
    return np.fromfile(path, dtype=np.float64).reshape(*count, 5)


# ---- checkpoints ---------------------------------------------------------------------------
def write_checkpoint(directory, case, W, time, step, dt):
    """Raw checkpoint of a run: the interior of the padded primitive state W (float64, i slowest)
    and {time, step, dt, grid} as JSON.  Returns the path of the state file."""
    os.makedirs(directory, exist_ok=True)
    base = os.path.join(directory, f"checkpoint_{step}")
    np.ascontiguousarray(W[NGC:-NGC, NGC:-NGC, NGC:-NGC, :], dtype=np.float64).tofile(base + ".dat")
    meta = dict(time=float(time), step=int(step), dt=float(dt),
                nx=case["nx"], ny=case["ny"], nz=case["nz"], L=list(case["L"]))
    with open(base + ".json", "w") as f:
        json.dump(meta, f)
    return base + ".dat"


def read_checkpoint(path, case):
    """(padded primitive state, time, step, dt) of a raw checkpoint written for `case`."""
    base = path[:-4] if path.endswith(".dat") else path
    with open(base + ".json") as f:
        meta = json.load(f)
    if (meta["nx"], meta["ny"], meta["nz"]) != (case["nx"], case["ny"], case["nz"]):
        raise ValueError("checkpoint grid differs from the case")
    W = np.zeros((case["nx"] + 6, case["ny"] + 6, case["nz"] + 6, 5))
    W[NGC:-NGC, NGC:-NGC, NGC:-NGC] = np.fromfile(base + ".dat", dtype=np.float64).reshape(
        case["nx"], case["ny"], case["nz"], 5)
    return W, meta["time"], meta["step"], meta["dt"]


# ---- HDF5 ----------------------------------------------------------------------------------
class Hdf5Error(RuntimeError):
    pass


_H5 = None
_H5F_ACC_RDONLY, _H5F_ACC_RDWR, _H5F_ACC_TRUNC = 0, 1, 2
_H5P_DEFAULT = 0
_H5S_ALL = 0


def _h5():
    global _H5
    if _H5 is not None:
        return _H5
    cands = [os.environ.get("NSGPU_HDF5_LIB"), ctypes.util.find_library("hdf5"),
             "/opt/conda/lib/libhdf5.so"]
    L = None
    for c in cands:
        if c:
            try:
                L = ctypes.CDLL(c)
                break
            except OSError:
                continue
    if L is None:
        raise Hdf5Error("the HDF5 C library was not found (set NSGPU_HDF5_LIB)")
    hid, herr, hs = ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_ulonglong)
    L.H5open.restype = herr
    L.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid, hid]
    L.H5Fcreate.restype = hid
    L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
    L.H5Fopen.restype = hid
    L.H5Gcreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid]
    L.H5Gcreate2.restype = hid
    L.H5Screate_simple.argtypes = [ctypes.c_int, hs, hs]
    L.H5Screate_simple.restype = hid
    L.H5Dcreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]
    L.H5Dcreate2.restype = hid
    L.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
    L.H5Dopen2.restype = hid
    L.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    L.H5Dwrite.restype = herr
    L.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    L.H5Dread.restype = herr
    L.H5Dget_space.argtypes = [hid]
    L.H5Dget_space.restype = hid
    L.H5Sget_simple_extent_dims.argtypes = [hid, hs, hs]
    L.H5Sget_simple_extent_dims.restype = ctypes.c_int
    L.H5Lexists.argtypes = [hid, ctypes.c_char_p, hid]
    L.H5Lexists.restype = ctypes.c_int
    for fn in ("H5Fclose", "H5Gclose", "H5Sclose", "H5Dclose"):
        getattr(L, fn).argtypes = [hid]
        getattr(L, fn).restype = herr
    if L.H5open() < 0:
        raise Hdf5Error("H5open failed")
    L.native_double = ctypes.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    _H5 = L
    return L


def _ok(v, what):
    if v < 0:
        raise Hdf5Error(what)
    return v


def _write_dataset(L, fid, name, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    dims = (ctypes.c_ulonglong * a.ndim)(*a.shape)
    sp = _ok(L.H5Screate_simple(a.ndim, dims, None), "H5Screate_simple")
    ds = _ok(L.H5Dcreate2(fid, name.encode(), L.native_double, sp, _H5P_DEFAULT, _H5P_DEFAULT,
                          _H5P_DEFAULT), f"H5Dcreate {name}")
    _ok(L.H5Dwrite(ds, L.native_double, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, a.ctypes.data),
        f"H5Dwrite {name}")
    L.H5Dclose(ds)
    L.H5Sclose(sp)


def _read_dataset(L, fid, name):
    ds = _ok(L.H5Dopen2(fid, name.encode(), _H5P_DEFAULT), f"H5Dopen {name}")
    sp = L.H5Dget_space(ds)
    dims = (ctypes.c_ulonglong * 8)()
    nd = _ok(L.H5Sget_simple_extent_dims(sp, dims, None), "H5Sget_simple_extent_dims")
    out = np.empty(tuple(dims[:nd]), dtype=np.float64)
    _ok(L.H5Dread(ds, L.native_double, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, out.ctypes.data),
        f"H5Dread {name}")
    L.H5Sclose(sp)
    L.H5Dclose(ds)
    return out


def hdf5_write_mesh(path, case):
    """Creates (truncates) the file and writes /mesh/{x,y,z}_coords of the global grid, as
    Hdf5Writer::write does the first time (hdf5_writer.cpp:13-22)."""
    L = _h5()
    fid = _ok(L.H5Fcreate(path.encode(), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), "H5Fcreate")
    try:
        L.H5Gclose(_ok(L.H5Gcreate2(fid, b"mesh", _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT),
                       "H5Gcreate mesh"))
        for name, X in zip(("x_coords", "y_coords", "z_coords"), _coords(case)):
            _write_dataset(L, fid, f"/mesh/{name}", X)
    finally:
        L.H5Fclose(fid)


def hdf5_write_checkpoint(path, index, W, time):
    """Group /checkpoint_<index> with the flow time and rho, p, u, v, w of the interior of the
    padded global primitive state W (hdf5_writer.cpp:24-43)."""
    L = _h5()
    fid = _ok(L.H5Fopen(path.encode(), _H5F_ACC_RDWR, _H5P_DEFAULT), "H5Fopen")
    try:
        g = f"checkpoint_{index}"
        L.H5Gclose(_ok(L.H5Gcreate2(fid, g.encode(), _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT),
                       f"H5Gcreate {g}"))
        _write_dataset(L, fid, f"/{g}/time", np.array([time], dtype=np.float64))
        Wi = W[NGC:-NGC, NGC:-NGC, NGC:-NGC]
        for q, name in enumerate(("rho", "p", "u", "v", "w")):
            _write_dataset(L, fid, f"/{g}/{name}", Wi[..., q])
    finally:
        L.H5Fclose(fid)


def hdf5_read_checkpoint(path, index):
    """(padded global primitive state, time) of /checkpoint_<index>: a run resumes from it."""
    L = _h5()
    fid = _ok(L.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen")
    try:
        g = f"checkpoint_{index}"
        if L.H5Lexists(fid, g.encode(), _H5P_DEFAULT) <= 0:
            raise Hdf5Error(f"{path} has no {g}")
        fields = [_read_dataset(L, fid, f"/{g}/{n}") for n in ("rho", "p", "u", "v", "w")]
        time = float(_read_dataset(L, fid, f"/{g}/time")[0])
    finally:
        L.H5Fclose(fid)
    nx, ny, nz = fields[0].shape
    W = np.zeros((nx + 6, ny + 6, nz + 6, 5))
    W[NGC:-NGC, NGC:-NGC, NGC:-NGC] = np.stack(fields, -1)
    return W, time


def hdf5_read_mesh(path):
    L = _h5()
    fid = _ok(L.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen")
    try:
        return [_read_dataset(L, fid, f"/mesh/{n}") for n in ("x_coords", "y_coords", "z_coords")]
    finally:
        L.H5Fclose(fid)
