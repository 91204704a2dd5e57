"""Raw binary output in the reference's per-rank format (BinaryWriter, io/io_binary.cpp:3-75):

  results/mesh_p_<rank>.dat           x, y, z of every interior cell (cell centres as
                                      Grid computes them, grid.cpp:40-56), i slowest, k fastest
  results/fields_<iter>_p_<rank>.dat  rho, p, u, v, w of every interior cell, same order

native-endian float64.  A block written here is rank 0 of a one-rank run.
"""
import os

import numpy as np

NGC = 3


def _coords(case):
    out = []
    for n, L in zip((case["nx"], case["ny"], case["nz"]), case["L"]):
        d = L / n
        out.append(d * 0.5 + np.arange(n) * d)
    return np.meshgrid(*out, indexing="ij")


def write_mesh(case, directory, rank=0):
    path = os.path.join(directory, "results", f"mesh_p_{rank}.dat")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.stack(_coords(case), axis=-1).astype(np.float64).tofile(path)
    return path


def write_fields(W, directory, it, rank=0):
    """W: primitive state, padded (nx+6, ny+6, nz+6, 5) as Solver.get_state returns it."""
    path = os.path.join(directory, "results", f"fields_{it}_p_{rank}.dat")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.ascontiguousarray(W[NGC:-NGC, NGC:-NGC, NGC:-NGC, :], dtype=np.float64).tofile(path)
    return path


def read_fields(path, case):
    """The interior primitive state (nx, ny, nz, 5) of a fields file."""
    return np.fromfile(path, dtype=np.float64).reshape(case["nx"], case["ny"], case["nz"], 5)
