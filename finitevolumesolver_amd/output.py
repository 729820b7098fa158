"""HDF5 output of an AMR hierarchy in the reference's layout.

The reference flattens the hierarchy onto the finest level's whole domain
(WriteHdf5UnRestricted, src/AMReX/output/WriteHdf5Impl.cpp:26-98): each
level's patches are copied into a zero-initialised array of its domain, and
a cell whose density is still zero takes every component of the coarser
level's cell above it; the finest array is written by WriteToHDF5
(src/AMReX/PatchHierarchy.cpp:442-607) as one frame of /data with /times and
/cycles.  Here the flattening runs on the host over the downloaded patch
interiors and libfvs_hdf5.so (io/fvs_hdf5.cpp) writes the file.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfvs_hdf5.so"
FIELDS_2D = ("Density", "Momentum0", "Momentum1", "Energy", "Pressure", "SpeedOfSound")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built (needs the HDF5 C library)")
        lib = C.CDLL(str(LIB_PATH))
        lib.fvsio_hdf5_write.restype = C.c_int
        lib.fvsio_hdf5_write.argtypes = [C.c_char_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                         C.c_int64]
        _lib = lib
    return _lib


def complete_2d(cons: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    """(4, ...) conservative -> (6, ...) complete state (density, momentum,
    energy, pressure, speed of sound) with PerfectGas::CompleteFromCons's
    array-form operations (src/equations/PerfectGas.cpp:112-124); cells of
    zero density stay all zero."""
    rho, m0, m1, E = cons
    out = np.zeros((6,) + rho.shape)
    ok = rho != 0.0
    r = np.where(ok, rho, 1.0)
    e_kin = 0.5 * (m0 * m0 + m1 * m1) / r
    p = (E - e_kin) / (1.0 / (gamma - 1.0))
    with np.errstate(invalid="ignore"):
        a = np.sqrt(gamma * p / r)
    out[:4] = cons
    out[4] = np.where(ok, p, 0.0)
    out[5] = np.where(ok, a, 0.0)
    return out


def composite(domain, layouts, interiors, ratio: int = 2, gamma: float = 1.4) -> np.ndarray:
    """Finest-level (6, Ny, Nx) composite of a 2D hierarchy: interiors[l][p]
    is the (4, ny, nx) conservative interior of patch p of level l."""
    fab = None
    for l, lay in enumerate(layouts):
        nx, ny = domain[0] * ratio ** l, domain[1] * ratio ** l
        cur = np.zeros((6, ny, nx))
        for (i0, j0, w, h), u in zip(lay, interiors[l]):
            cur[:, j0:j0 + h, i0:i0 + w] = complete_2d(np.asarray(u), gamma)
        if fab is not None:
            coarse = fab.repeat(ratio, axis=1).repeat(ratio, axis=2)
            empty = cur[0] == 0.0
            cur[:, empty] = coarse[:, empty]
        fab = cur
    return fab


def write_hdf5(path, fab: np.ndarray, xlower, cell_size, time: float, cycle: int,
               fields=FIELDS_2D) -> None:
    """One frame of a (ncomp, Ny, Nx) composite: a new file, or appended to
    an existing one (same shape)."""
    lib = _load()
    fab = np.ascontiguousarray(fab, dtype=np.float64)
    ncomp, ny, nx = fab.shape
    xl = np.ascontiguousarray(xlower, dtype=np.float64)
    cs = np.ascontiguousarray(cell_size, dtype=np.float64)
    names = [f.encode() for f in fields]
    arr = (C.c_char_p * len(names))(*names)
    rc = lib.fvsio_hdf5_write(str(path).encode(), fab.ctypes.data, ncomp, nx, ny,
                              xl.ctypes.data, cs.ctypes.data, C.cast(arr, C.c_void_p),
                              len(names), float(time), int(cycle))
    if rc != 0:
        raise OSError(f"HDF5 output to {path} failed ({rc})")
