"""Batched sweeps over all patches of an AMR level (fvs_batch_* of include/fvs.h).

The reference sweeps a level patch by patch (ForEachFab,
include/fub/AMReX/ForEachFab.hpp:48-57); here all patches of uneven size live
in one device pool and each split sweep is one persistent launch over a tile
list of every patch, with an autotuned tile shape and a hipGraph-captured
split step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .patch import FluxMethod

GHOSTS = 2


class PatchBatch:
    def __init__(self, flux_method: FluxMethod, sizes, dx):
        """sizes: list of (nx, ny); dx: one cell width per patch (or a scalar)."""
        self._lib = L.load()
        self.flux_method = flux_method
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        n = np.ascontiguousarray(np.array(self.sizes, dtype=np.int32).ravel())
        dxs = np.broadcast_to(np.asarray(dx, dtype=np.float64), (len(self.sizes),)).copy()
        b = C.c_void_p()
        L.check(self._lib.fvs_batch_create(flux_method.handle, len(self.sizes), n.ctypes.data,
                                           dxs.ctypes.data, C.byref(b)))
        self.b = b
        self.pool_size = self._lib.fvs_batch_pool_size(b)
        self.offsets = [self._lib.fvs_batch_offset(b, p) for p in range(len(self.sizes))]
        self.device = torch.device("cuda", torch.cuda.current_device())

    def __del__(self):
        b = getattr(self, "b", None)
        if b is not None and b.value:
            self._lib.fvs_batch_destroy(b)
            self.b = None

    # layout -------------------------------------------------------------------
    def patch_view(self, pool: torch.Tensor, p: int) -> torch.Tensor:
        """(4, ny + 4, nx + 4) view of patch p inside a pool-shaped tensor."""
        nx, ny = self.sizes[p]
        size = 4 * (nx + 2 * GHOSTS) * (ny + 2 * GHOSTS)
        o = self.offsets[p]
        return pool[o:o + size].view(4, ny + 2 * GHOSTS, nx + 2 * GHOSTS)

    def new_pool(self) -> torch.Tensor:
        return torch.zeros(self.pool_size, dtype=torch.float64, device=self.device)

    def upload(self, pool: torch.Tensor, which: int = 0) -> None:
        torch.cuda.synchronize()
        L.check(self._lib.fvs_batch_upload(self.b, which, pool.data_ptr()))

    def download(self, which: int = 0) -> torch.Tensor:
        out = self.new_pool()
        torch.cuda.synchronize()
        L.check(self._lib.fvs_batch_download(self.b, which, out.data_ptr()))
        return out

    # work ---------------------------------------------------------------------
    def set_dt(self, dt: float) -> None:
        L.check(self._lib.fvs_batch_set_dt(self.b, float(dt)))

    def fill_ghosts(self, which: int = 0) -> None:
        L.check(self._lib.fvs_batch_fill_ghosts(self.b, which))

    def sweep(self, dir: int, src: int = 0, dt_out: torch.Tensor | None = None) -> None:
        L.check(self._lib.fvs_batch_sweep(self.b, dir, src,
                                          dt_out.data_ptr() if dt_out is not None else None))

    def tune(self, dir: int, reps: int = 5):
        shape = C.c_int()
        ms = (C.c_float * 4)()
        L.check(self._lib.fvs_batch_tune(self.b, dir, reps, C.byref(shape), ms))
        return shape.value, list(ms)

    def set_shape(self, dir: int, shape: int) -> None:
        L.check(self._lib.fvs_batch_set_shape(self.b, dir, shape))

    def tile_shape(self, dir: int):
        tx, ty, nt = C.c_int(), C.c_int(), C.c_int()
        L.check(self._lib.fvs_batch_tile_shape(self.b, dir, C.byref(tx), C.byref(ty),
                                               C.byref(nt)))
        return tx.value, ty.value, nt.value

    def capture(self) -> None:
        L.check(self._lib.fvs_batch_capture(self.b))

    def replay(self, steps: int = 1) -> None:
        L.check(self._lib.fvs_batch_replay(self.b, steps))

    def cfl(self) -> np.ndarray:
        """CFL check values of the last replayed step's X and Y sweeps (the
        minimum stable dt where it lies below the time step, otherwise a
        value of at least the time step)."""
        out = np.zeros(2)
        L.check(self._lib.fvs_batch_read_cfl(self.b, out.ctypes.data))
        return out
