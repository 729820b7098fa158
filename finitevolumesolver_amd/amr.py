"""2D AMR hierarchy over batched levels (fvs_amr_* of include/fvs.h).

Mirrors the reference's subcycled hierarchy step (SubcycleFineFirstSolver::
AdvanceHierarchy, include/fub/solver/SubcycleFineFirstSolver.hpp:114-175,
driven by RunSimulation, include/fub/RunSimulation.hpp:86-191) with the
AMReX operations the IntegratorContext calls for it
(src/AMReX/IntegratorContext.cpp:380-533): coarse-fine ghost fill with
piecewise-constant interpolation, flux-register reflux and average_down.
Every level is a PatchBatch; the ghost, register and averaging link lists
are built once in the library and every operation is a launch on the
handle's stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .batch import GHOSTS, PatchBatch
from .patch import FluxMethod


class AmrHierarchy:
    def __init__(self, flux_method: FluxMethod, domain, layouts, dx0: float, ratio: int = 2):
        """domain: (nx, ny) cells of level 0; layouts: per level a list of
        patches (i0, j0, nx, ny) in that level's cell indices; level l has
        cell width dx0 / ratio**l."""
        self._lib = L.load()
        self.flux_method = flux_method
        self.ratio = int(ratio)
        self.domain = (int(domain[0]), int(domain[1]))
        self.layouts = [[tuple(int(v) for v in p) for p in lay] for lay in layouts]
        self.dx = [dx0 / self.ratio ** l for l in range(len(layouts))]
        self.levels = [PatchBatch(flux_method, [(p[2], p[3]) for p in lay], self.dx[l])
                       for l, lay in enumerate(self.layouts)]
        lo = np.ascontiguousarray(
            np.array([(p[0], p[1]) for lay in self.layouts for p in lay], dtype=np.int32).ravel())
        dom = np.array(self.domain, dtype=np.int64)
        handles = (C.c_void_p * len(self.levels))(*[b.b.value for b in self.levels])
        a = C.c_void_p()
        L.check(self._lib.fvs_amr_create(flux_method.handle, len(self.levels), handles,
                                         lo.ctypes.data, dom.ctypes.data, self.ratio, C.byref(a)))
        self.a = a
        self.time = 0.0  # time point of every level after a whole coarse step

    def __del__(self):
        a = getattr(self, "a", None)
        if a is not None and a.value:
            self._lib.fvs_amr_destroy(a)
            self.a = None

    @property
    def nlevels(self) -> int:
        return len(self.levels)

    def cell_centres(self, level: int, p: int):
        """x, y of the interior cell centres of patch p (domain [0, nx dx0] x
        [0, ny dx0] shifted by the caller)."""
        i0, j0, nx, ny = self.layouts[level][p]
        dx = self.dx[level]
        x = (i0 + np.arange(nx) + 0.5) * dx
        y = (j0 + np.arange(ny) + 0.5) * dx
        return np.meshgrid(x, y)

    def set_state(self, cons_of) -> None:
        """cons_of(level, p) -> (4, ny, nx) conservative interior of patch p;
        written to pool 0 of every level (ghosts are filled by the steps)."""
        for l, b in enumerate(self.levels):
            pool = b.new_pool()
            for p in range(len(b.sizes)):
                v = b.patch_view(pool, p)
                v[:, GHOSTS:-GHOSTS, GHOSTS:-GHOSTS] = torch.as_tensor(
                    np.asarray(cons_of(l, p), dtype=np.float64), device=pool.device)
            b.upload(pool, 0)
        torch.cuda.synchronize()

    def interiors(self, level: int) -> list:
        """Host copies (4, ny, nx) of every patch interior of a level (pool 0)."""
        b = self.levels[level]
        pool = b.download(0).cpu()
        return [b.patch_view(pool, p)[:, GHOSTS:-GHOSTS, GHOSTS:-GHOSTS].numpy().copy()
                for p in range(len(b.sizes))]

    def fill_ghosts(self, level: int, which: int = 0) -> None:
        L.check(self._lib.fvs_amr_fill_ghosts(self.a, level, which))

    def average_down(self, fine_level: int) -> None:
        L.check(self._lib.fvs_amr_average_down(self.a, fine_level))

    def tag_cells(self, level: int, tol_rho: float, tol_p: float) -> list:
        """Refinement tags (ScalarGradientDetector with the density and
        pressure projections) of every patch of a level: (ny, nx) uint8
        arrays."""
        sizes = self.levels[level].sizes
        n = sum(nx * ny for nx, ny in sizes)
        dev = torch.zeros(max(n, 1), dtype=torch.uint8, device=self.levels[level].device)
        L.check(self._lib.fvs_amr_tag_cells(self.a, level, float(tol_rho), float(tol_p),
                                            dev.data_ptr()))
        host = dev.cpu().numpy()
        out, o = [], 0
        for nx, ny in sizes:
            out.append(host[o:o + nx * ny].reshape(ny, nx).copy())
            o += nx * ny
        return out

    def transfer_from(self, old: "AmrHierarchy") -> None:
        """Moves old's state onto this hierarchy's layout (same domain,
        ratio and number of levels): old same-level cells where they exist,
        else this hierarchy's coarser cell above (piecewise constant)."""
        L.check(self._lib.fvs_amr_transfer(self.a, old.a))

    def regrid_finest(self, tol_rho: float, tol_p: float, tile: int = 8,
                      margin: int = 1) -> "AmrHierarchy":
        """A new hierarchy whose finest level is rebuilt from the gradient
        tags of the level below it (tile_layout), with the state moved onto
        it.  The other levels keep their patches."""
        lv = self.nlevels - 2
        tags = self.tag_cells(lv, tol_rho, tol_p)
        fine = tile_layout(self.domain, self.layouts[lv], tags, lv, tile, self.ratio, margin)
        if not fine:
            raise ValueError("no tagged tile nests inside the level below; nothing to refine")
        new = AmrHierarchy(self.flux_method, self.domain, self.layouts[:-1] + [fine],
                           self.dx[0], self.ratio)
        new.transfer_from(self)
        return new

    def regrid_all(self, base: int, tol_rho: float, tol_p: float, tile: int = 8,
                   margin: int = 1) -> "AmrHierarchy":
        """GriddingAlgorithm::RegridAllFinerlevels(base)
        (src/AMReX/GriddingAlgorithm.cpp:86-108): every level finer than
        `base` is rebuilt.  Level l + 1 (l = base .. finest - 1) is laid out
        from the gradient tags of the current level l (tile_layout), nested
        in the new level l; levels 0 .. base keep their patches.  Returns a
        new hierarchy with the state moved onto it (transfer_from); the
        number of levels stays (a level whose tiles all drop keeps the tiles
        of its old layout that still nest, or raises when none does)."""
        if not 0 <= base < self.nlevels - 1:
            raise ValueError(f"base level {base} has no finer level to regrid")
        tags = {lv: self.tag_cells(lv, tol_rho, tol_p) for lv in range(base, self.nlevels - 1)}
        layouts = [list(lay) for lay in self.layouts[:base + 1]]
        for lv in range(base, self.nlevels - 1):
            fine = tile_layout(self.domain, self.layouts[lv], tags[lv], lv, tile, self.ratio,
                               margin, owner_layout=layouts[lv])
            if not fine:
                ext = (self.domain[0] * self.ratio ** lv, self.domain[1] * self.ratio ** lv)
                fine = [p for p in self.layouts[lv + 1]
                        if nests(p, layouts[lv], ext, self.ratio, margin)]
            if not fine:
                raise ValueError(f"no patch of level {lv + 1} nests inside the new level {lv}")
            layouts.append(fine)
        new = AmrHierarchy(self.flux_method, self.domain, layouts, self.dx[0], self.ratio)
        new.transfer_from(self)
        new.cycle = self.cycle
        new.time = self.time
        return new

    def _adopt(self, new: "AmrHierarchy") -> None:
        """Takes over new's levels and library hierarchy (with this
        hierarchy's graph setting); new keeps (and frees) the old ones."""
        for k in ("a", "layouts", "levels", "dx", "time"):
            mine, theirs = getattr(self, k), getattr(new, k)
            setattr(self, k, theirs)
            setattr(new, k, mine)
        self.set_graphs(getattr(self, "_graphs", True))
        self.set_sequencing(getattr(self, "_sequencing", "reference"))

    def write_hdf5(self, path, time: float, cycle: int, xlower=(0.0, 0.0), box=None) -> None:
        """Appends (or starts) the HDF5 record of the composite state on the
        finest level's domain (finitevolumesolver_amd.output), or on the
        finest-level cell box ((i0, j0), (i1, j1)) inclusive, as the
        reference's WriteHdf5RestrictedToBox (src/AMReX/output/
        WriteHdf5Impl.cpp:100-175) does; xlower stays the domain's."""
        from . import output
        fab = output.composite(self.domain, self.layouts,
                               [self.interiors(l) for l in range(self.nlevels)], self.ratio,
                               self.flux_method.equation.gamma)
        if box is not None:
            (i0, j0), (i1, j1) = box
            fab = fab[:, j0:j1 + 1, i0:i1 + 1]
        dx = self.dx[-1]
        output.write_hdf5(path, fab, xlower, (dx, dx), time, cycle)

    SEQUENCINGS = {"reference": 0, "refill": 1}

    def set_sequencing(self, mode: str = "reference") -> None:
        """Ghost-layer sequencing of a level step: "reference" fills the
        ghost layers once per level step, lets the first split sweep advance
        the ghost rows across it and applies only the physical boundary
        before the second direction (DimensionalSplitLevelIntegrator,
        include/fub/solver/DimensionalSplitLevelIntegrator.hpp:169-190,
        230-233); "refill" refills every ghost layer before each sweep."""
        L.check(self._lib.fvs_amr_set_sequencing(self.a, self.SEQUENCINGS[mode]))
        self._sequencing = mode

    def set_graphs(self, on: bool = True) -> None:
        """Replay each coarse step from a captured HIP graph (default) or
        launch it eagerly."""
        L.check(self._lib.fvs_amr_set_graphs(self.a, 1 if on else 0))
        self._graphs = bool(on)

    def stable_dt(self) -> float:
        out = C.c_double()
        L.check(self._lib.fvs_amr_stable_dt(self.a, C.byref(out)))
        return out.value

    def advance(self, dt: float) -> None:
        L.check(self._lib.fvs_amr_advance(self.a, float(dt)))
        self.time += float(dt)

    @property
    def cycle(self) -> int:
        """Coarse steps taken (selects the next step's split order)."""
        out = C.c_int64()
        L.check(self._lib.fvs_amr_cycle(self.a, C.byref(out)))
        return out.value

    @cycle.setter
    def cycle(self, value: int) -> None:
        L.check(self._lib.fvs_amr_set_cycle(self.a, int(value)))

    def level_cfl(self) -> np.ndarray:
        """Per level, the CFL check value of the last advance's sweeps: the
        minimum stable dt where it lies below the level's time step (the
        reference's TimeStepTooLarge; run() retries with it), otherwise a
        value of at least that time step."""
        out = np.zeros(self.nlevels)
        L.check(self._lib.fvs_amr_level_cfl(self.a, out.ctypes.data))
        return out

    def run(self, steps: int, cfl: float = 0.8, max_restarts: int = 64,
            regrid_every: int = 0, tol_rho: float = 0.1, tol_p: float = 0.1,
            final_time: float = float("inf"), output_times=(), output=None,
            smallest_time_step_size: float = 1e-12) -> list:
        """RunSimulation-style loop (include/fub/RunSimulation.hpp:86-191):
        at most `steps` coarse steps while t + smallest_time_step_size <
        final_time; dt = min(final_time - t, next output time - t, cfl *
        stable dt) (limited_dt, :125-131), one hierarchy step, then each
        level's sweeps' stable dt checked against the level's own time step.
        A failed check is the reference's TimeStepTooLarge: the step is
        undone (the state before it is kept on the device, like the
        reference's data / scratch pair) and retried with cfl times the
        stable dt the sweeps found (capped the same way); more than
        max_restarts failures in a row raise L.TimeStepTooLarge.  Undoing a
        step restores the cycle count and time too (the reference restores
        its whole solver, RunSimulation.hpp:151), so the retry runs in the
        split order of the failed attempt.  output(hierarchy) is called when
        the run reaches each of output_times.
        With regrid_every = k > 0, the levels finer than 0 are rebuilt from
        gradient tags (regrid_all(0, tol_rho, tol_p)) before every k-th
        coarse step after the first, as the reference's PreAdvanceLevel
        regrids all finer levels at the start of a coarse step
        (src/AMReX/IntegratorContext.cpp:600-626; its regrid_frequency
        option, include/fub/AMReX/IntegratorContext.hpp:54).
        Returns the accepted dts."""
        eps = smallest_time_step_size
        outs = sorted(float(x) for x in output_times if x > self.time + eps)
        dts = []

        def limited(dt):
            cap = final_time - self.time
            if outs:
                cap = min(cap, outs[0] - self.time)
            return min(cap, dt)

        for step in range(steps):
            if not self.time + eps < final_time:
                break
            if regrid_every > 0 and step > 0 and step % regrid_every == 0:
                self._adopt(self.regrid_all(0, tol_rho, tol_p))
            dt = limited(cfl * self.stable_dt())
            saved = [b.download(0) for b in self.levels]
            cycle, time = self.cycle, self.time
            for attempt in range(max_restarts + 1):
                self.advance(dt)
                got = self.level_cfl()
                limit = min(got[l] * self.ratio ** l for l in range(self.nlevels))
                if all(got[l] >= dt / self.ratio ** l for l in range(self.nlevels)):
                    break
                for b, pool in zip(self.levels, saved):
                    b.upload(pool, 0)
                self.cycle, self.time = cycle, time
                dt = limited(cfl * limit)
            else:
                raise L.TimeStepTooLarge(L.FVS_TIME_STEP_TOO_LARGE,
                                         f"{max_restarts} restarts of one coarse step")
            dts.append(dt)
            while outs and self.time + eps >= outs[0]:
                if output is not None:
                    output(self)
                outs.pop(0)
        return dts

def nests(patch, coarse_layout, extent, ratio: int = 2, margin: int = 1) -> bool:
    """Whether a patch (i0, j0, nx, ny) of level l + 1, grown by `margin`
    cells of level l and clamped to level l's extent (nx_l, ny_l), lies
    inside the patches of level l (coarse_layout): the nesting rule
    tile_layout keeps."""
    i0, j0, nx, ny = patch
    ci0 = max(i0 // ratio - margin, 0)
    cj0 = max(j0 // ratio - margin, 0)
    ci1 = min(-(-(i0 + nx) // ratio) + margin, extent[0])
    cj1 = min(-(-(j0 + ny) // ratio) + margin, extent[1])
    cover = np.zeros((cj1 - cj0, ci1 - ci0), dtype=bool)
    for (a0, b0, an, bn) in coarse_layout:
        x0, x1 = max(a0, ci0), min(a0 + an, ci1)
        y0, y1 = max(b0, cj0), min(b0 + bn, cj1)
        if x0 < x1 and y0 < y1:
            cover[y0 - cj0:y1 - cj0, x0 - ci0:x1 - ci0] = True
    return bool(cover.all())


def tile_layout(domain, layout, tags, level: int, tile: int = 8, ratio: int = 2,
                margin: int = 1, max_run: int = 8, owner_layout=None) -> list:
    """Patches of level + 1 from the tags of a level's patches, on a fixed
    tile lattice (the GPU analogue of clustering the tagged cells into
    boxes): a tile (tile x tile cells of the level, aligned to multiples of
    `tile`) is refined when it holds a tagged cell and, grown by `margin`
    cells and clamped to the domain, lies inside the level's patches (proper
    nesting: the finer patches' ghost layers and the coarse cells beside them
    then have coarse data).  Tiles kept in one tile row merge into runs of
    at most `max_run` tiles.  The tags belong to the patches of `layout`;
    nesting is checked against `owner_layout` (default: the same layout),
    the level's new patches when it is being regridded too.  Returns
    (i0, j0, nx, ny) in level + 1 indices."""
    nx_l = domain[0] * ratio ** level
    ny_l = domain[1] * ratio ** level
    owner = np.zeros((ny_l, nx_l), dtype=bool)
    tagged = np.zeros((ny_l, nx_l), dtype=bool)
    for (i0, j0, nx, ny), t in zip(layout, tags):
        tagged[j0:j0 + ny, i0:i0 + nx] |= t.astype(bool)
    for (i0, j0, nx, ny) in (layout if owner_layout is None else owner_layout):
        owner[j0:j0 + ny, i0:i0 + nx] = True
    tx, ty = -(-nx_l // tile), -(-ny_l // tile)
    keep = np.zeros((ty, tx), dtype=bool)
    for b in range(ty):
        for a in range(tx):
            x0, y0 = a * tile, b * tile
            if not tagged[y0:y0 + tile, x0:x0 + tile].any():
                continue
            gx0, gy0 = max(x0 - margin, 0), max(y0 - margin, 0)
            gx1, gy1 = min(x0 + tile + margin, nx_l), min(y0 + tile + margin, ny_l)
            if x0 + tile > nx_l or y0 + tile > ny_l:
                continue
            keep[b, a] = owner[gy0:gy1, gx0:gx1].all()
    out = []
    f = tile * ratio
    for b in range(ty):
        a = 0
        while a < tx:
            if not keep[b, a]:
                a += 1
                continue
            run = 1
            while a + run < tx and keep[b, a + run] and run < max_run:
                run += 1
            out.append((a * f, b * f, run * f, f))
            a += run
    return out
