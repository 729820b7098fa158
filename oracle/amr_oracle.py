"""CPU restatement of the 2D AMR hierarchy step: test infrastructure only.

Only tests use this module.  It restates
the composition the reference drives for a subcycled hierarchy and checks
the device library against it; nothing in the product calls it.

Follows:
  * SubcycleFineFirstSolver::AdvanceLevel
    (include/fub/solver/SubcycleFineFirstSolver.hpp:114-175): reset the
    finer level's register, advance the finer level `ratio` times with
    dt / ratio, advance this level, reflux, average down.
  * DimensionalSplitLevelIntegrator::AdvanceLevelNonRecursively
    (include/fub/solver/DimensionalSplitLevelIntegrator.hpp:221-276): per
    split direction flux, update and AccumulateCoarseFineFluxes with
    scale = split dt; split order from MakeSplitDirections
    (include/fub/split_method/SplittingMethod.hpp:115-126): reversed when
    (cycle % 2 + subcycle) % 2 == 1.
  * IntegratorContext::AccumulateCoarseFineFluxes (src/AMReX/
    IntegratorContext.cpp:473-503): FineAdd with +vol * dt / dx (= face
    length * dt in 2D) on the fine level, CrseInit with -vol * dt / dx on
    the coarse level; ApplyFluxCorrection (:519-533): FluxRegister::Reflux.
  * CoarsenConservatively (:451-471): amrex::average_down, the coarse value
    (1 / ratio^2) * (sum of the fine block, rows outer, x inner).
  * FillGhostLayerTwoLevels (:380-418): FillPatchTwoLevels with pc_interp
    (a fine ghost takes the value of the coarse cell above it), same-level
    patches first, transmissive domain edges (a ghost outside the domain
    takes the value of the clamped in-domain cell).
AMReX itself is not part of the reference tree and cannot run here, so the
AMReX-side arithmetic (register sums, average_down order) is parity
unpinned beyond this restatement; the composition is pinned by the
conservation property the register exists for (tests/test_amr_oracle.py).

Two ghost-layer sequencings (`sequencing`):
  * "reference" (default): DimensionalSplitLevelIntegrator fills a level's
    ghost layers once per level step (PreAdvanceLevel,
    include/fub/solver/DimensionalSplitLevelIntegrator.hpp:169-190) and
    skips the boundary condition before the first split direction
    (:230-233); the first sweep updates its grown scratch
    (include/fub/AMReX/IntegratorContext.hpp:52-53: scratch_gcw 4,
    flux_gcw 2; src/AMReX/TimeIntegrator.cpp:30-56 updates every cell the
    flux box encloses), which advances the ghost rows across the sweep
    direction; before the second direction only the physical boundary in
    that direction is applied (ApplyBoundaryCondition -> TransmissiveBoundary,
    src/AMReX/boundary_condition/TransmissiveBoundary.cpp:60-93: a ghost
    cell outside the domain copies the patch's own cell on the domain's edge
    row).  With two ghost layers the first sweep here updates the interior
    columns of the two ghost rows on each side, the only ghost cells the
    second sweep's interior stencils read.
  * "refill": every split sweep refills all ghost layers (the device
    library's earlier sequencing, kept for comparison).

The per-patch sweep is injected: sweep(cons, dt, dx, d, grow=0) takes a
patch with two ghost layers (4, ny + 4, nx + 4) and returns (new values of
the interior grown by `grow` cells across d: (4, ny + 2 grow, nx) for d = 0,
(4, ny, nx + 2 grow) for d = 1; face fluxes along d of the interior rows /
columns (4, ny, nx + 1) or (4, ny + 1, nx); minimum stable dt of those faces).
"""
from __future__ import annotations

import numpy as np

G = 2  # ghost layers


class AmrOracle:
    def __init__(self, domain, layouts, dx0, ratio, sweep, sequencing="reference"):
        if sequencing not in ("reference", "refill"):
            raise ValueError(f"unknown sequencing {sequencing!r}")
        self.sequencing = sequencing
        self.domain = tuple(int(v) for v in domain)
        self.layouts = [[tuple(int(v) for v in p) for p in lay] for lay in layouts]
        self.r = int(ratio)
        self.dx = [dx0 / self.r ** l for l in range(len(layouts))]
        self.sweep = sweep
        self.U = [[np.zeros((4, p[3] + 2 * G, p[2] + 2 * G)) for p in lay] for lay in self.layouts]
        self.cycle = 0
        self.level_cfl = [np.inf] * len(layouts)

    # geometry -------------------------------------------------------------
    def extent(self, l):
        s = self.r ** l
        return self.domain[0] * s, self.domain[1] * s

    def find(self, l, I, J):
        """Patch of level l whose interior holds cell (I, J), or None."""
        for q, (i0, j0, nx, ny) in enumerate(self.layouts[l]):
            if i0 <= I < i0 + nx and j0 <= J < j0 + ny:
                return q
        return None

    # ghost layers -----------------------------------------------------------
    def fill_ghosts(self, l):
        Nx, Ny = self.extent(l)
        for p, (i0, j0, nx, ny) in enumerate(self.layouts[l]):
            u = self.U[l][p]
            for j in range(-G, ny + G):
                for i in range(-G, nx + G):
                    if 0 <= i < nx and 0 <= j < ny:
                        continue
                    I = min(max(i0 + i, 0), Nx - 1)
                    J = min(max(j0 + j, 0), Ny - 1)
                    q = self.find(l, I, J)
                    if q is not None:
                        a, b = self.layouts[l][q][:2]
                        u[:, j + G, i + G] = self.U[l][q][:, J - b + G, I - a + G]
                    else:
                        q = self.find(l - 1, I // self.r, J // self.r) if l > 0 else None
                        if q is None:
                            raise ValueError("ghost cell without a source")
                        a, b = self.layouts[l - 1][q][:2]
                        u[:, j + G, i + G] = self.U[l - 1][q][:, J // self.r - b + G,
                                                              I // self.r - a + G]

    def fill_physical(self, l, d):
        """TransmissiveBoundary in direction d: every ghost cell outside the
        level's domain across d takes the value of the patch's own cell in
        the domain's edge row (same index along the other direction)."""
        Nd = self.extent(l)[d]
        for p, (i0, j0, nx, ny) in enumerate(self.layouts[l]):
            u = self.U[l][p]
            lo, n = (i0, nx) if d == 0 else (j0, ny)
            for k in range(-G, n + G):
                K = lo + k
                if 0 <= K < Nd:
                    continue
                edge = (0 if K < 0 else Nd - 1) - lo
                if d == 0:
                    u[:, :, k + G] = u[:, :, edge + G]
                else:
                    u[:, k + G, :] = u[:, edge + G, :]

    # register -----------------------------------------------------------------
    def register_faces(self, f):
        """Coarse-fine faces of level f: (p, dir, side, ct, coarse patch q,
        coarse face index along dir)."""
        out = []
        Nx, Ny = self.extent(f)
        for p, (i0, j0, nx, ny) in enumerate(self.layouts[f]):
            for d in (0, 1):
                for side in (0, 1):
                    d0, nd = (i0, nx) if d == 0 else (j0, ny)
                    t0, nt = (j0, ny) if d == 0 else (i0, nx)
                    face = d0 if side == 0 else d0 + nd
                    if face == 0 or face == (Nx if d == 0 else Ny):
                        continue
                    outside = face - 1 if side == 0 else face
                    for ct in range(t0 // self.r, (t0 + nt) // self.r):
                        oi, oj = (outside, ct * self.r) if d == 0 else (ct * self.r, outside)
                        if self.find(f, oi, oj) is not None:
                            continue
                        cf = face // self.r
                        cc = cf - 1 if side == 0 else cf
                        q = self.find(f - 1, *((cc, ct) if d == 0 else (ct, cc)))
                        if q is None:
                            raise ValueError("coarse cell beside a fine patch is in no patch")
                        out.append((p, d, side, ct, q, cf))
        # reflux order of the device library: (dir, side) groups
        out.sort(key=lambda e: (e[1], e[2]))
        return out

    # one level ----------------------------------------------------------------
    def advance_level(self, l, dt, sub):
        L = len(self.layouts)
        self.fill_ghosts(l)
        finer = l + 1 < L
        if finer:
            faces = self.register_faces(l + 1)
            reg = np.zeros((len(faces), 4))
            self._reg[l + 1] = (faces, reg)
            for s in range(self.r):
                self.advance_level(l + 1, dt / self.r, s)
        rev = ((self.cycle % 2) + sub) % 2 == 1
        ref = self.sequencing == "reference"
        for w, d in enumerate((1, 0) if rev else (0, 1)):
            if w > 0:
                if ref:
                    self.fill_physical(l, d)
                else:
                    self.fill_ghosts(l)
            grow = G if (ref and w == 0) else 0
            fluxes = []
            for p, (i0, j0, nx, ny) in enumerate(self.layouts[l]):
                new, flux, sdt = self.sweep(self.U[l][p], dt, self.dx[l], d, grow)
                fluxes.append(flux)
                self.level_cfl[l] = min(self.level_cfl[l], sdt)
                self._new[p] = (new, grow, d)
            if finer:  # CrseInit: -(dx dt) F of the coarse face
                faces, reg = self._reg[l + 1]
                w_c = self.dx[l] * dt
                for e, (p, dd, side, ct, q, cf) in enumerate(faces):
                    if dd != d:
                        continue
                    a, b = self.layouts[l][q][:2]
                    fl = fluxes[q]
                    f = fl[:, ct - b, cf - a] if d == 0 else fl[:, cf - b, ct - a]
                    reg[e] = reg[e] + (-1.0) * (w_c * f)
            if l > 0:  # FineAdd: +(dx dt) F of each fine face, in order
                faces, reg = self._reg[l]
                w_f = self.dx[l] * dt
                for e, (p, dd, side, ct, q, cf) in enumerate(faces):
                    if dd != d:
                        continue
                    i0, j0, nx, ny = self.layouts[l][p]
                    fl = fluxes[p]
                    fi = (0 if side == 0 else nx) if d == 0 else (0 if side == 0 else ny)
                    for m in range(self.r):
                        t = ct * self.r - (j0 if d == 0 else i0) + m
                        f = fl[:, t, fi] if d == 0 else fl[:, fi, t]
                        reg[e] = reg[e] + 1.0 * (w_f * f)
            for p, (i0, j0, nx, ny) in enumerate(self.layouts[l]):
                new, grow, dd = self._new.pop(p)
                if dd == 0:
                    self.U[l][p][:, G - grow:G + ny + grow, G:-G] = new
                else:
                    self.U[l][p][:, G:-G, G - grow:G + nx + grow] = new
        if finer:
            self.reflux(l + 1)
            self.average_down(l + 1)

    def reflux(self, f):
        faces, reg = self._reg[f]
        c = f - 1
        vol = self.dx[c] * self.dx[c]
        for e, (p, d, side, ct, q, cf) in enumerate(faces):
            cc = cf - 1 if side == 0 else cf
            a, b = self.layouts[c][q][:2]
            i, j = ((cc, ct) if d == 0 else (ct, cc))
            cell = self.U[c][q][:, j - b + G, i - a + G]
            corr = reg[e] / vol
            cell[:] = cell - corr if side == 0 else cell + corr

    def average_down(self, f):
        c, r = f - 1, self.r
        volfrac = 1.0 / (r * r)
        for p, (i0, j0, nx, ny) in enumerate(self.layouts[f]):
            u = self.U[f][p]
            for cj in range(j0 // r, (j0 + ny) // r):
                for ci in range(i0 // r, (i0 + nx) // r):
                    q = self.find(c, ci, cj)
                    a, b = self.layouts[c][q][:2]
                    s = np.zeros(4)
                    for jj in range(r):
                        for ii in range(r):
                            s = s + u[:, cj * r - j0 + jj + G, ci * r - i0 + ii + G]
                    self.U[c][q][:, cj - b + G, ci - a + G] = volfrac * s

    # hierarchy ------------------------------------------------------------------
    def stable_dt(self):
        m = np.inf
        for l in range(len(self.layouts)):
            self.fill_ghosts(l)
            lv = np.inf
            for d in (0, 1):
                for p in range(len(self.layouts[l])):
                    lv = min(lv, self.sweep(self.U[l][p], 0.0, self.dx[l], d)[2])
            m = min(m, self.r ** l * lv)
        return m

    def advance(self, dt):
        self._reg = {}
        self._new = {}
        self.level_cfl = [np.inf] * len(self.layouts)
        self.advance_level(0, dt, 0)
        self.cycle += 1

    def total(self):
        """Sum over level 0 of U * cell area (the composite integral after
        average_down)."""
        return sum(u[:, G:-G, G:-G].sum(axis=(1, 2)) for u in self.U[0]) * self.dx[0] ** 2


def gradient_tags(rho, p, tol_rho, tol_p):
    """ScalarGradientDetector::TagCellsForRefinement
    (include/fub/tagging_method/GradientDetector.hpp:150-194) with the
    density and pressure projections, on one patch: rho and p with two
    ghost layers (ny + 4, nx + 4); returns (ny, nx) tags of the interior.
    A cell is tagged when along x or y a neighbour differs and
    |x_M - x_N| / (|x_M| + |x_N|) exceeds the tolerance on either side."""
    ny, nx = rho.shape[0] - 2 * G, rho.shape[1] - 2 * G
    tags = np.zeros((ny, nx), dtype=np.uint8)
    for j in range(ny):
        for i in range(nx):
            jj, ii = j + G, i + G
            t = False
            for (dj, di) in ((0, 1), (1, 0)):
                for x, tol in ((rho, tol_rho), (p, tol_p)):
                    xL, xM, xR = x[jj - dj, ii - di], x[jj, ii], x[jj + dj, ii + di]
                    if xM != xR or xM != xL:
                        with np.errstate(invalid="ignore", divide="ignore"):
                            left = abs(xM - xL) / (abs(xM) + abs(xL))
                            right = abs(xM - xR) / (abs(xM) + abs(xR))
                        t = t or left > tol or right > tol
            tags[j, i] = 1 if t else 0
    return tags
