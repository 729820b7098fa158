"""Host mirror of the reference's KBN cut-cell API over the C ABI.

  fub::CutCellData<Rank>                 include/fub/CutCellData.hpp:48-75
  fub::KbnCutCellMethod<FM, RS>          include/fub/cutcell_method/KbnStabilisation.hpp:53-429
  fub::amrex::cutcell::FillCutCellData   src/AMReX/cutcell/FillCutCellData.cpp:28-118
  fub::amrex::cutcell::TimeIntegrator(2) src/AMReX/cutcell/TimeIntegrator.cpp:65-508
  fub::amrex::cutcell::Reconstruction    include/fub/AMReX/cutcell/Reconstruction.hpp:35-200

Views are PatchView objects over torch device tensors (component-major,
x fastest); geometry lives in a CutCellData of device tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .patch import FluxMethod, PatchView, _dir


class CutCellData:
    """Device copy of a patch's embedded-boundary geometry.

    Cell arrays cover the cell box [lo, hi] (ghosts included); face arrays of
    direction d cover the cells' faces normal to d.  The shielded fractions
    are filled by FillCutCellData (two components: absolute, relative).
    """

    def __init__(self, alpha: torch.Tensor, betas, normals: torch.Tensor,
                 centroids: torch.Tensor, lo=(0, 0, 0)):
        self.dim = len(betas)
        lo = tuple(lo) + (0,) * (3 - len(lo))
        self.lo = lo[:3]
        dev = alpha.device
        # cell views: shape (1, nz, ny, nx)
        self.alpha = PatchView(alpha.reshape((1,) + self._cell_shape(alpha)), self.lo)
        self.betas = [PatchView(b.reshape((1,) + self._cell_shape(b)), self.lo) for b in betas]
        self.normals = PatchView(normals.reshape((self.dim,) + self._cell_shape(normals[0])),
                                 self.lo)
        self.centroids = PatchView(centroids.reshape((self.dim,) + self._cell_shape(centroids[0])),
                                   self.lo)
        self.fractions = {}
        for kind in ("unshielded", "shielded_left", "shielded_right", "doubly_shielded"):
            self.fractions[kind] = [
                PatchView(torch.zeros((2,) + tuple(b.data.shape[1:]), dtype=torch.float64,
                                      device=dev), self.lo) for b in self.betas]
        self._struct = None

    @staticmethod
    def _cell_shape(t: torch.Tensor):
        if t.dim() == 2:
            return (1,) + tuple(t.shape)
        return tuple(t.shape)

    @classmethod
    def from_geometry2d(cls, geom, device) -> "CutCellData":
        def dev(a):
            return torch.from_numpy(np.ascontiguousarray(a)).to(device)
        return cls(dev(geom.alpha), [dev(geom.beta_x), dev(geom.beta_y)], dev(geom.normals),
                   dev(geom.centroids), geom.lo)

    def c_struct(self) -> L.CutCellData:
        s = L.CutCellData()
        s.volume_fractions = self.alpha.c_view()
        for d in range(self.dim):
            s.face_fractions[d] = self.betas[d].c_view()
            s.unshielded[d] = self.fractions["unshielded"][d].c_view()
            s.shielded_left[d] = self.fractions["shielded_left"][d].c_view()
            s.shielded_right[d] = self.fractions["shielded_right"][d].c_view()
            s.doubly_shielded[d] = self.fractions["doubly_shielded"][d].c_view()
        s.boundary_normals = self.normals.c_view()
        s.boundary_centroids = self.centroids.c_view()
        self._struct = s  # keep alive while in use
        return s

    def cell_box(self):
        return self.alpha.lo, self.alpha.hi


def FillCutCellData(flux_method: FluxMethod, geom: CutCellData, dir, face_lo=None,
                    face_hi=None) -> None:
    """Shielded face fractions of every face of the box (default: all faces
    of direction dir whose two neighbours along dir exist)."""
    d = _dir(dir)
    flo, fhi = list(geom.betas[d].lo), list(geom.betas[d].hi)
    flo[d] += 1
    fhi[d] -= 1
    if face_lo is not None:
        flo = list(face_lo) + [0] * (3 - len(face_lo))
    if face_hi is not None:
        fhi = list(face_hi) + [0] * (3 - len(face_hi))
    s = geom.c_struct()
    L.check(L.load().fvs_kbn_fill_cutcell_data(flux_method.handle, C.byref(s), d, L.i64x3(flo),
                                               L.i64x3(fhi)))


def IsCutCell(flux_method: FluxMethod, geom: CutCellData, lo, hi) -> torch.Tensor:
    lo = list(lo) + [0] * (3 - len(lo))
    hi = list(hi) + [0] * (3 - len(hi))
    shape = (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
    flags = torch.zeros(shape, dtype=torch.int32, device=geom.alpha.data.device)
    s = geom.c_struct()
    L.check(L.load().fvs_kbn_is_cut_cell(flux_method.handle, C.byref(s), L.i64x3(lo),
                                         L.i64x3(hi), flags.data_ptr()))
    return flags


class KbnCutCellMethod:
    """KbnCutCellMethod<FM, RiemannSolver> over a FluxMethod handle; the
    boundary Riemann solver is the handle's boundary_solver (exact by
    default, as the reference's template default)."""

    def __init__(self, flux_method: FluxMethod):
        self.flux_method = flux_method

    @property
    def handle(self):
        return self.flux_method.handle

    def GetStencilWidth(self) -> int:
        return self.flux_method.GetStencilWidth()

    def PreAdvanceHierarchy(self, references: PatchView, states: PatchView,
                            geom: CutCellData) -> None:
        rv, sv, s = references.c_view(), states.c_view(), geom.c_struct()
        L.check(L.load().fvs_kbn_reference_states(self.handle, C.byref(rv), C.byref(sv),
                                                  C.byref(s)))

    def ComputeBoundaryFluxes(self, boundary_fluxes: PatchView, states: PatchView,
                              references: PatchView, geom: CutCellData, dt: float, dx: float,
                              dir) -> None:
        bv, sv, rv, s = (boundary_fluxes.c_view(), states.c_view(), references.c_view(),
                         geom.c_struct())
        L.check(L.load().fvs_kbn_boundary_fluxes(self.handle, C.byref(bv), C.byref(sv),
                                                 C.byref(rv), C.byref(s), float(dt), float(dx),
                                                 _dir(dir)))

    def ComputeRegularFluxes(self, fluxes: PatchView, states: PatchView, geom: CutCellData,
                             dt: float, dx: float, dir) -> None:
        fv, sv, s = fluxes.c_view(), states.c_view(), geom.c_struct()
        L.check(L.load().fvs_kbn_regular_fluxes(self.handle, C.byref(fv), C.byref(sv),
                                                C.byref(s), float(dt), float(dx), _dir(dir)))

    def ComputeStableDt(self, states: PatchView, geom: CutCellData, dx: float, dir,
                        out: torch.Tensor | None = None) -> torch.Tensor:
        if out is None:
            out = torch.empty(1, dtype=torch.float64, device=states.data.device)
        sv, s = states.c_view(), geom.c_struct()
        L.check(L.load().fvs_kbn_stable_dt(self.handle, C.byref(sv), C.byref(s), float(dx),
                                           _dir(dir), out.data_ptr()))
        return out

    def ComputeCutCellFluxes(self, stabilised: PatchView, shielded_left: PatchView,
                             shielded_right: PatchView, boundary_fluxes: PatchView,
                             regular_fluxes: PatchView, geom: CutCellData, dt: float,
                             dx: float, dir) -> None:
        views = [v.c_view() for v in (stabilised, shielded_left, shielded_right,
                                      boundary_fluxes, regular_fluxes)]
        s = geom.c_struct()
        L.check(L.load().fvs_kbn_cutcell_fluxes(self.handle, *[C.byref(v) for v in views],
                                                C.byref(s), float(dt), float(dx), _dir(dir)))


class TimeIntegrator:
    """cutcell::TimeIntegrator: h-box update with the shielded fluxes."""

    @staticmethod
    def UpdateConservatively(flux_method: FluxMethod, scratch: PatchView,
                             unshielded: PatchView, stabilised: PatchView,
                             shielded_left: PatchView, shielded_right: PatchView,
                             boundary_fluxes: PatchView, geom: CutCellData, dt: float,
                             dx: float, dir) -> None:
        views = [v.c_view() for v in (scratch, unshielded, stabilised, shielded_left,
                                      shielded_right, boundary_fluxes)]
        s = geom.c_struct()
        L.check(L.load().fvs_kbn_update_conservatively(flux_method.handle,
                                                       *[C.byref(v) for v in views], C.byref(s),
                                                       float(dt), float(dx), _dir(dir)))


class TimeIntegrator2:
    """cutcell::TimeIntegrator2: conservative update with the boundary flux."""

    @staticmethod
    def UpdateConservatively(flux_method: FluxMethod, dest: PatchView, src: PatchView,
                             stabilised: PatchView, boundary_fluxes: PatchView,
                             geom: CutCellData, dt: float, dx: float, dir) -> None:
        views = [v.c_view() for v in (dest, src, stabilised, boundary_fluxes)]
        s = geom.c_struct()
        L.check(L.load().fvs_kbn_update_conservatively2(flux_method.handle,
                                                        *[C.byref(v) for v in views], C.byref(s),
                                                        float(dt), float(dx), _dir(dir)))


class Reconstruction:
    """cutcell::Reconstruction: complete states, zero where alpha = 0."""

    @staticmethod
    def CompleteFromCons(flux_method: FluxMethod, complete: PatchView, cons: PatchView,
                         geom: CutCellData) -> None:
        cv, uv, s = complete.c_view(), cons.c_view(), geom.c_struct()
        L.check(L.load().fvs_kbn_complete_from_cons(flux_method.handle, C.byref(cv), C.byref(uv),
                                                    C.byref(s)))


class CutCellLevel:
    """Single-patch 2D cut-cell level on the GPU (fvs_kbn_level_*): the
    reference's cut-cell split sequence with transmissive boundaries.

    geom: CutCellData over the patch grown by two ghost layers."""

    INTEGRATORS = {"TimeIntegrator": 0, "TimeIntegrator2": 1}

    def __init__(self, flux_method: FluxMethod, n, dx, geom: CutCellData,
                 integrator="TimeIntegrator", splitting="Godunov", cfl=0.8):
        self._lib = L.load()
        self.flux_method = flux_method
        self.n = (int(n[0]), int(n[1]))
        pd = L.PatchDesc()
        pd.n = L.i64x3((self.n[0], self.n[1], 1))
        pd.dx = (C.c_double * 3)(float(dx[0]), float(dx[1]), 1.0)
        pd.splitting = L.SPLIT_STRANG if splitting == "Strang" else L.SPLIT_GODUNOV
        pd.cfl = cfl
        self._geom = geom
        s = geom.c_struct()
        lv = C.c_void_p()
        L.check(self._lib.fvs_kbn_level_create(flux_method.handle, C.byref(pd), C.byref(s),
                                               self.INTEGRATORS[integrator], C.byref(lv)))
        self.level = lv
        self.device = geom.alpha.data.device

    def __del__(self):
        lv = getattr(self, "level", None)
        if lv is not None and lv.value:
            self._lib.fvs_kbn_level_destroy(lv)
            self.level = None

    @property
    def ncut(self) -> int:
        return self._lib.fvs_kbn_level_ncut(self.level)

    def set_state(self, cons: torch.Tensor, t0: float = 0.0) -> None:
        cons = cons.reshape(4, self.n[1], self.n[0]).contiguous()
        self._cons_in = cons
        L.check(self._lib.fvs_kbn_level_set_state(
            self.level, L.ptr_array([cons[c].data_ptr() for c in range(4)]), t0))

    def get_state(self) -> torch.Tensor:
        out = torch.empty((4, self.n[1], self.n[0]), dtype=torch.float64, device=self.device)
        L.check(self._lib.fvs_kbn_level_get_state(
            self.level, L.ptr_array([out[c].data_ptr() for c in range(4)])))
        return out

    def advance(self, max_steps: int, final_time: float = float("inf"), batch: int = 16,
                return_dts: bool = False):
        dts = np.zeros(max(max_steps, 1))
        done, fails = C.c_int(), C.c_int()
        L.check(self._lib.fvs_kbn_level_advance(
            self.level, max_steps, final_time, batch,
            dts.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done), C.byref(fails)))
        if return_dts:
            return done.value, fails.value, dts[:done.value].tolist()
        return done.value, fails.value

    def time(self):
        t = C.c_double()
        cyc = C.c_int64()
        L.check(self._lib.fvs_kbn_level_time(self.level, C.byref(t), C.byref(cyc)))
        return t.value, cyc.value
