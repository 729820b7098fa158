"""3D KBN boundary problems on the GPU against the oracle: the reference
states of KbnCutCellMethod::PreAdvanceHierarchy
(include/fub/cutcell_method/KbnStabilisation.hpp:188-210) and the boundary
fluxes (:212-273) of a 3D patch, which rotate each cut cell's momentum into
the frame of its boundary normal with the 3D MakeRotation / Rotate /
Reflect (include/fub/ext/Eigen.hpp:104-130, src/equations/PerfectGas.cpp:
238-300).  The geometry is synthetic (random volume and face fractions,
random unit normals and centroids): both sides evaluate the same arithmetic
on any input.  With the HLLEM boundary solver the results are bit-exact;
with the exact solver (through pow) within 1e-12."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
from fields import rough_field
from tolerances import tolerance_ratio

import finitevolumesolver_amd as fvs
from finitevolumesolver_amd import _lib as L
from finitevolumesolver_amd import cutcell as K
from finitevolumesolver_amd.patch import PatchView

pytestmark = pytest.mark.gpu
N = (10, 8, 6)  # interior cells (x, y, z)
G = 2
LO = (-G, -G, -G)


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _geometry(seed=7):
    rng = np.random.default_rng(seed)
    nx, ny, nz = (n + 2 * G for n in N)
    u = rng.random((nz, ny, nx))
    alpha = np.where(u < 0.6, 1.0, np.where(u < 0.7, 0.0, 0.05 + 0.9 * rng.random((nz, ny, nx))))
    betas = []
    for d in range(3):
        shape = [nz, ny, nx]
        shape[2 - d] += 1
        b = rng.random(shape)
        betas.append(np.where(b < 0.7, 1.0, b))
    nrm = rng.normal(size=(3, nz, ny, nx))
    nrm /= np.linalg.norm(nrm, axis=0)
    centroids = rng.random((3, nz, ny, nx)) - 0.5
    return alpha, betas, nrm, centroids


class _Host3D:
    """CutCellData<3> of host arrays for the oracle."""

    def __init__(self, alpha, betas, normals, centroids):
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.alpha = c(alpha[None])
        self.betas = [c(b[None]) for b in betas]
        self.normals = c(normals)
        self.centroids = c(centroids)
        self.fr = [np.zeros((2,) + b.shape[1:]) for b in self.betas]

    def c_struct(self):
        s = L.CutCellData()
        s.volume_fractions = O.host_view(self.alpha, LO)
        for d in range(3):
            s.face_fractions[d] = O.host_view(self.betas[d], LO)
            for k in ("unshielded", "shielded_left", "shielded_right", "doubly_shielded"):
                getattr(s, k)[d] = O.host_view(self.fr[d], LO)
        s.boundary_normals = O.host_view(self.normals, LO)
        s.boundary_centroids = O.host_view(self.centroids, LO)
        self._s = s
        return s


def _states(alpha):
    nx, ny, nz = (n + 2 * G for n in N)
    cons = rough_field((nz, ny, nx), seed=23)
    comp = O.complete_from_cons(cons.reshape(5, -1), 3).reshape(7, nz, ny, nx)
    comp[:, alpha == 0.0] = 0.0
    return np.ascontiguousarray(comp)


@pytest.mark.parametrize("solver", ["hllem", "exact"])
def test_reference_states_and_boundary_fluxes_3d(cuda, solver):
    import torch
    _note(f"3D KBN boundary problems, {solver} solver")
    alpha, betas, nrm, cen = _geometry()
    comp = _states(alpha)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=3), base="HLLEM", reconstruction="MusclHancock",
                        boundary_solver=solver)
    desc = fm._h.desc
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    g = K.CutCellData(dev(alpha), [dev(b) for b in betas], dev(nrm), dev(cen), LO)
    h = _Host3D(alpha, betas, nrm, cen)
    s = h.c_struct()
    lo3 = np.array(LO, np.int64)
    hi3 = np.array([LO[d] + N[d] + 2 * G - 1 for d in range(3)], np.int64)
    flags = np.zeros(alpha.size, np.int32)
    O.load().oracle_kbn_is_cut_cell(3, C.byref(s), O.ptr(lo3), O.ptr(hi3), O.ptr(flags))
    gflags = K.IsCutCell(fm, g, LO, tuple(hi3)).cpu().numpy().ravel()
    np.testing.assert_array_equal(gflags, flags)
    kbn = K.KbnCutCellMethod(fm)
    t_comp = dev(comp)
    t_refs = torch.zeros_like(t_comp)
    kbn.PreAdvanceHierarchy(PatchView(t_refs, LO), PatchView(t_comp, LO), g)
    refs = np.zeros_like(comp)
    O.load().oracle_kbn_reference_states(C.byref(desc), C.byref(O.host_view(refs, LO)),
                                         C.byref(O.host_view(comp, LO)), C.byref(s))
    got = t_refs.cpu().numpy()
    _note(f"  {int(flags.sum())} cut cells of {flags.size}")
    assert int(flags.sum()) > 50
    if solver == "hllem":
        np.testing.assert_array_equal(got, refs)
    else:
        assert tolerance_ratio(got, refs) <= 1.0
    for d in range(3):
        t_fb = torch.zeros((5,) + comp.shape[1:], dtype=torch.float64, device=cuda)
        kbn.ComputeBoundaryFluxes(PatchView(t_fb, LO), PatchView(t_comp, LO),
                                  PatchView(dev(refs), LO), g, 0.0, 0.1, d)
        fb = np.zeros((5,) + comp.shape[1:])
        O.load().oracle_kbn_boundary_fluxes(C.byref(desc), C.byref(O.host_view(fb, LO)),
                                            C.byref(O.host_view(comp, LO)),
                                            C.byref(O.host_view(refs, LO)), C.byref(s), 0.0,
                                            0.1, d)
        gotf = t_fb.cpu().numpy()
        assert np.any(fb != 0.0)
        if solver == "hllem":
            np.testing.assert_array_equal(gotf, fb, err_msg=f"dir {d}")
        else:
            r = tolerance_ratio(gotf, fb)
            _note(f"  exact boundary fluxes dir {d}: tolerance ratio {r:.3f}")
            assert r <= 1.0
