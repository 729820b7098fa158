"""GPU KBN cut-cell operators against the FP64 oracle.

Geometry, regular fluxes (masked MUSCL-Hancock + HLLEM), CFL, stabilised
fluxes, both time integrators and the masked reconstruction are compared
bit for bit.  Reference states and boundary fluxes from the exact boundary
solver go through pow and meet the 1e-12 bar; with the HLLEM boundary
solver they are bit-exact too."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from kbn_cases import circle_case
from tolerances import tolerance_ratio

import finitevolumesolver_amd as fvs
from finitevolumesolver_amd import cutcell as K
from finitevolumesolver_amd.patch import PatchView

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)


def _dev(a, cuda):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


@pytest.fixture(scope="module")
def case(cuda):
    geom, comp, lo = circle_case()
    _note(f"kbn case: {geom.alpha.shape} cells, "
          f"{int(((geom.alpha > 0) & (geom.alpha < 1)).sum())} cut, "
          f"min alpha {geom.alpha[geom.alpha > 0].min():.2e}")
    return geom, comp, lo


def _fm(solver="exact"):
    return fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock",
                          limiter="MinMod", boundary_solver=solver)


def test_geometry_fill_and_classification(case, cuda):
    geom, _, lo = case
    fm = _fm()
    g = K.CutCellData.from_geometry2d(geom, cuda)
    for d in range(2):
        K.FillCutCellData(fm, g, d)
    h = O.HostCutCellData(geom).fill()
    for kind in h.fractions:
        for d in range(2):
            got = g.fractions[kind][d].data.cpu().numpy()
            np.testing.assert_array_equal(got, h.fractions[kind][d], err_msg=f"{kind} {d}")
    hi = (lo[0] + geom.alpha.shape[1] - 1, lo[1] + geom.alpha.shape[0] - 1)
    flags = K.IsCutCell(fm, g, lo, hi).cpu().numpy()
    ref = np.zeros(flags.shape, np.int32)
    s = h.c_struct()
    O.load().oracle_kbn_is_cut_cell(2, C.byref(s), O.ptr(np.array([lo[0], lo[1], 0], np.int64)),
                                    O.ptr(np.array([hi[0], hi[1], 0], np.int64)), O.ptr(ref))
    np.testing.assert_array_equal(flags, ref)
    _note(f"  fractions bit-exact; {int(flags.sum())} cut cells")


@pytest.mark.parametrize("solver", ["exact", "hllem"])
def test_reference_states_and_boundary_fluxes(case, cuda, solver):
    geom, comp, lo = case
    fm = _fm(solver)
    desc = fm._h.desc
    g = K.CutCellData.from_geometry2d(geom, cuda)
    h = O.HostCutCellData(geom)
    s = h.c_struct()
    kbn = K.KbnCutCellMethod(fm)
    t_comp = _dev(comp, cuda)
    import torch
    t_refs = torch.zeros_like(t_comp)
    kbn.PreAdvanceHierarchy(PatchView(t_refs, lo), PatchView(t_comp, lo), g)
    refs = np.zeros_like(comp)
    O.load().oracle_kbn_reference_states(C.byref(desc), C.byref(O.host_view(refs, lo)),
                                         C.byref(O.host_view(comp, lo)), C.byref(s))
    got = t_refs.cpu().numpy()
    if solver == "hllem":
        np.testing.assert_array_equal(got, refs)
    else:
        r = tolerance_ratio(got, refs)
        _note(f"  exact reference states: tolerance ratio {r:.3f}")
        assert r <= 1.0
    for d in range(2):
        t_fb = torch.zeros((4,) + comp.shape[1:], dtype=torch.float64, device=cuda)
        kbn.ComputeBoundaryFluxes(PatchView(t_fb, lo), PatchView(t_comp, lo),
                                  PatchView(_dev(refs, cuda), lo), g, 0.0, 1.0 / 32, d)
        fb = np.zeros((4,) + comp.shape[1:])
        O.load().oracle_kbn_boundary_fluxes(C.byref(desc), C.byref(O.host_view(fb, lo)),
                                            C.byref(O.host_view(comp, lo)),
                                            C.byref(O.host_view(refs, lo)), C.byref(s), 0.0,
                                            1.0 / 32, d)
        got = t_fb.cpu().numpy()
        if solver == "hllem":
            np.testing.assert_array_equal(got, fb)
        else:
            r = tolerance_ratio(got, fb)
            _note(f"  exact boundary fluxes dir {d}: tolerance ratio {r:.3f}")
            assert r <= 1.0


def test_regular_fluxes_stable_dt_and_update(case, cuda):
    """One full split step per direction on both sides: masked fluxes, CFL,
    stabilised fluxes, h-box update (TimeIntegrator and TimeIntegrator2) and
    the masked reconstruction, compared bit for bit."""
    import torch
    geom, comp, lo = case
    fm = _fm("hllem")
    desc = fm._h.desc
    kbn = K.KbnCutCellMethod(fm)
    g = K.CutCellData.from_geometry2d(geom, cuda)
    for d in range(2):
        K.FillCutCellData(fm, g, d)
    h = O.HostCutCellData(geom).fill()
    s = h.c_struct()
    dx = 1.0 / 32
    t_comp = _dev(comp, cuda)
    refs = np.zeros_like(comp)
    O.load().oracle_kbn_reference_states(C.byref(desc), C.byref(O.host_view(refs, lo)),
                                         C.byref(O.host_view(comp, lo)), C.byref(s))
    gy, gx = geom.alpha.shape
    for d in range(2):
        # CFL over the whole patch
        dt_dev = kbn.ComputeStableDt(PatchView(t_comp, lo), g, dx, d).item()
        dt_ref = C.c_double()
        O.load().oracle_kbn_stable_dt(C.byref(desc), C.byref(O.host_view(comp, lo)), C.byref(s),
                                      dx, d, C.byref(dt_ref))
        assert dt_dev == dt_ref.value
        dt = 0.5 * dt_ref.value
        # regular fluxes on all faces with a full stencil
        fshape = [4, 1, gy, gx]
        fshape[3 - d] -= 3
        flo = [lo[0], lo[1], 0]
        flo[d] += 2
        t_f = torch.zeros(fshape, dtype=torch.float64, device=cuda)
        kbn.ComputeRegularFluxes(PatchView(t_f, flo), PatchView(t_comp, lo), g, dt, dx, d)
        f = np.zeros(fshape)
        O.load().oracle_kbn_regular_fluxes(C.byref(desc), C.byref(O.host_view(f, flo)),
                                           C.byref(O.host_view(comp, lo)), C.byref(s), dt, dx, d)
        np.testing.assert_array_equal(t_f.cpu().numpy(), f)
        # boundary fluxes (HLLEM boundary solver, bit-exact) and stabilised fluxes
        fb = np.zeros((4,) + comp.shape[1:])
        O.load().oracle_kbn_boundary_fluxes(C.byref(desc), C.byref(O.host_view(fb, lo)),
                                            C.byref(O.host_view(comp, lo)),
                                            C.byref(O.host_view(refs, lo)), C.byref(s), dt, dx, d)
        # stabilised fluxes on the faces with both neighbours inside the flux box
        sshape = list(fshape)
        sshape[3 - d] -= 2
        slo = list(flo)
        slo[d] += 1
        # array axis a (2: y, 3: x) runs along direction 3 - a
        sl_ = tuple([slice(None), slice(None)] +
                    [slice(1, -1) if (3 - a) == d else slice(None) for a in (2, 3)])
        f_in = np.ascontiguousarray(f[sl_])
        outs = [np.zeros(sshape) for _ in range(3)]
        O.load().oracle_kbn_cutcell_fluxes(2, *[C.byref(O.host_view(o, slo)) for o in outs],
                                           C.byref(O.host_view(fb, lo)),
                                           C.byref(O.host_view(f_in, slo)), C.byref(s), d)
        t_outs = [torch.zeros(sshape, dtype=torch.float64, device=cuda) for _ in range(3)]
        t_fb = _dev(fb, cuda)
        kbn.ComputeCutCellFluxes(*[PatchView(t, slo) for t in t_outs], PatchView(t_fb, lo),
                                 PatchView(_dev(f_in, cuda), slo), g, dt, dx, d)
        for t, o in zip(t_outs, outs):
            np.testing.assert_array_equal(t.cpu().numpy(), o)
        stab, fsl, fsr = outs
        # h-box update of the cells enclosed by the stabilised face box
        cshape = list(sshape)
        cshape[3 - d] -= 1
        clo = list(slo)
        off = [clo[2] - 0, clo[1] - lo[1], clo[0] - lo[0]]
        cells = np.ascontiguousarray(comp[:4, :, off[1]:off[1] + cshape[2],
                                          off[2]:off[2] + cshape[3]])
        u_ref = cells.copy()
        O.load().oracle_kbn_update(C.byref(O.host_view(u_ref, clo)),
                                   C.byref(O.host_view(f_in, slo)), C.byref(O.host_view(stab, slo)),
                                   C.byref(O.host_view(fsl, slo)), C.byref(O.host_view(fsr, slo)),
                                   C.byref(O.host_view(fb, lo)), C.byref(s), dt, dx, d)
        t_u = _dev(cells, cuda)
        K.TimeIntegrator.UpdateConservatively(fm, PatchView(t_u, clo),
                                              PatchView(_dev(f_in, cuda), slo),
                                              PatchView(_dev(stab, cuda), slo),
                                              PatchView(_dev(fsl, cuda), slo),
                                              PatchView(_dev(fsr, cuda), slo),
                                              PatchView(t_fb, lo), g, dt, dx, d)
        np.testing.assert_array_equal(t_u.cpu().numpy(), u_ref)
        u2_ref = cells.copy()
        O.load().oracle_kbn_update2(C.byref(O.host_view(u2_ref, clo)),
                                    C.byref(O.host_view(cells, clo)),
                                    C.byref(O.host_view(stab, slo)), C.byref(O.host_view(fb, lo)),
                                    C.byref(s), dt, dx, d)
        t_u2 = _dev(cells, cuda)
        K.TimeIntegrator2.UpdateConservatively(fm, PatchView(t_u2, clo),
                                               PatchView(_dev(cells, cuda), clo),
                                               PatchView(_dev(stab, cuda), slo),
                                               PatchView(t_fb, lo), g, dt, dx, d)
        np.testing.assert_array_equal(t_u2.cpu().numpy(), u2_ref)
        # masked reconstruction of the updated cells
        cq_ref = np.zeros((6,) + tuple(cshape[1:]))
        O.load().oracle_kbn_complete_from_cons(2, 1.4, C.byref(O.host_view(cq_ref, clo)),
                                               C.byref(O.host_view(u_ref, clo)), C.byref(s))
        t_cq = torch.zeros_like(_dev(cq_ref, cuda))
        K.Reconstruction.CompleteFromCons(fm, PatchView(t_cq, clo), PatchView(_dev(u_ref, cuda), clo),
                                          g)
        np.testing.assert_array_equal(t_cq.cpu().numpy(), cq_ref)
        _note(f"  dir {d}: dt {dt_dev!r}, fluxes/update/reconstruction bit-exact")


def test_channel_geometry_full_size_bitexact(cuda):
    """The 2048 x 1024 channel (cylinder r = 0.25 and the 30-degree wedge,
    small cells shrunk to volume fractions down to 1e-4) at its full size:
    shielded fractions of both directions and IsCutCell over the whole grown
    box equal the oracle's bit for bit, and the cut-cell level lists exactly
    the oracle's interior cut cells."""
    import torch
    from finitevolumesolver_amd.geometry import channel_geometry
    from finitevolumesolver_amd.initial_data import shock_channel
    nx, ny = 2048, 1024
    _note(f"channel geometry {nx} x {ny}")
    geom = channel_geometry(nx, ny)
    fm = _fm("hllem")
    g = K.CutCellData.from_geometry2d(geom, cuda)
    for d in range(2):
        K.FillCutCellData(fm, g, d)
    h = O.HostCutCellData(geom).fill()
    for kind in h.fractions:
        for d in range(2):
            got = g.fractions[kind][d].data.cpu().numpy()
            np.testing.assert_array_equal(got, h.fractions[kind][d], err_msg=f"{kind} {d}")
    lo = geom.lo
    hi = (lo[0] + geom.alpha.shape[1] - 1, lo[1] + geom.alpha.shape[0] - 1)
    flags = K.IsCutCell(fm, g, lo, hi).cpu().numpy()
    ref = np.zeros(flags.shape, np.int32)
    s = h.c_struct()
    O.load().oracle_kbn_is_cut_cell(2, C.byref(s), O.ptr(np.array([lo[0], lo[1], 0], np.int64)),
                                    O.ptr(np.array([hi[0], hi[1], 0], np.int64)), O.ptr(ref))
    np.testing.assert_array_equal(flags, ref)
    interior = ref.reshape(geom.alpha.shape)[-lo[1]:-lo[1] + ny, -lo[0]:-lo[0] + nx]
    cons, dx = shock_channel(nx, ny, x_shock=-0.4)
    lvl = K.CutCellLevel(fm, (nx, ny), (dx, dx), g)
    assert lvl.ncut == int(interior.sum())
    _note(f"  fractions and classification bit-exact; {lvl.ncut} interior cut cells, "
          f"min alpha {geom.alpha[geom.alpha > 0].min():.2e}")
    del lvl, g
    torch.cuda.empty_cache()
