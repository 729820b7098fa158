"""GPU parity of the fused split sweep, the patch operators and the level
driver against the FP64 CPU oracle (bit-exact for the FP64 HLLEM path)."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
from fields import rough_field
from sweep_ref import oracle_sweep, padded_complete

import finitevolumesolver_amd as fvs
from finitevolumesolver_amd import _lib as L

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _torch():
    import torch
    return torch


def _max_rel(a, b):
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den))


CASES_2D = [
    # (shape (nz, ny, nx), base, limiter, quirks)
    ((1, 48, 64), "HLLEM", "MinMod", True),
    ((1, 37, 300), "HLLEM", "VanLeer", True),
    ((1, 40, 33), "HLLEM_plain", "MinMod", True),
    ((1, 24, 20), "HLLEM", "NoLimiter", True),
]
CASES_3D = [
    ((18, 20, 24), "HLLEM", "MinMod", True),
    ((17, 19, 23), "HLLEM", "MinMod", False),
    ((12, 14, 270), "HLLEM_plain", "VanLeer", True),
    ((10, 12, 16), "HLLEM_plain", "MinMod", False),
]


@pytest.mark.parametrize("case", CASES_2D + CASES_3D)
def test_split_sweep_bitexact(cuda, case):
    torch = _torch()
    shape, base, limiter, quirks = case
    D = 2 if shape[0] == 1 else 3
    _note(f"split sweep case shape={shape} base={base} lim={limiter} quirks={quirks}")
    # An unlimited central slope across the field's jumps produces negative
    # pressures, so NoLimiter runs on the smooth part of the field only.
    cons = rough_field(shape, seed=11 + shape[2], jumps=(limiter != "NoLimiter"))
    eq = fvs.PerfectGas(dim=D)
    fm = fvs.FluxMethod(eq, base=base, limiter=limiter, quirks=quirks)
    desc = fm._h.desc
    dx = 1.0 / shape[2]
    for d in range(D):
        dt = 0.2 * dx
        ref, ref_dt, _ = oracle_sweep(desc, cons, dt, dx, d)
        t_in = torch.from_numpy(cons).to(cuda)
        t_out = torch.empty_like(t_in)
        sdt = torch.empty(1, dtype=torch.float64, device=cuda)
        fvs.split_sweep(fm, t_in, t_out, dt, dx, d, sdt)
        torch.cuda.synchronize()
        got = t_out.cpu().numpy()
        assert np.all(np.isfinite(got))
        diff = _max_rel(got, ref)
        _note(f"  dir {d}: max rel diff {diff:.3e}")
        assert np.array_equal(got, ref), f"dir {d}: max rel diff {diff}"
        assert sdt.item() == ref_dt


@pytest.mark.parametrize("D", [2, 3])
def test_patch_stable_dt_matches_oracle(cuda, D):
    torch = _torch()
    shape = (1, 40, 52) if D == 2 else (14, 16, 18)
    cons = rough_field(shape, seed=5)
    eq = fvs.PerfectGas(dim=D)
    fm = fvs.FluxMethod(eq)
    dx = [1.0 / 52, 1.3 / 40, 0.9 / 14][:D]
    got = fvs.patch_stable_dt(fm, torch.from_numpy(cons).to(cuda), dx).cpu().numpy()
    desc = fm._h.desc
    for d in range(D):
        _, ref_dt, _ = oracle_sweep(desc, cons, 0.0, dx[d], d)
        assert got[d] == ref_dt, (d, got[d], ref_dt)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("recon", ["MusclHancock", "None", "MusclHancockPrim",
                                   "MusclHancockCharacteristic"])
def test_view_operators_bitexact(cuda, D, recon):
    """fvs_numeric_fluxes / fvs_stable_dt / fvs_update_conservatively /
    fvs_complete_from_cons / fvs_transmissive_fill on strided views with
    ghost cells, against the same oracle operators."""
    torch = _torch()
    shape = (1, 30, 41) if D == 2 else (9, 11, 13)
    cons = rough_field(shape, seed=23)
    eq = fvs.PerfectGas(dim=D)
    fm = fvs.FluxMethod(eq, reconstruction=recon)
    desc = fm._h.desc
    gw = fm.GetStencilWidth()
    comp, g = padded_complete(cons, D, gw)
    nz, ny, nx = shape
    lo = (-g[0], -g[1], -g[2])
    t_comp = torch.from_numpy(comp).to(cuda)
    full = fvs.PatchView(t_comp, lo)
    # 1. transmissive fill on the device from a state whose ghosts are garbage
    t_dirty = t_comp.clone()
    if g[0]:
        t_dirty[:, :, :, :g[0]] = -7.0
    bc_view = fvs.PatchView(t_dirty, lo)
    for d in range(D):
        fvs.TransmissiveBoundary(fm, d).FillBoundary(bc_view, (0, 0, 0), (nx - 1, ny - 1, nz - 1))
    torch.cuda.synchronize()
    assert np.array_equal(t_dirty.cpu().numpy(), comp)
    dx = 0.01
    dt = 0.002
    for d in range(D):
        n = [nx, ny, nz]
        flo = [0, 0, 0]
        fhi = [nx - 1, ny - 1, nz - 1]
        fhi[d] += 1
        fshape = (D + 2, fhi[2] + 1, fhi[1] + 1, fhi[0] + 1)
        t_flux = torch.zeros(fshape, dtype=torch.float64, device=cuda)
        fm.ComputeNumericFluxes(fvs.PatchView(t_flux, flo), full, dt, dx, d)
        ref_flux = np.zeros(fshape)
        fv = O.host_view(ref_flux)
        sv = O.host_view(comp, lo)
        O.load().oracle_numeric_fluxes(C.byref(desc), C.byref(fv), C.byref(sv), dt, dx, d)
        torch.cuda.synchronize()
        got = t_flux.cpu().numpy()
        assert np.array_equal(got, ref_flux), _max_rel(got, ref_flux)
        # stable dt over the whole (ghost-grown) box
        sdt = fm.ComputeStableDt(full, dx, d)
        rdt = C.c_double()
        O.load().oracle_stable_dt(C.byref(desc), C.byref(sv), dx, d, C.byref(rdt))
        assert sdt.item() == rdt.value
        # update + complete on the interior
        inner_lo = (0, 0, 0)
        inner_hi = (nx - 1, ny - 1, nz - 1)
        t_next = t_comp.clone()
        nv = fvs.PatchView(t_next, lo).subview(inner_lo, inner_hi).components(0, D + 2)
        integ = fvs.HyperbolicPatchIntegrator(fm)
        integ.UpdateConservatively(nv, nv, fvs.PatchView(t_flux, flo), dt, dx, d)
        cview = fvs.PatchView(t_next, lo).subview(inner_lo, inner_hi)
        fvs.CompleteFromCons(fm, cview, cview)
        ref_next = comp.copy()
        tmp = np.ascontiguousarray(ref_next[:, g[2]:g[2] + nz, g[1]:g[1] + ny, g[0]:g[0] + nx])
        cons_part = np.ascontiguousarray(tmp[:D + 2])
        cpv = O.host_view(cons_part)
        O.load().oracle_update_conservatively(C.byref(cpv), C.byref(cpv), C.byref(fv), dt, dx, d)
        tmp[:D + 2] = cons_part
        iv = O.host_view(tmp)
        O.load().oracle_complete_view(D, 1.4, C.byref(iv), C.byref(iv))
        torch.cuda.synchronize()
        got_next = t_next.cpu().numpy()[:, g[2]:g[2] + nz, g[1]:g[1] + ny, g[0]:g[0] + nx]
        assert np.array_equal(got_next, tmp), _max_rel(got_next, tmp)


@pytest.mark.parametrize("splitting", ["Godunov", "Strang"])
def test_level_2d_circular_riemann(cuda, splitting):
    """Config 1 family (2D circular Riemann problem, MUSCL-Hancock + HLLEM,
    transmissive boundaries) at 64^2 for 30 steps: identical dt sequence and
    state as the oracle's level driver."""
    torch = _torch()
    from finitevolumesolver_amd.initial_data import circular_riemann_2d
    cons, dx = circular_riemann_2d(64)
    eq = fvs.PerfectGas(dim=2)
    fm = fvs.FluxMethod(eq)
    lvl = fvs.SinglePatchLevel(fm, (64, 64), dx, splitting=splitting, cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    done, fails, dts = lvl.advance(30, batch=7, return_dts=True)
    got = lvl.get_state().cpu().numpy()
    ref, ref_dts, ref_fails = O.level_run(fm._h.desc, cons, dx, 30,
                                          L.SPLIT_STRANG if splitting == "Strang" else L.SPLIT_GODUNOV)
    assert done == 30 and fails == ref_fails
    assert dts == list(ref_dts)
    assert np.array_equal(got, ref), _max_rel(got, ref)


def test_level_3d_fourier(cuda):
    torch = _torch()
    from finitevolumesolver_amd.initial_data import fourier_modes_3d
    cons, dx = fourier_modes_3d(24)
    eq = fvs.PerfectGas(dim=3)
    fm = fvs.FluxMethod(eq)
    lvl = fvs.SinglePatchLevel(fm, (24, 24, 24), dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    done, fails, dts = lvl.advance(6, batch=4, return_dts=True)
    got = lvl.get_state().cpu().numpy()
    ref, ref_dts, _ = O.level_run(fm._h.desc, cons, dx, 6)
    assert dts == list(ref_dts)
    assert np.array_equal(got, ref), _max_rel(got, ref)


def test_level_final_time_and_cfl_failure(cuda):
    """final_time caps the last dt; cfl > 1 makes the split-step CFL check
    fail and the step restart from the step-start state with cfl*failure_dt,
    exactly as the oracle's RunSimulation restatement does."""
    torch = _torch()
    from finitevolumesolver_amd.initial_data import circular_riemann_2d
    cons, dx = circular_riemann_2d(32)
    eq = fvs.PerfectGas(dim=2)
    fm = fvs.FluxMethod(eq)
    for cfl, final_time in [(0.8, 0.05), (1.5, 1e300)]:
        lvl = fvs.SinglePatchLevel(fm, (32, 32), dx, splitting="Strang", cfl=cfl)
        lvl.set_state(torch.from_numpy(cons).to(cuda))
        done, fails, dts = lvl.advance(12, final_time=final_time, batch=5, return_dts=True)
        got = lvl.get_state().cpu().numpy()
        ref, ref_dts, ref_fails = O.level_run(fm._h.desc, cons, dx, 12, L.SPLIT_STRANG, cfl=cfl,
                                              final_time=final_time)
        _note(f"cfl={cfl} steps={done} fails={fails} (oracle {len(ref_dts)}, {ref_fails})")
        assert dts == list(ref_dts)
        assert fails == ref_fails
        assert np.array_equal(got, ref)
        if final_time < 1e300:
            t, _ = lvl.time()
            assert abs(t - final_time) < 1e-15 and done < 12
        else:
            assert fails > 0


def test_level_run_lands_on_output_times(cuda):
    """RunSimulation's loop (include/fub/RunSimulation.hpp:86-191): each step
    capped by the next output time and the final time, the output called at
    exactly those times, the run ending once t + 1e-12 reaches the final
    time; state and dt sequence equal the oracle run piecewise between the
    output times (each piece resumed at the previous piece's time and cycle
    count) bit for bit."""
    torch = _torch()
    from finitevolumesolver_amd.initial_data import circular_riemann_2d
    cons, dx = circular_riemann_2d(32)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2))
    lvl = fvs.SinglePatchLevel(fm, (32, 32), dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    seen = []
    dts = lvl.run(0.12, output_times=[0.03, 0.07, 0.1],
                  output=lambda level, t, cycle: seen.append((t, cycle)))
    got = lvl.get_state().cpu().numpy()
    ref, ref_dts, cyc, t = cons, [], 0, 0.0
    for stop in (0.03, 0.07, 0.1, 0.12):
        ref, d, _ = O.level_run(fm._h.desc, ref, dx, 1000, L.SPLIT_GODUNOV, final_time=stop,
                                t0=t, cycle0=cyc)
        ref_dts += list(d)
        cyc += len(d)
        t = stop
    _note(f"run: {len(dts)} steps, outputs at {seen}")
    assert [round(x[0], 15) for x in seen] == [0.03, 0.07, 0.1]
    assert dts == ref_dts and lvl.time()[1] == cyc
    assert np.array_equal(got, ref)
    assert abs(lvl.time()[0] - 0.12) < 1e-15


@pytest.mark.parametrize("case", ["2d_strang", "3d_godunov", "godunov_base"])
def test_level_graph_replay_matches_eager(cuda, case):
    """The level's captured step graphs (single steps and chains of eight),
    replayed over several batches of 11 steps with host synchronisation in
    between, give the eager launches' states and dt sequence bit for bit."""
    torch = _torch()
    from finitevolumesolver_amd.initial_data import circular_riemann_2d, fourier_modes_3d
    if case == "2d_strang":
        cons, dx = circular_riemann_2d(48)
        fm = fvs.FluxMethod(fvs.PerfectGas(dim=2))
        shape, split = (48, 48), "Strang"
    elif case == "3d_godunov":
        cons, dx = fourier_modes_3d(20)
        fm = fvs.FluxMethod(fvs.PerfectGas(dim=3), precision="fp64_rcp")
        shape, split = (20, 20, 20), "Godunov"
    else:
        cons, dx = circular_riemann_2d(40)
        fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="Godunov", reconstruction="None")
        shape, split = (40, 40), "Godunov"
    levels = []
    for graphs in (True, False):
        lvl = fvs.SinglePatchLevel(fm, shape, dx, splitting=split, cfl=0.8)
        lvl.set_graphs(graphs)
        lvl.set_state(torch.from_numpy(cons).to(cuda))
        levels.append(lvl)
    for batch in range(3):
        outs = []
        for lvl in levels:
            done, fails, dts = lvl.advance(11, batch=11, return_dts=True)
            assert done == 11 and fails == 0
            outs.append((lvl.get_state().cpu().numpy(), dts))
        assert outs[0][1] == outs[1][1], (case, batch)
        assert np.array_equal(outs[0][0], outs[1][0]), (case, batch)
    print(f"{case}: 3 batches of 11 graph-replayed steps equal the eager ones", flush=True)


def test_stream_copy_probe(cuda):
    """fvs_stream_copy (the bench's bandwidth probe) copies every element of
    each array once; odd lengths and misaligned arrays are rejected."""
    torch = _torch()
    import ctypes as C
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=3))
    src = torch.rand((5, 1 << 20), dtype=torch.float64, device=cuda)
    dst = torch.zeros_like(src)
    sp = L.ptr_array([src[c].data_ptr() for c in range(5)])
    dp = L.ptr_array([dst[c].data_ptr() for c in range(5)])
    L.check(L.load().fvs_stream_copy(fm.handle, sp, dp, 5, C.c_int64(1 << 20)))
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    assert L.load().fvs_stream_copy(fm.handle, sp, dp, 5, C.c_int64(3)) == L.FVS_BAD_ARGUMENT
    odd = L.ptr_array([src[c].data_ptr() + 8 for c in range(5)])
    assert L.load().fvs_stream_copy(fm.handle, odd, dp, 5, C.c_int64(16)) == L.FVS_BAD_ARGUMENT
