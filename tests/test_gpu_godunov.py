"""GPU exact-Godunov path against the FP64 oracle.

The device pow and the host libm pow may differ in the last bit, so these
comparisons use the FP64 bar of <= 1e-12 relative per cell instead of bit
equality (the HLLEM path, which uses no pow, is compared bit for bit)."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
from fields import rough_field
from sweep_ref import oracle_sweep, padded_complete
from tolerances import tolerance_ratio

import finitevolumesolver_amd as fvs
from finitevolumesolver_amd import _lib as L

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("shape", [(1, 30, 41), (9, 11, 13)])
@pytest.mark.parametrize("recon", ["None", "MusclHancock"])
def test_godunov_view_fluxes_and_dt(cuda, shape, recon):
    import torch
    D = 2 if shape[0] == 1 else 3
    _note(f"godunov view shape={shape} recon={recon}")
    cons = rough_field(shape, seed=71)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=D), base="Godunov", reconstruction=recon)
    desc = fm._h.desc
    gw = 2 * fm.GetStencilWidth()
    comp, g = padded_complete(cons, D, gw)
    nz, ny, nx = shape
    lo = (-g[0], -g[1], -g[2])
    full = fvs.PatchView(torch.from_numpy(comp).to(cuda), lo)
    dx, dt = 0.01, 0.002
    for d in range(D):
        fhi = [nx - 1, ny - 1, nz - 1]
        fhi[d] += 1
        fshape = (D + 2, fhi[2] + 1, fhi[1] + 1, fhi[0] + 1)
        t_flux = torch.zeros(fshape, dtype=torch.float64, device=cuda)
        fm.ComputeNumericFluxes(fvs.PatchView(t_flux, (0, 0, 0)), full, dt, dx, d)
        ref = np.zeros(fshape)
        fv, sv = O.host_view(ref), O.host_view(comp, lo)
        assert O.load().oracle_numeric_fluxes(C.byref(desc), C.byref(fv), C.byref(sv), dt, dx, d) == 0
        got = t_flux.cpu().numpy()
        scale = np.maximum(np.abs(ref), 1e-3)
        err = float(np.max(np.abs(got - ref) / scale))
        _note(f"  dir {d}: flux max rel {err:.2e}")
        assert err <= TOL
        sdt = fm.ComputeStableDt(full, dx, d).item()
        rdt = C.c_double()
        O.load().oracle_stable_dt(C.byref(desc), C.byref(sv), dx, d, C.byref(rdt))
        assert abs(sdt - rdt.value) <= TOL * rdt.value


@pytest.mark.parametrize("shape", [(1, 40, 52), (14, 16, 18)])
def test_godunov_fused_sweep(cuda, shape):
    import torch
    D = 2 if shape[0] == 1 else 3
    _note(f"godunov fused sweep shape={shape}")
    cons = rough_field(shape, seed=72)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=D), base="Godunov", reconstruction="None")
    desc = fm._h.desc
    dx = 1.0 / shape[2]
    fm.newton_stats(reset=True)
    for d in range(D):
        dt = 0.2 * dx
        ref, ref_dt, _ = oracle_sweep(desc, cons, dt, dx, d)
        t_in = torch.from_numpy(cons).to(cuda)
        t_out = torch.empty_like(t_in)
        sdt = torch.empty(1, dtype=torch.float64, device=cuda)
        fvs.split_sweep(fm, t_in, t_out, dt, dx, d, sdt)
        got = t_out.cpu().numpy()
        err = _rel(got, ref)
        _note(f"  dir {d}: state max rel {err:.2e}, dt {sdt.item()!r} vs {ref_dt!r}")
        assert err <= TOL
        assert abs(sdt.item() - ref_dt) <= TOL * ref_dt
    st = fm.newton_stats()
    faces = sum((shape[2 - d] + 1) * np.prod([shape[2 - e] for e in range(D) if e != d])
                for d in range(D))
    assert st["faces"] == faces
    assert st["not_converged"] == 0
    assert 0 < st["iterations"] < 20 * faces


def test_godunov_guess_at_root_needs_no_iterations(cuda):
    """Seeding Newton with the converged star pressure leaves nothing to do:
    the per-face guess array reaches the solver in the documented layout."""
    import torch
    shape = (1, 20, 24)
    cons = rough_field(shape, seed=73, jumps=False)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="Godunov", reconstruction="None")
    nz, ny, nx = shape
    # star pressures of every x face from the oracle (face box (nx+1) x ny)
    q = O.complete_from_cons(cons.reshape(4, -1), 2).reshape(6, ny, nx)
    left = np.concatenate([q[:, :, :1], q], axis=2)   # cell f-1 (clamped)
    right = np.concatenate([q, q[:, :, -1:]], axis=2)  # cell f (clamped)
    pM, _, _, _ = O.exact_middle_state(left.reshape(6, -1), right.reshape(6, -1), 2, 0)
    guess = torch.from_numpy(pM.reshape(ny, nx + 1).copy()).to(cuda)
    t_in = torch.from_numpy(cons).to(cuda)
    fm.newton_stats(reset=True)
    fvs.godunov_sweep(fm, t_in, None, 0.0, 1.0 / nx, 0, guess=guess)
    st = fm.newton_stats(reset=True)
    assert st["faces"] == (nx + 1) * ny
    assert st["iterations"] <= 2  # the oracle root satisfies |f| <= 1e-7 on the GPU too
    fvs.godunov_sweep(fm, t_in, None, 0.0, 1.0 / nx, 0)
    st0 = fm.newton_stats(reset=True)
    assert st0["iterations"] > st["iterations"]


@pytest.mark.parametrize("D", [2, 3])
def test_godunov_patch_stable_dt(cuda, D):
    import torch
    shape = (1, 28, 36) if D == 2 else (12, 14, 16)
    cons = rough_field(shape, seed=74)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=D), base="Godunov", reconstruction="None")
    dx = [1.0 / shape[2], 1.1 / shape[1], 0.9 / shape[0]][:D]
    got = fvs.patch_stable_dt(fm, torch.from_numpy(cons).to(cuda), dx).cpu().numpy()
    for d in range(D):
        _, ref_dt, _ = oracle_sweep(fm._h.desc, cons, 0.0, dx[d], d)
        assert abs(got[d] - ref_dt) <= TOL * ref_dt, (d, got[d], ref_dt)


def test_godunov_level_3d(cuda):
    """Config 5 family (3D exact Godunov, first order) at 20^3 for 5 steps."""
    import torch
    from finitevolumesolver_amd.initial_data import fourier_modes_3d
    cons, dx = fourier_modes_3d(20, seed0=11)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=3), base="Godunov", reconstruction="None")
    lvl = fvs.SinglePatchLevel(fm, (20, 20, 20), dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    done, fails, dts = lvl.advance(5, batch=5, return_dts=True)
    got = lvl.get_state().cpu().numpy()
    ref, ref_dts, ref_fails = O.level_run(fm._h.desc, cons, dx, 5)
    assert done == 5 and fails == ref_fails
    assert np.max(np.abs(np.array(dts) - ref_dts) / ref_dts) <= TOL
    ratio = tolerance_ratio(got, ref)
    comp = [_rel(got[c], ref[c]) for c in range(got.shape[0])]
    _note(f"godunov level: tolerance ratio {ratio:.2f}, per component max rel "
          f"{['%.1e' % e for e in comp]}")
    assert ratio <= 1.0
