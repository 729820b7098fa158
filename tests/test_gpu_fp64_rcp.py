"""FP64 with refined hardware reciprocals (precision "fp64_rcp") against the
IEEE FP64 oracle: every fused sweep, CFL reduction, level run and batched
sweep must agree to 1e-12 relative per cell (tolerances.tolerance_ratio),
and the time steps to 1e-13 relative."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
from fields import rough_field
from sweep_ref import oracle_sweep
from tolerances import tolerance_ratio

import finitevolumesolver_amd as fvs
from finitevolumesolver_amd import _lib as L

pytestmark = pytest.mark.gpu
DT_RTOL = 1e-13


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    ((1, 48, 64), "HLLEM", "MinMod", True),
    ((1, 37, 300), "HLLEM", "VanLeer", True),
    ((1, 40, 33), "HLLEM_plain", "MinMod", True),
    ((18, 20, 24), "HLLEM", "MinMod", True),
    ((12, 14, 270), "HLLEM_plain", "VanLeer", True),
    ((10, 12, 16), "HLLEM", "NoLimiter", False),
]


@pytest.mark.parametrize("case", CASES)
def test_rcp_sweep_within_1e12(cuda, case):
    import torch
    shape, base, limiter, quirks = case
    D = 2 if shape[0] == 1 else 3
    _note(f"fp64_rcp sweep shape={shape} base={base} lim={limiter}")
    cons = rough_field(shape, seed=31 + shape[2], jumps=(limiter != "NoLimiter"))
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=D), base=base, limiter=limiter, quirks=quirks,
                        precision="fp64_rcp")
    desc = fm._h.desc
    assert desc.precision == L.FP64_RCP
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
        r = tolerance_ratio(got, ref)
        rel = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300))
        _note(f"  dir {d}: tolerance ratio {r:.3e} (max rel {rel:.2e}), "
              f"dt rel {abs(sdt.item() - ref_dt) / ref_dt:.2e}")
        assert r <= 1.0
        assert abs(sdt.item() - ref_dt) <= DT_RTOL * ref_dt


@pytest.mark.parametrize("D", [2, 3])
def test_rcp_patch_stable_dt(cuda, D):
    import torch
    shape = (1, 40, 52) if D == 2 else (14, 16, 18)
    cons = rough_field(shape, seed=5)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=D), precision="fp64_rcp")
    dx = [1.0 / 52, 1.3 / 40, 0.9 / 14][:D]
    got = fvs.patch_stable_dt(fm, torch.from_numpy(cons).to(cuda), dx).cpu().numpy()
    for d in range(D):
        _, ref_dt, _ = oracle_sweep(fm._h.desc, cons, 0.0, dx[d], d)
        assert abs(got[d] - ref_dt) <= DT_RTOL * ref_dt, (d, got[d], ref_dt)


@pytest.mark.parametrize("splitting", ["Godunov", "Strang"])
def test_rcp_level_2d_circular_riemann(cuda, splitting):
    import torch
    from finitevolumesolver_amd.initial_data import circular_riemann_2d
    cons, dx = circular_riemann_2d(64)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), precision="fp64_rcp")
    lvl = fvs.SinglePatchLevel(fm, (64, 64), dx, splitting=splitting, cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    done, fails, dts = lvl.advance(30, batch=7, return_dts=True)
    got = lvl.get_state().cpu().numpy()
    ref, ref_dts, ref_fails = O.level_run(fm._h.desc, cons, dx, 30,
                                          L.SPLIT_STRANG if splitting == "Strang" else L.SPLIT_GODUNOV)
    assert done == 30 and fails == ref_fails
    dt_rel = np.max(np.abs(np.array(dts) - ref_dts) / ref_dts)
    r = tolerance_ratio(got, ref)
    _note(f"  {splitting}: 30 steps, dt rel {dt_rel:.2e}, tolerance ratio {r:.3e}")
    assert dt_rel <= DT_RTOL
    assert r <= 1.0


def test_rcp_level_3d_fourier(cuda):
    import torch
    from finitevolumesolver_amd.initial_data import fourier_modes_3d
    cons, dx = fourier_modes_3d(24)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=3), precision="fp64_rcp")
    lvl = fvs.SinglePatchLevel(fm, (24, 24, 24), dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    done, fails, dts = lvl.advance(6, batch=4, return_dts=True)
    got = lvl.get_state().cpu().numpy()
    ref, ref_dts, _ = O.level_run(fm._h.desc, cons, dx, 6)
    dt_rel = np.max(np.abs(np.array(dts) - ref_dts) / ref_dts)
    r = tolerance_ratio(got, ref)
    _note(f"  3D: 6 steps, dt rel {dt_rel:.2e}, tolerance ratio {r:.3e}")
    assert dt_rel <= DT_RTOL
    assert r <= 1.0


def test_rcp_batched_sweep(cuda):
    """Batched AMR sweeps in fp64_rcp against the per-patch IEEE oracle."""
    import torch
    from test_gpu_batch import _fill_pool, _oracle_sweep, _sizes
    from finitevolumesolver_amd.batch import PatchBatch
    sizes = _sizes(24, seed=12)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock",
                        precision="fp64_rcp")
    dx = [1.0 / 64] * len(sizes)
    batch = PatchBatch(fm, sizes, dx)
    pool = batch.new_pool()
    _fill_pool(batch, pool, seed=300)
    host = pool.cpu().numpy()
    dt = 0.002
    for d in range(2):
        batch.upload(pool, 0)
        batch.upload(pool, 1)
        batch.set_dt(dt)
        dt_dev = torch.empty(1, dtype=torch.float64, device=cuda)
        batch.sweep(d, 0, dt_dev)
        out = batch.download(1).cpu().numpy()
        worst, cfl = 0.0, np.inf
        for p, (nx, ny) in enumerate(batch.sizes):
            o = batch.offsets[p]
            cons = host[o:o + 4 * (nx + 4) * (ny + 4)].reshape(4, ny + 4, nx + 4)
            ref, pdt = _oracle_sweep(fm._h.desc, cons, dt, dx[p], d)
            got = out[o:o + 4 * (nx + 4) * (ny + 4)].reshape(4, ny + 4, nx + 4)[:, 2:-2, 2:-2]
            worst = max(worst, tolerance_ratio(got, ref))
            cfl = min(cfl, pdt)
        _note(f"  batched dir {d}: tolerance ratio {worst:.3e}")
        assert worst <= 1.0
        # the CFL check value: at least dt while the check passes (faces
        # whose certified bound passes report dt), the minimum otherwise
        got = dt_dev.item()
        if cfl > dt * (1 + DT_RTOL):
            assert dt <= got <= cfl * (1 + DT_RTOL), (got, dt, cfl)
        else:
            assert abs(got - cfl) <= DT_RTOL * cfl
        # a stable-dt pass (time step 0) evaluates every face
        batch.upload(pool, 0)
        batch.set_dt(0.0)
        batch.sweep(d, 0, dt_dev)
        assert abs(dt_dev.item() - cfl) <= DT_RTOL * cfl, (dt_dev.item(), cfl)
