"""The 3D level's step-start CFL pass with a speed floor.

The CFL pass evaluates the exact speed (HllemMethod::ComputeStableDt,
src/equations/perfect_gas/HllemMethod.hpp:1152-1216) only on faces whose cheap
upper bound reaches a floor, the previous pass's largest speed less a margin,
and repeats the step without the floor when no evaluated face reached it.
Every step's dt must still equal cfl times
the smallest stable dt of the full pass (fvs_patch_stable_dt, which evaluates
every face) on the step's start state, bit for bit, in every arithmetic
mode, on smooth and on shocked states; a floor above the true maximum (set
through fvs_level_hint_max_speed) must be caught and the step repeated
without changing any result."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from fields import rough_field

import finitevolumesolver_amd as fvs
from finitevolumesolver_amd.patch import patch_stable_dt

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _field(kind, n):
    from finitevolumesolver_amd.initial_data import fourier_modes_3d
    if kind == "fourier":
        return fourier_modes_3d(n)
    cons = rough_field((n, n + 2, n + 4), seed=5)
    return cons, (1.0 / n, 1.0 / n, 1.0 / n)


@pytest.mark.parametrize("prec,kind,split", [
    ("fp64", "fourier", "Godunov"),
    ("fp64", "rough", "Strang"),
    ("fp64_rcp", "fourier", "Strang"),
    ("fp64_rcp", "rough", "Godunov"),
    ("fp32", "rough", "Godunov"),
])
def test_level_dt_equals_full_cfl_pass(cuda, prec, kind, split):
    import torch
    _note(f"speed floor: {prec} {kind} {split}")
    cons, dx = _field(kind, 20)
    shape = (cons.shape[3], cons.shape[2], cons.shape[1])
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=3), precision=prec)
    lvl = fvs.SinglePatchLevel(fm, shape, dx, splitting=split, cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    for step in range(8):
        sdt = patch_stable_dt(fm, lvl.get_state(), dx).cpu().numpy()
        done, fails, dts = lvl.advance(1, batch=1, return_dts=True)
        assert done == 1 and fails == 0
        assert dts[0] == 0.8 * float(np.min(sdt)), (step, dts[0], sdt)
    _note(f"  8 steps, {lvl.cfl_redos()} repeated for an unreached floor")


@pytest.mark.parametrize("factor", [10.0, 0.5])
def test_wrong_speed_hint_repeats_the_step(cuda, factor):
    """A floor ten times the true largest speed hides every face: the pass
    reaches none, the first step is repeated without the floor and the run
    equals the unhinted one bit for bit; half the true speed is a valid
    floor that changes nothing either."""
    import torch
    cons, dx = _field("rough", 16)
    shape = (cons.shape[3], cons.shape[2], cons.shape[1])
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=3))
    runs = []
    for hint in (False, True):
        lvl = fvs.SinglePatchLevel(fm, shape, dx, splitting="Godunov", cfl=0.8)
        lvl.set_state(torch.from_numpy(cons).to(cuda))
        if hint:
            sdt = patch_stable_dt(fm, lvl.get_state(), dx).cpu().numpy()
            speed = [dx[d] / sdt[d] * factor for d in range(3)]
            lvl.hint_max_speed(speed)
        done, fails, dts = lvl.advance(5, batch=5, return_dts=True)
        assert done == 5 and fails == 0
        runs.append((lvl.get_state().cpu().numpy(), dts, lvl.cfl_redos()))
    assert runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][0], runs[1][0])
    if factor > 1.0:
        assert runs[1][2] >= 1
    _note(f"hint x{factor}: {runs[1][2]} repeated step(s) (unhinted {runs[0][2]}), "
          "results unchanged")


@pytest.mark.parametrize("prec,kind,dim", [
    ("fp64", "fourier", 3),
    ("fp64", "rough", 3),
    ("fp32", "rough", 3),
    ("fp64", "rough", 2),
])
def test_godunov_level_dt_equals_full_cfl_pass(cuda, prec, kind, dim):
    """First-order exact Godunov: the level's CFL pass solves Newton only on
    faces whose certified speed bound reaches the floor (and on every face
    of the first sweep's direction, whose star pressures the sweep reuses);
    every step's dt still equals cfl times the smallest stable dt of the pass
    that solves every face (fvs_patch_stable_dt) bit for bit, and fewer
    faces are solved than the three directions hold."""
    import torch
    _note(f"godunov speed floor: {prec} {kind} {dim}D")
    if dim == 3:
        cons, dx = _field(kind, 20)
        shape = (cons.shape[3], cons.shape[2], cons.shape[1])
    else:
        cons = rough_field((1, 40, 48), seed=9)
        dx = (1.0 / 48, 1.0 / 40)
        shape = (48, 40)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=dim), base="Godunov", reconstruction="None",
                        precision=prec)
    lvl = fvs.SinglePatchLevel(fm, shape, dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    solved = []
    for step in range(8):
        sdt = patch_stable_dt(fm, lvl.get_state(), dx).cpu().numpy()
        fm.newton_stats(reset=True)
        done, fails, dts = lvl.advance(1, batch=1, return_dts=True)
        assert done == 1 and fails == 0
        assert dts[0] == 0.8 * float(np.min(sdt)), (step, dts[0], sdt)
        solved.append(fm.newton_stats()["faces"])
    cells = int(np.prod(shape))
    _note(f"  faces solved per step {solved} ({cells} cells, {lvl.cfl_redos()} repeated)")
    # after the first step the floor skips faces: fewer than the 2D + 1 face
    # sets a step solves without it (CFL pass D sets, two sweeps)
    assert min(solved[1:]) < dim * cells + (dim - 1) * cells


def test_godunov_wrong_speed_hint_repeats_the_step(cuda):
    """A floor ten times the largest speed on the exact-Godunov level: the
    pass reaches none, the step is repeated and the run equals the unhinted
    one bit for bit."""
    import torch
    cons, dx = _field("rough", 16)
    shape = (cons.shape[3], cons.shape[2], cons.shape[1])
    runs = []
    for hint in (False, True):
        fm = fvs.FluxMethod(fvs.PerfectGas(dim=3), base="Godunov", reconstruction="None")
        lvl = fvs.SinglePatchLevel(fm, shape, dx, splitting="Godunov", cfl=0.8)
        lvl.set_state(torch.from_numpy(cons).to(cuda))
        if hint:
            sdt = patch_stable_dt(fm, lvl.get_state(), dx).cpu().numpy()
            lvl.hint_max_speed([10.0 * dx[d] / sdt[d] for d in range(3)])
        done, fails, dts = lvl.advance(5, batch=5, return_dts=True)
        assert done == 5 and fails == 0
        runs.append((lvl.get_state().cpu().numpy(), dts, lvl.cfl_redos()))
    assert runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[1][2] >= 1
    _note(f"godunov hint x10: {runs[1][2]} repeated step(s), results unchanged")
