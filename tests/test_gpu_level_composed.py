"""The single-patch level for the flux methods without a fused sweep:
MUSCL-Hancock over the exact-Godunov base (the reference's default
MusclHancockMethod, include/fub/flux_method/MusclHancockMethod.hpp:277-285),
the HLL base (first order and MUSCL-Hancock), MusclHancockPrim and
MusclHancockCharacteristic.  fvs_level_* composes the patch operators in the
reference's DimensionalSplitLevelIntegrator sequence; every step's dt and
the state must equal the oracle's level driver (oracle_level_run_from): bit
for bit where the operators are bit-exact (HLL, Prim, Characteristic), within
the exact solver's 1e-12 bar per cell (1e-13 for dt) over the Godunov base,
in 2D and 3D, Godunov and Strang splitting, and through a CFL failure and its
restart."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
from fields import rough_field
from tolerances import tolerance_ratio

import finitevolumesolver_amd as fvs

pytestmark = pytest.mark.gpu

DT_RTOL = 1e-13


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # (base, reconstruction, dim, splitting, exact)
    ("Godunov", "MusclHancock", 2, "Strang", False),
    ("Godunov", "MusclHancock", 3, "Godunov", False),
    ("HLL", "MusclHancock", 2, "Godunov", True),
    ("HLL", "None", 3, "Strang", True),
    ("HLLEM", "MusclHancockPrim", 2, "Godunov", True),
    ("HLLEM", "MusclHancockCharacteristic", 2, "Strang", True),
]


def _field(dim, seed):
    if dim == 2:
        cons = rough_field((1, 20, 24), seed=seed)
        return cons, (1.0 / 24, 1.0 / 20), (24, 20)
    cons = rough_field((10, 11, 12), seed=seed)
    return cons, (1.0 / 12, 1.0 / 11, 1.0 / 10), (12, 11, 10)


def _run(cuda, base, recon, dim, split, cfl, steps, seed):
    import torch
    cons, dx, shape = _field(dim, seed)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=dim), base=base, reconstruction=recon)
    lvl = fvs.SinglePatchLevel(fm, shape, dx, splitting=split, cfl=cfl)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    done, fails, dts = lvl.advance(steps, batch=steps, return_dts=True)
    got = lvl.get_state().cpu().numpy()
    sp = O.L.SPLIT_STRANG if split == "Strang" else O.L.SPLIT_GODUNOV
    ref, ref_dts, ref_fails = O.level_run(fm._h.desc, cons, dx, steps, splitting=sp, cfl=cfl)
    t, cycles = lvl.time()
    return done, fails, np.array(dts), got, ref, ref_dts, ref_fails, t, cycles


@pytest.mark.parametrize("base,recon,dim,split,exact", CASES)
def test_composed_level_matches_oracle(cuda, base, recon, dim, split, exact):
    _note(f"composed level: {base} + {recon}, {dim}D, {split}")
    done, fails, dts, got, ref, ref_dts, ref_fails, t, cycles = _run(
        cuda, base, recon, dim, split, 0.8, 4, seed=21)
    assert done == 4 and fails == ref_fails == 0
    t_ref = 0.0
    for d in (ref_dts if exact else dts):
        t_ref += d
    assert cycles == 4 and t == t_ref, (cycles, t, t_ref)
    if exact:
        assert list(dts) == list(ref_dts)
        np.testing.assert_array_equal(got, ref)
        _note(f"  4 steps bit-exact, dts {list(dts)}")
    else:
        assert np.max(np.abs(dts - ref_dts) / ref_dts) <= DT_RTOL
        ratio = tolerance_ratio(got, ref)
        _note(f"  4 steps, tolerance ratio {ratio:.2f}")
        assert ratio <= 1.0


def test_composed_level_final_time_and_restarts_like_the_oracle(cuda):
    """final_time caps the last dt; cfl 1.5 with Strang half steps makes some
    split-step CFL checks fail, and the step is repeated from its start state
    with the failing dt, as the oracle's RunSimulation loop does; dts,
    failure count and state equal the oracle's bit for bit."""
    import torch
    from finitevolumesolver_amd.initial_data import circular_riemann_2d
    cons, dx = circular_riemann_2d(32)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLL", reconstruction="MusclHancock")
    for cfl, final_time in [(0.8, 0.05), (1.5, 1e300)]:
        lvl = fvs.SinglePatchLevel(fm, (32, 32), dx, splitting="Strang", cfl=cfl)
        lvl.set_state(torch.from_numpy(cons).to(cuda))
        done, fails, dts = lvl.advance(12, final_time=final_time, batch=5, return_dts=True)
        got = lvl.get_state().cpu().numpy()
        ref, ref_dts, ref_fails = O.level_run(fm._h.desc, cons, dx, 12, O.L.SPLIT_STRANG,
                                              cfl=cfl, final_time=final_time)
        _note(f"composed level cfl={cfl}: steps {done}, fails {fails} (oracle {len(ref_dts)}, "
              f"{ref_fails})")
        assert dts == list(ref_dts)
        assert fails == ref_fails
        np.testing.assert_array_equal(got, ref)
        if final_time < 1e300:
            t, _ = lvl.time()
            assert abs(t - final_time) < 1e-15 and done < 12
        else:
            assert fails > 0
