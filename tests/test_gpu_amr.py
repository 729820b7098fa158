"""AMR hierarchy on the GPU (fvs_amr_*) against the AMR oracle.

Three levels, ratio 2 (tests/amr_cases.py): the device hierarchy step
(ghost fill from same-level patches and piecewise-constant coarse data,
subcycled batched sweeps, flux-register reflux, average_down) must equal
the oracle's composition of the per-patch FP64 oracle sweeps bit for bit on
every level after every coarse step, with the same stable dt and per-level
CFL check values (the oracle's minimum where it lies below the level's time
step, otherwise a value between that time step and the minimum); with no
flux through the domain edges, mass and energy of the composite are
conserved to round-off."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
from amr_oracle import AmrOracle  # noqa: E402

import amr_cases as A  # noqa: E402
import finitevolumesolver_amd as fvs  # noqa: E402
from finitevolumesolver_amd import _lib as L  # noqa: E402
from finitevolumesolver_amd.amr import AmrHierarchy  # noqa: E402

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)


def _pair(state, graphs=True, sequencing="reference"):
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock")
    h = AmrHierarchy(fm, A.DOMAIN, A.LAYOUTS, A.DX0, 2)
    h.set_graphs(graphs)
    h.set_sequencing(sequencing)
    h.set_state(state)
    o = AmrOracle(A.DOMAIN, A.LAYOUTS, A.DX0, 2, A.oracle_sweep(fm._h.desc), sequencing)
    for l, lay in enumerate(A.LAYOUTS):
        for p in range(len(lay)):
            o.U[l][p][:, 2:-2, 2:-2] = state(l, p)
    for f in range(len(A.LAYOUTS) - 1, 0, -1):
        h.average_down(f)
        o.average_down(f)
    return fm, h, o


def _compare(h, o, what):
    for l in range(h.nlevels):
        got = h.interiors(l)
        for p, g in enumerate(got):
            np.testing.assert_array_equal(g, o.U[l][p][:, 2:-2, 2:-2],
                                          err_msg=f"{what}: level {l} patch {p}")


@pytest.mark.parametrize("state,graphs,sequencing", [
    ("moving", True, "reference"), ("moving", False, "reference"), ("resting", True, "reference"),
    ("shock_bubble", True, "reference"), ("moving", True, "refill"),
    ("shock_bubble", False, "refill")])
def test_amr_steps_match_oracle(cuda, state, graphs, sequencing):
    fn = {"moving": A.moving_state, "resting": A.resting_blob,
          "shock_bubble": A.shock_bubble}[state]
    _note(f"amr parity: {state} state, 3 levels, graphs {graphs}, {sequencing} sequencing")
    fm, h, o = _pair(fn, graphs, sequencing)
    _compare(h, o, "after average_down")
    for step in range(4):
        if step == 2:
            # retuned tiles of a level: the captured steps must be replaced
            h.levels[1].set_shape(0, 2)
            h.levels[2].set_shape(1, 3)
        dt_g, dt_o = h.stable_dt(), o.stable_dt()
        assert dt_g == dt_o, (step, dt_g, dt_o)
        dt = 0.8 * dt_g
        h.advance(dt)
        o.advance(dt)
        _compare(h, o, f"step {step}")
        # the check value: the exact minimum where it is below the level's
        # time step, else at least that time step (and at most the minimum)
        for l, (g, e) in enumerate(zip(h.level_cfl(), o.level_cfl)):
            dt_l = dt / 2 ** l
            assert (g == e) if e < dt_l else (dt_l <= g <= e), (step, l, g, e, dt_l)
        _note(f"  step {step}: dt {dt!r}, all levels bit-exact, level cfl {h.level_cfl()}")


def test_amr_conserves_composite_integrals(cuda):
    _note("amr conservation: resting blob, 3 levels, 3 steps")
    fm, h, o = _pair(A.resting_blob)

    def total():
        return sum(u.sum(axis=(1, 2)) for u in h.interiors(0)) * A.DX0 ** 2

    before = total()
    dts = h.run(3)
    after = total()
    rel = np.abs(after[[0, 3]] - before[[0, 3]]) / np.abs(before[[0, 3]])
    _note(f"  dts {dts}, relative change of mass and energy {rel}")
    assert rel.max() < 1e-14


def test_amr_rejects_bad_layouts(cuda):
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock")
    bad = [
        [A.LAYOUTS[0], [(17, 16, 16, 24)]],                       # not aligned to the ratio
        [A.LAYOUTS[0], [(16, 16, 16, 24), (24, 16, 16, 16)]],     # overlapping
        [[(0, 0, 16, 32)], [(16, 16, 16, 24)]],                   # level 0 misses half the domain
        [A.LAYOUTS[0], A.LAYOUTS[1], [(30, 36, 24, 40)]],         # level 2 not nested in level 1
    ]
    for lay in bad:
        with pytest.raises(L.FvsError):
            AmrHierarchy(fm, A.DOMAIN, lay, A.DX0, 2)


def test_amr_tags_match_oracle_and_regrid_nests(cuda):
    """GPU gradient tags of level 1 equal the oracle's on the same state
    (ghosts filled), and the tile regrid built from them is a layout the
    hierarchy accepts as its level 2."""
    import oracle_lib as O
    from amr_oracle import gradient_tags
    from finitevolumesolver_amd.amr import tile_layout
    _note("amr tagging: moving state, level 1")
    fm, h, o = _pair(A.moving_state)
    tags = h.tag_cells(1, 0.004, 0.004)
    b = h.levels[1]
    pool = b.download(0).cpu()
    total = 0
    for p, (nx, ny) in enumerate(b.sizes):
        cons = b.patch_view(pool, p).numpy()
        comp = O.complete_from_cons(np.ascontiguousarray(cons).reshape(4, -1), 2)
        comp = comp.reshape(6, ny + 4, nx + 4)
        ref = gradient_tags(comp[0], comp[4], 0.004, 0.004)
        np.testing.assert_array_equal(tags[p], ref, err_msg=f"patch {p}")
        total += int(ref.sum())
    assert total > 0
    new2 = tile_layout(A.DOMAIN, A.LAYOUTS[1], tags, level=1, tile=4, ratio=2, margin=1)
    _note(f"  {total} tagged cells, {len(new2)} level-2 patches from 4x4 tiles")
    assert new2
    AmrHierarchy(fm, A.DOMAIN, [A.LAYOUTS[0], A.LAYOUTS[1], new2], A.DX0, 2)


def test_amr_regrid_transfer(cuda):
    """Regrid of the finest level from level-1 tags: levels 0 and 1 keep
    their state bit for bit; a new level-2 cell holds the old level-2 cell
    where the old layout had one, else the level-1 cell above it; the new
    hierarchy then steps."""
    _note("amr regrid: moving state, one step, regrid level 2, one step")
    fm, h, o = _pair(A.moving_state)
    h.advance(0.8 * h.stable_dt())
    new = h.regrid_finest(0.004, 0.004, tile=4)
    _note(f"  new level 2: {len(new.layouts[2])} patches")
    for l in (0, 1):
        for a, b in zip(h.interiors(l), new.interiors(l)):
            np.testing.assert_array_equal(a, b)
    old2 = h.interiors(2)
    lvl1 = h.interiors(1)
    for p, got in enumerate(new.interiors(2)):
        i0, j0, nx, ny = new.layouts[2][p]
        for j in range(ny):
            for i in range(nx):
                I, J = i0 + i, j0 + j
                src = None
                for q, (a, b, w, hh) in enumerate(h.layouts[2]):
                    if a <= I < a + w and b <= J < b + hh:
                        src = old2[q][:, J - b, I - a]
                if src is None:
                    for q, (a, b, w, hh) in enumerate(h.layouts[1]):
                        if a <= I // 2 < a + w and b <= J // 2 < b + hh:
                            src = lvl1[q][:, J // 2 - b, I // 2 - a]
                np.testing.assert_array_equal(got[:, j, i], src)
    new.advance(0.8 * new.stable_dt())
    assert np.all(np.isfinite(new.level_cfl()))


def _expected_transfer(old, new, level, coarse_new):
    """Per new patch of `level`: the old same-level cell where the old
    layout had one, else the new coarser level's cell above it."""
    olds = old.interiors(level)
    out = []
    for (i0, j0, nx, ny) in new.layouts[level]:
        exp = np.zeros((4, ny, nx))
        for j in range(ny):
            for i in range(nx):
                I, J = i0 + i, j0 + j
                src = None
                for q, (a, b, w, hh) in enumerate(old.layouts[level]):
                    if a <= I < a + w and b <= J < b + hh:
                        src = olds[q][:, J - b, I - a]
                if src is None:
                    for q, (a, b, w, hh) in enumerate(new.layouts[level - 1]):
                        if a <= I // 2 < a + w and b <= J // 2 < b + hh:
                            src = coarse_new[q][:, J // 2 - b, I // 2 - a]
                exp[:, j, i] = src
        out.append(exp)
    return out


def test_amr_regrid_all_finer_levels(cuda):
    """RegridAllFinerlevels(0) on the shock-bubble state after two steps:
    levels 1 and 2 are rebuilt from the tags of the current levels 0 and 1,
    level 2 nested in the new level 1; level 0 keeps its state bit for bit,
    every new cell holds the old same-level cell or, where the old layout had
    none, the new coarser cell above it; the new hierarchy keeps the cycle
    and steps on."""
    from finitevolumesolver_amd.amr import nests
    _note("amr regrid all: shock-bubble state, two steps, regrid levels 1 and 2")
    fm, h, o = _pair(A.shock_bubble)
    for _ in range(2):
        h.advance(0.8 * h.stable_dt())
    new = h.regrid_all(0, 0.05, 0.05, tile=4)
    _note(f"  level 1: {len(h.layouts[1])} -> {len(new.layouts[1])} patches, "
          f"level 2: {len(h.layouts[2])} -> {len(new.layouts[2])} patches")
    assert new.layouts[1] != h.layouts[1] or new.layouts[2] != h.layouts[2]
    for l in (1, 2):
        ext = (A.DOMAIN[0] * 2 ** (l - 1), A.DOMAIN[1] * 2 ** (l - 1))
        assert all(nests(p, new.layouts[l - 1], ext, 2, 1) for p in new.layouts[l])
    for a, b in zip(h.interiors(0), new.interiors(0)):
        np.testing.assert_array_equal(a, b)
    for l in (1, 2):
        exp = _expected_transfer(h, new, l, new.interiors(l - 1))
        for p, got in enumerate(new.interiors(l)):
            np.testing.assert_array_equal(got, exp[p], err_msg=f"level {l} patch {p}")
    assert new.cycle == h.cycle == 2
    new.advance(0.8 * new.stable_dt())
    assert np.all(np.isfinite(new.level_cfl()))


def test_amr_run_regrids_on_a_cadence(cuda):
    """run() with regrid_every = 2 rebuilds the finer levels before steps 2
    and 4 of the shock-bubble run: the layouts follow the shock, every step
    passes its CFL checks, and the state stays finite and positive."""
    _note("amr run with regridding: shock-bubble state, 5 steps, every 2")
    fm, h, o = _pair(A.shock_bubble)
    before = [list(l) for l in h.layouts]
    dts = h.run(5, cfl=0.8, regrid_every=2, tol_rho=0.05, tol_p=0.05)
    assert len(dts) == 5 and h.cycle == 5
    assert h.layouts != before
    for l in range(h.nlevels):
        for u in h.interiors(l):
            assert np.all(np.isfinite(u)) and np.all(u[0] > 0)
    _note(f"  dts {dts}, level sizes {[len(l) for l in h.layouts]}")


def test_amr_checkpoint_restart_continues_bit_for_bit(cuda, tmp_path):
    """WriteCheckpointFile after two steps of the shock-bubble run, then
    ReadCheckpointFile into a new hierarchy: two more steps from the
    restart equal two more steps of the original run bit for bit (dts,
    every level, cycle and time)."""
    from finitevolumesolver_amd import amrex_files as ck
    _note("amr checkpoint/restart: shock-bubble state, 2 + 2 steps")
    fm, h, o = _pair(A.shock_bubble)
    h.run(2, cfl=0.8)
    ck.write_checkpoint(tmp_path / "chk00002", h)
    r = ck.read_checkpoint(tmp_path / "chk00002", fm, A.DOMAIN, A.DX0, 2)
    assert r.layouts == h.layouts and r.cycle == h.cycle == 2 and r.time == h.time
    for l in range(h.nlevels):
        for a, b in zip(h.interiors(l), r.interiors(l)):
            np.testing.assert_array_equal(a, b)
    dts_a = h.run(2, cfl=0.8)
    dts_b = r.run(2, cfl=0.8)
    assert dts_a == dts_b and r.time == h.time and r.cycle == h.cycle == 4
    for l in range(h.nlevels):
        for a, b in zip(h.interiors(l), r.interiors(l)):
            np.testing.assert_array_equal(a, b)


def test_amr_run_lands_on_output_times_and_final_time(cuda):
    """run() caps each coarse step at the next output time and the final
    time (RunSimulation's limited_dt) and calls the output there; dts and
    every level equal the oracle's run loop with the same caps bit for bit,
    and the run ends at the final time."""
    _note("amr run with output times: moving state")
    fm, h, o = _pair(A.moving_state)
    dt0 = 0.8 * h.stable_dt()
    outs = [1.5 * dt0, 2.2 * dt0]
    final = 3.7 * dt0
    seen = []
    dts = h.run(20, cfl=0.8, final_time=final, output_times=outs,
                output=lambda hh: seen.append(hh.time))
    ref_dts, _ = A.oracle_run(o, 20, 0.8, final_time=final, output_times=outs)
    _note(f"  dts {dts}, outputs at {seen}")
    assert dts == ref_dts
    assert len(seen) == 2 and all(abs(a - b) <= 1e-15 for a, b in zip(seen, outs))
    assert abs(h.time - final) <= 1e-15
    _compare(h, o, "after the capped run")


def test_amr_run_restarts_on_cfl_failure(cuda):
    """With cfl 1.6 every attempt of a step exceeds a level's stable dt (as
    in the reference, a cfl above 1 never passes the check): run() undoes
    each attempt, retries with 1.6 times the stable dt the sweeps found, and
    after max_restarts raises TimeStepTooLarge with the state of before the
    step restored bit for bit."""
    _note("amr restart: moving state, cfl 1.6, 3 restarts")
    fm, h, o = _pair(A.moving_state)
    before = [h.interiors(l) for l in range(h.nlevels)]
    with pytest.raises(L.TimeStepTooLarge):
        h.run(1, cfl=1.6, max_restarts=3)
    for l in range(h.nlevels):
        for a, b in zip(before[l], h.interiors(l)):
            np.testing.assert_array_equal(a, b)
    # and the hierarchy steps normally afterwards
    assert len(h.run(1, cfl=0.8)) == 1


def test_amr_run_retry_keeps_split_order(cuda):
    """cfl 0.99 on the resting bump: the first attempt of each of the first
    two coarse steps fails the CFL check and its retry passes.  run()
    restores the cycle count with the state, so the retry keeps the failed
    attempt's split order: dts, attempts and every level equal the oracle's
    run loop bit for bit (which differs when the cycle is not restored,
    tests/test_amr_oracle.py)."""
    _note("amr retry: resting bump, cfl 0.99, 2 steps")
    fm, h, o = _pair(A.resting_blob)
    ref_dts, attempts = A.oracle_run(o, 2, 0.99)
    assert attempts[0] == 2
    dts = h.run(2, cfl=0.99)
    assert dts == ref_dts
    assert h.cycle == o.cycle == 2
    _compare(h, o, "after two steps with retries")


@pytest.mark.parametrize("precision,tol", [("fp64_rcp", 1e-12), ("fp32", 1e-6)])
def test_amr_tolerance_modes(cuda, precision, tol):
    """The refined-reciprocal FP64 and the FP32-flux forms of the hierarchy
    step against the FP64 oracle over three steps (same dts) of the moving
    blob, the resting bump and the shock-bubble state: relative L1 of
    density and energy per level within tol (1e-12 and 1e-6; momentum, whose
    sums are small, is reported);
    composite mass and energy of the resting bump conserved to tol."""
    _note(f"amr {precision}: moving state, 3 steps")
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock",
                        precision=precision)
    for state in (A.moving_state, A.resting_blob, A.shock_bubble):
        h = AmrHierarchy(fm, A.DOMAIN, A.LAYOUTS, A.DX0, 2)
        h.set_state(state)
        ref = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock")
        o = AmrOracle(A.DOMAIN, A.LAYOUTS, A.DX0, 2, A.oracle_sweep(ref._h.desc))
        for l, lay in enumerate(A.LAYOUTS):
            for p in range(len(lay)):
                o.U[l][p][:, 2:-2, 2:-2] = state(l, p)
        for f in (2, 1):
            h.average_down(f)
            o.average_down(f)
        before = o.total()
        for _ in range(3):
            dt = 0.8 * o.stable_dt()
            h.advance(dt)
            o.advance(dt)
        worst = 0.0
        mom = 0.0
        for l in range(h.nlevels):
            for g, p in zip(h.interiors(l), range(len(A.LAYOUTS[l]))):
                r = o.U[l][p][:, 2:-2, 2:-2]
                with np.errstate(invalid="ignore", divide="ignore"):
                    err = np.abs(g - r).sum(axis=(1, 2)) / np.abs(r).sum(axis=(1, 2))
                worst = max(worst, float(err[[0, 3]].max()))
                m = err[[1, 2]]
                mom = max(mom, float(m[np.isfinite(m)].max(initial=0.0)))
        _note(f"  {state.__name__}: worst relative L1 density/energy {worst:.3e}, "
              f"momentum {mom:.3e}")
        assert worst <= tol
        if state is A.resting_blob:
            tot = sum(u.sum(axis=(1, 2)) for u in h.interiors(0)) * A.DX0 ** 2
            rel = np.abs(tot[[0, 3]] - before[[0, 3]]) / np.abs(before[[0, 3]])
            _note(f"  conservation of mass and energy {rel}")
            assert rel.max() <= tol


def test_amr_write_hdf5_roundtrip_frame(cuda, tmp_path):
    """The hierarchy's HDF5 record: two frames of the composite, the second
    after a step; the file's frame count and times follow the calls."""
    import shutil
    import subprocess
    from finitevolumesolver_amd import output
    try:
        output._load()
    except (ImportError, OSError) as e:
        pytest.skip(f"HDF5 writer unavailable: {e}")
    fm, h, o = _pair(A.moving_state)
    path = tmp_path / "amr.h5"
    h.write_hdf5(path, 0.0, 0, xlower=(-1.0, -1.0))
    dt = 0.8 * h.stable_dt()
    h.advance(dt)
    h.write_hdf5(path, dt, 1, xlower=(-1.0, -1.0))
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if shutil.which(h5dump):
        head = subprocess.run([h5dump, "-H", str(path)], capture_output=True, text=True).stdout
        assert "( 2, 6, 128, 128 )" in head
