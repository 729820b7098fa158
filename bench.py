"""Benchmark: 3D uniform-grid split sweeps of the perfect-gas Euler equations.

Workload (the 256^3 uniform-grid configuration): one 256^3 patch on [0,1]^3
with transmissive boundaries, MUSCL-Hancock (MinMod) + HLLEM (Larrouturou),
X-Y-Z Godunov splitting with direction reversal on odd cycles, CFL 0.8, FP64.
Initial state: rho = 1 + 0.2 * (8 seeded Fourier modes, |k| <= 4),
p = 1 + 0.1 * (same form), velocities N(0, 0.1) smoothed (synthetic data).

A step is one full time step: the CFL reduction over the state plus three
fused split sweeps (each with its own CFL check).  `value` is interior
cell-updates per second summed over the sweeps (Gcell-updates/s), with the
state resident in HBM before the timed region starts.

The JSON line also carries `variants`: the same workload with FP32
reconstruction/flux, the 384^3 exact-Godunov workload with and without the
learned Newton guess, the 2048x1024 cut-cell channel and the batched AMR
levels of the shock-bubble hierarchy.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 2] [--precision fp64]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
BYTES_PER_CELL_UPDATE_3D = 80  # 5 conservative FP64 read + write


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(precision: str, kernel: str):
    """HBM bytes per launch of the dominant sweep kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each counter scaled by what it
    reads for the streaming probe's known traffic), or None when the summary
    is absent."""
    path = ROOT / "profiles" / "sweep3d_256_hbm_traffic.txt"
    if not path.exists():
        return None
    names = {"X": "SweepXKernel", "Y": "SweepPencilKernel<3, 1", "Z": "SweepPencilKernel<3, 2"}
    section = None
    for line in path.read_text().splitlines():
        if line.startswith("-- "):
            section = line[3:].strip()
        elif section == precision and line.startswith(names[kernel]):
            # kernel, grid, launches, read MiB, write MiB, bytes/launch, vs alg
            return int(line.split(">")[-1].split()[4])
    return None


def measured_copy_bandwidth(torch) -> float:
    """Achievable HBM bandwidth on this GPU: a 2 GiB device-to-device copy
    (bytes read + written per second, best of five)."""
    a = torch.empty(1 << 28, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    best = 0.0
    for _ in range(5):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = max(best, 2 * a.numel() * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9)
    del a, b
    torch.cuda.empty_cache()
    return best


def measured_stream_bandwidth(torch, fvs, fm, state) -> float:
    """Achievable HBM bandwidth of a sweep's traffic: the library's streaming
    probe (fvs_stream_copy) copies the five 256^3 state arrays, every element
    read and written once (bytes read + written per second, best of five,
    HIP events on the handle's stream, the current torch stream)."""
    import ctypes as C
    from finitevolumesolver_amd import _lib as L
    out = torch.empty_like(state)
    nc, n = state.shape[0], state[0].numel()
    src = L.ptr_array([state[c].data_ptr() for c in range(nc)])
    dst = L.ptr_array([out[c].data_ptr() for c in range(nc)])
    best = 0.0
    for _ in range(5):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        L.check(L.load().fvs_stream_copy(fm.handle, src, dst, nc, C.c_int64(n)))
        e1.record()
        e1.synchronize()
        best = max(best, 2 * nc * n * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9)
    del out
    torch.cuda.empty_cache()
    return best


def bench_c1(torch, fvs) -> dict:
    """C1: 2D 256^2 circular Riemann problem on [-1, 1]^2, MUSCL-Hancock
    (MinMod) + HLLEM, Strang split, CFL 0.8, 100 steps, FP64 (IEEE), with the
    oracle's final state compared bit for bit (the state is cache resident:
    launch-bound, reported without a roofline)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np
    import oracle_lib as O
    from finitevolumesolver_amd import _lib as L
    from finitevolumesolver_amd.initial_data import circular_riemann_2d
    cons, dx = circular_riemann_2d(256)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2))
    res = {}
    for graphs in (False, True):
        lvl = fvs.SinglePatchLevel(fm, (256, 256), dx, splitting="Strang", cfl=0.8)
        lvl.set_graphs(graphs)
        # untimed run: the first advance captures the level's step graphs
        # (single steps and 8-step chains per start buffer and cycle parity);
        # set_state then restarts the run at t = 0
        lvl.set_state(torch.from_numpy(cons).cuda())
        lvl.advance(6, batch=6)
        lvl.set_state(torch.from_numpy(cons).cuda())
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        done, fails = lvl.advance(100, batch=100)
        e1.record()
        e1.synchronize()
        res[graphs] = (e0.elapsed_time(e1), done, lvl.get_state().cpu().numpy())
    ms, done, got = res[True]
    ref, _, _ = O.level_run(fm._h.desc, cons, dx, 100, L.SPLIT_STRANG)
    return {"workload": "2D 256^2 circular Riemann, MUSCL-Hancock+HLLEM, Strang, 100 steps, fp64",
            "value": round(256 * 256 * 4 * done / (ms * 1e-3) / 1e9, 4),
            "unit": "Gcell-updates/s", "steps": done, "ms_per_step": round(ms / max(done, 1), 4),
            "steps_as": "replayed HIP graphs (captured in an untimed 6-step run)",
            "eager_ms_per_step": round(res[False][0] / max(res[False][1], 1), 4),
            "bitexact_vs_oracle": bool(np.array_equal(got, ref) and np.array_equal(res[False][2], ref))}


def cpu_baseline(n: int, steps: int, precision: str) -> dict:
    """Times the FP64 CPU oracle (this project's restatement of the reference
    algorithm; the reference itself cannot be built without its libraries)
    on a bounded sample of the same workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    import finitevolumesolver_amd as fvs
    from finitevolumesolver_amd.initial_data import fourier_modes_3d
    from finitevolumesolver_amd.patch import make_desc
    # Every CPU this process may run on, unless the job's thread share
    # (OMP_NUM_THREADS, set by the GPU host for its per-GPU CPU share) is
    # smaller; FVS_CPU_THREADS overrides both.
    affinity = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", affinity) or affinity)
    threads = int(os.environ.get("FVS_CPU_THREADS", min(affinity, share)))
    threads = O.load().oracle_set_threads(threads)
    cons, dx = fourier_modes_3d(n)
    desc = make_desc(fvs.PerfectGas(3))
    t0 = time.perf_counter()
    _, dts, _ = O.level_run(desc, cons, dx, steps)
    el = time.perf_counter() - t0
    updates = n ** 3 * 3 * len(dts)
    # one core on a smaller sample of the same workload
    O.load().oracle_set_threads(1)
    n1, s1 = 64, 4
    c1, dx1 = fourier_modes_3d(n1)
    t1 = time.perf_counter()
    _, d1, _ = O.level_run(desc, c1, dx1, s1)
    el1 = time.perf_counter() - t1
    O.load().oracle_set_threads(threads)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": updates / el / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle level driver, {n}^3 patch, {len(dts)} steps x 3 sweeps "
                      f"({updates / 1e6:.1f} M cell-updates, {el:.1f} s, {threads} threads)",
            "single_core": {"value": n1 ** 3 * 3 * len(d1) / el1 / 1e9,
                            "sample": f"{n1}^3 patch, {len(d1)} steps, 1 thread, {el1:.1f} s"},
            "cpu_model": model, "host_cpus": os.cpu_count(), "affinity_cpus": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def l1_vs_oracle(precision: str, torch) -> dict:
    """Error of the GPU path against the FP64 oracle on a 48^3 sample of the
    same workload (4 steps): L1 / Linf of density, relative to the oracle's
    density, and the largest per-cell relative error over all components
    (with the 1e-13 x component-maximum floor of the parity tests)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np
    import oracle_lib as O
    import finitevolumesolver_amd as fvs
    from finitevolumesolver_amd.initial_data import fourier_modes_3d
    from tolerances import tolerance_ratio
    n = 48
    cons, dx = fourier_modes_3d(n)
    fm = fvs.FluxMethod(fvs.PerfectGas(3), precision=precision)
    lvl = fvs.SinglePatchLevel(fm, (n, n, n), dx)
    lvl.set_state(torch.from_numpy(cons).cuda())
    lvl.advance(4, batch=4)
    got = lvl.get_state().cpu().numpy()
    fm64 = fvs.FluxMethod(fvs.PerfectGas(3))
    ref, _, _ = O.level_run(fm64._h.desc, cons, dx, 4)
    d = np.abs(got[0] - ref[0])
    return {"l1_density": float(d.sum() / np.abs(ref[0]).sum()),
            "linf_density": float((d / np.abs(ref[0])).max()),
            "max_rel_per_cell": float(tolerance_ratio(got, ref) * 1e-12),
            "bitexact": bool(np.array_equal(got, ref)), "sample": "48^3, 4 steps"}


def kernels_only(args, torch, fvs) -> int:
    """Each fused kernel (X, Y, Z sweep without and with the CFL check, the
    CFL pass; IEEE FP64, refined-reciprocal FP64 and FP32 flux) --kernel-reps
    times on the 256^3 state, then one device copy of
    a 1 GiB buffer: the copy's FETCH_SIZE/WRITE_SIZE calibrate the counters
    (it moves exactly 1 GiB each way)."""
    from finitevolumesolver_amd.initial_data import fourier_modes_3d_device
    n = args.size
    state, dx = fourier_modes_3d_device(n, seed0=1)
    out = torch.empty_like(state)
    sdt = torch.empty(1, dtype=torch.float64, device="cuda")
    dt = 0.5 * 0.8 / n
    for prec in ("fp64", "fp64_rcp", "fp32"):
        fm = fvs.FluxMethod(fvs.PerfectGas(3), precision=prec)
        for d, name in enumerate("XYZ"):
            for check in (False, True):
                log(f"[kernels] {prec} sweep {name} check={check}")
                for _ in range(args.kernel_reps):
                    fvs.split_sweep(fm, state, out, dt, dx[d], d, sdt if check else None)
                torch.cuda.synchronize()
        log(f"[kernels] {prec} CFL")
        for _ in range(args.kernel_reps):
            fvs.patch_stable_dt(fm, state, dx)
        torch.cuda.synchronize()
    if not args.no_variants:
        # first-order exact-Godunov sweeps (plain guess) on the same state
        fg = fvs.FluxMethod(fvs.PerfectGas(3), base="Godunov", reconstruction="None")
        for d, name in enumerate("XYZ"):
            log(f"[kernels] godunov sweep {name}")
            for _ in range(args.kernel_reps):
                fvs.godunov_sweep(fg, state, out, dt, dx[d], d)
            torch.cuda.synchronize()
        del out
        torch.cuda.empty_cache()
        log("[kernels] learned-guess inference (matrix cores)")
        from finitevolumesolver_amd.newton_guess import LearnedNewtonGuess
        net = LearnedNewtonGuess(seed=1)
        npairs = 1 << 22
        net.generate(npairs, seed=42)
        prim = net.dataset()["prim"]
        def complete(rho, u, p):
            q = torch.zeros((6, npairs), dtype=torch.float64, device=prim.device)
            q[0], q[1], q[4] = rho, rho * u, p
            q[3] = 0.5 * rho * u * u + p / 0.4
            q[5] = torch.sqrt(1.4 * p / rho)
            return q
        left = complete(prim[0], prim[1], prim[2])
        right = complete(prim[3], prim[4], prim[5])
        for _ in range(args.kernel_reps):
            net.infer_pairs(2, 0, left, right)
        torch.cuda.synchronize()
        del net, prim, left, right
        torch.cuda.empty_cache()
        log("[kernels] cut-cell level steps")
        bench_cutcell(2048, 1024, args.kernel_reps, torch, fvs)
        log("[kernels] batched AMR level steps")
        bench_amr(args.kernel_reps, torch, fvs)
    log("[kernels] copy 1 GiB")
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    torch.cuda.empty_cache()
    log("[kernels] stream probe: five state arrays copied (fvs_stream_copy), five times")
    measured_stream_bandwidth(torch, fvs, fvs.FluxMethod(fvs.PerfectGas(3)), state)
    log("[kernels] done")
    return 0


_DTYPES = {"fp64": "f64 (IEEE division and square root)",
           "fp64_rcp": "f64 (refined hardware reciprocals)",
           "fp32": "f32 reconstruction and flux, f64 update"}


def bench_level(state, dx, n, steps, precision, torch, fvs) -> dict:
    """The same 3D workload in another arithmetic mode."""
    fm = fvs.FluxMethod(fvs.PerfectGas(3), precision=precision)
    lvl = fvs.SinglePatchLevel(fm, (n, n, n), dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(state)
    lvl.advance(1, batch=1)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    done, fails = lvl.advance(steps, batch=steps)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    out = {"value": round(n ** 3 * 3 * done / (ms * 1e-3) / 1e9, 4), "unit": "Gcell-updates/s",
           "steps": done, "ms_per_step": round(ms / max(done, 1), 4),
           "dtype": _DTYPES[precision]}
    try:
        out["accuracy"] = l1_vs_oracle(precision, torch)
    except Exception as e:
        out["accuracy"] = {"error": str(e)}
    return out


def bench_composed(state, dx, n, steps, torch, fvs) -> dict:
    """The reference's default method, MUSCL-Hancock over the exact Godunov
    flux (FP64 Newton), on the same 3D field: a level composed from the patch
    operators (ghost fill, stable dt, face fluxes, conservative update and
    completion per split step), with no fused sweep."""
    if state is None:
        from finitevolumesolver_amd.initial_data import fourier_modes_3d_device
        state, dx = fourier_modes_3d_device(n, seed0=1)
    fm = fvs.FluxMethod(fvs.PerfectGas(3), base="Godunov", reconstruction="MusclHancock")
    lvl = fvs.SinglePatchLevel(fm, (n, n, n), dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(state)
    lvl.advance(1, batch=1)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    done, fails = lvl.advance(steps, batch=steps)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    del lvl
    return {"workload": f"3D {n}^3, MUSCL-Hancock(MinMod) + exact Godunov (FP64 Newton), "
                        "Godunov split, CFL 0.8, composed from the patch operators",
            "value": round(n ** 3 * 3 * done / (ms * 1e-3) / 1e9, 4), "unit": "Gcell-updates/s",
            "steps": done, "cfl_restarts": fails, "ms_per_step": round(ms / max(done, 1), 4),
            "dtype": "f64"}


def bench_godunov(n: int, steps: int, torch, fvs, epochs: int = 12) -> dict:
    """3D first-order exact Godunov at n^3 on two fields: the smooth
    Fourier-mode field (seeds 11-15) and a strong-wave field of 4^3-cell
    blocks (density and pressure log-uniform in [0.1, 10], velocities in
    [-0.5, 0.5]).  Per form one untimed step, then `steps` timed steps.
    Four forms per field: plain FP64 (mean-pressure Newton
    guess), learned guess (network trained here on the GPU, inferred inline
    on the matrix cores), mixed precision (FP64 Newton with exp/log powers,
    float sample and flux, FP64 update) and mixed precision with the learned
    guess.  L1 density error of each form against the plain FP64 GPU result
    (itself within 1e-12 of the CPU oracle)."""
    from finitevolumesolver_amd.initial_data import blocky_field_3d_device, fourier_modes_3d_device
    from finitevolumesolver_amd.newton_guess import LearnedNewtonGuess
    res = {"workload": f"3D {n}^3 exact Godunov (first order), Godunov split, CFL 0.8, "
                       f"FP64 Newton; smooth Fourier-mode field and strong-wave block field"}
    t0 = time.perf_counter()
    net = LearnedNewtonGuess(seed=1)
    counts = net.generate(1 << 22, seed=42)
    hist = net.train(epochs=epochs, batch=1024, lr=5e-3, huber=0.25, seed=3)
    torch.cuda.synchronize()
    res["training"] = {"pairs": 1 << 22, "usable": counts["usable"], "vacuum": counts["vacuum"],
                       "epochs": epochs, "seconds": round(time.perf_counter() - t0, 2),
                       "final_loss": float(hist[-1])}
    forms = (("plain", "fp64", False), ("learned", "fp64", True),
             ("mixed", "fp32", False), ("mixed_learned", "fp32", True))
    for field in ("smooth", "strong"):
        log(f"[godunov] {n}^3 {field} field")
        if field == "smooth":
            state, dx = fourier_modes_3d_device(n, seed0=11)
        else:
            state, dx = blocky_field_3d_device(n, seed=5)
        out = {}
        finals = {}
        for form, prec, learned in forms:
            fm = fvs.FluxMethod(fvs.PerfectGas(3), base="Godunov", reconstruction="None",
                                precision=prec)
            if learned:
                net.attach(fm)
            lvl = fvs.SinglePatchLevel(fm, (n, n, n), dx, splitting="Godunov", cfl=0.8)
            lvl.set_state(state)
            # one untimed step: kernel loading, graph capture, and the first
            # CFL pass, which has no speed floor yet
            lvl.advance(1, batch=1)
            fm.newton_stats(reset=True)
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            ev0.record()
            done, fails = lvl.advance(steps, batch=steps)
            ev1.record()
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1)
            st = fm.newton_stats()
            out[form] = {"value": round(n ** 3 * 3 * done / (ms * 1e-3) / 1e9, 4),
                         "unit": "Gcell-updates/s", "steps": done,
                         "ms_per_step": round(ms / max(done, 1), 3),
                         "newton_iterations_per_face":
                             round(st["iterations"] / max(st["faces"], 1), 4),
                         "newton_max_iterations": st["max_iterations"],
                         "fallback_rate": round(st["fallbacks"] / max(st["faces"], 1), 6),
                         "not_converged": st["not_converged"]}
            finals[form] = lvl.get_state()[0].clone()
            log(f"[godunov] {field} {form}: {out[form]}")
            del lvl
        ref = finals["plain"]
        for form in ("learned", "mixed", "mixed_learned"):
            d = (finals[form] - ref).abs()
            out[form]["l1_density_vs_fp64"] = float(d.sum() / ref.abs().sum())
            out[form]["linf_density_vs_fp64"] = float((d / ref.abs()).max())
        it0 = out["plain"]["newton_iterations_per_face"]
        out["iterations_saved_learned"] = round(1.0 - out["learned"]["newton_iterations_per_face"]
                                                / max(it0, 1e-30), 4)
        out["speedup_mixed_learned_over_plain"] = round(
            out["mixed_learned"]["value"] / out["plain"]["value"], 4)
        res[field] = out
        del state, finals
        torch.cuda.empty_cache()
    return res


def bench_cutcell(nx: int, ny: int, steps: int, torch, fvs) -> dict:
    """2D cut-cell channel: Mach 2 shock (x = -0.4) towards the cylinder and
    the 30-degree wedge, small cells shrunk to volume fractions in
    [1e-4, 1e-2]; MUSCL-Hancock(MinMod) + HLLEM regular fluxes, KBN h-box
    stabilisation with the exact boundary solver, Godunov split, CFL 0.8."""
    from finitevolumesolver_amd import cutcell as K
    from finitevolumesolver_amd.geometry import channel_geometry
    from finitevolumesolver_amd.initial_data import shock_channel
    t0 = time.perf_counter()
    geom = channel_geometry(nx, ny)
    cons, dx = shock_channel(nx, ny, x_shock=-0.4)
    log(f"[cutcell] geometry {nx}x{ny} in {time.perf_counter() - t0:.1f} s")
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock")
    g = K.CutCellData.from_geometry2d(geom, torch.device("cuda:0"))
    lvl = K.CutCellLevel(fm, (nx, ny), (dx, dx), g)
    lvl.set_state(torch.from_numpy(cons).cuda())
    lvl.advance(1, batch=1)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    done, fails = lvl.advance(steps, batch=steps)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    value = nx * ny * 2 * done / (ms * 1e-3) / 1e9
    alg = 80 * nx * ny * 2 * done / (ms * 1e-3) / 1e9
    return {"workload": f"2D {nx}x{ny} cut-cell channel, cylinder + 30-degree wedge, "
                        f"Mach 2 shock, MUSCL-Hancock+HLLEM with KBN h-box",
            "value": round(value, 4), "unit": "Gcell-updates/s", "steps": done,
            "cut_cells": lvl.ncut, "shrunk_cells": geom.meta.get("perturbed"),
            "ms_per_step": round(ms / max(done, 1), 4),
            "hbm_frac_at_80B": round(alg / HBM_PEAK_GBS, 4)}


def bench_amr(steps: int, torch, fvs) -> dict:
    """Batched AMR levels of the 2D shock-bubble hierarchy: level 0 one 512^2
    patch, levels 1 and 2 with 300 patches each (sides from {16..64},
    seeded, non-overlapping); each level's split step (ghost fill, X sweep,
    ghost fill, Y sweep) is one hipGraph replay of two persistent launches
    with autotuned tiles.  FP32 reconstruction/flux with FP64 update (and
    FP64 for comparison).  A coarse step advances level 1 twice and level 2
    four times (refinement ratio 2)."""
    import numpy as np
    from finitevolumesolver_amd.batch import GHOSTS, PatchBatch
    from finitevolumesolver_amd.initial_data import (amr_patch_layout, cons_from_prim,
                                                     shock_bubble_state)
    layouts = [([(0, 0, 512, 512)], 2.0 / 512)]
    for lev in (1, 2):
        layouts.append(amr_patch_layout(lev))
    res = {"workload": "2D shock-bubble AMR levels: 512^2 base + 2 x 300 patches "
                       "(ratio 2), per-patch split sweeps, transmissive patch ghosts"}
    for prec in ("fp32", "fp64"):
        fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock",
                            precision=prec)
        per_level = []
        total_updates = 0.0
        total_ms = 0.0
        for lev, (lay, dx) in enumerate(layouts):
            sizes = [(a[2], a[3]) for a in lay]
            batch = PatchBatch(fm, sizes, dx)
            pool = batch.new_pool()
            for p, (i0, j0, nx, ny) in enumerate(lay):
                xs = -1.0 + (np.arange(i0 - GHOSTS, i0 + nx + GHOSTS) + 0.5) * dx
                ys = -1.0 + (np.arange(j0 - GHOSTS, j0 + ny + GHOSTS) + 0.5) * dx
                X, Y = np.meshgrid(xs, ys)
                rho, u, v, pr = shock_bubble_state(X, Y)
                cons = cons_from_prim(rho, [u, v], pr).reshape(4, ny + 4, nx + 4)
                batch.patch_view(pool, p).copy_(torch.from_numpy(cons))
            batch.upload(pool, 0)
            batch.set_dt(0.3 * dx / 3.0)
            shapes = [batch.tune(d, reps=3)[0] for d in (0, 1)]
            batch.upload(pool, 0)
            batch.capture()
            batch.replay(1)
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            ev0.record()
            batch.replay(steps)
            ev1.record()
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / steps
            cells = sum(nx * ny for nx, ny in sizes)
            sub = 2 ** lev
            total_updates += 2.0 * cells * sub
            total_ms += ms * sub
            per_level.append({"level": lev, "patches": len(sizes), "cells": cells,
                              "ms_per_step": round(ms, 4),
                              "gcell_updates_per_s": round(2 * cells / (ms * 1e-3) / 1e9, 3),
                              "tiles": [batch.tile_shape(d)[:2] for d in (0, 1)],
                              "tuned_shapes": shapes})
            del batch, pool
        res[prec] = {"levels": per_level,
                     "coarse_step_ms": round(total_ms, 4),
                     "value": round(total_updates / (total_ms * 1e-3) / 1e9, 4),
                     "unit": "Gcell-updates/s"}
    return res


def bench_amr_hierarchy(steps: int, torch, fvs) -> dict:
    """End-to-end subcycled hierarchy step of the 2D shock-bubble problem:
    level 0 512^2, level 1 the 300 seeded patches, level 2 nested inside them
    (ratio 2).  One coarse step = ghost fills (same-level copies and
    piecewise-constant coarse data), level 2 four times and level 1 twice
    with their flux-register contributions, level 0 once, reflux and
    average_down; the sweeps use the autotuned tile shapes.  Cell-updates
    count each level's cells times its sweeps per coarse step."""
    import numpy as np
    from finitevolumesolver_amd.amr import AmrHierarchy
    from finitevolumesolver_amd.initial_data import (amr_nested_levels, cons_from_prim,
                                                     shock_bubble_state)
    layouts = amr_nested_levels()
    res = {"workload": "2D shock-bubble AMR hierarchy, 3 levels (512^2, 300 patches, "
                       f"{len(layouts[2])} nested patches), ratio 2, fine-first subcycling "
                       "with ghost fill, reflux and average_down"}
    for prec in ("fp64", "fp32"):
        fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock",
                            precision=prec)
        h = AmrHierarchy(fm, (512, 512), layouts, 2.0 / 512, 2)

        def state(l, p):
            X, Y = h.cell_centres(l, p)
            rho, u, v, pr = shock_bubble_state(X - 1.0, Y - 1.0)
            return cons_from_prim(rho, [u, v], pr).reshape(4, X.shape[0], X.shape[1])

        h.set_state(state)
        for f in (2, 1):
            h.average_down(f)
        for b in h.levels:
            for d in (0, 1):
                b.tune(d, reps=3)
        dt = 0.8 * h.stable_dt()
        h.advance(dt)  # warm-up
        ok = all(h.level_cfl()[l] >= dt / 2 ** l for l in range(3))
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(steps):
            h.advance(dt)
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / steps
        cells = [sum(p[2] * p[3] for p in lay) for lay in layouts]
        updates = sum(2.0 * c * 2 ** l for l, c in enumerate(cells))
        res[prec] = {"coarse_step_ms": round(ms, 4), "cells_per_level": cells,
                     "cell_updates_per_coarse_step": int(updates),
                     "value": round(updates / (ms * 1e-3) / 1e9, 4), "unit": "Gcell-updates/s",
                     "dt": dt, "cfl_ok": bool(ok)}
        del h
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", default="fp64_rcp", choices=["fp64", "fp64_rcp", "fp32"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--kernels-only", action="store_true",
                    help="launch each fused kernel --kernel-reps times (and a plain "
                         "copy of known size) for hardware-counter runs; no timing line")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP32-flux, exact-Godunov and cut-cell variants")
    ap.add_argument("--godunov-size", type=int, default=384)
    ap.add_argument("--godunov-steps", type=int, default=5)
    ap.add_argument("--cutcell-steps", type=int, default=20)
    ap.add_argument("--amr-steps", type=int, default=20)
    ap.add_argument("--godunov-epochs", type=int, default=12)
    ap.add_argument("--only", choices=["godunov", "c1", "cutcell", "amr", "amr_hier", "composed"],
                    help="run one variant alone and print its JSON (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path "
              "(the patch shards as z-slabs with 2-plane halos).",
              file=sys.stderr)
        return 2

    import torch
    import finitevolumesolver_amd as fvs
    from finitevolumesolver_amd.initial_data import fourier_modes_3d

    n = args.size
    if args.kernels_only:
        return kernels_only(args, torch, fvs)
    if args.only:
        run = {"godunov": lambda: bench_godunov(args.godunov_size, args.godunov_steps, torch, fvs,
                                                epochs=args.godunov_epochs),
               "c1": lambda: bench_c1(torch, fvs),
               "cutcell": lambda: bench_cutcell(2048, 1024, args.cutcell_steps, torch, fvs),
               "amr": lambda: bench_amr(args.amr_steps, torch, fvs),
               "amr_hier": lambda: bench_amr_hierarchy(args.amr_steps, torch, fvs),
               "composed": lambda: bench_composed(None, None, n, 3, torch, fvs)}[args.only]
        print(json.dumps({args.only: run()}), flush=True)
        return 0
    from finitevolumesolver_amd.initial_data import fourier_modes_3d_device
    log(f"[bench] generating {n}^3 initial state")
    dev = torch.device("cuda:0")
    t_cons, dx = fourier_modes_3d_device(n, seed0=1)
    fm = fvs.FluxMethod(fvs.PerfectGas(3), precision=args.precision)
    lvl = fvs.SinglePatchLevel(fm, (n, n, n), dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(t_cons)
    log("[bench] warmup")
    lvl.advance(args.warmup, batch=max(args.warmup, 1))
    torch.cuda.synchronize()

    log(f"[bench] timing {args.steps} steps")
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    done, fails = lvl.advance(args.steps, batch=args.steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    updates = n ** 3 * 3 * done
    value = updates / (ms * 1e-3) / 1e9
    log(f"[bench] {done} steps ({fails} CFL restarts) in {ms:.2f} ms -> {value:.3f} Gcell-updates/s")

    # Per-kernel durations, measured live: the level repeats the timed steps
    # with HIP events recorded on its stream around every launch.
    lvl.set_timing(True)
    lvl.advance(args.steps, batch=args.steps)
    times = lvl.kernel_times()
    lvl.set_timing(False)
    per = {"X": times["sweep_x"], "Y": times["sweep_y"], "Z": times["sweep_z"],
           "CFL": times["cfl"]}
    per = {k: ms / max(cnt, 1) for k, (ms, cnt) in per.items()}
    dominant = max(("X", "Y", "Z"), key=lambda k: per[k])
    alg_bytes = BYTES_PER_CELL_UPDATE_3D * n ** 3
    achieved = alg_bytes / (per[dominant] * 1e-3) / 1e9
    log("[bench] per-launch ms: " + ", ".join(f"{k}={v:.3f}" for k, v in per.items()))
    traffic = measured_traffic(args.precision, dominant)
    hbm_copy = measured_copy_bandwidth(torch)
    hbm_stream = measured_stream_bandwidth(torch, fvs, fm, t_cons)

    result = {
        "metric": "Gcell-updates/s per split sweep, % HBM roofline, L1 err vs FP64 CPU oracle",
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / max(done, 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision != "fp32" else "f32-flux/f64-update",
        "data": "synthetic",
        "config": {
            "workload": f"3D uniform grid {n}^3 (one patch), MUSCL-Hancock(MinMod)+HLLEM "
                        f"(Larrouturou), X-Y-Z Godunov split, CFL 0.8, transmissive, "
                        f"{args.precision}",
            "arithmetic": _DTYPES[args.precision],
            "cells": n ** 3, "sweeps_per_step": 3, "steps_completed": done,
            "cfl_restarts": fails, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "kernel": f"fused sweep {dominant}",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_cell_update": BYTES_PER_CELL_UPDATE_3D,
            "per_launch_ms": {k: round(v, 4) for k, v in per.items()},
            "hbm_copy_GBs": round(hbm_copy, 1),
            "frac_of_copy": round(achieved / hbm_copy, 4),
            "hbm_stream_GBs": round(hbm_stream, 1),
            "frac_of_stream": round(achieved / hbm_stream, 4),
            "stream_probe": "fvs_stream_copy of the five state arrays (the sweep's 80 B/cell "
                            "without its arithmetic)",
        },
    }
    try:
        result["accuracy"] = l1_vs_oracle(args.precision, torch)
    except Exception as e:  # accuracy is reported, never required for timing
        result["accuracy"] = {"error": str(e)}
    if not args.no_variants:
        variants = {}
        del lvl
        torch.cuda.empty_cache()
        for prec in ("fp64", "fp64_rcp", "fp32"):
            if prec == args.precision:
                continue
            log(f"[bench] variant: {prec}")
            variants[f"{prec}_{n}"] = bench_level(t_cons, dx, n, args.steps, prec, torch, fvs)
        log("[bench] variant: MUSCL-Hancock + exact Godunov, composed level")
        torch.cuda.empty_cache()
        variants["muscl_godunov_composed"] = bench_composed(t_cons, dx, n, 3, torch, fvs)
        del t_cons
        torch.cuda.empty_cache()
        log("[bench] variant: exact Godunov, plain and learned guess")
        variants["godunov_exact"] = bench_godunov(args.godunov_size, args.godunov_steps, torch, fvs)
        torch.cuda.empty_cache()
        log("[bench] variant: C1 2D circular Riemann")
        variants["c1_2d_256"] = bench_c1(torch, fvs)
        torch.cuda.empty_cache()
        log("[bench] variant: cut-cell channel")
        variants["cutcell_2048x1024"] = bench_cutcell(2048, 1024, args.cutcell_steps, torch, fvs)
        torch.cuda.empty_cache()
        log("[bench] variant: batched AMR levels")
        variants["batched_amr"] = bench_amr(args.amr_steps, torch, fvs)
        variants["amr_hierarchy"] = bench_amr_hierarchy(args.amr_steps, torch, fvs)
        result["variants"] = variants
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        result["cpu_baseline"] = cpu_baseline(192, 16, args.precision)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
