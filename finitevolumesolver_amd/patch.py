"""Host mirror of the reference's patch-level API over the C ABI.

Class and method names follow FiniteVolumeSolver so that code written
against the reference reads the same:

  fub::PerfectGas<Dim>                    include/fub/equations/PerfectGas.hpp:163-222
  fub::FluxMethod<Base>                   include/fub/flux_method/FluxMethod.hpp:57-382
  fub::HyperbolicPatchIntegrator<Tag>     include/fub/HyperbolicPatchIntegrator.hpp:34-59
  fub::CompleteFromConsFn                 include/fub/CompleteFromCons.hpp:141-204
  fub::amrex::TransmissiveBoundary        src/AMReX/boundary_condition/TransmissiveBoundary.cpp:31-93

Device memory and streams come from torch; every computation runs in
libfvs_hip.so.  Views are component-major with x fastest: a tensor of shape
(ncomp, nz, ny, nx) (nz = 1 in 2D) covering the index box [lo, hi].
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import torch

from . import _lib as L

DIRECTIONS = {"X": 0, "Y": 1, "Z": 2}


def _dir(d) -> int:
    return DIRECTIONS[d] if isinstance(d, str) else int(d)


@dataclass
class PerfectGas:
    """PerfectGas<Dim>: gamma = 1.4, Rspec = 287.058 (PerfectGas.hpp:216-222)."""
    dim: int = 2
    gamma: float = 1.4
    Rspec: float = 287.058

    @property
    def gamma_minus_1_inv(self) -> float:
        return 1.0 / (self.gamma - 1.0)

    def n_cons(self) -> int:
        return self.dim + 2

    def n_comp(self) -> int:
        return self.dim + 4


@dataclass
class PatchView:
    """PatchDataView<double, R+1, layout_stride> over a torch tensor.

    `data` has shape (ncomp, nz, ny, nx); its element [c, k, j, i] is the
    value of component c at index (lo[0]+i, lo[1]+j, lo[2]+k).
    """
    data: torch.Tensor
    lo: tuple = (0, 0, 0)

    def __post_init__(self):
        if self.data.dim() != 4 or self.data.dtype != torch.float64:
            raise ValueError("PatchView needs a float64 tensor of shape (ncomp, nz, ny, nx)")
        lo = tuple(self.lo) + (0,) * (3 - len(self.lo))
        self.lo = lo[:3]

    @property
    def hi(self) -> tuple:
        nc, nz, ny, nx = self.data.shape
        return (self.lo[0] + nx - 1, self.lo[1] + ny - 1, self.lo[2] + nz - 1)

    def subview(self, lo, hi) -> "PatchView":
        lo = tuple(lo) + (0,) * (3 - len(lo))
        hi = tuple(hi) + (0,) * (3 - len(hi))
        s = [slice(lo[d] - self.lo[d], hi[d] - self.lo[d] + 1) for d in range(3)]
        if any(sl.start < 0 for sl in s) or any(hi[d] > self.hi[d] for d in range(3)):
            raise ValueError("subview box outside the view")
        return PatchView(self.data[:, s[2], s[1], s[0]], lo)

    def components(self, first: int, count: int) -> "PatchView":
        return PatchView(self.data[first:first + count], self.lo)

    def c_view(self) -> L.View:
        t = self.data
        st = t.stride()
        v = L.View()
        v.data = t.data_ptr()
        v.ncomp = t.shape[0]
        v.lo = L.i64x3(self.lo)
        v.hi = L.i64x3(self.hi)
        v.stride = L.i64x3((st[3], st[2], st[1]))
        v.comp_stride = st[0]
        return v


def _stream_ptr(stream) -> int | None:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


class _Handle:
    def __init__(self, desc: L.MethodDesc, stream=None):
        self._lib = L.load()
        self.desc = desc
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = C.c_void_p()
        L.check(self._lib.fvs_create(C.byref(desc), _stream_ptr(self.stream), C.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self._lib.fvs_destroy(h)
            self.h = None


_BASES = {"HLLEM": L.BASE_HLLEM_LARROUTUROU, "HLLEM_Larrouturou": L.BASE_HLLEM_LARROUTUROU,
          "HLLEM_plain": L.BASE_HLLEM, "Godunov": L.BASE_GODUNOV_EXACT,
          "HLL": L.BASE_HLL_EINFELDT, "HLL_Einfeldt": L.BASE_HLL_EINFELDT}
_LIMITERS = {"MinMod": L.LIMITER_MINMOD, "VanLeer": L.LIMITER_VANLEER,
             "NoLimiter": L.LIMITER_NONE, "NoGradient": L.LIMITER_NO_GRADIENT}


def make_desc(equation: PerfectGas, base="HLLEM", reconstruction="MusclHancock",
              limiter="MinMod", precision="fp64", quirks=True,
              max_newton_iterations=5000, newton_tolerance=1e-7,
              newton_guess="mean", boundary_solver="exact") -> L.MethodDesc:
    d = L.MethodDesc()
    d.dim = equation.dim
    d.gamma = equation.gamma
    d.base = _BASES[base] if isinstance(base, str) else int(base)
    if reconstruction in ("MusclHancock", 1):
        d.reconstruction = L.RECON_MUSCL_HANCOCK
    elif reconstruction in ("MusclHancockPrim", 2):
        d.reconstruction = L.RECON_MUSCL_HANCOCK_PRIM
    elif reconstruction in ("MusclHancockCharacteristic", 3):
        d.reconstruction = L.RECON_MUSCL_HANCOCK_CHARACTERISTIC
    else:
        d.reconstruction = L.RECON_NONE
    d.limiter = _LIMITERS[limiter] if isinstance(limiter, str) else int(limiter)
    if precision in ("fp32", "fp32_flux", 1):
        d.precision = L.FP32_FLUX
    elif precision in ("fp64_rcp", 2):
        d.precision = L.FP64_RCP
    elif precision in ("fp64", 0):
        d.precision = L.FP64
    else:
        raise ValueError(f"unknown precision {precision!r} (fp64, fp64_rcp, fp32_flux)")
    d.quirks = 1 if quirks else 0
    d.newton_guess = L.GUESS_LEARNED if newton_guess in ("learned", 1) else L.GUESS_MEAN_PRESSURE
    d.max_newton_iterations = max_newton_iterations
    d.newton_tolerance = newton_tolerance
    d.boundary_solver = (L.BOUNDARY_HLLEM if boundary_solver in ("hllem", "HLLEM", 1)
                         else L.BOUNDARY_EXACT)
    return d


class FluxMethod:
    """FluxMethod<MusclHancock<PerfectGas<D>, Hllem, Limiter>> and friends."""

    def __init__(self, equation: PerfectGas, base="HLLEM", reconstruction="MusclHancock",
                 limiter="MinMod", precision="fp64", quirks=True, stream=None, **kw):
        self.equation = equation
        self._h = _Handle(make_desc(equation, base, reconstruction, limiter, precision,
                                    quirks, **kw), stream)
        self.reconstruction = reconstruction

    @property
    def handle(self):
        return self._h.h

    def GetStencilWidth(self) -> int:
        return 2 if self.reconstruction in ("MusclHancock", 1, "MusclHancockPrim", 2,
                                            "MusclHancockCharacteristic", 3) else 1

    def GetStencilSize(self) -> int:
        return 2 * self.GetStencilWidth()

    def ComputeNumericFluxes(self, fluxes: PatchView, states: PatchView, dt: float,
                             dx: float, dir) -> None:
        fv, sv = fluxes.c_view(), states.c_view()
        L.check(L.load().fvs_numeric_fluxes(self.handle, C.byref(fv), C.byref(sv),
                                            float(dt), float(dx), _dir(dir)))

    def newton_stats(self, reset: bool = False) -> dict:
        """Newton counters of the exact Riemann solver on this handle."""
        arr = (C.c_uint64 * 5)()
        rc = L.load().fvs_newton_stats(self.handle, arr, 1 if reset else 0)
        if rc != L.FVS_NEWTON_NOT_CONVERGED:  # that status still fills the counters
            L.check(rc)
        return {"iterations": arr[0], "max_iterations": arr[1],
                "not_converged": arr[2], "faces": arr[3], "fallbacks": arr[4]}

    def check_newton(self, reset: bool = False) -> None:
        """Raises NewtonNotConverged when a Riemann problem solved on this
        handle hit the Newton iteration cap, where the reference throws
        (include/fub/NewtonIteration.hpp:44-50); the patch operators run
        asynchronously, so their failures surface here."""
        arr = (C.c_uint64 * 5)()
        L.check(L.load().fvs_newton_stats(self.handle, arr, 1 if reset else 0))

    def ComputeStableDt(self, states: PatchView, dx: float, dir,
                        out: torch.Tensor | None = None) -> torch.Tensor:
        """Minimum face stable dt; returns a 1-element device tensor."""
        if out is None:
            out = torch.empty(1, dtype=torch.float64, device=states.data.device)
        sv = states.c_view()
        L.check(L.load().fvs_stable_dt(self.handle, C.byref(sv), float(dx), _dir(dir),
                                       out.data_ptr()))
        return out


class HyperbolicPatchIntegrator:
    """HyperbolicPatchIntegrator<Tag>::UpdateConservatively."""

    def __init__(self, flux_method: FluxMethod):
        self._fm = flux_method

    def UpdateConservatively(self, next: PatchView, prev: PatchView, fluxes: PatchView,
                             dt: float, dx: float, dir) -> None:
        nv, pv, fv = next.c_view(), prev.c_view(), fluxes.c_view()
        L.check(L.load().fvs_update_conservatively(self._fm.handle, C.byref(nv), C.byref(pv),
                                                   C.byref(fv), float(dt), float(dx),
                                                   _dir(dir)))


def CompleteFromCons(flux_method: FluxMethod, complete: PatchView, cons: PatchView) -> None:
    cv, uv = complete.c_view(), cons.c_view()
    L.check(L.load().fvs_complete_from_cons(flux_method.handle, C.byref(cv), C.byref(uv)))


class TransmissiveBoundary:
    """amrex::TransmissiveBoundary for one direction (both sides)."""

    def __init__(self, flux_method: FluxMethod, dir):
        self._fm = flux_method
        self.dir = _dir(dir)

    def FillBoundary(self, states: PatchView, domain_lo, domain_hi) -> None:
        v = states.c_view()
        L.check(L.load().fvs_transmissive_fill(self._fm.handle, C.byref(v), L.i64x3(domain_lo),
                                               L.i64x3(domain_hi), self.dir))


def split_sweep(flux_method: FluxMethod, cons_in: torch.Tensor, cons_out: torch.Tensor,
                dt: float, dx: float, dir, stable_dt_out: torch.Tensor | None = None) -> None:
    """Fused split sweep on a ghost-free patch (transmissive boundaries).

    cons_in / cons_out: float64 tensors of shape (D+2, nz, ny, nx).
    """
    D = flux_method.equation.dim
    _check_cons(cons_in, D)
    _check_cons(cons_out, D)
    n = (cons_in.shape[3], cons_in.shape[2], cons_in.shape[1])
    pin = L.ptr_array([cons_in[c].data_ptr() for c in range(D + 2)])
    pout = L.ptr_array([cons_out[c].data_ptr() for c in range(D + 2)])
    L.check(L.load().fvs_split_sweep(flux_method.handle, pin, pout, L.i64x3(n), float(dt),
                                     float(dx), _dir(dir),
                                     stable_dt_out.data_ptr() if stable_dt_out is not None else None))


def godunov_sweep(flux_method: FluxMethod, cons_in: torch.Tensor,
                  cons_out: torch.Tensor | None, dt: float, dx: float, dir,
                  guess: torch.Tensor | None = None,
                  stable_dt_out: torch.Tensor | None = None) -> None:
    """First-order exact-Godunov fused sweep; `guess` (float64, direction
    dir's face box, x fastest) seeds the Newton star-pressure iteration."""
    D = flux_method.equation.dim
    _check_cons(cons_in, D)
    n = (cons_in.shape[3], cons_in.shape[2], cons_in.shape[1])
    pin = L.ptr_array([cons_in[c].data_ptr() for c in range(D + 2)])
    pout = None
    if cons_out is not None:
        _check_cons(cons_out, D)
        pout = L.ptr_array([cons_out[c].data_ptr() for c in range(D + 2)])
    if guess is not None and (guess.dtype != torch.float64 or not guess.is_contiguous()):
        raise ValueError("guess must be a contiguous float64 tensor")
    L.check(L.load().fvs_godunov_sweep(
        flux_method.handle, pin, pout, L.i64x3(n), float(dt), float(dx), _dir(dir),
        guess.data_ptr() if guess is not None else None,
        stable_dt_out.data_ptr() if stable_dt_out is not None else None))


def patch_stable_dt(flux_method: FluxMethod, cons: torch.Tensor, dx) -> torch.Tensor:
    """Per-direction stable dt of a ghost-free transmissive patch (device tensor)."""
    D = flux_method.equation.dim
    _check_cons(cons, D)
    out = torch.empty(3, dtype=torch.float64, device=cons.device)
    n = (cons.shape[3], cons.shape[2], cons.shape[1])
    p = L.ptr_array([cons[c].data_ptr() for c in range(D + 2)])
    dxa = (C.c_double * 3)(*(list(dx) + [1.0] * (3 - len(dx))))
    L.check(L.load().fvs_patch_stable_dt(flux_method.handle, p, L.i64x3(n), dxa, out.data_ptr()))
    return out[:D]


def _check_cons(t: torch.Tensor, D: int) -> None:
    if t.dtype != torch.float64 or t.dim() != 4 or t.shape[0] != D + 2:
        raise ValueError(f"expected a float64 tensor of shape ({D + 2}, nz, ny, nx)")
    if not t.is_contiguous():
        raise ValueError("fused sweeps need contiguous component arrays")
    if D == 2 and t.shape[1] != 1:
        raise ValueError("2D patches have nz == 1")


class SinglePatchLevel:
    """Single-patch level integrator: DimensionalSplitLevelIntegrator with a
    NoSubcycleSolver and RunSimulation's dt logic, driven on the device."""

    def __init__(self, flux_method: FluxMethod, n, dx, splitting="Godunov", cfl=0.8,
                 smallest_time_step_size: float = 1e-12):
        self.fm = flux_method
        D = flux_method.equation.dim
        self.D = D
        pd = L.PatchDesc()
        pd.n = L.i64x3(list(n) + [1] * (3 - len(n)))
        pd.dx = (C.c_double * 3)(*(list(dx) + [1.0] * (3 - len(dx))))
        pd.splitting = L.SPLIT_STRANG if splitting == "Strang" else L.SPLIT_GODUNOV
        pd.cfl = cfl
        pd.smallest_dt = smallest_time_step_size
        self.eps = smallest_time_step_size
        self.n = tuple(n)
        h = C.c_void_p()
        L.check(L.load().fvs_level_create(flux_method.handle, C.byref(pd), C.byref(h)))
        self.h = h
        self._lib = L.load()

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self._lib.fvs_level_destroy(h)
            self.h = None

    def shape(self):
        nx, ny = self.n[0], self.n[1]
        nz = self.n[2] if self.D == 3 else 1
        return (self.D + 2, nz, ny, nx)

    def set_state(self, cons: torch.Tensor, t0: float = 0.0) -> None:
        _check_cons(cons, self.D)
        p = L.ptr_array([cons[c].data_ptr() for c in range(self.D + 2)])
        L.check(self._lib.fvs_level_set_state(self.h, p, float(t0)))

    def get_state(self, out: torch.Tensor | None = None) -> torch.Tensor:
        if out is None:
            out = torch.empty(self.shape(), dtype=torch.float64, device="cuda")
        p = L.ptr_array([out[c].data_ptr() for c in range(self.D + 2)])
        L.check(self._lib.fvs_level_get_state(self.h, p))
        return out

    def advance(self, max_steps: int, final_time: float = float("inf"), batch: int = 16,
                return_dts: bool = False):
        dts = (C.c_double * max(max_steps, 1))()
        done = C.c_int()
        fails = C.c_int()
        L.check(self._lib.fvs_level_advance(self.h, int(max_steps), float(final_time), int(batch),
                                            dts if return_dts else None, C.byref(done),
                                            C.byref(fails)))
        if return_dts:
            return done.value, fails.value, [dts[i] for i in range(done.value)]
        return done.value, fails.value

    def run(self, final_time: float, output_times=(), output=None, max_cycles: int = -1,
            batch: int = 16) -> list:
        """RunSimulation (include/fub/RunSimulation.hpp:86-191): steps until
        t + smallest_time_step_size reaches final_time (or max_cycles steps
        are taken), each step's dt capped by the final time and by the next
        output time (limited_dt, :125-131), so the run lands exactly on every
        output time in output_times; output(level, t, cycle) is called there
        (the reference's output functor).  CFL failures restart the step as
        advance() does.  Returns the accepted dts."""
        dts = []
        t, cycle = self.time()
        outs = sorted(float(x) for x in output_times if x > t + self.eps)
        while t + self.eps < final_time and (max_cycles < 0 or cycle < max_cycles):
            stop = min(outs[0], final_time) if outs else final_time
            # the device driver ends a call when t + smallest dt reaches `stop`
            while True:
                k = 256 if max_cycles < 0 else min(256, max_cycles - cycle)
                done, _, got = self.advance(k, final_time=stop, batch=batch, return_dts=True)
                dts += got
                t, cycle = self.time()
                if done < k or k == 0 or (max_cycles >= 0 and cycle >= max_cycles):
                    break
            while outs and t + self.eps >= outs[0]:
                if output is not None:
                    output(self, t, cycle)
                outs.pop(0)
            if done == 0 and t + self.eps < stop:
                break
        return dts

    def time(self):
        t = C.c_double()
        cyc = C.c_int64()
        L.check(self._lib.fvs_level_time(self.h, C.byref(t), C.byref(cyc)))
        return t.value, cyc.value

    def hint_max_speed(self, speed) -> None:
        """Speed floor of the next step-start CFL pass of a 3D HLLEM or an
        exact-Godunov level (per direction; by default the previous pass's
        largest speed less a margin that follows how fast that speed falls).
        A floor above
        the true largest signal speed costs one repeated step and never
        changes a result."""
        arr = (C.c_double * 3)(*(list(speed) + [0.0] * (3 - len(speed))))
        L.check(self._lib.fvs_level_hint_max_speed(self.h, arr))

    def cfl_redos(self) -> int:
        """Steps repeated because the CFL pass's speed floor was not reached."""
        n = C.c_int64()
        L.check(self._lib.fvs_level_cfl_redos(self.h, C.byref(n)))
        return n.value

    def set_timing(self, on: bool = True) -> None:
        """Brackets every kernel launch of advance() with HIP events."""
        L.check(self._lib.fvs_level_set_timing(self.h, 1 if on else 0))

    def set_graphs(self, on: bool = True) -> None:
        """Replay captured HIP graphs of the time steps (default) or launch
        every kernel eagerly."""
        L.check(self._lib.fvs_level_set_graphs(self.h, 1 if on else 0))

    def kernel_times(self) -> dict:
        """{class: (summed ms, launches)} since the last call (see set_timing)."""
        ms = (C.c_double * L.LEVEL_TIMERS)()
        cnt = (C.c_int64 * L.LEVEL_TIMERS)()
        L.check(self._lib.fvs_level_kernel_times(self.h, ms, cnt))
        return {name: (ms[k], cnt[k]) for k, name in enumerate(L.TIMER_NAMES)}
