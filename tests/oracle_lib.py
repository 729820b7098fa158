"""ctypes access to the CPU oracle (oracle/libfvs_oracle.so) for tests.

Test infrastructure only: the oracle is the checker, never the product.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from finitevolumesolver_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
ORACLE_PATH = ROOT / "oracle" / "libfvs_oracle.so"

_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_oracle = None


def load():
    global _oracle
    if _oracle is None:
        if not ORACLE_PATH.exists():
            raise FileNotFoundError(f"{ORACLE_PATH} missing; run make -C oracle")
        lib = C.CDLL(str(ORACLE_PATH))
        I, D, I64, P = C.c_int, C.c_double, C.c_int64, C.c_void_p
        sig = {
            "oracle_set_threads": (I, [I]),
            "oracle_plant": (I, [I]),
            "oracle_complete_from_cons": (I, [I, D, I64, P, P]),
            "oracle_complete_from_prim": (I, [I, D, I64, P, P, P, P]),
            "oracle_physical_flux": (I, [I, I, I64, P, P]),
            "oracle_limited_slope": (I, [I, I64, P, P, P, P]),
            "oracle_hllem_flux": (I, [I, D, I, I, I, I64, P, P, P, P]),
            "oracle_hllem_stable_dt": (I, [I, D, I, D, I64, P, P, P]),
            "oracle_face_flux": (I, [C.POINTER(L.MethodDesc), I, D, D, I64, P, P]),
            "oracle_exact_middle_state": (I, [I, D, I, I64, P, P, P, D, I, P, P, P, P]),
            "oracle_exact_riemann_state": (I, [I, D, I, I64, P, P, P]),
            "oracle_exact_signals": (I, [I, D, I, I64, P, P, P, P]),
            "oracle_stable_dt": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.View), D, I, _PD]),
            "oracle_numeric_fluxes": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.View),
                                          C.POINTER(L.View), D, D, I]),
            "oracle_update_conservatively": (I, [C.POINTER(L.View), C.POINTER(L.View),
                                                 C.POINTER(L.View), D, D, I]),
            "oracle_complete_view": (I, [I, D, C.POINTER(L.View), C.POINTER(L.View)]),
            "oracle_transmissive_fill": (I, [C.POINTER(L.View), C.POINTER(C.c_int64),
                                             C.POINTER(C.c_int64), I]),
            "oracle_level_run": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.PatchDesc), P, D, I,
                                     D, _PD, _PI, _PI]),
            "oracle_level_run_from": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.PatchDesc), P, D,
                                          C.c_int64, I, D, _PD, _PI, _PI]),
            "oracle_kbn_fill_cutcell_data": (I, [P, I, P, P]),
            "oracle_kbn_is_cut_cell": (I, [I, P, P, P, P]),
            "oracle_kbn_reference_states": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.View),
                                                C.POINTER(L.View), P]),
            "oracle_kbn_boundary_fluxes": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.View),
                                               C.POINTER(L.View), C.POINTER(L.View), P, D, D, I]),
            "oracle_kbn_regular_fluxes": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.View),
                                              C.POINTER(L.View), P, D, D, I]),
            "oracle_kbn_stable_dt": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.View), P, D, I,
                                         _PD]),
            "oracle_kbn_cutcell_fluxes": (I, [I, C.POINTER(L.View), C.POINTER(L.View),
                                              C.POINTER(L.View), C.POINTER(L.View),
                                              C.POINTER(L.View), P, I]),
            "oracle_kbn_update": (I, [C.POINTER(L.View)] * 6 + [P, D, D, I]),
            "oracle_kbn_update2": (I, [C.POINTER(L.View)] * 4 + [P, D, D, I]),
            "oracle_kbn_complete_from_cons": (I, [I, D, C.POINTER(L.View), C.POINTER(L.View), P]),
            "oracle_kbn_rotate": (I, [I, P, P, P, P]),
            "oracle_hllem_riemann_state": (I, [I, D, I, I, I64, P, P, P]),
            "oracle_kbn_level_run": (I, [C.POINTER(L.MethodDesc), C.POINTER(L.PatchDesc), P, I, P,
                                         D, I, D, _PD, _PI, _PI]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _oracle = lib
    return _oracle


class planted:
    """Context manager: runs the oracle with one planted mistake (the codes
    of enum Plant in oracle/fvs_oracle.hpp), restoring the clean oracle on
    exit, so a pinning test can show that its closed form notices it."""

    MUSCL_HALF_STEP_SIGN = 1
    MUSCL_HALF_STEP_DROP = 2
    HLLEM_NO_ANTIDIFFUSION = 3
    HLLEM_L21 = 4
    HLLEM_L2U_SIGN = 5
    HLLEM_DELTA = 6
    EXACT_FAN_LEFT_U = 7
    EXACT_FAN_RIGHT_RHO = 8
    HBOX_RIGHT_DISTANCE = 9
    HBOX_RIGHT_DB_SIGN = 10

    def __init__(self, which: int):
        self.which = which

    def __enter__(self):
        load().oracle_plant(self.which)
        return self

    def __exit__(self, *exc):
        load().oracle_plant(0)
        return False


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def host_view(arr: np.ndarray, lo=(0, 0, 0)) -> L.View:
    """fvs_view over a numpy array of shape (ncomp, nz, ny, nx)."""
    assert arr.dtype == np.float64 and arr.ndim == 4
    st = [s // 8 for s in arr.strides]
    v = L.View()
    v.data = arr.ctypes.data
    v.ncomp = arr.shape[0]
    lo = tuple(lo) + (0,) * (3 - len(lo))
    v.lo = L.i64x3(lo)
    v.hi = L.i64x3((lo[0] + arr.shape[3] - 1, lo[1] + arr.shape[2] - 1, lo[2] + arr.shape[1] - 1))
    v.stride = L.i64x3((st[3], st[2], st[1]))
    v.comp_stride = st[0]
    return v


def complete_from_cons(cons: np.ndarray, dim: int, gamma: float = 1.4) -> np.ndarray:
    """cons: (D+2, n) -> complete (D+4, n), array-path arithmetic."""
    cons = np.ascontiguousarray(cons, dtype=np.float64)
    n = cons.shape[1]
    out = np.empty((dim + 4, n))
    load().oracle_complete_from_cons(dim, gamma, n, ptr(cons), ptr(out))
    return out


def hllem_flux(left, right, dim, dir, larrouturou=True, quirk=True, gamma=1.4):
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    n = left.shape[1]
    f = np.empty((dim + 2, n))
    ps = np.empty(n)
    load().oracle_hllem_flux(dim, gamma, dir, int(larrouturou), int(quirk), n, ptr(left),
                             ptr(right), ptr(f), ptr(ps))
    return f, ps


def hllem_stable_dt(left, right, dim, dir, dx, gamma=1.4):
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    n = left.shape[1]
    dt = np.empty(n)
    load().oracle_hllem_stable_dt(dim, gamma, dir, dx, n, ptr(left), ptr(right), ptr(dt))
    return dt


def physical_flux(complete, dim, dir):
    complete = np.ascontiguousarray(complete)
    n = complete.shape[1]
    f = np.empty((dim + 2, n))
    load().oracle_physical_flux(dim, dir, n, ptr(complete), ptr(f))
    return f


def face_flux(desc: L.MethodDesc, stencil: np.ndarray, dir, dt, dx):
    """stencil: (S, D+4, n)."""
    stencil = np.ascontiguousarray(stencil)
    n = stencil.shape[2]
    f = np.empty((desc.dim + 2, n))
    rc = load().oracle_face_flux(C.byref(desc), dir, dt, dx, n, ptr(stencil), ptr(f))
    return f, rc


def exact_middle_state(left, right, dim, dir, x0=None, tol=1e-7, max_it=5000, gamma=1.4):
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    n = left.shape[1]
    pM, uM = np.empty(n), np.empty(n)
    it = np.empty(n, dtype=np.int32)
    conv = np.empty(n, dtype=np.int32)
    x0p = ptr(np.ascontiguousarray(x0, dtype=np.float64)) if x0 is not None else None
    load().oracle_exact_middle_state(dim, gamma, dir, n, ptr(left), ptr(right), x0p, tol, max_it,
                                     ptr(pM), ptr(uM), ptr(it), ptr(conv))
    return pM, uM, it, conv


def exact_riemann_state(left, right, dim, dir, gamma=1.4):
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    n = left.shape[1]
    s = np.empty((dim + 4, n))
    load().oracle_exact_riemann_state(dim, gamma, dir, n, ptr(left), ptr(right), ptr(s))
    return s


def exact_signals(left, right, dim, dir, gamma=1.4):
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    n = left.shape[1]
    sL, sR = np.empty(n), np.empty(n)
    load().oracle_exact_signals(dim, gamma, dir, n, ptr(left), ptr(right), ptr(sL), ptr(sR))
    return sL, sR


def level_run(desc: L.MethodDesc, cons: np.ndarray, dx, n_steps, splitting=L.SPLIT_GODUNOV,
              cfl=0.8, final_time=1e300, t0=0.0, cycle0=0):
    """Runs the oracle's single-patch level driver from time t0 and cycle
    count cycle0 (which sets the first step's split order); returns (cons,
    dts, failures)."""
    D = desc.dim
    cons = np.ascontiguousarray(cons.copy())
    pd = L.PatchDesc()
    shape = cons.shape  # (D+2, nz, ny, nx)
    pd.n = L.i64x3((shape[3], shape[2], shape[1] if D == 3 else 1))
    pd.dx = (C.c_double * 3)(*(list(dx) + [1.0] * (3 - len(dx))))
    pd.splitting = splitting
    pd.cfl = cfl
    dts = np.zeros(max(n_steps, 1))
    done = C.c_int()
    fails = C.c_int()
    rc = load().oracle_level_run_from(C.byref(desc), C.byref(pd), ptr(cons), t0, cycle0, n_steps,
                                      final_time, dts.ctypes.data_as(_PD), C.byref(done),
                                      C.byref(fails))
    assert rc == 0, rc
    return cons, dts[:done.value], fails.value


class HostCutCellData:
    """Host (numpy) CutCellData<2> for the oracle, mirroring
    finitevolumesolver_amd.cutcell.CutCellData: arrays (ncomp, 1, ny, nx)."""

    def __init__(self, geom):
        self.lo = (geom.lo[0], geom.lo[1], 0)
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.alpha = c(geom.alpha[None, None])
        self.betas = [c(geom.beta_x[None, None]), c(geom.beta_y[None, None])]
        self.normals = c(geom.normals[:, None])
        self.centroids = c(geom.centroids[:, None])
        self.fractions = {k: [np.zeros((2,) + b.shape[1:]) for b in self.betas]
                          for k in ("unshielded", "shielded_left", "shielded_right",
                                    "doubly_shielded")}

    def c_struct(self) -> L.CutCellData:
        s = L.CutCellData()
        s.volume_fractions = host_view(self.alpha, self.lo)
        for d in range(2):
            s.face_fractions[d] = host_view(self.betas[d], self.lo)
            for k in self.fractions:
                getattr(s, k)[d] = host_view(self.fractions[k][d], self.lo)
        s.boundary_normals = host_view(self.normals, self.lo)
        s.boundary_centroids = host_view(self.centroids, self.lo)
        self._s = s
        return s

    def fill(self):
        """FillCutCellData on every face with both neighbours present."""
        s = self.c_struct()
        for d in range(2):
            b = self.betas[d]
            lo = [self.lo[0], self.lo[1], 0]
            hi = [self.lo[0] + b.shape[3] - 1, self.lo[1] + b.shape[2] - 1, 0]
            lo[d] += 1
            hi[d] -= 1
            assert load().oracle_kbn_fill_cutcell_data(C.byref(s), d, ptr(np.array(lo, np.int64)),
                                                       ptr(np.array(hi, np.int64))) == 0
        return self


def hllem_riemann_state(left, right, dim, dir, larrouturou=True, gamma=1.4):
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    n = left.shape[1]
    out = np.empty((dim + 4, n))
    load().oracle_hllem_riemann_state(dim, gamma, dir, int(larrouturou), n, ptr(left),
                                      ptr(right), ptr(out))
    return out


def kbn_rotate(frm, to, mom):
    dim = len(frm)
    out = np.empty(dim)
    load().oracle_kbn_rotate(dim, ptr(np.ascontiguousarray(frm, dtype=np.float64)),
                             ptr(np.ascontiguousarray(to, dtype=np.float64)),
                             ptr(np.ascontiguousarray(mom, dtype=np.float64)), ptr(out))
    return out


def kbn_level_run(desc: L.MethodDesc, geom, cons: np.ndarray, dx, n_steps, integrator=0,
                  splitting=L.SPLIT_GODUNOV, cfl=0.8, final_time=1e300):
    """The oracle's single-patch cut-cell level; geom: Geometry2D over the
    patch grown by two ghost layers.  Returns (cons, dts, failures)."""
    h = HostCutCellData(geom)
    s = h.c_struct()
    cons = np.ascontiguousarray(cons.copy())
    ny, nx = cons.shape[-2:]
    pd = L.PatchDesc()
    pd.n = L.i64x3((nx, ny, 1))
    pd.dx = (C.c_double * 3)(dx[0], dx[1], 1.0)
    pd.splitting = splitting
    pd.cfl = cfl
    dts = np.zeros(max(n_steps, 1))
    done, fails = C.c_int(), C.c_int()
    rc = load().oracle_kbn_level_run(C.byref(desc), C.byref(pd), C.byref(s), integrator, ptr(cons),
                                     0.0, n_steps, final_time, dts.ctypes.data_as(_PD),
                                     C.byref(done), C.byref(fails))
    assert rc == 0, rc
    return cons, dts[:done.value], fails.value
