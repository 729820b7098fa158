"""Learned initial guess for the exact Riemann solver's Newton iteration.

The reference seeds Newton with the mean pressure 0.5 (pL + pR)
(include/fub/equations/perfect_gas/ExactRiemannSolver.hpp:91,
src/equations/perfect_gas/ExactRiemannSolver.cpp); the iteration itself
(include/fub/NewtonIteration.hpp:28-52) is unchanged here.  A 6 -> 64 -> 64 -> 1
ReLU network predicts r = ln p* - (ln pL + ln pR) / 2 from
(ln rhoL / ln 1000, uL / 5, ln pL / ln 1000, ln rhoR / ln 1000, uR / 5,
ln pR / ln 1000); the guess is sqrt(pL pR) exp(r).  Data generation (labels
from the FP64 Newton solve), FP32 training (forward, backward, Adam) and BF16
matrix-core inference all run in libfvs_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .patch import FluxMethod, _stream_ptr

N_IN = 6
N_HIDDEN = 64
# FP32 parameter layout: W1 [64][8] (inputs padded to 8), b1 [64], W2 [64][64],
# b2 [64], W3 [64], b3.
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HIDDEN * 8
OFF_W2 = OFF_B1 + N_HIDDEN
OFF_B2 = OFF_W2 + N_HIDDEN * N_HIDDEN
OFF_W3 = OFF_B2 + N_HIDDEN
OFF_B3 = OFF_W3 + N_HIDDEN
NUM_PARAMS = OFF_B3 + 1


def unpack_params(p: np.ndarray) -> dict:
    """Splits the flat parameter vector into W1 (64x6), b1, W2, b2, W3, b3."""
    p = np.asarray(p)
    return {
        "W1": p[OFF_W1:OFF_B1].reshape(N_HIDDEN, 8)[:, :N_IN],
        "b1": p[OFF_B1:OFF_W2],
        "W2": p[OFF_W2:OFF_B2].reshape(N_HIDDEN, N_HIDDEN),
        "b2": p[OFF_B2:OFF_W3],
        "W3": p[OFF_W3:OFF_B3],
        "b3": p[OFF_B3],
    }


class LearnedNewtonGuess:
    """Owner of one network on the device (fvs_mlp_* of include/fvs.h)."""

    def __init__(self, seed: int = 0, stream=None):
        self._lib = L.load()
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        m = C.c_void_p()
        L.check(self._lib.fvs_mlp_create(_stream_ptr(self.stream), C.c_uint64(seed), C.byref(m)))
        self.m = m
        assert self._lib.fvs_mlp_num_params() == NUM_PARAMS

    def __del__(self):
        m = getattr(self, "m", None)
        if m is not None and m.value:
            self._lib.fvs_mlp_destroy(m)
            self.m = None

    # parameters -------------------------------------------------------------
    def get_params(self) -> np.ndarray:
        p = np.zeros(NUM_PARAMS, dtype=np.float32)
        L.check(self._lib.fvs_mlp_get_params(self.m, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != NUM_PARAMS:
            raise ValueError(f"expected {NUM_PARAMS} parameters")
        L.check(self._lib.fvs_mlp_set_params(self.m, p.ctypes.data))

    # data ---------------------------------------------------------------------
    def generate(self, n: int, seed: int = 1, gamma: float = 1.4) -> dict:
        counts = (C.c_int64 * 3)()
        L.check(self._lib.fvs_mlp_generate(self.m, n, C.c_uint64(seed), gamma, counts))
        return {"usable": counts[0], "vacuum": counts[1], "not_converged": counts[2]}

    def dataset(self) -> dict:
        n = C.c_int64()
        L.check(self._lib.fvs_mlp_dataset(self.m, C.byref(n), None, None, None, None, None))
        n = n.value
        dev = torch.device("cuda", torch.cuda.current_device())
        out = {
            "prim": torch.empty((6, n), dtype=torch.float64, device=dev),
            "features": torch.empty((N_IN, n), dtype=torch.float32, device=dev),
            "target": torch.empty(n, dtype=torch.float32, device=dev),
            "flag": torch.empty(n, dtype=torch.int32, device=dev),
            "iterations": torch.empty(n, dtype=torch.int32, device=dev),
        }
        torch.cuda.synchronize()
        L.check(self._lib.fvs_mlp_dataset(self.m, None, out["prim"].data_ptr(),
                                          out["features"].data_ptr(), out["target"].data_ptr(),
                                          out["flag"].data_ptr(), out["iterations"].data_ptr()))
        return out

    # FP32 path ----------------------------------------------------------------
    def forward(self, features: torch.Tensor) -> torch.Tensor:
        x = features.contiguous()
        if x.dtype != torch.float32 or x.dim() != 2 or x.shape[0] != N_IN:
            raise ValueError("features must be float32 of shape (6, n)")
        y = torch.empty(x.shape[1], dtype=torch.float32, device=x.device)
        L.check(self._lib.fvs_mlp_forward(self.m, x.data_ptr(), x.shape[1], y.data_ptr()))
        return y

    def gradients(self, features: torch.Tensor, target: torch.Tensor,
                  flag: torch.Tensor, huber: float = 0.0) -> tuple[np.ndarray, float]:
        """Loss and gradient on one batch (squared error, or Huber when
        huber > 0), from the FP32 forward and backward kernels."""
        x = features.contiguous()
        g = np.zeros(NUM_PARAMS, dtype=np.float32)
        loss = C.c_float()
        torch.cuda.synchronize()
        L.check(self._lib.fvs_mlp_gradients(self.m, x.data_ptr(), target.contiguous().data_ptr(),
                                            flag.contiguous().data_ptr(), x.shape[1],
                                            C.c_float(huber), g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def train(self, epochs: int, batch: int = 1024, lr: float = 5e-3, huber: float = 0.25,
              seed: int = 7) -> np.ndarray:
        """Adam with a cosine-decayed rate on the generated set."""
        hist = np.zeros(max(epochs, 1), dtype=np.float32)
        L.check(self._lib.fvs_mlp_train(self.m, epochs, batch, C.c_float(lr), C.c_float(huber),
                                        C.c_uint64(seed), hist.ctypes.data))
        return hist[:epochs]

    # matrix-core inference ----------------------------------------------------
    def infer_pairs(self, dim: int, dir: int, left: torch.Tensor, right: torch.Tensor):
        """left/right: complete states, float64 (dim + 4, n).  Returns (r, guess)."""
        left, right = left.contiguous(), right.contiguous()
        n = left.shape[1]
        r = torch.empty(n, dtype=torch.float32, device=left.device)
        guess = torch.empty(n, dtype=torch.float64, device=left.device)
        L.check(self._lib.fvs_mlp_infer_pairs(self.m, dim, dir, left.data_ptr(), right.data_ptr(),
                                              n, r.data_ptr(), guess.data_ptr()))
        return r, guess

    def face_guess(self, dim: int, gamma: float, cons: torch.Tensor, dir: int) -> torch.Tensor:
        """Guesses of every face of direction dir of a ghost-free patch
        (cons: (dim + 2, nz, ny, nx)), in the face-box layout."""
        nz, ny, nx = cons.shape[1:]
        n = [nx, ny, nz]
        shape = [nz + (dir == 2), ny + (dir == 1), nx + (dir == 0)]
        out = torch.empty(shape, dtype=torch.float64, device=cons.device)
        ptrs = L.ptr_array([cons[c].data_ptr() for c in range(dim + 2)])
        L.check(self._lib.fvs_mlp_face_guess(self.m, dim, gamma, ptrs, L.i64x3(n), dir,
                                             out.data_ptr()))
        return out

    def attach(self, flux_method: FluxMethod) -> None:
        """Seeds every Newton iteration of a first-order exact-Godunov flux
        method's level driver with this network (`detach` removes it)."""
        L.check(self._lib.fvs_set_mlp(flux_method.handle, self.m))
        flux_method._mlp = self  # keeps the network alive while attached


def detach(flux_method: FluxMethod) -> None:
    L.check(L.load().fvs_set_mlp(flux_method.handle, None))
    flux_method._mlp = None


def exact_middle_state(flux_method: FluxMethod, dir: int, left: torch.Tensor,
                       right: torch.Tensor, guess: torch.Tensor | None = None):
    """ExactRiemannSolver::ComputeMiddleState on complete state pairs
    (float64 (dim + 4, n) device tensors), Newton seeded with `guess` or the
    mean pressure.  Returns (pM, uM, iterations, converged) device tensors."""
    left, right = left.contiguous(), right.contiguous()
    n = left.shape[1]
    dev = left.device
    pM = torch.empty(n, dtype=torch.float64, device=dev)
    uM = torch.empty(n, dtype=torch.float64, device=dev)
    it = torch.empty(n, dtype=torch.int32, device=dev)
    conv = torch.empty(n, dtype=torch.int32, device=dev)
    g = guess.contiguous().data_ptr() if guess is not None else None
    L.check(L.load().fvs_exact_middle_state(flux_method.handle, int(dir), left.data_ptr(),
                                            right.data_ptr(), n, g, pM.data_ptr(), uM.data_ptr(),
                                            it.data_ptr(), conv.data_ptr()))
    return pM, uM, it, conv
