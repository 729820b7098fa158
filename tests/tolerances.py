"""Comparison bars shared by the GPU parity tests."""
from __future__ import annotations

import numpy as np

FP64_RTOL = 1e-12


def tolerance_ratio(a, b, rtol=FP64_RTOL, atol_scale=1e-13):
    """max over cells of |a - b| / (rtol |b| + atol_scale s_c), s_c the
    largest magnitude of component c (axis 0): relative 1e-12 per cell, with
    an absolute floor for momenta passing through zero (there the error of
    an exact-Riemann flux follows the sound speed through the star velocity,
    not the momentum).  A ratio <= 1 passes."""
    a = np.asarray(a)
    b = np.asarray(b)
    s_c = np.max(np.abs(b).reshape(b.shape[0], -1), axis=1)
    den = rtol * np.abs(b) + atol_scale * s_c.reshape((-1,) + (1,) * (b.ndim - 1))
    return float(np.max(np.abs(a - b) / den))
