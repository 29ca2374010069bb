"""Point-wise spectral operations (src/periodicflow/math.py:4-100).

Same names and argument order as the reference's Numba kernels, acting in
place on device tensors.  The solver's time step never calls these: its
curl, cross product, projection and linear operator are fused into the HIP
passes of libpfhip.so.  They are kept for scripts that build or inspect
fields with the reference's vocabulary.  Arguments may be torch tensors on
the GPU (k, k_k2 may be NumPy arrays; they are moved to the device).
"""
from __future__ import annotations

import numpy as np
import torch


def _dev(a, like):
    if isinstance(a, torch.Tensor):
        return a.to(device=like.device)
    return torch.as_tensor(np.asarray(a), device=like.device)


def cross(c, a, b):
    """c = a x b (math.py:5-28); 2D: c = (a1 b, -a0 b) with scalar b."""
    if a.shape[0] == 2:
        c[0] = a[1] * b
        c[1] = -a[0] * b
        return
    c[0] = a[1] * b[2] - a[2] * b[1]
    c[1] = a[2] * b[0] - a[0] * b[2]
    c[2] = a[0] * b[1] - a[1] * b[0]


def curl(c, k, a):
    """c = i k x a (math.py:32-49); scalar in 2D."""
    k = _dev(k, a).to(a.real.dtype)
    if a.shape[0] == 2:
        c[...] = 1j * (k[0] * a[1] - k[1] * a[0])
        return
    c[0] = 1j * (k[1] * a[2] - k[2] * a[1])
    c[1] = 1j * (k[2] * a[0] - k[0] * a[2])
    c[2] = 1j * (k[0] * a[1] - k[1] * a[0])


def apply_linear_operator(Lu, L, u):
    """Lu = L u with L broadcast over the component axis (math.py:53-68)."""
    Lu[...] = _dev(L, u).to(u.dtype) * u


def leray_projection(u, k, k_k2, p_hat):
    """p = (k / k^2) . u; u -= k p, in place (math.py:72-100)."""
    k = _dev(k, u).to(u.real.dtype)
    kk = _dev(k_k2, u).to(u.real.dtype)
    p = (kk * u).sum(dim=0)
    p_hat[...] = p
    u -= k * p
