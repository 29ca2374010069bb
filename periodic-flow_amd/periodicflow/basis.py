"""Fourier space set-up (src/periodicflow/basis.py:10-92).

Host-side geometry of the box: domain parsing (:50-60), broadcast scaled
wavenumbers ``k`` (:65, :75), ``k2`` (:76), ``k_over_k2`` (:77), the mesh
``x`` (:74), the dealias table (:79-87) and the Nyquist mask (:72).  These
are small NumPy arrays kept for user scripts (initial conditions,
correlators); the kernels regenerate wavenumbers from indices and never read
them.  ``fft_plan`` and ``decomposition`` are accepted for compatibility:
one GPU holds the whole box, and transform plans are fixed per length.
"""
from __future__ import annotations

import numpy as np

DEALIAS_OPTIONS = {
    "3/2": {"padding_factor": 1.5, "dealias_direct": False},
    "pad": {"padding_factor": 1.5, "dealias_direct": False},
    "2/3": {"padding_factor": 1.0, "dealias_direct": True},
    "truncate": {"padding_factor": 1.0, "dealias_direct": True},
}


def parse_domain(domain, dim):
    """basis.py:50-60: a float is (0, L) on every axis; a list of floats is one
    interval for all axes; a list of intervals gives one per axis."""
    if isinstance(domain, (float, np.floating)):
        return [(0.0, float(domain)) for _ in range(dim)]
    if isinstance(domain, (list, tuple)):
        if all(isinstance(i, (list, tuple)) for i in domain):
            return [tuple(float(v) for v in i) for i in domain]
        if all(isinstance(i, (float, np.floating)) for i in domain):
            return [tuple(float(v) for v in domain) for _ in range(dim)]
        raise ValueError("Invalid domain.")
    raise ValueError("Input domain must be either float, list or tuple.")


class FourierSpace:
    """Geometry of a periodic box with the last axis real-to-complex."""

    def __init__(self, comm=None, mpi_rank=0, N=None, domain=2 * np.pi, dealias="3/2", mask_nyquist=False,
                 fft_plan="FFTW_MEASURE", decomposition="slab"):
        self.comm = comm
        self.mpi_rank = mpi_rank
        self.N = list(N)
        self.dim = len(self.N)
        self.domain = parse_domain(domain, self.dim)
        opt = DEALIAS_OPTIONS.get(dealias)
        if opt is None:
            raise ValueError(f"FourierSpace: Invalid dealiasing rule: {dealias}.")
        self.dealias = dealias
        self.padding_factor = opt["padding_factor"]
        self.dealias_direct = opt["dealias_direct"]
        self.fft_plan = fft_plan
        self.decomposition = decomposition
        self.lengths = [b - a for a, b in self.domain]
        d = self.dim
        axes = []
        for i, n in enumerate(self.N):
            s = 2 * np.pi / self.lengths[i]
            ki = (np.fft.fftfreq(n, 1.0 / n) if i < d - 1 else np.fft.rfftfreq(n, 1.0 / n)) * s
            shp = [1] * d
            shp[i] = ki.size
            axes.append(ki.reshape(shp))
        self.k = np.array(np.broadcast_arrays(*axes))
        self.k2 = sum(ki**2 for ki in self.k)
        self.k_over_k2 = np.array([ki / np.where(self.k2 == 0, 1, self.k2) for ki in self.k])
        mesh = []
        for i, n in enumerate(self.N):
            a = self.domain[i][0]
            shp = [1] * d
            shp[i] = n
            mesh.append((a + np.arange(n) * self.lengths[i] / n).reshape(shp))
        self.x = mesh
        if mask_nyquist:
            m = np.ones(self.k2.shape)
            for i, n in enumerate(self.N):
                if n % 2 == 0:
                    idx = [slice(None)] * d
                    idx[i] = n // 2
                    m[tuple(idx)] = 0
            self.nyquist_mask = m
        else:
            self.nyquist_mask = None
        self.shape_physical_scalar = tuple(self.N)
        self.shape_fourier_scalar = tuple(self.N[:-1]) + (self.N[-1] // 2 + 1,)
