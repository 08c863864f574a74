"""Host-side grid and shell geometry (reference src/Utils.jl, src/ShellBinning.jl).

These run once per plan on the host; the resulting integer shell map is uploaded to the device.
"""
import math

import numpy as np

from .types import (CustomBinning, DyadicBinning, LinearBinning, LogarithmicBinning,
                    ShellMagnitude)


def wavenumber_grid(ns, Ls, dtype=np.float64):
    """Utils.jl:51-63: FFTW index order, k_idx <= n÷2 -> k_idx, else k_idx - n, times 2π/L."""
    out = []
    for n, L in zip(ns, Ls):
        dk = dtype(2 * math.pi) / dtype(L)
        i = np.arange(n)
        out.append((np.where(i <= n // 2, i, i - n).astype(dtype) * dk).astype(dtype))
    return tuple(out)


def wavenumber_magnitude_grid(ks):
    """Utils.jl:76-88."""
    mesh = np.meshgrid(*ks, indexing="ij")
    return np.sqrt(sum(m * m for m in mesh))


def as_component_field(f, nd):
    """Utils.jl:19-25."""
    if f.ndim == nd:
        return f[..., None]
    if f.ndim == nd + 1:
        return f
    raise ValueError(f"field has {f.ndim} dims; expected nd={nd} (shape (ns...)) or "
                     f"nd+1={nd + 1} (shape (ns...,1)).")


def shell_coordinate(geometry, ks):
    """ShellBinning.jl:19-36."""
    nd = len(ks)
    dims = tuple(range(1, nd + 1)) if (geometry is None or geometry.dims is None) else tuple(
        geometry.dims)
    if not all(1 <= d <= nd for d in dims):
        raise ValueError(f"ShellMagnitude dims={dims} out of range for nd={nd} spatial dimensions.")
    mesh = np.meshgrid(*ks, indexing="ij")
    ft = np.result_type(ks[0].dtype, np.float32)
    s = np.zeros(tuple(len(k) for k in ks), dtype=ft)
    for d in dims:
        s = s + mesh[d - 1].astype(ft) ** 2
    return np.sqrt(s)


def shell_edges(b, k_max):
    """ShellBinning.jl:50-79."""
    if isinstance(b, LinearBinning):
        if not b.dk > 0:
            raise ValueError("LinearBinning: Δk must be positive.")
        if not k_max > 0:
            raise ValueError("k_max must be positive.")
        k_max = float(k_max)
        n = int(math.floor(k_max / b.dk * (1 + 1e-12))) + 1
        edges = [i * float(b.dk) for i in range(n)]
        while len(edges) > 1 and edges[-1] > k_max:
            edges.pop()
        if edges[-1] < k_max:
            edges.append(k_max)
        return np.asarray(edges)
    if isinstance(b, LogarithmicBinning):
        if not b.k0 > 0:
            raise ValueError("LogarithmicBinning: k₀ must be positive.")
        if not b.lam > 1:
            raise ValueError("LogarithmicBinning: λ must be > 1.")
        if not k_max >= b.k0:
            raise ValueError("k_max must be >= k₀.")
        n_max = int(math.floor(math.log(float(k_max) / b.k0) / math.log(b.lam)))
        edges = [float(b.k0) * float(b.lam) ** n for n in range(n_max + 2)]
        while len(edges) > 2 and edges[-2] > k_max:
            edges.pop()
        return np.asarray(edges)
    if isinstance(b, DyadicBinning):
        return shell_edges(LogarithmicBinning(b.k0, 2.0), k_max)
    if isinstance(b, CustomBinning):
        e = np.asarray(b.edges, dtype=np.float64)
        if np.any(np.diff(e) < 0):
            raise ValueError("CustomBinning: edges must be monotonically increasing.")
        return e
    raise ValueError(f"unsupported binning {b!r}")


def shell_centers(b, k_max):
    """ShellBinning.jl:91-106."""
    e = shell_edges(b, k_max)
    if len(e) < 2:
        raise ValueError(f"No shells within k_max={k_max} for this binning.")
    if isinstance(b, (LogarithmicBinning, DyadicBinning)):
        return np.sqrt(e[:-1] * e[1:])
    return (e[:-1] + e[1:]) / 2


def n_shells(b, k_max):
    """ShellBinning.jl:114-116."""
    return len(shell_edges(b, k_max)) - 1


def assign_shells(k_mag, edges):
    """ShellBinning.jl:126-140: n with edges[n] <= k < edges[n+1] (1-based), 0 outside."""
    e = np.asarray(edges, dtype=np.float64)
    j = np.searchsorted(e, np.asarray(k_mag, dtype=np.float64), side="right")
    return np.where((j >= 1) & (j <= len(e) - 1), j, 0).astype(np.int64)


def default_binning(ks):
    """SpectralFlux.jl:302-313."""
    dk = math.inf
    for k in ks:
        a = np.abs(np.asarray(k, dtype=np.float64))
        a = a[a > 0]
        if a.size:
            dk = min(dk, float(a.min()))
    return LinearBinning(dk if math.isfinite(dk) else 1.0)


def shell_setup(ks, binning, geometry):
    """(centers, edges, shell map) exactly as SpectralFlux.jl:77-81 builds them."""
    binning = binning if binning is not None else default_binning(ks)
    kc = shell_coordinate(geometry if geometry is not None else ShellMagnitude(None), ks)
    kmax = float(kc.max())
    edges = shell_edges(binning, kmax)
    centers = shell_centers(binning, kmax)
    return centers, edges, assign_shells(kc, edges)
