"""FP64 CPU restatement of FlowInvariantTransfer.jl's pseudospectral transfer path.

TEST INFRASTRUCTURE ONLY.  This module is the oracle the HIP backend is checked against.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it, and
only as the checker / CPU baseline -- never as the thing measured or shipped.  The product path
(``flowinvarianttransfer.jl_amd``) never imports it.

It follows the reference's Julia sources function by function (citations are ``file:line`` in the
reference repository).  The reference computes every transform as a full complex-to-complex FFT
followed by ``real.()``; this restatement does the same with ``scipy.fft`` (pocketfft) in float64,
so it is also the "CPU restatement of the reference's FFT path" that is timed as the CPU baseline.

Array convention: fields are numpy arrays of shape ``(n_1, ..., n_nd, C)`` whose first axes are
the spatial FFT axes in FFTW index order and whose last axis is the component axis (the
reference's ``(ns..., C)``).  Memory order does not matter here (Fortran order matches Julia).

Parity pins (what checks each function; see tests/test_oracle.py):
  * nonlinear_term_fft        -- DirectSum twin (explicit DFT matrices), single-cosine-mode zero
                                 transfer, energy / scalar-variance / 2D-enstrophy conservation.
  * nonlinear_term_padded_fft -- equals the 2/3 term on |k|<N/3 for band-limited fields, conserves,
                                 brute-force alias-free convolution on tiny grids.
  * transfer_density          -- per-mode closed forms (helicity = k·(û×û*) identities, enstrophy
                                 = |k|^2 KE for 2D solenoidal fields).
  * shell binning             -- the reference's own known answers (test/runtests.jl:56-98,718-735)
                                 plus brute-force membership.
  * spectral / shell-to-shell / mode-to-mode fluxes -- Pi = cumsum(T), antisymmetry, the
                                 reductions S -> T(n,m) -> T(k), and the independent closed-form
                                 triad sum (mode_to_mode_closed_form).
  * helical decomposition     -- reconstruction, orthonormality, divergence, helicity split.
  * band-to-band transfer     -- the DirectSum twin, the bilinear identity T = W S Wᵀ with the
                                 independent mode-to-mode closed form S, partition of unity of
                                 the weights, antisymmetry / conservation / net = row sums.
No function of this module is "parity unpinned".
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.fft as sfft

# ---------------------------------------------------------------------------------------------
# Vocabulary (duck-typed on class name so the package's own types are accepted as well)
# ---------------------------------------------------------------------------------------------


@dataclass(frozen=True)
class LinearBinning:          # Types.jl:370-381
    dk: float


@dataclass(frozen=True)
class LogarithmicBinning:     # Types.jl:383-396
    k0: float
    lam: float


@dataclass(frozen=True)
class DyadicBinning:          # Types.jl:398-409
    k0: float


@dataclass(frozen=True)
class CustomBinning:          # Types.jl:411-421
    edges: tuple


@dataclass(frozen=True)
class ShellMagnitude:         # Types.jl:504-506 (dims=None -> isotropic)
    dims: tuple | None = None


def IsotropicShells():        # Types.jl:513
    return ShellMagnitude(None)


def PerpendicularShells(dims=(1, 2)):   # Types.jl:521 (1-based dims like the reference)
    return ShellMagnitude(tuple(dims))


def ParallelShells(dims=(3,)):          # Types.jl:529
    return ShellMagnitude(tuple(dims))


def _kind(x) -> str:
    """Normalise a dealiasing / invariant / backend tag (string or singleton object)."""
    if isinstance(x, str):
        return x
    return type(x).__name__


_DEALIAS = {"NoDealiasing": "none", "none": "none",
            "OrszagTwoThirds": "23", "23": "23",
            "PaddedThreeHalves": "pad", "pad": "pad"}
_INVARIANT = {"KineticEnergy": "ke", "ke": "ke", "Helicity": "helicity", "helicity": "helicity",
              "Enstrophy": "enstrophy", "enstrophy": "enstrophy",
              "PassiveScalar": "scalar", "scalar": "scalar"}


def dealias_tag(d) -> str:
    return _DEALIAS[_kind(d)]


def invariant_tag(i) -> str:
    return _INVARIANT[_kind(i)]


# ---------------------------------------------------------------------------------------------
# Grid utilities
# ---------------------------------------------------------------------------------------------

def wavenumber_grid(ns, Ls):
    """Utils.jl:51-63 -- FFTW-order physical wavenumbers, Nyquist taken POSITIVE (+N/2)."""
    out = []
    for n, L in zip(ns, Ls):
        dk = 2.0 * math.pi / float(L)
        idx = np.arange(n)
        kk = np.where(idx <= n // 2, idx, idx - n).astype(np.float64)
        out.append(kk * dk)
    return tuple(out)


def as_component_field(f, nd):
    """Utils.jl:19-25 -- (ns...) -> (ns...,1); (ns...,M) unchanged; else ArgumentError."""
    if f.ndim == nd:
        return f.reshape(f.shape + (1,), order="F") if np.isfortran(f) else f[..., None]
    if f.ndim == nd + 1:
        return f
    raise ValueError(f"field has {f.ndim} dims; expected nd={nd} or nd+1={nd + 1}")


def _mesh(ks):
    return np.meshgrid(*ks, indexing="ij")


def wavenumber_magnitude_grid(ks):
    """Utils.jl:76-88."""
    s = np.zeros(tuple(len(k) for k in ks))
    for K in _mesh(ks):
        s = s + K * K
    return np.sqrt(s)


def dealiased_mask(ns):
    """NonlinearTerm.jl:134-141 -- True where a mode is DISCARDED by the integer 2/3 rule:
    folded index |k_d| >= n_d ÷ 3 along any dimension."""
    disc = np.zeros(tuple(ns), dtype=bool)
    for d, n in enumerate(ns):
        idx = np.arange(n)
        kabs = np.where(idx <= n // 2, idx, n - idx)
        sh = [1] * len(ns)
        sh[d] = n
        disc = disc | (kabs >= n // 3).reshape(sh)
    return disc


def keep_mask(ns):
    """FFTWExt.jl:47 -- keepmask = !_is_dealiased."""
    return ~dealiased_mask(ns)


def k_component(ks, d):
    """FFTWExt.jl:301-309 -- dense (ns...) array holding ks[d] along axis d."""
    return _mesh(ks)[d]


# ---------------------------------------------------------------------------------------------
# FFT helpers (c2c, like the reference)
# ---------------------------------------------------------------------------------------------

def _fft(a, nd, workers):
    return sfft.fftn(a, axes=tuple(range(nd)), workers=workers)


def _ifft(a, nd, workers):
    return sfft.ifftn(a, axes=tuple(range(nd)), workers=workers)


def padded_len(n: int) -> int:
    """FFTWExt.jl:128-131 -- smallest m >= 3n/2 with (m - n) even."""
    m = -(-3 * n // 2)
    return m if (m - n) % 2 == 0 else m + 1


def embed_spectrum(A, Ms):
    """FFTWExt.jl:134-141 -- fftshift, centre in a zero Ms grid, ifftshift.  (numpy's fftshift /
    ifftshift shift by n//2 / -(n//2), exactly Julia's circshift by ±div(n,2).)"""
    nd = A.ndim
    c = np.fft.fftshift(A)
    B = np.zeros(tuple(Ms), dtype=A.dtype)
    sl = tuple(slice((Ms[d] - A.shape[d]) // 2, (Ms[d] - A.shape[d]) // 2 + A.shape[d])
               for d in range(nd))
    B[sl] = c
    return np.fft.ifftshift(B)


def truncate_spectrum(A, ns):
    """FFTWExt.jl:144-149 -- inverse of embed_spectrum."""
    nd = A.ndim
    c = np.fft.fftshift(A)
    sl = tuple(slice((A.shape[d] - ns[d]) // 2, (A.shape[d] - ns[d]) // 2 + ns[d])
               for d in range(nd))
    return np.fft.ifftshift(c[sl])


# ---------------------------------------------------------------------------------------------
# Nonlinear term N̂ = FFT[(u_adv·∇) f]
# ---------------------------------------------------------------------------------------------

def nonlinear_term_fft(f_hat, ks, truncate=True, advecting_hat=None, workers=1):
    """FFTWExt.jl:61-121 -- c2c pseudospectral (u·∇)f with optional Orszag 2/3 truncation of the
    inputs (advecting velocity AND advected gradient) and of the output; ifft = bfft/Np and
    N̂ = fft(N)/Np."""
    if advecting_hat is None:
        advecting_hat = f_hat
    nd = len(ks)
    ns = f_hat.shape[:nd]
    M = f_hat.shape[nd]
    Np = float(np.prod(ns))
    keep = keep_mask(ns) if truncate else np.ones(ns, dtype=bool)
    u = [np.real(_ifft(keep * advecting_hat[..., j], nd, workers)) for j in range(nd)]
    kc = [k_component(ks, d) for d in range(nd)]
    out = np.zeros(f_hat.shape, dtype=np.complex128)
    for c in range(M):
        Nc = np.zeros(ns)
        for j in range(nd):
            g = np.real(_ifft(1j * kc[j] * keep * f_hat[..., c], nd, workers))
            Nc += u[j] * g
        out[..., c] = _fft(Nc, nd, workers) * keep / Np
    return out


def nonlinear_term_padded_fft(f_hat, ks, advecting_hat=None, workers=1):
    """FFTWExt.jl:160-183 -- exact 3/2 zero padding.  The advecting field is NOT 2/3-masked; the
    spectral derivative uses ks at the original index (Nyquist +N/2) and the fftshift embedding
    then places that bin at -N/2 on the padded grid (the reference's Nyquist placement)."""
    if advecting_hat is None:
        advecting_hat = f_hat
    nd = len(ks)
    ns = f_hat.shape[:nd]
    M = f_hat.shape[nd]
    Ms = tuple(padded_len(n) for n in ns)
    Ntot = float(np.prod(ns))
    Mtot = float(np.prod(Ms))
    scale = Mtot / Ntot ** 2
    kc = [k_component(ks, d) for d in range(nd)]
    uM = [np.real(_ifft(embed_spectrum(advecting_hat[..., j], Ms), nd, workers))
          for j in range(nd)]
    out = np.zeros(f_hat.shape, dtype=np.complex128)
    for c in range(M):
        NcM = np.zeros(Ms)
        for j in range(nd):
            g = np.real(_ifft(embed_spectrum(1j * kc[j] * f_hat[..., c], Ms), nd, workers))
            NcM += uM[j] * g
        out[..., c] = truncate_spectrum(_fft(NcM, nd, workers), ns) * scale
    return out


def _dft_matrix(ns):
    """Explicit DFT phases of NonlinearTerm.jl:189-194: E[x, k] = exp(i 2π Σ_d km_d x_d / n_d),
    with x and k enumerated in column-major (first index fastest) order."""
    grids_x = np.meshgrid(*[np.arange(n) for n in ns], indexing="ij")
    grids_k = np.meshgrid(*[np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
                            for n in ns], indexing="ij")
    X = [g.reshape(-1, order="F").astype(np.float64) for g in grids_x]
    K = [g.reshape(-1, order="F").astype(np.float64) for g in grids_k]
    phase = np.zeros((X[0].size, K[0].size))
    for d, n in enumerate(ns):
        phase += 2.0 * math.pi * np.outer(X[d] / n, K[d])
    return np.exp(1j * phase)


def nonlinear_term_direct(f_hat, ks, truncate=True, advecting_hat=None):
    """NonlinearTerm.jl:160-265 -- the O(N^{2D}) DirectSum backend (explicit DFT sums).  Used as
    an independent twin of nonlinear_term_fft on tiny grids (different algorithm, same math)."""
    if advecting_hat is None:
        advecting_hat = f_hat
    nd = len(ks)
    ns = f_hat.shape[:nd]
    M = f_hat.shape[nd]
    Np = int(np.prod(ns))
    E = _dft_matrix(ns)                              # (x, k)
    keepv = (~dealiased_mask(ns)).reshape(-1, order="F") if truncate else np.ones(Np, bool)

    def flat(a):
        return a.reshape(-1, order="F")

    u = [np.real(E @ (flat(advecting_hat[..., j]) * keepv)) / Np for j in range(nd)]
    kcf = [flat(k_component(ks, d)) for d in range(nd)]
    out = np.zeros(f_hat.shape, dtype=np.complex128)
    for c in range(M):
        Nc = np.zeros(Np)
        for j in range(nd):
            g = np.real(E @ (1j * kcf[j] * flat(f_hat[..., c]) * keepv)) / Np
            Nc += u[j] * g
        Nh = (E.conj().T @ Nc) / Np
        if truncate:
            Nh = Nh * keepv
        out[..., c] = Nh.reshape(ns, order="F")
    return out


def compute_nonlinear_term(f_hat, ks, dealiasing="23", spectral="FFTBackend",
                           advecting_hat=None, workers=1):
    """NonlinearTerm.jl:65-116 -- dispatch on (spectral, dealiasing)."""
    d = dealias_tag(dealiasing)
    s = _kind(spectral)
    if s == "DirectSumBackend":
        if d == "pad":
            raise ValueError("PaddedThreeHalves dealiasing requires the FFT path")
        return nonlinear_term_direct(f_hat, ks, truncate=(d == "23"), advecting_hat=advecting_hat)
    if d == "pad":
        return nonlinear_term_padded_fft(f_hat, ks, advecting_hat=advecting_hat, workers=workers)
    return nonlinear_term_fft(f_hat, ks, truncate=(d == "23"), advecting_hat=advecting_hat,
                              workers=workers)


# ---------------------------------------------------------------------------------------------
# Per-mode transfer density
# ---------------------------------------------------------------------------------------------

def vorticity(u_hat, ks):
    """ω̂ = i k × û (3D) -- Invariants.jl:68-70."""
    kx, ky, kz = _mesh(ks)
    ux, uy, uz = u_hat[..., 0], u_hat[..., 1], u_hat[..., 2]
    return (1j * (ky * uz - kz * uy), 1j * (kz * ux - kx * uz), 1j * (kx * uy - ky * ux))


def transfer_density(invariant, c_hat, N_hat, ks):
    """Invariants.jl:40-104 -- real per-mode density t(k) for each invariant."""
    inv = invariant_tag(invariant)
    nd = len(ks)
    if inv in ("ke", "scalar"):
        return np.sum(np.real(np.conj(c_hat) * N_hat), axis=nd)
    if inv == "helicity":
        if nd != 3:
            raise ValueError(f"Helicity transfer is defined in 3D only (got nd={nd}).")
        w = vorticity(c_hat, ks)
        return sum(np.real(np.conj(w[c]) * N_hat[..., c]) for c in range(3))
    if inv == "enstrophy":
        if nd == 2:
            kx, ky = _mesh(ks)
            w = 1j * (kx * c_hat[..., 1] - ky * c_hat[..., 0])
            Nw = 1j * (kx * N_hat[..., 1] - ky * N_hat[..., 0])
            return np.real(np.conj(w) * Nw)
        if nd == 3:
            w = vorticity(c_hat, ks)
            Nw = vorticity(N_hat, ks)
            return sum(np.real(np.conj(w[c]) * Nw[c]) for c in range(3))
        raise ValueError(f"Enstrophy transfer is defined in 2D or 3D; got nd={nd}.")
    raise ValueError(f"unknown invariant {invariant!r}")


# ---------------------------------------------------------------------------------------------
# Shell binning
# ---------------------------------------------------------------------------------------------

def shell_coordinate(geometry, ks):
    """ShellBinning.jl:19-36 -- sqrt(Σ_{d∈dims} k_d²) (dims 1-based, None = all)."""
    nd = len(ks)
    dims = tuple(range(1, nd + 1)) if (geometry is None or geometry.dims is None) else geometry.dims
    if not all(1 <= d <= nd for d in dims):
        raise ValueError(f"ShellMagnitude dims={dims} out of range for nd={nd}")
    mesh = _mesh(ks)
    s = np.zeros(tuple(len(k) for k in ks))
    for d in dims:
        s = s + mesh[d - 1] * mesh[d - 1]
    return np.sqrt(s)


def shell_edges(b, k_max):
    """ShellBinning.jl:50-79."""
    name = type(b).__name__
    k_max = float(k_max)
    if name == "LinearBinning":
        if not b.dk > 0:
            raise ValueError("LinearBinning: Δk must be positive.")
        if not k_max > 0:
            raise ValueError("k_max must be positive.")
        # collect(0:Δk:k_max): Julia ranges compute element i as i*Δk (no accumulation)
        n = int(math.floor(k_max / b.dk + 1e-12 * (k_max / b.dk))) + 1
        edges = [i * float(b.dk) for i in range(n)]
        while edges and edges[-1] > k_max:
            edges.pop()
        if edges[-1] < k_max:
            edges.append(k_max)
        return np.asarray(edges, dtype=np.float64)
    if name == "LogarithmicBinning":
        if not b.k0 > 0:
            raise ValueError("LogarithmicBinning: k₀ must be positive.")
        if not b.lam > 1:
            raise ValueError("LogarithmicBinning: λ must be > 1.")
        if not k_max >= b.k0:
            raise ValueError("k_max must be >= k₀.")
        n_max = int(math.floor(math.log(k_max / b.k0) / math.log(b.lam)))
        edges = [float(b.k0) * float(b.lam) ** n for n in range(0, n_max + 2)]
        while len(edges) > 2 and edges[-2] > k_max:
            edges.pop()
        return np.asarray(edges, dtype=np.float64)
    if name == "DyadicBinning":
        return shell_edges(LogarithmicBinning(b.k0, 2.0), k_max)
    if name == "CustomBinning":
        e = np.asarray(b.edges, dtype=np.float64)
        if np.any(np.diff(e) < 0):
            raise ValueError("CustomBinning: edges must be monotonically increasing.")
        return e
    raise ValueError(f"unknown binning {b!r}")


def shell_centers(b, k_max):
    """ShellBinning.jl:91-106."""
    e = shell_edges(b, k_max)
    if len(e) - 1 <= 0:
        raise ValueError(f"No shells within k_max={k_max} for this binning.")
    if type(b).__name__ in ("LogarithmicBinning", "DyadicBinning"):
        return np.sqrt(e[:-1] * e[1:])
    return (e[:-1] + e[1:]) / 2


def assign_shells(k_mag, edges):
    """ShellBinning.jl:126-140 -- brute-force first-match search, 0 = outside every shell."""
    idx = np.zeros(k_mag.shape, dtype=np.int64)
    flat = k_mag.reshape(-1)
    out = idx.reshape(-1)
    nsh = len(edges) - 1
    for i, k in enumerate(flat):
        for n in range(nsh):
            if edges[n] <= k < edges[n + 1]:
                out[i] = n + 1
                break
    return idx


def assign_shells_fast(k_mag, edges):
    """Vectorised assign_shells (same rule) for large grids; checked against the brute force."""
    e = np.asarray(edges)
    j = np.searchsorted(e, k_mag, side="right")
    nsh = len(e) - 1
    idx = np.where((j >= 1) & (j <= nsh), j, 0).astype(np.int64)
    # empty shells [a, a) never match: searchsorted already lands past equal edges
    return idx


def default_binning(ks):
    """SpectralFlux.jl:302-313."""
    dk = math.inf
    for k in ks:
        a = np.abs(k)
        a = a[a > 0]
        if a.size:
            dk = min(dk, float(a.min()))
    return LinearBinning(dk if math.isfinite(dk) else 1.0)


def shell_setup(ks, binning=None, geometry=None):
    binning = binning if binning is not None else default_binning(ks)
    kc = shell_coordinate(geometry, ks)
    kmax = float(kc.max())
    edges = shell_edges(binning, kmax)
    centers = shell_centers(binning, kmax)
    sidx = assign_shells_fast(kc, edges) if kc.size > 4096 else assign_shells(kc, edges)
    return centers, edges, sidx


def shell_sum(t, sidx, nsh):
    """SpectralFlux.jl:164-169 -- T[n] = Σ_{shell(I)=n} t[I], shell 0 dropped."""
    w = t.reshape(-1)
    s = sidx.reshape(-1)
    m = s > 0
    return np.bincount(s[m] - 1, weights=w[m], minlength=nsh)[:nsh]


# ---------------------------------------------------------------------------------------------
# Diagnostics
# ---------------------------------------------------------------------------------------------

def calculate_spectral_flux(u_hat, ks, binning=None, dealiasing="23", invariant="ke",
                            spectral="FFTBackend", advecting_hat=None, geometry=None,
                            decomposition=None, workers=1):
    """SpectralFlux.jl:49-180 -- returns (k_shells, T, Pi) with Pi = +cumsum(T) (:176); with a
    helical decomposition returns {name: (k_shells, T, Pi)} sharing one N̂ (:96-122)."""
    centers, edges, sidx = shell_setup(ks, binning, geometry)
    nsh = len(centers)
    N_hat = compute_nonlinear_term(u_hat, ks, dealiasing, spectral, advecting_hat, workers)
    if decomposition is not None and _kind(decomposition) != "NoDecomposition":
        comps = decompose(decomposition, u_hat, ks)
        res = {}
        for name, comp in comps.items():
            T = shell_sum(transfer_density(invariant, comp, N_hat, ks), sidx, nsh)
            res[name] = (centers, T, np.cumsum(T))
        return res
    T = shell_sum(transfer_density(invariant, u_hat, N_hat, ks), sidx, nsh)
    return centers, T, np.cumsum(T)


def calculate_scalar_flux(u_hat, theta_hat, ks, binning=None, dealiasing="23",
                          spectral="FFTBackend", geometry=None, workers=1):
    """SpectralFlux.jl:203-215."""
    th = as_component_field(theta_hat, len(ks))
    return calculate_spectral_flux(th, ks, binning=binning, dealiasing=dealiasing,
                                   invariant="scalar", spectral=spectral, advecting_hat=u_hat,
                                   geometry=geometry, workers=workers)


def _dealias_copy(a, ns):
    """FFTWExt.jl:186-196."""
    return a * keep_mask(ns)[..., None]


def shell_to_shell_columns(u_hat, ks, sidx, nsh, givers, dealiasing="23", invariant="ke",
                           advecting_hat=None, spectral="FFTBackend", workers=1):
    """FFTWExt.jl:209-275 (FFT) / ShellToShellTransfer.jl:176-224 (DirectSum) -- the giver loop
    restricted to the given giver shells (1-based); returns T[:, givers] (n_sh × len(givers))."""
    if advecting_hat is None:
        advecting_hat = u_hat
    nd = len(ks)
    ns = u_hat.shape[:nd]
    M = u_hat.shape[nd]
    d = dealias_tag(dealiasing)
    s = _kind(spectral)
    Np = float(np.prod(ns))
    T = np.zeros((nsh, len(givers)))
    if s == "DirectSumBackend":
        for col, m in enumerate(givers):
            um = u_hat * (sidx == m)[..., None]
            Nm = compute_nonlinear_term(um, ks, dealiasing, "DirectSumBackend", advecting_hat)
            T[:, col] = shell_sum(transfer_density(invariant, u_hat, Nm, ks), sidx, nsh)
        return T
    trunc = d == "23"
    padded = d == "pad"
    vhat = _dealias_copy(u_hat, ns) if trunc else u_hat
    if not padded:
        ahat = _dealias_copy(advecting_hat, ns) if trunc else advecting_hat
        kc = [k_component(ks, dd) for dd in range(nd)]
        u_full = [np.real(_ifft(ahat[..., j], nd, workers)) for j in range(nd)]
        keep = keep_mask(ns)
    for col, m in enumerate(givers):
        um = vhat * (sidx == m)[..., None]
        if padded:
            Nm = nonlinear_term_padded_fft(um, ks, advecting_hat=advecting_hat, workers=workers)
        else:
            Nm = np.zeros(u_hat.shape, dtype=np.complex128)
            for c in range(M):
                Nc = np.zeros(ns)
                for j in range(nd):
                    Nc += u_full[j] * np.real(_ifft(1j * kc[j] * um[..., c], nd, workers))
                Nm[..., c] = _fft(Nc, nd, workers) / Np
                if trunc:
                    Nm[..., c] *= keep
        T[:, col] = shell_sum(transfer_density(invariant, u_hat, Nm, ks), sidx, nsh)
    return T


def _giver_column_task(args):
    """One task of the threaded giver loop: column m of T(n,m) from memory-mapped inputs, with its
    own workspace and single-threaded transforms (OhMyThreadsExt.jl:44-73)."""
    u_path, s_path, ks, nsh, m, dealiasing, invariant = args
    u_hat = np.load(u_path, mmap_mode="r")
    sidx = np.load(s_path, mmap_mode="r")
    return m, shell_to_shell_columns(np.asarray(u_hat), ks, np.asarray(sidx), nsh, [m], dealiasing, invariant,
                                     workers=1)[:, 0]


def shell_to_shell_threaded(u_hat, ks, sidx, nsh, givers, dealiasing="23", invariant="ke", workers=1,
                            tmpdir=None):
    """OhMyThreadsExt.jl:24-97 -- the reference's threaded giver loop: one task per giver column m
    (each computing the full nonlinear term of the band-m field with its own workspace), run on
    `workers` spawned processes with single-threaded transforms.  û and the shell map reach the
    workers as memory-mapped .npy files (written under tmpdir).  Returns T[:, givers]."""
    import multiprocessing as mp
    import os
    import tempfile
    d = tempfile.mkdtemp(dir=tmpdir)
    u_path, s_path = os.path.join(d, "u.npy"), os.path.join(d, "s.npy")
    try:
        np.save(u_path, np.asarray(u_hat))
        np.save(s_path, np.asarray(sidx))
        tasks = [(u_path, s_path, [np.asarray(k) for k in ks], nsh, int(m), dealiasing, invariant) for m in givers]
        T = np.zeros((nsh, len(givers)))
        col = {int(m): i for i, m in enumerate(givers)}
        with mp.get_context("spawn").Pool(processes=max(1, int(workers))) as pool:
            for m, c in pool.imap_unordered(_giver_column_task, tasks):
                T[:, col[m]] = c
        return T
    finally:
        for p in (u_path, s_path):
            if os.path.exists(p):
                os.remove(p)
        os.rmdir(d)


def calculate_shell_to_shell_transfer(u_hat, ks, binning=None, dealiasing="23",
                                      verify_antisymmetry=True, invariant="ke",
                                      spectral="FFTBackend", advecting_hat=None, geometry=None,
                                      workers=1):
    """ShellToShellTransfer.jl:78-104 + FFTWExt.jl:209-295 -- (centers, edges, T, net, asym)."""
    centers, edges, sidx = shell_setup(ks, binning, geometry)
    nsh = len(centers)
    if _kind(spectral) == "DirectSumBackend" and dealias_tag(dealiasing) == "pad":
        raise ValueError("PaddedThreeHalves dealiasing requires the FFT path")
    T = shell_to_shell_columns(u_hat, ks, sidx, nsh, list(range(1, nsh + 1)), dealiasing,
                               invariant, advecting_hat, spectral, workers)
    net = T.sum(axis=1)
    asym = float(np.max(np.abs(T + T.T))) if verify_antisymmetry else float("nan")
    return centers, edges, T, net, asym


def calculate_scalar_shell_to_shell_transfer(u_hat, theta_hat, ks, **kw):
    """ShellToShellTransfer.jl:143-147."""
    th = as_component_field(theta_hat, len(ks))
    return calculate_shell_to_shell_transfer(th, ks, invariant="scalar", advecting_hat=u_hat, **kw)


def calculate_mode_to_mode_transfer(u_hat, ks, invariant="ke", dealiasing="23",
                                    spectral="FFTBackend", max_modes=1024, force=False,
                                    advecting_hat=None):
    """ScaleToScaleTransfer.jl:49-91 -- one full nonlinear term per giver mode p.  Returns
    (net (ns), S (ns..., ns...)) with S[k..., p...]."""
    nd = len(ks)
    ns = u_hat.shape[:nd]
    M = u_hat.shape[nd]
    Nm = int(np.prod(ns))
    if not (force or Nm <= max_modes):
        raise ValueError(f"N_modes={Nm} exceeds max_modes={max_modes}")
    if advecting_hat is None:
        advecting_hat = u_hat
    S = np.zeros(tuple(ns) + tuple(ns))
    for p in np.ndindex(*ns):
        up = np.zeros_like(u_hat)
        up[p] = u_hat[p]
        N = compute_nonlinear_term(up, ks, dealiasing, spectral, advecting_hat)
        S[(Ellipsis,) + p] = transfer_density(invariant, u_hat, N, ks)
    net = S.reshape(tuple(ns) + (Nm,), order="C").sum(axis=-1)
    return net, S


def effective_carrier(invariant, c_hat, ks):
    """The per-mode vector e(k) with t(k) = Re Σ_c conj(e_c(k)) N̂_c(k) for each invariant
    (KE/scalar: ĉ; helicity: i k×û; enstrophy 3D: i k×(i k×û); 2D: (i k_y ω̂, -i k_x ω̂))."""
    inv = invariant_tag(invariant)
    nd = len(ks)
    if inv in ("ke", "scalar"):
        return c_hat
    if inv == "helicity":
        return np.stack(vorticity(c_hat, ks), axis=-1)
    if nd == 2:
        kx, ky = _mesh(ks)
        w = 1j * (kx * c_hat[..., 1] - ky * c_hat[..., 0])
        return np.stack([1j * ky * w, -1j * kx * w], axis=-1)
    w = np.stack(vorticity(c_hat, ks), axis=-1)
    return np.stack(vorticity(w, ks), axis=-1)


def mode_to_mode_closed_form(u_hat, ks, invariant="ke", dealiasing="23", advecting_hat=None):
    """Independent closed form of the mode-resolved triad transfer (derived from the definition
    of the reference's loop, ScaleToScaleTransfer.jl:75-88 with FFTWExt.jl:61-121): for each
    giver p the single-mode field enters ifft through its Hermitian part, so

      N̂_p(k) = keep(k)keep(p)/(2Np²) Σ_j p_j [ i f̂(p) A_j(k-p) - i conj(f̂(p)) A_j(k+p) ],
      A_j = Hermitian part of keep·â_j,   S(k|p) = Re Σ_c conj(e_c(k)) N̂_{p,c}(k).

    Evaluated by index arithmetic (no FFT), so it checks the FFT loop independently.

    PaddedThreeHalves (FFTWExt.jl:160-183): the same sum on the padded grid G = padded_len(n),
    with keep = 1, A_j the Hermitian part (on G) of the fftshift-embedded â_j, and k, p placed by
    the embedding E(i) = i for i < ceil(n/2), i - n + G otherwise (the Nyquist bin of an even n at
    -n/2); k ∓ p are taken mod G, so a pair whose sum reaches the Nyquist bin of the other side
    picks up A_j there exactly as the padded transform does.  The scale Mtot/Ntot² of the
    reference's truncation and the 1/Mtot² of the two padded transforms leave 1/(2 Ntot²)."""
    d = dealias_tag(dealiasing)
    if advecting_hat is None:
        advecting_hat = u_hat
    nd = len(ks)
    ns = u_hat.shape[:nd]
    M = u_hat.shape[nd]
    Np = float(np.prod(ns))
    keep = keep_mask(ns) if d == "23" else np.ones(ns, bool)
    e = effective_carrier(invariant, u_hat, ks)
    Gs = tuple(padded_len(n) for n in ns) if d == "pad" else tuple(ns)
    A = []
    for j in range(nd):
        g = keep * advecting_hat[..., j]
        if d == "pad":
            g = embed_spectrum(g, Gs)
        neg = g[tuple(np.s_[::-1] for _ in range(nd))]
        neg = np.roll(neg, 1, axis=tuple(range(nd)))          # g(-k)
        A.append(0.5 * (g + np.conj(neg)))
    # transform-grid index of every original index (identity without padding)
    emb = [np.where(np.arange(n) < (n + 1) // 2, np.arange(n), np.arange(n) - n + G)
           for n, G in zip(ns, Gs)]
    S = np.zeros(tuple(ns) + tuple(ns))
    kidx = np.meshgrid(*[emb[dd] for dd in range(nd)], indexing="ij")
    for p in np.ndindex(*ns):
        if not keep[p]:
            continue
        pk = [ks[j][p[j]] for j in range(nd)]
        kmp = tuple((kidx[dd] - emb[dd][p[dd]]) % Gs[dd] for dd in range(nd))
        kpp = tuple((kidx[dd] + emb[dd][p[dd]]) % Gs[dd] for dd in range(nd))
        dot_m = sum(pk[j] * A[j][kmp] for j in range(nd))
        dot_p = sum(pk[j] * A[j][kpp] for j in range(nd))
        tot = np.zeros(ns)
        for c in range(M):
            fp = u_hat[p + (c,)]
            Nc = keep / (2 * Np * Np) * (1j * dot_m * fp - 1j * dot_p * np.conj(fp))
            tot += np.real(np.conj(e[..., c]) * Nc)
        S[(Ellipsis,) + p] = tot
    return S


# ---------------------------------------------------------------------------------------------
# Smooth band-to-band transfer T(K,Q)
# ---------------------------------------------------------------------------------------------

@dataclass(frozen=True)
class SmoothBands:            # Types.jl:414-430
    centers: tuple
    logwidth: float = 0.6


def band_weights(centers, logwidth, k_coord):
    """BandTransfer.jl:20-44 -- w_n(κ) = exp(-(ln(κ/c_n))²/(2σ²)) renormalised over the bands to a
    partition of unity; modes with κ <= 0 (and modes whose weights all underflow) get weight 0."""
    nb = len(centers)
    pos = k_coord > 0
    kk = np.where(pos, k_coord, 1.0)
    W = np.zeros((nb,) + k_coord.shape)
    for n, c in enumerate(centers):
        r = np.log(kk / float(c))
        W[n] = np.where(pos, np.exp(-(r * r) / (2.0 * float(logwidth) ** 2)), 0.0)
    ssum = W.sum(axis=0)
    return np.where(ssum > 0, W / np.where(ssum > 0, ssum, 1.0), W)


def calculate_band_to_band_transfer(u_hat, ks, bands, dealiasing="23", invariant="ke",
                                    spectral="FFTBackend", advecting_hat=None, geometry=None,
                                    workers=1):
    """BandTransfer.jl:66-113 -- T(n,m) = Σ_k w_n(k) t_m(k), t_m the density of the nonlinear term
    of the band-m field w_m ⊙ û advected by the full velocity.  Returns
    (centers, T, net = Σ_m T(n,m), max|T + Tᵀ|)."""
    if advecting_hat is None:
        advecting_hat = u_hat
    kc = shell_coordinate(geometry, ks)
    W = band_weights(bands.centers, bands.logwidth, kc)
    nb = len(bands.centers)
    T = np.zeros((nb, nb))
    for m in range(nb):
        fm = W[m][..., None] * u_hat
        N = compute_nonlinear_term(fm, ks, dealiasing, spectral, advecting_hat, workers)
        d = transfer_density(invariant, u_hat, N, ks)
        for n in range(nb):
            T[n, m] = float(np.sum(W[n] * d))
    return np.asarray(bands.centers, dtype=float), T, T.sum(axis=1), float(np.max(np.abs(T + T.T)))


# ---------------------------------------------------------------------------------------------
# Helical decomposition and partial fluxes
# ---------------------------------------------------------------------------------------------

def helical_decompose(u_hat, ks):
    """Decomposition.jl:63-98 -- Craya–Herring helical basis h± = (e1 ± i e2)/√2."""
    nd = len(ks)
    if nd != 3:
        raise ValueError(f"HelicalDecomposition is defined in 3D only (got nd={nd}).")
    if u_hat.shape[nd] < 3:
        raise ValueError("HelicalDecomposition needs ≥3 velocity components")
    kx, ky, kz = _mesh(ks)
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    zero = kk == 0
    kks = np.where(zero, 1.0, kk)
    kh = (kx / kks, ky / kks, kz / kks)
    use_z = np.abs(kh[2]) < 0.9
    rx = np.where(use_z, 0.0, 1.0)
    ry = np.zeros_like(rx)
    rz = np.where(use_z, 1.0, 0.0)
    e1 = (kh[1] * rz - kh[2] * ry, kh[2] * rx - kh[0] * rz, kh[0] * ry - kh[1] * rx)
    n1 = np.sqrt(e1[0] ** 2 + e1[1] ** 2 + e1[2] ** 2)
    n1 = np.where(zero, 1.0, n1)
    e1 = tuple(c / n1 for c in e1)
    e2 = (kh[1] * e1[2] - kh[2] * e1[1], kh[2] * e1[0] - kh[0] * e1[2],
          kh[0] * e1[1] - kh[1] * e1[0])
    u = [u_hat[..., c] for c in range(3)]
    ue1 = u[0] * e1[0] + u[1] * e1[1] + u[2] * e1[2]
    ue2 = u[0] * e2[0] + u[1] * e2[1] + u[2] * e2[2]
    s = 1.0 / math.sqrt(2.0)
    upc = (ue1 - 1j * ue2) * s
    umc = (ue1 + 1j * ue2) * s
    up = np.zeros(u_hat.shape, dtype=np.complex128)
    um = np.zeros(u_hat.shape, dtype=np.complex128)
    for c in range(3):
        up[..., c] = np.where(zero, 0, upc * (e1[c] + 1j * e2[c]) * s)
        um[..., c] = np.where(zero, 0, umc * (e1[c] - 1j * e2[c]) * s)
    return {"positive": up, "negative": um}


def toroidal_poloidal_decompose(u_hat, ks):
    """Decomposition.jl:107-141 -- Craya–Herring toroidal / poloidal split: e1 = k̂ × ẑ / |k̂ × ẑ|
    (horizontal, toroidal direction), e2 = k̂ × e1 (poloidal); for k ∥ ẑ (k_⊥ = 0) the pair
    (x̂, ŷ); tor = (û·e1) e1, pol = (û·e2) e2, zero at k = 0 and in components past the third."""
    nd = len(ks)
    if nd != 3:
        raise ValueError(f"ToroidalPoloidalDecomposition is defined in 3D only (got nd={nd}).")
    if u_hat.shape[nd] < 3:
        raise ValueError("ToroidalPoloidalDecomposition needs ≥3 velocity components")
    kx, ky, kz = _mesh(ks)
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    zero = kk == 0
    kks = np.where(zero, 1.0, kk)
    kh = (kx / kks, ky / kks, kz / kks)
    kperp = np.sqrt(kx * kx + ky * ky)
    deg = kperp == 0
    e1 = (kh[1], -kh[0], np.zeros_like(kk))                       # k̂ × ẑ
    n1 = np.sqrt(e1[0] ** 2 + e1[1] ** 2)
    n1 = np.where(deg, 1.0, n1)
    e1 = (np.where(deg, 1.0, e1[0] / n1), np.where(deg, 0.0, e1[1] / n1), np.zeros_like(kk))
    e2 = (kh[1] * e1[2] - kh[2] * e1[1], kh[2] * e1[0] - kh[0] * e1[2], kh[0] * e1[1] - kh[1] * e1[0])
    e2 = (np.where(deg, 0.0, e2[0]), np.where(deg, 1.0, e2[1]), np.where(deg, 0.0, e2[2]))
    u = [u_hat[..., c] for c in range(3)]
    c1 = u[0] * e1[0] + u[1] * e1[1] + u[2] * e1[2]
    c2 = u[0] * e2[0] + u[1] * e2[1] + u[2] * e2[2]
    tor = np.zeros(u_hat.shape, dtype=np.complex128)
    pol = np.zeros(u_hat.shape, dtype=np.complex128)
    for c in range(3):
        tor[..., c] = np.where(zero, 0, c1 * e1[c])
        pol[..., c] = np.where(zero, 0, c2 * e2[c])
    return {"toroidal": tor, "poloidal": pol}


def decompose(decomposition, u_hat, ks):
    """Decomposition.jl:42-141 -- the spectral decompositions on the path: NoDecomposition returns
    û itself; the helical and toroidal/poloidal splits return named components in order."""
    k = _kind(decomposition)
    if k == "NoDecomposition":
        return u_hat
    if k == "HelicalDecomposition":
        return helical_decompose(u_hat, ks)
    if k == "ToroidalPoloidalDecomposition":
        return toroidal_poloidal_decompose(u_hat, ks)
    raise ValueError(f"decomposition {decomposition!r} is not a spectral decomposition of this path")


def calculate_partial_fluxes(u_hat, ks, binning=None, dealiasing="23", spectral="FFTBackend",
                             geometry=None, workers=1, decomposition="HelicalDecomposition"):
    """SpectralFlux.jl:240-285: channels (s_k, s_p, s_q) of Re{û_{s_k}* · [(u_{s_p}·∇) u_{s_q}]} over
    the named components of the decomposition (8 for the helical and toroidal/poloidal splits)."""
    comps = decompose(decomposition, u_hat, ks)
    centers, edges, sidx = shell_setup(ks, binning, geometry)
    nsh = len(centers)
    ch = {}
    for sp in comps:
        for sq in comps:
            N = compute_nonlinear_term(comps[sq], ks, dealiasing, spectral, comps[sp], workers)
            for sk in comps:
                T = shell_sum(transfer_density("ke", comps[sk], N, ks), sidx, nsh)
                ch[(sk, sp, sq)] = (centers, T, np.cumsum(T))
    totT = sum(v[1] for v in ch.values())
    return ch, (centers, totT, np.cumsum(totT)), centers


# ---------------------------------------------------------------------------------------------
# Synthetic inputs (the seeded solenoidal fields the benchmarks and tests use)
# ---------------------------------------------------------------------------------------------

def spectrum_shape(kind, kmag, n=None):
    """Per-mode |û(k)|² shapes.  'k-5/3', 'k-3', 'k-1' are power laws (0 at k=0);
    'pk8' is k^4 exp(-2(k/8)^2) for k<=8 joined continuously to k^-5/3 above, zero for
    |k| >= n/3."""
    k = np.where(kmag > 0, kmag, 1.0)
    if kind == "k-5/3":
        s = k ** (-5.0 / 3.0)
    elif kind == "k-3":
        s = k ** (-3.0)
    elif kind == "k-1":
        s = k ** (-1.0)
    elif kind == "pk8":
        kp = 8.0
        low = k ** 4 * np.exp(-2.0 * (k / kp) ** 2)
        join = kp ** 4 * math.exp(-2.0) / kp ** (-5.0 / 3.0)
        s = np.where(k <= kp, low, join * k ** (-5.0 / 3.0))
        if n is not None:
            s = np.where(kmag >= n / 3.0, 0.0, s)
    else:
        raise ValueError(kind)
    return np.where(kmag > 0, s, 0.0)


def synthetic_field(ns, D, seed, spectrum="k-5/3", solenoidal=True, workers=1):
    """Seeded real N(0,1) field -> û = fft(u)/Np (reference convention, test/runtests.jl:187),
    per-mode magnitude reshaped to the requested |û|² shape keeping the phases, Leray-projected
    (if solenoidal), DC and every Nyquist plane zeroed, normalised so Σ|û|² = 1 (rms u = 1 for
    u = bfft(û)).  Hermitian by construction.  Returns complex128 (ns..., D) in Fortran order."""
    nd = len(ns)
    rng = np.random.default_rng(seed)
    Np = float(np.prod(ns))
    ks = wavenumber_grid(ns, (2 * math.pi,) * nd)
    kmag = wavenumber_magnitude_grid(ks)
    shape = np.sqrt(spectrum_shape(spectrum, kmag, n=min(ns)))
    out = np.zeros(tuple(ns) + (D,), dtype=np.complex128, order="F")
    for c in range(D):
        u = rng.standard_normal(tuple(ns))
        uh = _fft(u, nd, workers) / Np
        a = np.abs(uh)
        out[..., c] = np.where(a > 0, uh / np.where(a > 0, a, 1.0), 0.0) * shape
    if solenoidal and D >= nd:
        mesh = _mesh(ks)
        k2 = np.where(kmag > 0, kmag * kmag, 1.0)
        div = sum(mesh[j] * out[..., j] for j in range(nd))
        for j in range(nd):
            out[..., j] -= mesh[j] * div / k2
    nyq = np.zeros(tuple(ns), dtype=bool)
    for d, n in enumerate(ns):
        if n % 2 == 0:
            sl = [slice(None)] * nd
            sl[d] = n // 2
            nyq[tuple(sl)] = True
    out[nyq] = 0
    out[(0,) * nd] = 0
    tot = float(np.sum(np.abs(out) ** 2))
    if tot > 0:
        out /= math.sqrt(tot)
    return out


def streamfunction_velocity_2d(n, seed):
    """Divergence-free 2D velocity from a random real streamfunction, as the reference's tests
    build it (test/runtests.jl:184-191): ψ̂ = fft(ψ)/N², û = (i k_y ψ̂, -i k_x ψ̂)."""
    rng = np.random.default_rng(seed)
    ks = wavenumber_grid((n, n), (2 * math.pi, 2 * math.pi))
    kx, ky = _mesh(ks)
    psi = rng.standard_normal((n, n))
    ph = sfft.fftn(psi) / n ** 2
    u = np.zeros((n, n, 2), dtype=np.complex128, order="F")
    u[..., 0] = 1j * ky * ph
    u[..., 1] = -1j * kx * ph
    return u, ks


def curl_velocity_3d(n, seed, scale=1.0):
    """Divergence-free 3D velocity û = i k × Â with complex Gaussian Â (test/runtests.jl:263-270);
    NOT Hermitian (exercises the real-part semantics)."""
    rng = np.random.default_rng(seed)
    ks = wavenumber_grid((n, n, n), (2 * math.pi,) * 3)
    kx, ky, kz = _mesh(ks)
    A = (rng.standard_normal((n, n, n, 3)) + 1j * rng.standard_normal((n, n, n, 3))) * scale
    u = np.zeros((n, n, n, 3), dtype=np.complex128, order="F")
    u[..., 0] = 1j * (ky * A[..., 2] - kz * A[..., 1])
    u[..., 1] = 1j * (kz * A[..., 0] - kx * A[..., 2])
    u[..., 2] = 1j * (kx * A[..., 1] - ky * A[..., 0])
    return u, ks
