"""Write the committed golden vectors of the pseudospectral transfer path (tests/golden/*.npz).

The reference ships no golden vectors for this path (its tests draw fields from Julia RNG streams
and check properties), so these fixtures are made here, from seeded inputs, and every expected
array is produced by the algorithm that is independent of the FFT code path wherever one exists:

  N_hat (NoDealiasing / OrszagTwoThirds) : explicit DFT sums, the DirectSum backend
                                           (NonlinearTerm.jl:160-265), not an FFT
  N_hat (PaddedThreeHalves, no Nyquist)  : brute-force alias-free convolution over integer
                                           wavevectors (what exact 3/2 padding computes)
  N_hat (PaddedThreeHalves, Nyquist set) : the FFT restatement of FFTWExt.jl:128-183 (the
                                           reference's fftshift placement of the Nyquist bin has
                                           no other closed form)
  T(k), Pi(K)                            : shell sums of the DirectSum density (SpectralFlux.jl:
                                           148-180), Pi = cumsum(T)
  T(n,m)                                 : the DirectSum giver loop (ShellToShellTransfer.jl:
                                           176-248)
  S(k|p)                                 : the reference's one-mode-per-giver loop
                                           (ScaleToScaleTransfer.jl:75-88) over DirectSum terms
  helical u+ / u-                        : Decomposition.jl:63-98 restated

Before writing, every case is cross-checked against the FFT restatement (oracle/fit_oracle.py)
and the mode-to-mode closed form; a disagreement above 1e-12 (normwise) aborts.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import fit_oracle as O  # noqa: E402

L2PI = 2 * math.pi


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def ks_of(ns):
    return O.wavenumber_grid(ns, (L2PI,) * len(ns))


def rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def brute_padded(f, ks, adv):
    """Exact (alias-free) (u·∇)f on integer wavevectors: N̂(k) = Np^-2 Σ_{p+q=k} Σ_j â_j(q) i p_j f̂(p)
    for every k with -n/2 <= k_d < n/2 (no wrap), for fields whose Nyquist planes are empty."""
    nd = len(ks)
    ns = f.shape[:nd]
    Np = float(np.prod(ns))
    freqs = [np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n) for n in ns]
    out = np.zeros(f.shape, dtype=np.complex128)
    idx = list(np.ndindex(*ns))
    for pi in idx:
        p = [freqs[d][pi[d]] for d in range(nd)]
        fp = f[pi]
        if not np.any(fp):
            continue
        for qi in idx:
            q = [freqs[d][qi[d]] for d in range(nd)]
            k = [p[d] + q[d] for d in range(nd)]
            if not all(-ns[d] // 2 <= k[d] < ns[d] // 2 for d in range(nd)):
                continue
            adv_dot = sum(adv[qi + (j,)] * 1j * ks[j][pi[j]] for j in range(nd))
            out[tuple(k[d] % ns[d] for d in range(nd))] += adv_dot * fp
    return out / Np ** 2


def direct_flux(u, ks, inv, binning, dealias, adv=None):
    centers, edges, sidx = O.shell_setup(ks, binning)
    N = O.nonlinear_term_direct(u, ks, truncate=(dealias == "23"), advecting_hat=adv)
    T = O.shell_sum(O.transfer_density(inv, u, N, ks), sidx, len(centers))
    return centers, T, np.cumsum(T), N


def direct_s2s(u, ks, inv, binning, dealias, adv=None):
    centers, edges, sidx = O.shell_setup(ks, binning)
    nsh = len(centers)
    return O.shell_to_shell_columns(u, ks, sidx, nsh, list(range(1, nsh + 1)), dealias, inv,
                                    adv, "DirectSumBackend")


CASES = []


def case(fn):
    CASES.append(fn)
    return fn


@case
def ke2d_12x10_none_nonhermitian():
    """2D 12x10, non-Hermitian random coefficients, NoDealiasing, kinetic energy."""
    ns = (12, 10)
    ks = ks_of(ns)
    u = rand_c(np.random.default_rng(101), ns + (2,))
    c, T, Pi, N = direct_flux(u, ks, "ke", O.LinearBinning(1.0), "none")
    assert rel(N, O.nonlinear_term_fft(u, ks, truncate=False)) < 1e-12
    Tm = direct_s2s(u, ks, "ke", O.LinearBinning(1.0), "none")
    assert rel(Tm, O.calculate_shell_to_shell_transfer(u, ks, O.LinearBinning(1.0), "none")[2]) < 1e-12
    return dict(ns=ns, dealias="none", invariant="ke", binning="linear1"), \
        dict(u_hat=u, N_hat=N, centers=c, T=T, Pi=Pi, Tnm=Tm)


@case
def ke2d_16_23():
    """2D 16², seeded k^-5/3 solenoidal field, OrszagTwoThirds, kinetic energy."""
    ns = (16, 16)
    ks = ks_of(ns)
    u = O.synthetic_field(ns, 2, seed=102)
    c, T, Pi, N = direct_flux(u, ks, "ke", O.LinearBinning(1.0), "23")
    assert rel(N, O.nonlinear_term_fft(u, ks)) < 1e-12
    Tm = direct_s2s(u, ks, "ke", O.LinearBinning(1.0), "23")
    assert rel(Tm, O.calculate_shell_to_shell_transfer(u, ks, O.LinearBinning(1.0), "23")[2]) < 1e-12
    return dict(ns=ns, dealias="23", invariant="ke", binning="linear1"), \
        dict(u_hat=u, N_hat=N, centers=c, T=T, Pi=Pi, Tnm=Tm)


@case
def enstrophy2d_12_23():
    """2D 12², velocity from a random streamfunction, OrszagTwoThirds, enstrophy."""
    u, ks = O.streamfunction_velocity_2d(12, 103)
    c, T, Pi, N = direct_flux(u, ks, "enstrophy", O.LinearBinning(1.0), "23")
    assert rel(T, O.calculate_spectral_flux(u, ks, O.LinearBinning(1.0), "23", "enstrophy")[1]) < 1e-12
    return dict(ns=(12, 12), dealias="23", invariant="enstrophy", binning="linear1"), \
        dict(u_hat=u, N_hat=N, centers=c, T=T, Pi=Pi)


@case
def scalar2d_16_23():
    """2D 16², solenoidal k^-3 velocity advecting a k^-1 passive scalar, OrszagTwoThirds."""
    ns = (16, 16)
    ks = ks_of(ns)
    u = O.synthetic_field(ns, 2, seed=104, spectrum="k-3")
    th = O.synthetic_field(ns, 1, seed=105, spectrum="k-1", solenoidal=False)
    c, T, Pi, N = direct_flux(th, ks, "scalar", O.LinearBinning(1.0), "23", adv=u)
    assert rel(Pi, O.calculate_scalar_flux(u, th, ks, O.LinearBinning(1.0), "23")[2]) < 1e-12
    Tm = direct_s2s(th, ks, "scalar", O.LinearBinning(1.0), "23", adv=u)
    assert rel(Tm, O.calculate_scalar_shell_to_shell_transfer(
        u, th, ks, binning=O.LinearBinning(1.0), dealiasing="23")[2]) < 1e-12
    return dict(ns=ns, dealias="23", invariant="scalar", binning="linear1"), \
        dict(u_hat=u, theta_hat=th, N_hat=N, centers=c, T=T, Pi=Pi, Tnm=Tm)


@case
def ke3d_12_23():
    """3D 12³, seeded k^-5/3 solenoidal field, OrszagTwoThirds; kinetic energy, helicity and
    enstrophy fluxes from one N̂, the kinetic-energy T(n,m)."""
    ns = (12, 12, 12)
    ks = ks_of(ns)
    u = O.synthetic_field(ns, 3, seed=106)
    c, T, Pi, N = direct_flux(u, ks, "ke", O.LinearBinning(1.0), "23")
    _, sidx = O.shell_setup(ks, O.LinearBinning(1.0))[1:]
    Th = O.shell_sum(O.transfer_density("helicity", u, N, ks), sidx, len(c))
    Te = O.shell_sum(O.transfer_density("enstrophy", u, N, ks), sidx, len(c))
    assert rel(N, O.nonlinear_term_fft(u, ks)) < 1e-12
    Tm = direct_s2s(u, ks, "ke", O.LinearBinning(1.0), "23")
    assert rel(Tm, O.calculate_shell_to_shell_transfer(u, ks, O.LinearBinning(1.0), "23")[2]) < 1e-12
    return dict(ns=ns, dealias="23", invariant="ke", binning="linear1"), \
        dict(u_hat=u, N_hat=N, centers=c, T=T, Pi=Pi, T_helicity=Th, Pi_helicity=np.cumsum(Th),
             T_enstrophy=Te, Pi_enstrophy=np.cumsum(Te), Tnm=Tm)


@case
def ke3d_6_none_curl():
    """3D 6³, û = i k × A with complex Gaussian A (not Hermitian), NoDealiasing."""
    u, ks = O.curl_velocity_3d(6, 107, scale=0.1)
    c, T, Pi, N = direct_flux(u, ks, "ke", O.LinearBinning(1.0), "none")
    assert rel(N, O.nonlinear_term_fft(u, ks, truncate=False)) < 1e-12
    return dict(ns=(6, 6, 6), dealias="none", invariant="ke", binning="linear1"), \
        dict(u_hat=u, N_hat=N, centers=c, T=T, Pi=Pi)


@case
def ke3d_8_padded():
    """3D 8³ (12³ padded grid), seeded solenoidal field with empty Nyquist planes,
    PaddedThreeHalves; N̂ by brute-force alias-free convolution."""
    ns = (8, 8, 8)
    ks = ks_of(ns)
    u = O.synthetic_field(ns, 3, seed=108)
    N = brute_padded(u, ks, u)
    assert rel(N, O.nonlinear_term_padded_fft(u, ks)) < 1e-12
    c, _, sidx = O.shell_setup(ks, O.LinearBinning(1.0))
    T = O.shell_sum(O.transfer_density("ke", u, N, ks), sidx, len(c))
    Th = O.shell_sum(O.transfer_density("helicity", u, N, ks), sidx, len(c))
    return dict(ns=ns, dealias="pad", invariant="ke", binning="linear1"), \
        dict(u_hat=u, N_hat=N, centers=c, T=T, Pi=np.cumsum(T), T_helicity=Th,
             Pi_helicity=np.cumsum(Th))


@case
def ke2d_8_padded_nyquist():
    """2D 8² (12² padded grid), Hermitian field WITH populated Nyquist planes, PaddedThreeHalves:
    the reference's fftshift embedding places the Nyquist derivative at -n/2 (FFTWExt.jl:134-149)."""
    ns = (8, 8)
    ks = ks_of(ns)
    rng = np.random.default_rng(109)
    u = np.zeros(ns + (2,), dtype=np.complex128)
    for c_ in range(2):
        u[..., c_] = np.fft.fft2(rng.standard_normal(ns)) / 64.0
    N = O.nonlinear_term_padded_fft(u, ks)
    c, _, sidx = O.shell_setup(ks, O.LinearBinning(1.0))
    T = O.shell_sum(O.transfer_density("ke", u, N, ks), sidx, len(c))
    return dict(ns=ns, dealias="pad", invariant="ke", binning="linear1"), \
        dict(u_hat=u, N_hat=N, centers=c, T=T, Pi=np.cumsum(T))


@case
def m2m3d_6_23():
    """3D 6³ mode-to-mode S(k|p) (kinetic energy, OrszagTwoThirds) by the reference's one-mode
    giver loop over DirectSum terms; S[k..., p...]."""
    ns = (6, 6, 6)
    ks = ks_of(ns)
    u = O.synthetic_field(ns, 3, seed=110)
    net, S = O.calculate_mode_to_mode_transfer(u, ks, "ke", "23", spectral="DirectSumBackend")
    netc, Sc = O.calculate_mode_to_mode_transfer(u, ks, "ke", "23")
    assert rel(S, Sc) < 1e-12
    Sf = O.mode_to_mode_closed_form(u, ks, "ke", "23")
    assert rel(S, np.asarray(Sf[1] if isinstance(Sf, tuple) else Sf)) < 1e-12
    return dict(ns=ns, dealias="23", invariant="ke"), dict(u_hat=u, S=S, net=net)


@case
def m2m2d_6_scalar_none():
    """2D 6² mode-to-mode S(k|p) of a passive scalar (NoDealiasing), one-mode DirectSum loop."""
    ns = (6, 6)
    ks = ks_of(ns)
    u = O.synthetic_field(ns, 2, seed=111)
    th = O.synthetic_field(ns, 1, seed=112, solenoidal=False)
    net, S = O.calculate_mode_to_mode_transfer(th, ks, "scalar", "none", spectral="DirectSumBackend",
                                               advecting_hat=u)
    netc, Sc = O.calculate_mode_to_mode_transfer(th, ks, "scalar", "none", advecting_hat=u)
    assert rel(S, Sc) < 1e-12
    return dict(ns=ns, dealias="none", invariant="scalar"), dict(u_hat=u, theta_hat=th, S=S, net=net)


@case
def helical3d_8():
    """3D 8³ helical (Craya-Herring) split u = u+ + u- (Decomposition.jl:63-98)."""
    ns = (8, 8, 8)
    ks = ks_of(ns)
    u = O.synthetic_field(ns, 3, seed=113)
    h = O.helical_decompose(u, ks)
    assert rel(h["positive"] + h["negative"], u) < 1e-12
    return dict(ns=ns), dict(u_hat=u, u_plus=h["positive"], u_minus=h["negative"])


def main():
    manifest = {}
    for fn in CASES:
        name = fn.__name__
        meta, arrays = fn()
        for key in ("T", "Tnm", "S"):
            if key in arrays:   # a case whose transfer is roundoff pins nothing
                assert np.max(np.abs(arrays[key])) > 1e-14, (name, key)
        meta = dict(meta, ns=list(meta["ns"]), doc=" ".join(fn.__doc__.split()))
        manifest[name] = dict(meta, arrays=sorted(arrays))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
        print(f"{name}: {os.path.getsize(path)} B", flush=True)
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
        fh.write("\n")


if __name__ == "__main__":
    main()
