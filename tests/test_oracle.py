"""Pins the FP64 oracle before it is trusted as the checker of the HIP path.

Each check is one the reference's own test suite makes (test/runtests.jl line cited), or one the
mathematics fixes independently of the oracle's FFT code path (explicit DFT sums, brute-force
convolutions, per-mode closed forms).  A dropped term, a wrong sign, a transposed operand or a
wrong normalisation in the oracle fails at least one of these.
"""
import math

import numpy as np
import pytest


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ----------------------------------------------------------------------------- grid and masks

def test_wavenumber_grid_known_answer(oracle):
    # runtests.jl:24-41: FFTW order 0,1,2,3,-4,-3,-2,-1 for N=8, L=2π (Nyquist +N/2 per Utils.jl:59)
    ks = oracle.wavenumber_grid((8, 8), (2 * math.pi, 2 * math.pi))
    assert list(ks[0]) == [0, 1, 2, 3, 4, -3, -2, -1]
    km = oracle.wavenumber_magnitude_grid(ks)
    assert km.shape == (8, 8) and km[0, 0] == 0 and np.all(km >= 0)


def test_integer_two_thirds_rule(oracle):
    # NonlinearTerm.jl:134-141: discard iff folded |k_d| >= n_d ÷ 3 (integer division)
    keep = oracle.keep_mask((12, 12))
    assert keep[0, 0] and keep[1, 0] and keep[3, 0] and not keep[4, 0]
    k64 = oracle.keep_mask((64,))
    kept = [i if i <= 32 else i - 64 for i in range(64) if k64[i]]
    assert max(abs(k) for k in kept) == 20                     # 64 ÷ 3 = 21 -> keep |k| <= 20
    k512 = oracle.keep_mask((512,))
    assert int(k512.sum()) == 2 * 169 + 1                      # keep |k| <= 169


def test_padded_len(oracle):
    assert [oracle.padded_len(n) for n in (6, 8, 12, 16, 24, 256, 512, 2048)] == \
        [10, 12, 18, 24, 36, 384, 768, 3072]


# ----------------------------------------------------------------------------- shell binning

def test_shell_binning_known_answers(oracle):
    e = oracle.shell_edges(oracle.LinearBinning(1.0), 5.0)          # runtests.jl:56-64
    assert e[0] == 0.0 and e[-1] >= 5.0
    c = oracle.shell_centers(oracle.LinearBinning(1.0), 5.0)
    assert len(c) == len(e) - 1 and np.all(np.diff(c) > 0)
    el = oracle.shell_edges(oracle.LogarithmicBinning(1.0, 2.0), 16.0)   # :66-72
    assert el[0] == 1.0 and np.allclose(el[1:] / el[:-1], 2.0)
    assert np.array_equal(oracle.shell_edges(oracle.DyadicBinning(1.0), 16.0), el)   # :74-79
    cb = oracle.CustomBinning((0.0, 1.0, 3.0, 6.0, 10.0))                # :81-86
    assert list(oracle.shell_edges(cb, 10.0)) == [0.0, 1.0, 3.0, 6.0, 10.0]
    # 256^3 and 512^3 LinearBinning(1): floor(√3 n/2)+1 shells
    for n, nsh in ((256, 222), (512, 444), (64, 46)):
        nd = 3 if n > 64 else 2
        kmax = math.sqrt(nd) * n / 2
        assert len(oracle.shell_centers(oracle.LinearBinning(1.0), kmax)) == nsh


def test_assign_shells_rule(oracle):
    # runtests.jl:718-735: every assigned mode satisfies edges[n] <= k < edges[n+1]
    ks = oracle.wavenumber_grid((8,), (2 * math.pi,))
    km = oracle.wavenumber_magnitude_grid(ks)
    edges = oracle.shell_edges(oracle.LinearBinning(2 * math.pi / 8), km.max())
    idx = oracle.assign_shells(km, edges)
    assert np.all((idx >= 0) & (idx <= len(edges) - 1)) and np.any(idx == 1)
    for i, n in enumerate(idx):
        if n:
            assert edges[n - 1] <= km[i] < edges[n]
    # the vectorised rule equals the brute force on a 3D grid with a custom partition
    ks3 = oracle.wavenumber_grid((12, 10, 8), (2 * math.pi,) * 3)
    km3 = oracle.wavenumber_magnitude_grid(ks3)
    for e in (np.arange(0.0, 6.0), np.array([0.0, 1.0, 1.0, 2.5, 7.0, 9.0])):
        assert np.array_equal(oracle.assign_shells(km3, e), oracle.assign_shells_fast(km3, e))


# ----------------------------------------------------------------------------- nonlinear term

@pytest.mark.parametrize("ns", [(8,), (12,), (8, 8), (12, 10), (6, 6, 6), (4, 6, 8)])
@pytest.mark.parametrize("dealias", ["none", "23"])
def test_fft_equals_direct_sum(oracle, ns, dealias):
    # runtests.jl:151-172, 194-199, 296-302: DirectSum ≡ FFT.  Non-Hermitian complex input so the
    # real-part semantics (and Nyquist handling) are exercised; the direct path is an explicit DFT.
    rng = np.random.default_rng(sum(ns) + len(dealias))
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    D = max(nd, 2)
    u = rng.standard_normal(ns + (D,)) + 1j * rng.standard_normal(ns + (D,))
    th = rng.standard_normal(ns + (1,)) + 1j * rng.standard_normal(ns + (1,))
    a = oracle.compute_nonlinear_term(u, ks, dealias, "FFTBackend")
    b = oracle.compute_nonlinear_term(u, ks, dealias, "DirectSumBackend")
    assert rel(a, b) < 1e-12
    a = oracle.compute_nonlinear_term(th, ks, dealias, "FFTBackend", advecting_hat=u)
    b = oracle.compute_nonlinear_term(th, ks, dealias, "DirectSumBackend", advecting_hat=u)
    assert rel(a, b) < 1e-12


def test_single_cosine_mode_zero_transfer(oracle):
    # runtests.jl:129-148
    N = 8
    u = np.zeros((N, 1), complex)
    u[1, 0] = 0.5
    u[N - 1, 0] = 0.5
    ks = oracle.wavenumber_grid((N,), (2 * math.pi,))
    _, T, Pi = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "none")
    assert np.all(np.abs(T) < 1e-10) and np.all(np.abs(Pi) < 1e-10)


def test_two_mode_1d_direct_vs_fft_and_flux_sign(oracle):
    # runtests.jl:151-172 and 203-214
    N = 8
    x = np.arange(N) * 2 * math.pi / N
    u = np.cos(x) + 0.3 * np.sin(2 * x) + 0.1 * np.cos(3 * x)
    uh = (np.fft.fft(u) / N).reshape(N, 1)
    ks = oracle.wavenumber_grid((N,), (2 * math.pi,))
    b = oracle.LinearBinning(1.0)
    _, Td, Pd = oracle.calculate_spectral_flux(uh, ks, b, "none", spectral="DirectSumBackend")
    _, Tf, Pf = oracle.calculate_spectral_flux(uh, ks, b, "none")
    assert np.max(np.abs(Td - Tf)) < 1e-10 and np.max(np.abs(Pd - Pf)) < 1e-10
    assert np.max(np.abs(Pf - np.cumsum(Tf))) < 1e-12
    assert np.max(np.abs(Tf)) > 1e-6                     # non-degenerate


def test_two_mode_exact_triad_value(oracle):
    # Independent closed form for u = cos x + a sin 2x (1D, no dealiasing): (u u_x) has
    # Fourier content only at |k| <= 4; T(k=1) = Re{conj(û_1) N̂_1}·2 (±1 modes), with
    # N̂_1 = (1/2)(û_{-1} i2 û_2 + û_2 i(-1) û_{-1}) in the û = fft(u)/N convention.
    N, a = 16, 0.3
    x = np.arange(N) * 2 * math.pi / N
    uh = (np.fft.fft(np.cos(x) + a * np.sin(2 * x)) / N).reshape(N, 1)
    ks = oracle.wavenumber_grid((N,), (2 * math.pi,))
    Nh = oracle.compute_nonlinear_term(uh, ks, "none")
    u1, um1, u2 = uh[1, 0], uh[N - 1, 0], uh[2, 0]
    # (u ∂x u)^(k=1) = Np^-2 Σ_{p+q=1} û_q (i p) û_p  over (q,p) in {(-1,2),(2,-1)}
    # (u = Re ifft(û) is u_true/Np and N̂ = fft(N)/Np: two factors of 1/Np)
    N1 = (um1 * (1j * 2) * u2 + u2 * (1j * -1) * um1) / N ** 2
    assert abs(Nh[1, 0] - N1) < 1e-14


@pytest.mark.parametrize("inv", ["ke", "enstrophy"])
def test_2d_conservation(oracle, inv):
    # runtests.jl:175-200 (energy) and 246-258 (2D enstrophy)
    u, ks = oracle.streamfunction_velocity_2d(16, 7)
    Nh = oracle.compute_nonlinear_term(u, ks, "23")
    t = oracle.transfer_density(inv, u, Nh, ks)
    assert abs(t.sum()) < 1e-10 * np.abs(t).sum()


def test_scalar_variance_conserved_and_flux_returns_to_zero(oracle):
    # runtests.jl:283-307 and 347-368
    rng = np.random.default_rng(31)
    u, ks = oracle.streamfunction_velocity_2d(16, 32)
    th = np.fft.fft2(rng.standard_normal((16, 16))) / 256
    _, T, Pi = oracle.calculate_scalar_flux(u, th, ks, oracle.LinearBinning(1.0), "23")
    sT = math.sqrt(np.sum(T ** 2))
    assert sT > 1e-9 and abs(T.sum()) < 1e-9 * sT and abs(Pi[-1]) < 1e-9 * sT
    _, T1, _ = oracle.calculate_scalar_flux(u, th[..., None], ks, oracle.LinearBinning(1.0), "23")
    assert np.allclose(T1, T)


def test_3d_energy_and_helicity_conservation(oracle):
    # Σ_k T = 0 for energy and helicity on a solenoidal 2/3-dealiased 3D field (both are inviscid
    # invariants); enstrophy is not conserved in 3D (runtests.jl:260-280) but stays finite.
    u = oracle.synthetic_field((12, 12, 12), 3, seed=3)
    ks = oracle.wavenumber_grid((12, 12, 12), (2 * math.pi,) * 3)
    Nh = oracle.compute_nonlinear_term(u, ks, "23")
    for inv in ("ke", "helicity"):
        t = oracle.transfer_density(inv, u, Nh, ks)
        assert abs(t.sum()) < 1e-10 * np.abs(t).sum()
    te = oracle.transfer_density("enstrophy", u, Nh, ks)
    assert np.all(np.isfinite(te))
    with pytest.raises(ValueError):
        oracle.transfer_density("helicity", u[:, :, 0, :2], Nh[:, :, 0, :2], ks[:2])


def test_effective_carrier_matches_densities(oracle):
    rng = np.random.default_rng(4)
    for ns in ((6, 6, 6), (8, 6)):
        nd = len(ns)
        ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
        u = rng.standard_normal(ns + (nd,)) + 1j * rng.standard_normal(ns + (nd,))
        N = rng.standard_normal(ns + (nd,)) + 1j * rng.standard_normal(ns + (nd,))
        invs = ("ke", "helicity", "enstrophy") if nd == 3 else ("ke", "enstrophy")
        for inv in invs:
            e = oracle.effective_carrier(inv, u, ks)
            t = np.sum(np.real(np.conj(e) * N), axis=-1)
            assert rel(t, oracle.transfer_density(inv, u, N, ks)) < 1e-13


# ----------------------------------------------------------------------------- padding

def _band_limited_2d(oracle, N, kcut, seed):
    # runtests.jl:313-327: Hermitian streamfunction band-limited to 0<|k|<kcut
    rng = np.random.default_rng(seed)
    ks = oracle.wavenumber_grid((N, N), (2 * math.pi,) * 2)
    kx, ky = np.meshgrid(*ks, indexing="ij")
    km = np.sqrt(kx ** 2 + ky ** 2)
    psi = np.zeros((N, N), complex)
    sel = (km > 0) & (km < kcut)
    psi[sel] = (rng.standard_normal(sel.sum()) + 1j * rng.standard_normal(sel.sum())) / km[sel] ** 2
    for i in range(N):
        for j in range(N):
            ci, cj = (-i) % N, (-j) % N
            if (ci, cj) > (i, j):
                psi[ci, cj] = np.conj(psi[i, j])
    for i in range(N):
        for j in range(N):
            if ((-i) % N, (-j) % N) == (i, j):
                psi[i, j] = psi[i, j].real
    u = np.stack([1j * ky * psi, -1j * kx * psi], axis=-1)
    return u, ks, km


def test_padded_matches_two_thirds_on_retained_band(oracle):
    # runtests.jl:310-344
    u, ks, km = _band_limited_2d(oracle, 24, 8, 81)
    N23 = oracle.compute_nonlinear_term(u, ks, "23")
    Npad = oracle.compute_nonlinear_term(u, ks, "pad")
    low = np.repeat((km < 8)[..., None], 2, axis=-1)
    mid = np.repeat(((km >= 8) & (km < 12))[..., None], 2, axis=-1)
    assert np.max(np.abs(N23 - Npad)[low]) < 1e-10 * np.max(np.abs(Npad)[low])
    assert np.sum(np.abs(Npad[mid]) ** 2) > 3 * np.sum(np.abs(N23[mid]) ** 2)
    _, T, _ = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "pad")
    assert abs(T.sum()) < 1e-9 * np.abs(T).sum()
    with pytest.raises(ValueError):
        oracle.compute_nonlinear_term(u, ks, "pad", "DirectSumBackend")


def test_padded_equals_exact_convolution(oracle):
    # Brute force: for a Hermitian field with no Nyquist content, exact (alias-free) padding gives
    # N̂(k) = Np^-2 Σ_{p+q=k, integer vectors, no wrap} Σ_j â_j(q) (i p_j) f̂(p), for every
    # -N/2 <= k_d < N/2 (the truncation returns padded frequency -N/2 into the Nyquist bin).
    N = 8
    u, ks, km = _band_limited_2d(oracle, N, 4, 5)
    Npad = oracle.compute_nonlinear_term(u, ks, "pad")
    freqs = [i if i < N // 2 else i - N for i in range(N)]
    ref = np.zeros_like(Npad)
    for i1 in range(N):
        for j1 in range(N):
            p = (freqs[i1], freqs[j1])
            for i2 in range(N):
                for j2 in range(N):
                    q = (freqs[i2], freqs[j2])
                    k = (p[0] + q[0], p[1] + q[1])
                    if not all(-N // 2 <= kk < N // 2 for kk in k):
                        continue
                    kk = (k[0] % N, k[1] % N)
                    adv = u[i2, j2, 0] * 1j * p[0] + u[i2, j2, 1] * 1j * p[1]
                    ref[kk[0], kk[1], :] += adv * u[i1, j1, :]
    ref /= float(N * N) ** 2
    assert rel(Npad, ref) < 1e-12


# ----------------------------------------------------------------------------- shell-to-shell

def test_shell_to_shell_properties(oracle):
    # runtests.jl:371-438: antisymmetry, direct ≡ FFT, Σ_m T(n,m) = T(k)
    u, ks = oracle.streamfunction_velocity_2d(16, 13)
    b = oracle.LinearBinning(1.0)
    _, _, Tf, netf, asym = oracle.calculate_shell_to_shell_transfer(u, ks, b, "23")
    _, _, Td, _, asymd = oracle.calculate_shell_to_shell_transfer(
        u, ks, b, "23", spectral="DirectSumBackend")
    nrm = math.sqrt(np.sum(Td ** 2))
    assert nrm > 0 and np.max(np.abs(Tf - Td)) < 1e-9 * nrm
    assert asym < 1e-12 * nrm and asymd < 1e-10 * nrm
    _, T, _ = oracle.calculate_spectral_flux(u, ks, b, "23")
    assert np.max(np.abs(netf - T)) < 1e-9 * nrm


def test_scalar_shell_to_shell_properties(oracle):
    # runtests.jl:441-464
    rng = np.random.default_rng(15)
    u, ks = oracle.streamfunction_velocity_2d(16, 14)
    th = np.fft.fft2(rng.standard_normal((16, 16))) / 256
    b = oracle.LinearBinning(1.0)
    _, _, Tf, netf, _ = oracle.calculate_scalar_shell_to_shell_transfer(u, th, ks, binning=b)
    _, _, Td, _, asym = oracle.calculate_scalar_shell_to_shell_transfer(
        u, th, ks, binning=b, spectral="DirectSumBackend")
    nrm = math.sqrt(np.sum(Td ** 2))
    assert nrm > 1e-12 and np.max(np.abs(Tf - Td)) < 1e-9 * nrm and asym < 1e-10 * nrm
    _, T, _ = oracle.calculate_scalar_flux(u, th, ks, b, "23")
    assert np.max(np.abs(netf - T)) < 1e-9 * nrm


def test_padded_shell_to_shell_reduces(oracle):
    u, ks, _ = _band_limited_2d(oracle, 12, 5, 9)
    b = oracle.LinearBinning(1.0)
    _, _, T, net, asym = oracle.calculate_shell_to_shell_transfer(u, ks, b, "pad")
    _, Tk, _ = oracle.calculate_spectral_flux(u, ks, b, "pad")
    nrm = math.sqrt(np.sum(T ** 2))
    assert nrm > 0 and np.max(np.abs(net - Tk)) < 1e-10 * nrm and asym < 1e-10 * nrm


def _pair_density_half_grid(N, e, sh, pad):
    """Shell sums of t = Re Σ_c conj(e_c) N̂_c from the original half grid x <= n1/2 only, as the
    device pair density forms them: t(k) alone on the planes x = 0 and x = n1/2 (−k is a half-grid
    bin there), t(k) + t(−k) elsewhere with N̂(−k) = conj N̂(k), except on the Nyquist rows of an
    even y / z axis of a padded grid, where N̂(−k) is its own stored bin."""
    n1, n2, n3 = N.shape[:3]
    out = np.zeros(sh.max() + 1)
    for ix in range(n1 // 2 + 1):
        mx = (n1 - ix) % n1
        for iy in range(n2):
            my = (n2 - iy) % n2
            for iz in range(n3):
                mz = (n3 - iz) % n3
                ek, em, nk = e[ix, iy, iz], e[mx, my, mz], N[ix, iy, iz]
                own = pad and ((n2 % 2 == 0 and 2 * iy == n2) or (n3 % 2 == 0 and 2 * iz == n3))
                if ix == 0 or 2 * ix == n1:
                    t = np.sum(np.real(np.conj(ek) * nk))
                elif own:
                    t = np.sum(np.real(np.conj(ek) * nk)) + np.sum(np.real(np.conj(em) * N[mx, my, mz]))
                else:
                    t = np.sum((ek.real + em.real) * nk.real + (ek.imag - em.imag) * nk.imag)
                out[sh[ix, iy, iz]] += t
    return out


@pytest.mark.parametrize("ns,dea,inv", [((12, 12), "pad", "ke"), ((10, 10), "23", "ke"), ((8, 8, 8), "pad", "ke"),
                                        ((8, 8, 8), "pad", "helicity"), ((6, 8, 10), "pad", "ke"),
                                        ((7, 9, 8), "pad", "helicity"), ((9, 8, 7), "none", "ke")])
def test_half_grid_pair_density_equals_full_grid(oracle, ns, dea, inv):
    # the flux's device density reads N̂ on the half grid only; its shell sums equal the full-grid
    # density's for symmetric shells, including the padded Nyquist rows where N̂(−k) is not
    # conj N̂(k) (the padded grid holds both at frequency −n/2); the plain pair formula there fails
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    if nd == 3:
        # a real random field: content at every bin, the Nyquist rows included
        rng = np.random.default_rng(3)
        u = np.fft.fftn(rng.standard_normal(ns + (3,)), axes=(0, 1, 2)) / np.prod(ns)
    else:
        u, _ = oracle.streamfunction_velocity_2d(ns[0], 3)
    N = oracle.compute_nonlinear_term(u, ks, dea)
    K = np.meshgrid(*ks, indexing="ij")
    kz = K[2] if nd == 3 else np.zeros_like(K[0])
    if inv == "ke":
        e = u
    else:
        ux, uy, uz = u[..., 0], u[..., 1], u[..., 2]
        e = 1j * np.stack([uz * K[1] - uy * kz, ux * kz - uz * K[0], uy * K[0] - ux * K[1]], -1)
    sh = np.floor(np.sqrt(K[0] ** 2 + K[1] ** 2 + kz ** 2) + 0.5).astype(int)
    full = np.bincount(sh.ravel(), weights=np.real(np.sum(np.conj(e) * N, -1)).ravel(), minlength=sh.max() + 1)
    if nd == 2:
        N, e, sh = N[:, :, None, :], e[:, :, None, :], sh[:, :, None]
    half = _pair_density_half_grid(N, e, sh, dea == "pad")
    assert rel(half, full) < 1e-12
    if dea == "pad" and (ns[1] % 2 == 0 or (nd == 3 and ns[2] % 2 == 0)):
        # without the Nyquist-row rule the half-grid sums are wrong
        plain = _pair_density_half_grid(N, e, sh, False)
        assert rel(plain, full) > 1e-6


# ----------------------------------------------------------------------------- mode-to-mode

@pytest.mark.parametrize("case", ["2d_ke_23", "2d_ke_none", "2d_ens", "3d_ke", "3d_hel", "3d_ens",
                                  "2d_scalar"])
def test_mode_to_mode_loop_equals_closed_form(oracle, case):
    # The reference's giver-mode loop (FFT based) against the independent index-arithmetic form;
    # non-Hermitian primary field, Hermitian advecting field.
    rng = np.random.default_rng(len(case))
    if case.startswith("2d"):
        ns = (6, 6)
        u, ks = oracle.streamfunction_velocity_2d(6, 3)
    else:
        ns = (4, 4, 4)
        u = oracle.synthetic_field(ns, 3, seed=2)
        ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 3)
    dea = "none" if case.endswith("none") else "23"
    if ns == (4, 4, 4):
        dea = "none"                   # 4 ÷ 3 = 1 would keep only the DC mode
    inv = {"ke": "ke", "hel": "helicity", "ens": "enstrophy", "scalar": "scalar"}[
        case.split("_")[1]]
    prim = u
    adv = None
    if inv == "scalar":
        prim = rng.standard_normal(ns + (1,)) + 1j * rng.standard_normal(ns + (1,))
        adv = u
    else:
        prim = u + 0.3 * (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape))
        adv = u
    net, S = oracle.calculate_mode_to_mode_transfer(prim, ks, inv, dea, force=True,
                                                    advecting_hat=adv)
    Sc = oracle.mode_to_mode_closed_form(prim, ks, inv, dea, advecting_hat=adv)
    assert rel(S, Sc) < 1e-12


@pytest.mark.parametrize("ns,inv", [((6, 6), "ke"), ((8, 8), "ke"), ((8, 8), "enstrophy"),
                                    ((5, 7), "ke"), ((4, 4, 4), "ke"), ((4, 4, 4), "helicity"),
                                    ((6,), "ke")])
def test_padded_mode_to_mode_loop_equals_closed_form(oracle, ns, inv):
    # PaddedThreeHalves: the reference's giver loop through the padded nonlinear term (embed,
    # padded transforms, truncate) against the index-arithmetic sum on the padded grid.  The
    # primary field is non-Hermitian with Nyquist content; (8, 8) pads to exactly 3n/2 = 12, where
    # k + p of two Nyquist bins lands on the opposite Nyquist bin of the padded grid.
    rng = np.random.default_rng(sum(ns) + len(inv))
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    prim = rng.standard_normal(ns + (nd,)) + 1j * rng.standard_normal(ns + (nd,))
    adv = (oracle.synthetic_field(ns, nd, seed=3) if nd > 1
           else rng.standard_normal(ns + (1,)) + 1j * rng.standard_normal(ns + (1,)))
    net, S = oracle.calculate_mode_to_mode_transfer(prim, ks, inv, "pad", force=True,
                                                    advecting_hat=adv)
    Sc = oracle.mode_to_mode_closed_form(prim, ks, inv, "pad", advecting_hat=adv)
    assert rel(S, Sc) < 1e-12
    S23 = oracle.mode_to_mode_closed_form(prim, ks, inv, "none", advecting_hat=adv)
    assert rel(S, S23) > 1e-3          # padding changes the triads (not the unpadded form)


def test_mode_to_mode_reductions(oracle):
    # runtests.jl:1072-1111: antisymmetry, conservation, Σ_p S -> T(k), shell sums -> T(n,m)
    u, ks = oracle.streamfunction_velocity_2d(12, 21)
    net, S = oracle.calculate_mode_to_mode_transfer(u, ks, "ke", "23", force=True)
    Nm = 144
    S2 = S.reshape(Nm, Nm)
    nrm = math.sqrt(np.sum(S ** 2))
    assert nrm > 0 and np.max(np.abs(S2 + S2.T)) < 1e-10 * nrm and abs(S.sum()) < 1e-10 * nrm
    b = oracle.LinearBinning(1.0)
    centers, edges, sidx = oracle.shell_setup(ks, b)
    _, T, _ = oracle.calculate_spectral_flux(u, ks, b, "23")
    netshell = oracle.shell_sum(net, sidx, len(centers))
    assert np.max(np.abs(netshell - T)) < 1e-9 * math.sqrt(np.sum(T ** 2))
    _, _, Tnm, _, _ = oracle.calculate_shell_to_shell_transfer(u, ks, b, "23")
    s = sidx.reshape(-1)
    TKQ = np.zeros_like(Tnm)
    for i in range(Nm):
        for j in range(Nm):
            if s[i] and s[j]:
                TKQ[s[i] - 1, s[j] - 1] += S2[i, j]
    assert np.max(np.abs(TKQ - Tnm)) < 1e-9 * math.sqrt(np.sum(Tnm ** 2))
    with pytest.raises(ValueError):                                    # ScaleToScale.jl:64-67
        oracle.calculate_mode_to_mode_transfer(u, ks, max_modes=100)


# ----------------------------------------------------------------------------- helical

def test_helical_decomposition_properties(oracle):
    # runtests.jl:540-578
    u, ks = oracle.curl_velocity_3d(8, 51)
    dec = oracle.helical_decompose(u, ks)
    up, um = dec["positive"], dec["negative"]
    assert np.max(np.abs(up + um - u)) < 1e-12 * np.max(np.abs(u))
    assert abs(np.sum(np.abs(up) ** 2) + np.sum(np.abs(um) ** 2) - np.sum(np.abs(u) ** 2)) \
        < 1e-12 * np.sum(np.abs(u) ** 2)
    kx, ky, kz = np.meshgrid(*ks, indexing="ij")
    for w in (up, um):
        assert np.max(np.abs(kx * w[..., 0] + ky * w[..., 1] + kz * w[..., 2])) < 1e-10 * np.max(np.abs(u))
    wv = oracle.vorticity(u, ks)
    Hd = sum(np.sum(np.real(np.conj(u[..., c]) * wv[c])) for c in range(3))
    km = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    Hh = np.sum(km * (np.sum(np.abs(up) ** 2, -1) - np.sum(np.abs(um) ** 2, -1)))
    assert abs(Hd - Hh) < 1e-10 * abs(Hd)


def test_helical_fluxes_sum(oracle):
    # runtests.jl:540-604: Π⁺+Π⁻ = Π and the 8 channels sum to the total
    u, ks = oracle.curl_velocity_3d(16, 91)
    b = oracle.LinearBinning(1.0)
    _, T, _ = oracle.calculate_spectral_flux(u, ks, b, "23")
    hel = oracle.calculate_spectral_flux(u, ks, b, "23", decomposition="HelicalDecomposition")
    assert np.max(np.abs(hel["positive"][1] + hel["negative"][1] - T)) < 1e-9 * math.sqrt(np.sum(T ** 2))
    ch, tot, _ = oracle.calculate_partial_fluxes(u, ks, b, "23")
    assert len(ch) == 8
    scale = max(math.sqrt(np.sum(v[1] ** 2)) for v in ch.values())
    assert np.max(np.abs(tot[1] - T)) < 1e-10 * scale


# ----------------------------------------------------------------------------- toroidal / poloidal

def test_toroidal_poloidal_properties(oracle):
    # runtests.jl:634-660: û = i k × Â (solenoidal); tor + pol = û, orthogonal, tor has no vertical
    # velocity, both divergence-free
    M = 8
    ks = oracle.wavenumber_grid((M, M, M), (2 * math.pi,) * 3)
    kx, ky, kz = np.meshgrid(*ks, indexing="ij")
    rng = np.random.default_rng(52)
    A = rng.standard_normal((M, M, M, 3)) + 1j * rng.standard_normal((M, M, M, 3))
    u = np.stack([1j * (ky * A[..., 2] - kz * A[..., 1]), 1j * (kz * A[..., 0] - kx * A[..., 2]),
                  1j * (kx * A[..., 1] - ky * A[..., 0])], axis=-1)
    dec = oracle.decompose("ToroidalPoloidalDecomposition", u, ks)
    tor, pol = dec["toroidal"], dec["poloidal"]
    assert list(dec) == ["toroidal", "poloidal"]
    um = np.max(np.abs(u))
    assert np.max(np.abs(tor + pol - u)) < 1e-12 * um
    assert abs(np.sum(np.abs(tor) ** 2) + np.sum(np.abs(pol) ** 2) - np.sum(np.abs(u) ** 2)) < 1e-12 * np.sum(np.abs(u) ** 2)
    assert np.max(np.abs(tor[..., 2])) < 1e-12 * um
    for w in (tor, pol):
        assert np.max(np.abs(kx * w[..., 0] + ky * w[..., 1] + kz * w[..., 2])) < 1e-10 * um
    # the toroidal part carries all of the vertical vorticity ω_z = i(k_x u_y - k_y u_x)
    wz = lambda w: 1j * (kx * w[..., 1] - ky * w[..., 0])
    assert np.max(np.abs(wz(pol))) < 1e-10 * np.max(np.abs(wz(u)))


def test_toroidal_poloidal_known_answers(oracle):
    # k = (1, 0, 0): e1 = k̂ × ẑ = -ŷ, e2 = k̂ × e1 = -ẑ -> tor = (0, u_y, 0), pol = (0, 0, u_z);
    # k = (0, 0, 1) (k_⊥ = 0): (x̂, ŷ) -> tor = (u_x, 0, 0), pol = (0, u_y, 0); k = 0 -> both zero;
    # k = (1, 1, 0): e1 = (1, -1, 0)/√2, e2 = -ẑ
    M = 4
    ks = oracle.wavenumber_grid((M, M, M), (2 * math.pi,) * 3)
    u = np.zeros((M, M, M, 4), dtype=np.complex128)
    vals = {(1, 0, 0): (0.3, 1 + 2j, -0.5j), (0, 0, 1): (2 - 1j, 0.25, 7.0), (0, 0, 0): (1, 1, 1),
            (1, 1, 0): (1.5, 0.5j, 2.0)}
    for idx, v in vals.items():
        u[idx + (slice(0, 3),)] = v
        u[idx + (3,)] = 9.0                     # a fourth component: zero in both parts
    d = oracle.toroidal_poloidal_decompose(u, ks)
    t, p = d["toroidal"], d["poloidal"]
    assert np.allclose(t[1, 0, 0], [0, 1 + 2j, 0, 0]) and np.allclose(p[1, 0, 0], [0, 0, -0.5j, 0])
    assert np.allclose(t[0, 0, 1], [2 - 1j, 0, 0, 0]) and np.allclose(p[0, 0, 1], [0, 0.25, 0, 0])
    assert np.allclose(t[0, 0, 0], 0) and np.allclose(p[0, 0, 0], 0)
    c1 = (1.5 - 0.5j) / math.sqrt(2)
    assert np.allclose(t[1, 1, 0], [c1 / math.sqrt(2), -c1 / math.sqrt(2), 0, 0])
    assert np.allclose(p[1, 1, 0], [0, 0, 2.0, 0])


def test_toroidal_poloidal_partial_fluxes_sum(oracle):
    # SpectralFlux.jl:240-285 with the toroidal/poloidal split: 8 channels summing to the total
    # kinetic-energy flux of the (solenoidal) field; per-component fluxes sum to it as well
    u, ks = oracle.curl_velocity_3d(16, 93)
    b = oracle.LinearBinning(1.0)
    _, T, _ = oracle.calculate_spectral_flux(u, ks, b, "23")
    ch, tot, _ = oracle.calculate_partial_fluxes(u, ks, b, "23", decomposition="ToroidalPoloidalDecomposition")
    assert len(ch) == 8 and set(k[0] for k in ch) == {"toroidal", "poloidal"}
    scale = max(math.sqrt(np.sum(v[1] ** 2)) for v in ch.values())
    assert np.max(np.abs(tot[1] - T)) < 1e-10 * scale
    per = oracle.calculate_spectral_flux(u, ks, b, "23", decomposition="ToroidalPoloidalDecomposition")
    assert np.max(np.abs(per["toroidal"][1] + per["poloidal"][1] - T)) < 1e-9 * math.sqrt(np.sum(T ** 2))


def test_synthetic_field_properties(oracle):
    u = oracle.synthetic_field((16, 12, 8), 3, seed=1)
    ks = oracle.wavenumber_grid((16, 12, 8), (2 * math.pi,) * 3)
    kx, ky, kz = np.meshgrid(*ks, indexing="ij")
    assert abs(np.sum(np.abs(u) ** 2) - 1.0) < 1e-12
    assert np.max(np.abs(kx * u[..., 0] + ky * u[..., 1] + kz * u[..., 2])) < 1e-13
    # Hermitian: u(-k) = conj u(k)
    neg = np.roll(u[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))
    assert np.max(np.abs(neg - np.conj(u))) < 1e-15


# ----------------------------------------------------------------------------- band-to-band

def test_band_weights_partition_of_unity(oracle):
    ks = oracle.wavenumber_grid((12, 10), (2 * math.pi,) * 2)
    kc = oracle.shell_coordinate(None, ks)
    W = oracle.band_weights((1.0, 2.0, 4.0), 0.5, kc)
    s = W.sum(axis=0)
    assert np.allclose(s[kc > 0], 1.0, atol=1e-14) and np.all(W[:, kc == 0] == 0)
    # a mode at a band centre carries that band's largest share
    i = np.argmin(np.abs(kc - 2.0).reshape(-1) + (kc == 0).reshape(-1))
    assert np.argmax(W.reshape(3, -1)[:, i]) == 1


@pytest.mark.parametrize("dea,inv", [("23", "ke"), ("none", "ke"), ("23", "enstrophy")])
def test_band_to_band_equals_w_s_wt(oracle, dea, inv):
    """T(n,m) = Σ_k Σ_p w_n(k) w_m(p) S(k|p): N̂ is linear in the advected field, so the band
    transfer is the band-weighted mode-to-mode matrix, S from the index-arithmetic closed form."""
    ns = (12, 12)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 2)
    u, _ = oracle.streamfunction_velocity_2d(12, 31)
    bands = oracle.SmoothBands((1.0, 2.0, 3.0), 0.5)
    _, T, net, asym = oracle.calculate_band_to_band_transfer(u, ks, bands, dea, inv)
    assert np.max(np.abs(T)) > 1e-9
    S = oracle.mode_to_mode_closed_form(u, ks, inv, dea)
    W = oracle.band_weights(bands.centers, bands.logwidth, oracle.shell_coordinate(None, ks))
    Np = 144
    S2 = S.reshape(Np, Np, order="F")          # [k, p], both column-major linear indices
    Wf = W.reshape(3, -1, order="F")
    ref = Wf @ S2 @ Wf.T
    assert rel(T, ref) < 1e-12
    assert rel(net, T.sum(axis=1)) < 1e-14


def test_band_to_band_direct_sum_and_properties(oracle):
    """runtests.jl:661-681: antisymmetric, conserves, net = Σ_m T(n,m), Σ net = 0; the FFT path
    equals the DirectSum twin."""
    ns = (16, 16)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 2)
    u, _ = oracle.streamfunction_velocity_2d(16, 71)
    bands = oracle.SmoothBands((1.0, 2.0, 3.0, 4.0), 0.5)
    _, T, net, asym = oracle.calculate_band_to_band_transfer(u, ks, bands, "23")
    Tn = np.sqrt(np.sum(T ** 2))
    assert Tn > 0 and asym < 1e-10 * Tn and abs(T.sum()) < 1e-10 * Tn
    assert abs(net.sum()) < 1e-10 * Tn
    _, Td, _, _ = oracle.calculate_band_to_band_transfer(u, ks, bands, "23", spectral="DirectSumBackend")
    assert rel(T, Td) < 1e-12
