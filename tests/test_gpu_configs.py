"""Full-size GPU checks of the benchmark configurations, through size-independent properties and
an FP64 device run of the same kernels (itself pinned to the oracle on small grids):

  C2  256³ kinetic-energy T(k), Π(K) and the 32-shell T(n,m) (2/3), FP32 vs FP64
  C5  32³ mode-to-mode S(k|p) over all 32768² pairs: antisymmetry, Σ_p S = T(k), shell sums =
      T(n,m), and columns against the oracle's closed form
  C5' 256³ helical 8-channel partial fluxes: the channels sum to the total flux
  C4  an uneven 2D ensemble (scalar + kinetic-energy fluxes, results in input order)
  C1  64² FP64 Π(K) against the oracle
  C2  256³ FP32 Π(K) and T(n,m) columns against the FP64 oracle
  C3  512³ KE + helicity Π(K) with 3/2 padding and the 48-shell T(n,m)
"""
import math
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def nerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def device_field(oracle, ns, D, seed, spectrum="k-5/3"):
    import torch
    u = oracle.synthetic_field(ns, D, seed=seed, spectrum=spectrum)
    rev = tuple(reversed(range(u.ndim)))
    return torch.from_numpy(np.ascontiguousarray(u.transpose(rev))).to("cuda").permute(rev), u


def test_c2_256_flux_and_shell_to_shell(fit, oracle):
    import torch
    n = 256
    say("case C2", n)
    ud, _ = device_field(oracle, (n,) * 3, 3, seed=2)
    ks = fit.wavenumber_grid((n,) * 3, (2 * math.pi,) * 3)
    u32 = ud.to(torch.complex64)
    t0 = time.perf_counter()
    r32 = fit.calculate_spectral_flux(u32, ks, binning=fit.LinearBinning(1.0))
    s32 = fit.calculate_shell_to_shell_transfer(u32, ks, binning=fit.CustomBinning(range(0, 33)))
    torch.cuda.synchronize()
    say("  FP32 Π + 32-giver T(n,m) s:", time.perf_counter() - t0)
    r64 = fit.calculate_spectral_flux(ud, ks, binning=fit.LinearBinning(1.0))
    s64 = fit.calculate_shell_to_shell_transfer(ud, ks, binning=fit.CustomBinning(range(0, 33)))
    assert len(r32.flux) == 222
    eT, eP, eS = nerr(r32.transfer_spectrum, r64.transfer_spectrum), nerr(r32.flux, r64.flux), \
        nerr(s32.transfer_matrix, s64.transfer_matrix)
    say("  errors T, Π, T(n,m):", eT, eP, eS)
    assert eT < 1e-5 and eP < 1e-6 and eS < 1e-5
    # conservation (FP64): Σ_k T = 0; antisymmetry of T(n,m); rows reduce to T(k) on the block
    T = r64.transfer_spectrum
    assert abs(T.sum()) < 1e-9 * np.abs(T).sum()
    nrm = np.sqrt(np.sum(s64.transfer_matrix ** 2))
    assert s64.max_antisymmetry_error < 1e-12 * nrm


def test_c5_mode_to_mode_32(fit, oracle):
    import torch
    n = 32
    say("case C5", n)
    ud, u = device_field(oracle, (n,) * 3, 3, seed=5)
    ks = fit.wavenumber_grid((n,) * 3, (2 * math.pi,) * 3)
    Np = n ** 3
    t0 = time.perf_counter()
    r = fit.calculate_mode_to_mode_transfer(ud.to(torch.complex64), ks, force=True)
    torch.cuda.synchronize()
    say("  FP32 S(k|p) over", Np * Np, "pairs, s:", time.perf_counter() - t0)
    S = r.transfer                                            # device (n,n,n,n,n,n) FP32
    S2 = S.permute(5, 4, 3, 2, 1, 0).reshape(Np, Np)           # [p, k] in memory order
    # columns against the oracle's closed form (3 givers)
    for p_lin in (1, 777, Np - 5):
        p = (p_lin % n, (p_lin // n) % n, p_lin // (n * n))
        col = S2[p_lin].double().cpu().numpy().reshape((n, n, n), order="F")
        ref = closed_form_column(oracle, u, ks, p)
        assert nerr(col, ref) < 2e-5
    # antisymmetry S(k|p) = -S(p|k)
    asym = (S2 + S2.t()).abs().max().item()
    smax = S2.abs().max().item()
    say("  antisymmetry", asym / smax)
    assert asym < 1e-5 * smax
    # Σ_p S(k|p) shell-summed == T(k) (FP64 device flux) ; shell sums of S == T(n,m)
    net = r.net_transfer.double().cpu().numpy()
    from fit_amd.grid import shell_setup
    c, e, sidx = shell_setup(ks, fit.LinearBinning(1.0), None)
    T = fit.calculate_spectral_flux(ud, ks, binning=fit.LinearBinning(1.0)).transfer_spectrum
    netshell = oracle.shell_sum(net, sidx, len(c))
    assert nerr(netshell, T) < 1e-5
    s64 = fit.calculate_shell_to_shell_transfer(ud, ks, binning=fit.LinearBinning(1.0))
    nsh = len(c)
    lab = torch.from_numpy(sidx.reshape(-1, order="F")).to("cuda")
    onehot = torch.zeros((nsh, Np), dtype=torch.float64, device="cuda")
    keep = lab > 0
    onehot[(lab[keep] - 1), torch.nonzero(keep).squeeze(1)] = 1.0
    TKQ = onehot @ S2.t().double() @ onehot.t()              # [n (receiver), m (giver)]
    say("  shell-sum err", nerr(TKQ.cpu().numpy(), s64.transfer_matrix))
    assert nerr(TKQ.cpu().numpy(), s64.transfer_matrix) < 1e-5


def closed_form_column(oracle, u, ks, p):
    """S(·|p) for one giver from the oracle's closed form (same formula, one column)."""
    nd = 3
    ns = u.shape[:3]
    Np = float(np.prod(ns))
    keep = oracle.keep_mask(ns)
    if not keep[p]:
        return np.zeros(ns)
    A = []
    for j in range(nd):
        g = keep * u[..., j]
        neg = np.roll(g[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))
        A.append(0.5 * (g + np.conj(neg)))
    kidx = np.meshgrid(*[np.arange(n) for n in ns], indexing="ij")
    pk = [ks[j][p[j]] for j in range(nd)]
    kmp = tuple((kidx[d] - p[d]) % ns[d] for d in range(nd))
    kpp = tuple((kidx[d] + p[d]) % ns[d] for d in range(nd))
    dm = sum(pk[j] * A[j][kmp] for j in range(nd))
    dp = sum(pk[j] * A[j][kpp] for j in range(nd))
    tot = np.zeros(ns)
    for c in range(3):
        fp = u[p + (c,)]
        Nc = keep / (2 * Np * Np) * (1j * dm * fp - 1j * dp * np.conj(fp))
        tot += np.real(np.conj(u[..., c]) * Nc)
    return tot


def test_c5_helical_channels_256(fit, oracle):
    import torch
    n = 256
    say("case C5 helical", n)
    ud, _ = device_field(oracle, (n,) * 3, 3, seed=6)
    ks = fit.wavenumber_grid((n,) * 3, (2 * math.pi,) * 3)
    u32 = ud.to(torch.complex64)
    t0 = time.perf_counter()
    hp = fit.calculate_helical_partial_fluxes(u32, ks, binning=fit.LinearBinning(1.0))
    torch.cuda.synchronize()
    say("  8 channels s:", time.perf_counter() - t0)
    tot = fit.calculate_spectral_flux(ud, ks, binning=fit.LinearBinning(1.0))
    scale = max(np.sqrt(np.sum(np.asarray(c.transfer_spectrum, np.float64) ** 2)) for c in hp["channels"].values())
    err = np.max(np.abs(hp["total"].transfer_spectrum - tot.transfer_spectrum)) / scale
    say("  channel-sum err", err)
    assert len(hp["channels"]) == 8 and err < 1e-5


def device_field_2d(n, D, seed, power, solenoidal):
    """Seeded Hermitian 2D field built on the device: complex Gaussian g, (g(k) + conj g(-k))/2,
    per-mode |û|² ∝ k^power, Leray projection when solenoidal, DC and Nyquist lines zeroed,
    Σ|û|² = 1; (n, n, D) column-major complex128 view."""
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    g = torch.complex(torch.randn((D, n, n), generator=gen, device="cuda", dtype=torch.float64),
                      torch.randn((D, n, n), generator=gen, device="cuda", dtype=torch.float64))
    u = 0.5 * (g + torch.roll(torch.flip(g, dims=(1, 2)), shifts=(1, 1), dims=(1, 2)).conj())
    i = torch.arange(n, device="cuda")
    k1 = torch.where(i <= n // 2, i, i - n).to(torch.float64)
    kx, ky = k1.view(1, n), k1.view(n, 1)
    kk = torch.sqrt(kx * kx + ky * ky)
    ks = torch.where(kk > 0, kk, torch.ones_like(kk))
    amp = torch.where(kk > 0, ks ** (0.5 * power), torch.zeros_like(kk))
    u = torch.where(u.abs() > 0, u / torch.where(u.abs() > 0, u.abs(), torch.ones_like(u.abs())), u) * amp
    if solenoidal:
        div = kx * u[0] + ky * u[1]
        u = torch.stack([u[0] - kx * div / (ks * ks), u[1] - ky * div / (ks * ks)])
    u[:, :, n // 2] = 0
    u[:, n // 2, :] = 0
    u[:, 0, 0] = 0
    u = u / torch.sqrt(torch.sum(u.abs() ** 2))
    return u.permute(2, 1, 0)


def test_c4_uneven_ensemble(fit, oracle):
    """The C4 ensemble at full size: 64 2D snapshots with sizes drawn by seed 4 from
    {256, 512, 1024, 2048}², k^-3 velocity + k^-1 scalar, FP32 batched graph vs an FP64 run of the
    same ensemble (itself pinned to the oracle here on the 256² snapshots), conservation of both
    fluxes, input order kept."""
    import torch
    say("case C4")
    rng = np.random.default_rng(4)
    sizes = [int(n) for n in rng.choice([256, 512, 1024, 2048], 64)]
    vel = [device_field_2d(n, 2, 4000 + j, -3.0, True) for j, n in enumerate(sizes)]
    th = [device_field_2d(n, 1, 5000 + j, -1.0, False)[..., 0] for j, n in enumerate(sizes)]
    kss = [fit.wavenumber_grid((n, n), (2 * math.pi,) * 2) for n in sizes]
    v32 = [v.to(torch.complex64) for v in vel]
    t32 = [t.to(torch.complex64) for t in th]
    t0 = time.perf_counter()
    out = fit.calculate_ensemble_fluxes(v32, kss, t32, binning=fit.LinearBinning(1.0))
    torch.cuda.synchronize()
    say("  FP32 ensemble of", len(sizes), "(eager + capture) s:", time.perf_counter() - t0)
    t0 = time.perf_counter()
    out = fit.calculate_ensemble_fluxes(v32, kss, t32, binning=fit.LinearBinning(1.0))
    say("  FP32 ensemble replay s:", time.perf_counter() - t0)
    o64 = fit.calculate_ensemble_fluxes(vel, kss, th, binning=fit.LinearBinning(1.0))
    assert len(out) == len(sizes) == len(o64)
    worst = 0.0
    for j, n in enumerate(sizes):
        ke, sc = out[j]
        ke64, sc64 = o64[j]
        assert len(ke.flux) == int(math.sqrt(2) * n / 2) + 1
        worst = max(worst, nerr(ke.flux, ke64.flux), nerr(sc.flux, sc64.flux),
                    nerr(ke.transfer_spectrum, ke64.transfer_spectrum))
        for T in (ke64.transfer_spectrum, sc64.transfer_spectrum):       # conservation (FP64)
            assert abs(T.sum()) < 1e-9 * np.abs(T).sum()
    say("  FP32 vs FP64 worst normwise error:", worst)
    assert worst < 1e-5
    checked = 0
    for j, n in enumerate(sizes):                                # FP64 run vs oracle, in order
        if n != 256 or checked == 3:
            continue
        u = vel[j].permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0)
        t = th[j].permute(1, 0).cpu().numpy().transpose(1, 0)
        ks = kss[j]
        _, Tk, Pk = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23")
        _, Ts, Ps = oracle.calculate_scalar_flux(u, t, ks, oracle.LinearBinning(1.0), "23")
        assert nerr(o64[j][0].flux, Pk) < 1e-11 and nerr(o64[j][1].flux, Ps) < 1e-11
        assert nerr(out[j][0].flux, Pk) < 1e-5 and nerr(out[j][1].flux, Ps) < 1e-5
        checked += 1
    assert checked == 3


def test_c1_64_2d_fp64_flux(fit, oracle):
    """C1: 2D 64² FP64 kinetic-energy Π(K), LinearBinning(1) (46 shells), OrszagTwoThirds, through
    the HIP path against the oracle."""
    n = 64
    say("case C1", n)
    u = oracle.synthetic_field((n, n), 2, seed=1)
    ks = fit.wavenumber_grid((n, n), (2 * math.pi,) * 2)
    c, T, Pi = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23")
    r = fit.calculate_spectral_flux(u, ks, binning=fit.LinearBinning(1.0))
    assert len(r.flux) == 46 and np.allclose(r.k_shells, c)
    eT, eP = nerr(r.transfer_spectrum, T), nerr(r.flux, Pi)
    say("  T err", eT, "Π err", eP)
    assert eT < 1e-11 and eP < 1e-11
    assert abs(T.sum()) < 1e-10 * np.abs(T).sum()


def test_c2_256_against_oracle(fit, oracle):
    """C2 against the FP64 oracle (not only an FP64 run of the same kernels): FP32 Π(K), T(k) on
    the device vs the oracle's c2c restatement at 256³, and FP32 / FP64 T(n,m) giver columns
    1, 2, 16 and 32 of the 32-shell matrix vs the oracle's giver loop for those columns (the full
    32-column oracle loop is ~400 transforms of 256³ on the host)."""
    import torch
    n = 256
    say("case C2 vs oracle", n)
    ud, u = device_field(oracle, (n,) * 3, 3, seed=2)
    ks = fit.wavenumber_grid((n,) * 3, (2 * math.pi,) * 3)
    u32 = ud.to(torch.complex64)
    r32 = fit.calculate_spectral_flux(u32, ks, binning=fit.LinearBinning(1.0))
    s32 = fit.calculate_shell_to_shell_transfer(u32, ks, binning=fit.CustomBinning(range(0, 33)))
    s64 = fit.calculate_shell_to_shell_transfer(ud, ks, binning=fit.CustomBinning(range(0, 33)))
    t0 = time.perf_counter()
    workers = min(16, os.cpu_count() or 1)
    _, T, Pi = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23", workers=workers)
    say("  oracle Π s:", time.perf_counter() - t0)
    eT, eP = nerr(r32.transfer_spectrum, T), nerr(r32.flux, Pi)
    say("  FP32 vs oracle: T", eT, "Π", eP)
    assert eT < 1e-5 and eP < 1e-6
    t0 = time.perf_counter()
    kc = oracle.shell_coordinate(None, ks)
    sidx = oracle.assign_shells_fast(kc, np.arange(0.0, 33.0))
    cols = [1, 2, 16, 32]
    Tc = oracle.shell_to_shell_columns(u, ks, sidx, 32, cols, "23", "ke", workers=workers)
    say("  oracle columns s:", time.perf_counter() - t0)
    got32 = s32.transfer_matrix[:, [m - 1 for m in cols]]
    got64 = s64.transfer_matrix[:, [m - 1 for m in cols]]
    e32, e64 = nerr(got32, Tc), nerr(got64, Tc)
    say("  T(n,m) columns vs oracle: FP32", e32, "FP64", e64)
    assert e32 < 1e-5 and e64 < 1e-11


def test_c3_512_full_size(fit, oracle):
    """C3 at full size: 512³ seeded solenoidal k⁴exp(-2(k/8)²) / k^-5/3 field (zero at |k| >= n/3).
    Π(K) of kinetic energy and helicity with PaddedThreeHalves (768³ transforms, 444 shells, one
    shared N̂): FP32 vs FP64 ≤ 1e-6, Σ_k T ≈ 0 for both invariants (FP64).  48-shell T(n,m)
    (OrszagTwoThirds): FP64 antisymmetry < 1e-12 ‖T‖, the band-limited sub-grid (FP64) against the
    plain full-grid giver loop (FP64) ≤ 1e-12, FP32 vs FP64 ≤ 1e-5."""
    import torch
    import bench
    from fit_amd.grid import shell_setup
    n, ng = 512, 48
    say("case C3", n)
    # the generator projects in FP32 (divergence ~1e-7 relative, which breaks the antisymmetry of
    # T(n,m) at ~1e-8): project again in FP64, so the FP64 checks see a solenoidal field
    u64 = bench.synthetic_velocity_gpu(n, 3, torch).to(torch.complex128)
    i = torch.arange(n, device="cuda")
    k1 = torch.where(i <= n // 2, i, i - n).to(torch.float64)
    kx, ky, kz = k1.view(n, 1, 1), k1.view(1, n, 1), k1.view(1, 1, n)    # (x, y, z) view axes
    k2 = kx * kx + ky * ky + kz * kz
    k2 = torch.where(k2 > 0, k2, torch.ones_like(k2))
    div = kx * u64[..., 0] + ky * u64[..., 1] + kz * u64[..., 2]
    for j, kj in enumerate((kx, ky, kz)):
        u64[..., j] -= kj * div / k2
    del div, k2
    u = u64.to(torch.complex64)
    ks = fit.wavenumber_grid((n,) * 3, (2 * math.pi,) * 3)
    c_pi, _, s_pi = shell_setup(ks, fit.LinearBinning(1.0), None)
    c_tm, _, s_tm = shell_setup(ks, fit.CustomBinning(range(0, ng + 1)), None)
    assert len(c_pi) == 444 and len(c_tm) == ng
    KE, HEL = fit.KineticEnergy(), fit.Helicity()
    res = {}
    for prec, uu in ((np.float32, u), (np.float64, u64)):
        with fit.Plan((n,) * 3, 3, 3, ks, fit.PaddedThreeHalves(), prec, s_pi, len(c_pi)) as pp:
            N = pp.nonlinear_term(uu)
            (Tk, Pk), (Th, Ph) = pp.spectral_fluxes_with_nhat((KE, HEL), uu, N)
            del N
        with fit.Plan((n,) * 3, 3, 3, ks, fit.OrszagTwoThirds(), prec, s_tm, len(c_tm)) as tp:
            assert tp.subgrid is not None
            Tm, net, asym = tp.shell_to_shell(KE, uu)
            if prec == np.float64:
                tp.set_subgrid(False)
                tp.set_plain(True)
                Tm_plain, _, _ = tp.shell_to_shell(KE, uu)
                res["plain"] = Tm_plain
        res[prec] = (Tk, Pk, Th, Ph, Tm, asym)
        say("  ", prec.__name__, "done")
    Tk, Pk, Th, Ph, Tm, asym = res[np.float64]
    Tk32, Pk32, Th32, Ph32, Tm32, _ = res[np.float32]
    e = {"pi_ke": nerr(Pk32, Pk), "pi_hel": nerr(Ph32, Ph), "t_ke": nerr(Tk32, Tk), "t_nm": nerr(Tm32, Tm),
         "sub_vs_plain": nerr(Tm, res["plain"])}
    say("  errors", e)
    assert e["pi_ke"] <= 1e-6 and e["pi_hel"] <= 1e-6 and e["t_nm"] < 1e-5
    assert e["sub_vs_plain"] <= 1e-12
    for T in (Tk, Th):
        assert abs(T.sum()) < 1e-9 * np.abs(T).sum()
    nrm = math.sqrt(np.sum(Tm ** 2))
    say("  antisymmetry", asym / nrm)
    assert asym < 1e-12 * nrm
