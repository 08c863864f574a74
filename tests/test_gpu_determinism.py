"""Bitwise reproducibility of the shell sums: every FP64 shell bin adds its contributions in a fixed
order inside a workgroup (per-wave private bins, or waves committing in turn when the private
copies would not fit) and across workgroups (fixed-order reduction), so a rerun on the same input
returns the same bits.  Covers each density kernel (full-grid density, the fused last pass in both
density forms and both bin modes, the LDS-engine last pass, band weights) against the oracle too."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(float(np.max(np.abs(b))), 1e-300))


def say(*a):
    print(*a, flush=True)


def to_dev(a, dtype):
    import torch
    rev = tuple(reversed(range(a.ndim)))
    return torch.from_numpy(np.ascontiguousarray(a.transpose(rev))).to("cuda").permute(rev).to(dtype)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,dea", [((24, 24, 24), "pad"), ((32, 32, 32), "23"), ((2048, 2048), "23")])
def test_flux_rerun_bitwise(fit, oracle, ns, dea, prec):
    # (2048, 2048): 1449 shells x 2 invariants exceed the private-copy budget of the full-grid
    # density pass, so its waves commit in turn
    say("case flux rerun", ns, dea, prec)
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    if nd == 2:
        u, _ = oracle.streamfunction_velocity_2d(ns[0], 7)
        invs = (fit.KineticEnergy(), fit.Enstrophy())
        tags = ("ke", "enstrophy")
    else:
        u = oracle.synthetic_field(ns, 3, seed=7)
        invs = (fit.KineticEnergy(), fit.Helicity())
        tags = ("ke", "helicity")
    dt = np.float64 if prec == "f64" else np.float32
    deal = {"23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    from fit_amd.grid import shell_setup
    c, e, s = shell_setup(ks, fit.LinearBinning(1.0), fit.IsotropicShells())
    with fit.Plan(ns, nd, nd, ks, deal, dt, s, len(c)) as plan:
        r1 = plan.spectral_fluxes(invs, u.astype(np.complex128 if prec == "f64" else np.complex64))
        r2 = plan.spectral_fluxes(invs, u.astype(np.complex128 if prec == "f64" else np.complex64))
    for (T1, P1), (T2, P2) in zip(r1, r2):
        assert np.array_equal(T1, T2) and np.array_equal(P1, P2)
    if nd == 3:
        for (T1, P1), tag in zip(r1, tags):
            _, To, Po = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), dea, tag)
            e = rel(P1, Po)
            say("  Pi err", tag, e)
            assert e < (1e-11 if prec == "f64" else 2e-5)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,nsh", [((24, 24, 24), 8), ((36, 30, 40), 10), ((64, 64), 21)])
def test_shell_to_shell_rerun_bitwise(fit, oracle, ns, nsh, prec):
    # (36, 30, 40): lengths without a register FFT take the LDS-engine last pass
    say("case s2s rerun", ns, nsh, prec)
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    u = oracle.synthetic_field(ns, nd, seed=9)
    dt = np.complex128 if prec == "f64" else np.complex64
    b = fit.CustomBinning(range(0, nsh + 1))
    r1 = fit.calculate_shell_to_shell_transfer(u.astype(dt), ks, binning=b)
    r2 = fit.calculate_shell_to_shell_transfer(u.astype(dt), ks, binning=b)
    assert np.array_equal(r1.transfer_matrix, r2.transfer_matrix)
    assert np.array_equal(r1.net_transfer, r2.net_transfer)
    _, _, Tm, _, _ = oracle.calculate_shell_to_shell_transfer(u, ks, oracle.CustomBinning(tuple(range(0, nsh + 1))), "23")
    e = rel(r1.transfer_matrix, Tm)
    say("  T(n,m) err", e)
    assert e < (1e-11 if prec == "f64" else 2e-5)


def test_ensemble_ordered_commit_rerun_bitwise(fit, oracle):
    # a 1024² snapshot: with a 1 KiB bin budget its 725 shells exceed the private copies, so the
    # waves commit in turn; replays of the captured graph return the same bits
    import torch
    from fit_amd import _lib
    shapes = [(1024, 1024), (64, 64), (1024, 1024)]
    us, kss = [], []
    for j, s in enumerate(shapes):
        u, ks = oracle.streamfunction_velocity_2d(s[0], 300 + j)
        us.append(u)
        kss.append(ks)
    ud = [to_dev(u, torch.complex64) for u in us]
    ens = fit.Ensemble(shapes, kss, binning=fit.LinearBinning(1.0), precision=np.float32)
    ens.set_plan_option(_lib.OPT_LAST_BINS, 1)
    a = ens.fluxes(ud)
    b = ens.fluxes(ud)
    c = ens.fluxes(ud)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x[0].flux, y[0].flux) and np.array_equal(y[0].flux, z[0].flux)
    _, _, Po = oracle.calculate_spectral_flux(us[1], kss[1], oracle.LinearBinning(1.0), "23")
    e = rel(a[1][0].flux, Po)
    say("  64² Pi err", e)
    assert e < 2e-5
    # the default 64 KiB budget gives the 1024-point last pass private per-wave bins at half the
    # columns per workgroup: the same fluxes to rounding, again bitwise on replay
    ens.set_plan_option(_lib.OPT_LAST_BINS, 64)
    d = ens.fluxes(ud)
    f = ens.fluxes(ud)
    for x, y, z in zip(a, d, f):
        assert np.array_equal(y[0].flux, z[0].flux)
        assert rel(y[0].flux, x[0].flux) < 1e-5
    ens.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_band_to_band_rerun_bitwise(fit, oracle, prec):
    ns = (32, 32)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 2)
    u, _ = oracle.streamfunction_velocity_2d(32, 11)
    dt = np.complex128 if prec == "f64" else np.complex64
    bands = fit.SmoothBands((1.0, 2.0, 4.0, 8.0), 0.4)
    r1 = fit.calculate_band_to_band_transfer(u.astype(dt), ks, bands=bands)
    r2 = fit.calculate_band_to_band_transfer(u.astype(dt), ks, bands=bands)
    assert np.array_equal(r1.transfer_matrix, r2.transfer_matrix)
