"""GPU regression tests for defects that were fixed:

  * component counts other than nd: kinetic energy with M = 4 (3D, advected by a separate
    3-component velocity) and 2D enstrophy with M = 3 through the batched and sub-grid giver
    loops and mode-to-mode, against the FP64 oracle (the carrier of components the invariant
    does not use must be zero, not left over from an earlier call);
  * stream ordering: a cached plan / ensemble reads device inputs that torch produced on its own
    stream right before the call (fit_plan_wait_stream / fit_ensemble_wait_stream);
  * one plan per host thread: two threads running the same grid concurrently get the results a
    single thread gets.
"""
import math
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def ks_of(oracle, ns):
    return oracle.wavenumber_grid(ns, (2 * math.pi,) * len(ns))


def fields_m4(oracle, ns):
    """3D: a 3-component solenoidal velocity and a 4-component advected field (the velocity plus a
    fourth, independent component)."""
    u = oracle.synthetic_field(ns, 3, seed=41)
    extra = oracle.synthetic_field(ns, 1, seed=42, solenoidal=False)
    return u, np.concatenate([u, extra], axis=3)


@pytest.mark.parametrize("binning", ["lin", "custom"])
def test_shell_to_shell_ke_four_components(fit, oracle, binning):
    """Kinetic energy of a 4-component field: Σ_c over all four components in every giver loop
    (batched on the plan's grid with LinearBinning(1); band-limited sub-grid with 0:5 shells)."""
    ns = (24, 24, 24) if binning == "custom" else (12, 12, 12)
    say("case s2s KE M=4", ns, binning)
    ks = ks_of(oracle, ns)
    u, f = fields_m4(oracle, ns)
    bf = fit.CustomBinning(range(0, 6)) if binning == "custom" else fit.LinearBinning(1.0)
    bo = oracle.CustomBinning(tuple(range(0, 6))) if binning == "custom" else oracle.LinearBinning(1.0)
    # kinetic energy, then helicity (whose fourth carrier component is zero), then kinetic energy
    # again, on the same cached plan: stale carrier memory between calls would show
    for inv, I in (("ke", fit.KineticEnergy()), ("helicity", fit.Helicity()), ("ke", fit.KineticEnergy())):
        _, _, T, net, _ = oracle.calculate_shell_to_shell_transfer(f, ks, bo, "23", invariant=inv,
                                                                   advecting_hat=u)
        r = fit.calculate_shell_to_shell_transfer(f, ks, binning=bf, advecting_hat=u, invariant=I)
        assert np.max(np.abs(T)) > 1e-12
        e = rel(r.transfer_matrix, T)
        say("  ", inv, "T(n,m) err", e)
        assert e < TOL64 and rel(r.net_transfer, net) < TOL64


def test_shell_to_shell_enstrophy_2d_three_components(fit, oracle):
    """2D enstrophy of a 3-component field (self-advected by its first two components): only the
    two velocity components carry vorticity."""
    ns = (16, 16)
    say("case s2s 2D enstrophy M=3", ns)
    ks = ks_of(oracle, ns)
    u, _ = oracle.streamfunction_velocity_2d(16, 43)
    th = np.fft.fft2(np.random.default_rng(44).standard_normal(ns)) / 256.0
    f = np.concatenate([u, th[..., None]], axis=2)
    # a kinetic-energy call on the same grid first leaves a nonzero third carrier component behind
    fit.calculate_shell_to_shell_transfer(f, ks, binning=fit.LinearBinning(1.0))
    _, _, T, net, _ = oracle.calculate_shell_to_shell_transfer(f, ks, oracle.LinearBinning(1.0), "23",
                                                               invariant="enstrophy")
    r = fit.calculate_shell_to_shell_transfer(f, ks, binning=fit.LinearBinning(1.0), invariant=fit.Enstrophy())
    assert np.max(np.abs(T)) > 1e-12
    e = rel(r.transfer_matrix, T)
    say("  T(n,m) err", e)
    assert e < TOL64 and rel(r.net_transfer, net) < TOL64


@pytest.mark.parametrize("case", ["ke4", "ens3", "hel4"])
def test_mode_to_mode_component_counts(fit, oracle, case):
    say("case m2m", case)
    if case == "ens3":
        ns = (6, 6)
        ks = ks_of(oracle, ns)
        u, _ = oracle.streamfunction_velocity_2d(6, 45)
        th = np.fft.fft2(np.random.default_rng(46).standard_normal(ns)) / 36.0
        f, adv, inv, I = np.concatenate([u, th[..., None]], axis=2), None, "enstrophy", fit.Enstrophy()
    else:
        ns = (4, 4, 4)
        ks = ks_of(oracle, ns)
        u, f = fields_m4(oracle, ns)
        adv = u
        inv, I = ("ke", fit.KineticEnergy()) if case == "ke4" else ("helicity", fit.Helicity())
    net_o, S_o = oracle.calculate_mode_to_mode_transfer(f, ks, inv, "none", force=True, advecting_hat=adv)
    r = fit.calculate_mode_to_mode_transfer(f, ks, invariant=I, dealiasing=fit.NoDealiasing(),
                                            advecting_hat=adv)
    eS, eN = rel(r.transfer, S_o), rel(r.net_transfer, net_o)
    say("  S err", eS, "net err", eN)
    assert eS < TOL64 and eN < TOL64


def _fresh_input_behind_heavy_work(torch, src, dtype):
    """A new device array of src's shape and (column-major) layout, first filled with NaN and
    synchronised, then -- on torch's current stream, behind several large matrix products --
    overwritten with src.  A consumer that does not wait for that stream reads NaN."""
    rev = tuple(reversed(range(src.dim())))
    buf = torch.empty(tuple(reversed(src.shape)), dtype=dtype, device=src.device).permute(rev)
    buf.fill_(float("nan"))
    torch.cuda.synchronize()
    a = torch.randn(4096, 4096, device=src.device)
    for _ in range(6):
        a = torch.tanh(a @ a * 1e-3)
    buf.copy_(src)
    return buf, a


def test_cached_plan_waits_for_torch_stream(fit, oracle):
    import torch
    ns = (32, 32, 32)
    say("case stream ordering (plan)", ns)
    u = oracle.synthetic_field(ns, 3, seed=47)
    ks = ks_of(oracle, ns)
    ud = torch.from_numpy(np.ascontiguousarray(u.transpose(3, 2, 1, 0))).to("cuda").permute(3, 2, 1, 0)
    _, T, Pi = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23")
    _, _, Tm, _, _ = oracle.calculate_shell_to_shell_transfer(u, ks, oracle.CustomBinning(tuple(range(0, 6))), "23")
    b = fit.LinearBinning(1.0)
    fit.calculate_spectral_flux(ud, ks, binning=b)                       # plans are cached now
    fit.calculate_shell_to_shell_transfer(ud, ks, binning=fit.CustomBinning(range(0, 6)))
    for trial in range(2):
        buf, _ = _fresh_input_behind_heavy_work(torch, ud, torch.complex128)
        r = fit.calculate_spectral_flux(buf, ks, binning=b)
        assert rel(r.flux, Pi) < TOL64
        buf, _ = _fresh_input_behind_heavy_work(torch, ud, torch.complex128)
        s = fit.calculate_shell_to_shell_transfer(buf, ks, binning=fit.CustomBinning(range(0, 6)))
        assert rel(s.transfer_matrix, Tm) < TOL64
        buf, _ = _fresh_input_behind_heavy_work(torch, ud, torch.complex128)
        N = fit.compute_nonlinear_term(buf, ks)
        assert not torch.isnan(N).any().item()


def test_ensemble_replay_waits_for_torch_stream(fit, oracle):
    import torch
    say("case stream ordering (ensemble)")
    sizes = [32, 64, 32]
    us = [oracle.synthetic_field((n, n), 2, seed=60 + j) for j, n in enumerate(sizes)]
    kss = [ks_of(oracle, (n, n)) for n in sizes]
    dev = [torch.from_numpy(np.ascontiguousarray(u.transpose(2, 1, 0))).to("cuda").permute(2, 1, 0) for u in us]
    refs = [oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23")[2] for u, ks in zip(us, kss)]
    fit.calculate_ensemble_fluxes(dev, kss, binning=fit.LinearBinning(1.0))   # eager run + capture
    for trial in range(2):
        bufs = [_fresh_input_behind_heavy_work(torch, d, torch.complex128)[0] for d in dev]
        out = fit.calculate_ensemble_fluxes(bufs, kss, binning=fit.LinearBinning(1.0))
        for (ke, _), Pi in zip(out, refs):
            assert rel(ke.flux, Pi) < TOL64


def test_threads_get_their_own_plans(fit, oracle):
    """Two host threads on the same grid at once (the reference's threaded backend runs one
    workspace per task): each gets the single-threaded result."""
    ns = (24, 24, 24)
    say("case threads", ns)
    ks = ks_of(oracle, ns)
    fields = [oracle.synthetic_field(ns, 3, seed=70 + i) for i in range(2)]
    b = fit.CustomBinning(range(0, 7))
    ref = [fit.calculate_shell_to_shell_transfer(f, ks, binning=b).transfer_matrix for f in fields]
    out = [[None] * 4 for _ in range(2)]
    errors = []

    def work(i):
        try:
            for r in range(4):
                out[i][r] = fit.calculate_shell_to_shell_transfer(fields[i], ks, binning=b).transfer_matrix
        except Exception as exc:  # surfaced by the assertion below
            errors.append(exc)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i in range(2):
        for r in range(4):
            assert rel(out[i][r], ref[i]) < 1e-13
