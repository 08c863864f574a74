"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the same seeded inputs.

Tolerances (normwise, max|x - x_ref| / max|x_ref|):
  FP64 compute on the device : 1e-11  (roundoff of a different FFT algorithm)
  FP32 compute on the device : 2e-5 per-mode quantities, 1e-5 shell sums
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-11
TOL32 = 2e-5


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def ks_of(oracle, ns):
    return oracle.wavenumber_grid(ns, (2 * math.pi,) * len(ns))


CASES_NL = [((8,), "none"), ((12,), "23"), ((16,), "pad"), ((16, 16), "23"), ((12, 10), "none"),
            ((8, 12), "pad"), ((8, 8, 8), "23"), ((6, 8, 10), "none"), ((8, 6, 4), "pad"),
            ((16, 12, 8), "23"), ((6, 6, 6), "pad")]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,dea", CASES_NL)
def test_nonlinear_term_matches_oracle(fit, oracle, ns, dea, prec):
    say("case nonlinear", ns, dea, prec)
    rng = np.random.default_rng(hash((ns, dea)) % 2 ** 32)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    D = max(nd, 2)
    u = rand_c(rng, ns + (D,))
    th = rand_c(rng, ns + (1,))
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    dt = np.complex128 if prec == "f64" else np.complex64
    tol = TOL64 if prec == "f64" else TOL32
    ref = oracle.compute_nonlinear_term(u, ks, dea)
    got = fit.compute_nonlinear_term(u.astype(dt), ks, dealiasing=deal)
    e = rel(got, ref)
    say("  self-advection err", e)
    assert e < tol
    ref = oracle.compute_nonlinear_term(th, ks, dea, advecting_hat=u)
    got = fit.compute_nonlinear_term(th.astype(dt), ks, dealiasing=deal, advecting_hat=u.astype(dt))
    e = rel(got, ref)
    say("  scalar err", e)
    assert e < tol


@pytest.mark.parametrize("ns,inv", [((8, 6, 10), "ke"), ((8, 6, 10), "helicity"), ((8, 6, 10), "enstrophy"),
                                    ((12, 10), "enstrophy"), ((12, 10), "ke"), ((16,), "ke")])
def test_transfer_density_matches_oracle(fit, oracle, ns, inv):
    say("case density", ns, inv)
    rng = np.random.default_rng(7)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    u = rand_c(rng, ns + (nd,))
    N = rand_c(rng, ns + (nd,))
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy()}[inv]
    got = fit.transfer_density(I, u, N, ks)
    assert rel(got, oracle.transfer_density(inv, u, N, ks)) < TOL64


FLUX_CASES = [
    ((16, 16), "23", "ke", "lin", None),
    ((16, 16), "23", "enstrophy", "lin", None),
    ((24, 24), "pad", "ke", "lin", None),
    ((12, 12, 12), "23", "ke", "lin", None),
    ((12, 12, 12), "23", "helicity", "log", None),
    ((12, 12, 12), "none", "enstrophy", "lin", None),
    ((16, 16, 16), "23", "ke", "lin", "perp"),
    ((16, 16, 16), "pad", "helicity", "lin", "par"),
    ((8,), "none", "ke", "lin", None),
]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,dea", [((8, 8, 8), "pad"), ((6, 8, 10), "pad"), ((5, 9, 8), "pad"), ((10, 8, 6), "23"),
                                    ((9, 6, 5), "none"), ((12, 12), "pad"), ((9, 10), "pad")])
def test_flux_full_band_field_nyquist_rows(fit, oracle, ns, dea, prec):
    """Kinetic-energy and helicity fluxes of a real random field with content at every bin: the
    flux stores N̂ on the half grid and sums pair densities, which on a padded grid needs the
    Nyquist rows' own N̂(−k) (there N̂(−k) != conj N̂(k))."""
    say("case full-band flux", ns, dea, prec)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    rng = np.random.default_rng(17)
    u = np.fft.fftn(rng.standard_normal(ns + (nd,)), axes=tuple(range(nd))) / np.prod(ns)
    invs = ("ke", "helicity") if nd == 3 else ("ke",)
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity()}
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    dt = np.complex128 if prec == "f64" else np.complex64
    for name in invs:
        r = fit.calculate_spectral_flux(u.astype(dt), ks, binning=fit.LinearBinning(1.0), dealiasing=deal,
                                        invariant=I[name])
        _, T, Pi = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), dea, name)
        e = max(rel(r.transfer_spectrum, T), rel(r.flux, Pi))
        say("  ", name, "err", e)
        assert e < (TOL64 if prec == "f64" else 2e-5)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,dea,inv,binn,geom", FLUX_CASES)
def test_spectral_flux_matches_oracle(fit, oracle, ns, dea, inv, binn, geom, prec):
    say("case flux", ns, dea, inv, binn, geom, prec)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    if nd == 1:
        x = np.arange(ns[0]) * 2 * math.pi / ns[0]
        u = (np.fft.fft(np.cos(x) + 0.3 * np.sin(2 * x) + 0.1 * np.cos(3 * x)) / ns[0]).reshape(ns[0], 1)
    elif nd == 2:
        u, _ = oracle.streamfunction_velocity_2d(ns[0], 11)
    else:
        u = oracle.synthetic_field(ns, 3, seed=5)
    b_f = {"lin": fit.LinearBinning(1.0), "log": fit.LogarithmicBinning(1.0, 1.5)}[binn]
    b_o = {"lin": oracle.LinearBinning(1.0), "log": oracle.LogarithmicBinning(1.0, 1.5)}[binn]
    g_f = {None: None, "perp": fit.PerpendicularShells(), "par": fit.ParallelShells()}[geom]
    g_o = {None: None, "perp": oracle.PerpendicularShells(), "par": oracle.ParallelShells()}[geom]
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy()}[inv]
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    dt = np.complex128 if prec == "f64" else np.complex64
    c, T, Pi = oracle.calculate_spectral_flux(u, ks, b_o, dea, inv, geometry=g_o)
    r = fit.calculate_spectral_flux(u.astype(dt), ks, binning=b_f, dealiasing=deal, invariant=I,
                                    geometry=g_f)
    eT, eP = rel(r.transfer_spectrum, T), rel(r.flux, Pi)
    say("  T err", eT, "Pi err", eP)
    tol = TOL64 if prec == "f64" else 1e-5
    assert np.allclose(r.k_shells, c)
    assert eT < tol and eP < tol
    assert r.transfer_spectrum.dtype == (np.float64 if prec == "f64" else np.float32)


def test_scalar_flux_matches_oracle(fit, oracle):
    say("case scalar flux")
    rng = np.random.default_rng(33)            # not the streamfunction's stream (θ = ψ ⇒ N ≡ 0)
    u, ks = oracle.streamfunction_velocity_2d(16, 32)
    th = np.fft.fft2(rng.standard_normal((16, 16))) / 256
    _, T, Pi = oracle.calculate_scalar_flux(u, th, ks, oracle.LinearBinning(1.0), "23")
    assert np.max(np.abs(T)) > 1e-12                                     # non-degenerate
    r = fit.calculate_scalar_flux(u, th, ks, binning=fit.LinearBinning(1.0))
    assert rel(r.transfer_spectrum, T) < TOL64 and rel(r.flux, Pi) < TOL64
    sT = math.sqrt(np.sum(r.transfer_spectrum ** 2))
    assert abs(r.flux[-1]) < 1e-9 * sT                                   # runtests.jl:347-368


# (at 8³ the 2/3 rule keeps |k_d| <= 1: every kept mode is in one shell and T ≡ 0)
S2S_CASES = [((16, 16), "23", "ke"), ((12, 12), "none", "ke"), ((12, 12), "pad", "ke"),
             ((12, 12, 12), "23", "ke"), ((12, 12, 12), "23", "helicity"), ((16, 16), "23", "enstrophy"),
             ((8, 8, 8), "none", "ke")]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,dea,inv", S2S_CASES)
def test_shell_to_shell_matches_oracle(fit, oracle, ns, dea, inv, prec):
    say("case s2s", ns, dea, inv, prec)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    if nd == 2:
        u, _ = oracle.streamfunction_velocity_2d(ns[0], 13)
    else:
        u = oracle.synthetic_field(ns, 3, seed=8)
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy()}[inv]
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    dt = np.complex128 if prec == "f64" else np.complex64
    c, e, T, net, asym = oracle.calculate_shell_to_shell_transfer(u, ks, oracle.LinearBinning(1.0), dea,
                                                                   invariant=inv)
    r = fit.calculate_shell_to_shell_transfer(u.astype(dt), ks, binning=fit.LinearBinning(1.0),
                                              dealiasing=deal, invariant=I)
    assert np.max(np.abs(T)) > 1e-12            # non-degenerate reference
    eT = rel(r.transfer_matrix, T)
    say("  T(n,m) err", eT, "asym", r.max_antisymmetry_error)
    tol = TOL64 if prec == "f64" else TOL32
    assert eT < tol and rel(r.net_transfer, net) < tol
    assert np.allclose(r.shell_edges, e)


def test_scalar_shell_to_shell(fit, oracle):
    say("case scalar s2s")
    rng = np.random.default_rng(15)            # not the streamfunction's stream (θ = ψ ⇒ N ≡ 0)
    u, ks = oracle.streamfunction_velocity_2d(16, 14)
    th = np.fft.fft2(rng.standard_normal((16, 16))) / 256
    _, _, T, net, _ = oracle.calculate_scalar_shell_to_shell_transfer(u, th, ks,
                                                                       binning=oracle.LinearBinning(1.0))
    r = fit.calculate_scalar_shell_to_shell_transfer(u, th, ks, binning=fit.LinearBinning(1.0))
    nrm = math.sqrt(np.sum(T ** 2))
    assert nrm > 1e-12
    assert rel(r.transfer_matrix, T) < TOL64
    assert r.max_antisymmetry_error < 1e-10 * nrm                         # runtests.jl:441-464


M2M_CASES = [((8, 8), "ke", "23"), ((6, 6), "ke", "none"), ((8, 8), "enstrophy", "23"),
             ((4, 4, 4), "ke", "none"), ((4, 4, 4), "helicity", "none"), ((4, 4, 4), "enstrophy", "none"),
             ((8, 8), "scalar", "23"),
             # PaddedThreeHalves: (8, 8) pads to exactly 3n/2, (6, 6) to 10, (5, 6) to (9, 10)
             ((8, 8), "ke", "pad"), ((6, 6), "enstrophy", "pad"), ((5, 6), "ke", "pad"),
             ((4, 4, 4), "ke", "pad"), ((4, 4, 4), "helicity", "pad"), ((8, 8), "scalar", "pad")]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,inv,dea", M2M_CASES)
def test_mode_to_mode_matches_oracle(fit, oracle, ns, inv, dea, prec):
    say("case m2m", ns, inv, dea, prec)
    rng = np.random.default_rng(3)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    if nd == 2 and ns[0] == ns[1]:
        u, _ = oracle.streamfunction_velocity_2d(ns[0], 21)
    else:
        u = oracle.synthetic_field(ns, nd, seed=2)
    dt = np.complex128 if prec == "f64" else np.complex64
    tol = TOL64 if prec == "f64" else TOL32
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    if inv == "scalar":
        th = rand_c(rng, ns + (1,))
        net_o, S_o = oracle.calculate_mode_to_mode_transfer(th, ks, "scalar", dea, force=True,
                                                            advecting_hat=u)
        r = fit.calculate_scalar_mode_to_mode_transfer(u.astype(dt), th.astype(dt), ks,
                                                       dealiasing=deal)
    else:
        prim = u + 0.2 * rand_c(rng, u.shape)        # non-Hermitian primary, Hermitian advecting
        I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy()}[inv]
        net_o, S_o = oracle.calculate_mode_to_mode_transfer(prim, ks, inv, dea, force=True,
                                                            advecting_hat=u)
        r = fit.calculate_mode_to_mode_transfer(prim.astype(dt), ks, invariant=I, dealiasing=deal,
                                                advecting_hat=u.astype(dt))
    eS, eN = rel(r.transfer, S_o), rel(r.net_transfer, net_o)
    say("  S err", eS, "net err", eN)
    assert r.transfer.shape == ns + ns
    assert eS < tol and eN < tol


def test_helical_decomposition_and_partial_fluxes(fit, oracle):
    say("case helical")
    u, ks = oracle.curl_velocity_3d(16, 91)
    dec_o = oracle.helical_decompose(u, ks)
    dec = fit.decompose_field(fit.HelicalDecomposition(), u, ks)
    assert rel(dec["positive"], dec_o["positive"]) < TOL64
    assert rel(dec["negative"], dec_o["negative"]) < TOL64
    b = fit.LinearBinning(1.0)
    ch_o, tot_o, _ = oracle.calculate_partial_fluxes(u, ks, oracle.LinearBinning(1.0), "23")
    hp = fit.calculate_helical_partial_fluxes(u, ks, binning=b)
    assert len(hp["channels"]) == 8
    scale = max(math.sqrt(np.sum(v[1] ** 2)) for v in ch_o.values())
    for key, (c, T, P) in ch_o.items():
        assert np.max(np.abs(hp["channels"][key].transfer_spectrum - T)) < 1e-10 * scale
    hel = fit.calculate_spectral_flux(u, ks, binning=b, decomposition=fit.HelicalDecomposition())
    hel_o = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23",
                                           decomposition="HelicalDecomposition")
    for name in ("positive", "negative"):
        assert rel(hel[name].transfer_spectrum, hel_o[name][1]) < TOL64


def test_zero_field_gives_zero(fit, oracle):
    # runtests.jl:738-776: !-variants on a zero field return zeros
    ks = ks_of(oracle, (6, 6))
    r = fit.calculate_shell_to_shell_transfer(np.zeros((6, 6, 2), complex), ks,
                                              binning=fit.LinearBinning(1.0),
                                              dealiasing=fit.NoDealiasing(), verify_antisymmetry=False)
    assert np.all(np.abs(r.transfer_matrix) < 1e-14) and math.isnan(r.max_antisymmetry_error)


def test_device_arrays_roundtrip(fit, oracle):
    # torch tensors on the HIP device go through FIT_PTR_DEVICE without host copies
    import torch
    say("case device arrays")
    ns = (16, 12, 8)
    u = oracle.synthetic_field(ns, 3, seed=4)
    ks = ks_of(oracle, ns)
    ud = torch.from_numpy(np.ascontiguousarray(u.transpose(3, 2, 1, 0))).to("cuda").permute(3, 2, 1, 0)
    N = fit.compute_nonlinear_term(ud, ks)
    assert N.is_cuda
    ref = oracle.compute_nonlinear_term(u, ks, "23")
    assert rel(N.cpu().numpy(), ref) < TOL64
    r = fit.calculate_spectral_flux(ud.to(torch.complex64), ks, binning=fit.LinearBinning(1.0))
    _, T, _ = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23")
    assert rel(r.transfer_spectrum, T) < 1e-5


@pytest.mark.parametrize("ns,dea,inv,batch,xopt", [((16, 16, 16), "23", "ke", 0, None),
                                                   ((16, 16, 16), "23", "helicity", 3, None),
                                                   ((24, 20), "none", "enstrophy", 2, None),
                                                   ((12, 12, 12), "none", "ke", 5, None),
                                                   ((256, 24, 20), "23", "ke", 0, None),
                                                   ((256, 24, 20), "23", "ke", 0, "block"),
                                                   ((512, 12, 10), "23", "helicity", 4, 3),
                                                   ((384, 16), "none", "ke", 0, 2),
                                                   ((768, 12), "23", "enstrophy", 0, 1)])
def test_fused_pipeline_equals_plain(fit, oracle, ns, dea, inv, batch, xopt):
    """The fused / pruned / batched giver loop against the plain one (same plan, FP64) and the
    oracle, with a giver-batch cap that splits the givers into several launches."""
    say("case fused vs plain", ns, dea, inv, batch)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    u = oracle.synthetic_field(ns, nd, seed=12)
    u = u / np.max(np.abs(u))                  # O(1) amplitudes: T well above roundoff at any size
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds()}[dea]
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy()}[inv]
    from fit_amd.grid import shell_setup
    c, e, sidx = shell_setup(ks, fit.CustomBinning(range(0, 7)), None)
    from fit_amd import _lib
    with fit.Plan(ns, nd, nd, ks, deal, np.float64, sidx, len(c)) as plan:
        plan.set_giver_batch(batch)
        if xopt == "block":
            plan.set_option(_lib.OPT_XPASS_BLOCK, 1)
        elif xopt is not None:
            plan.set_option(_lib.OPT_XPASS_GIVERS, xopt)
        Tf, netf, _ = plan.shell_to_shell(I, u)
        plan.set_plain(True)
        Tp, netp, _ = plan.shell_to_shell(I, u)
    _, _, To, _, _ = oracle.calculate_shell_to_shell_transfer(u, ks, oracle.CustomBinning(tuple(range(0, 7))), dea,
                                                               invariant=inv)
    assert np.max(np.abs(To)) > 1e-12
    e1, e2 = rel(Tf, To), rel(Tp, To)
    say("  fused err", e1, "plain err", e2)
    assert e1 < TOL64 and e2 < TOL64


@pytest.mark.parametrize("ns,dea", [((16, 16, 16), "23"), ((12, 10, 8), "pad"), ((20, 24), "23")])
def test_fused_nonlinear_equals_plain(fit, oracle, ns, dea):
    say("case fused NL vs plain", ns, dea)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    rng = np.random.default_rng(2)
    u = rand_c(rng, ns + (nd,))
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    with fit.Plan(ns, nd, nd, ks, deal, np.float64) as plan:
        Nf = plan.nonlinear_term(u)
        plan.set_plain(True)
        Np_ = plan.nonlinear_term(u)
    ref = oracle.compute_nonlinear_term(u, ks, dea)
    assert rel(Nf, ref) < TOL64 and rel(Np_, ref) < TOL64


@pytest.mark.parametrize("ns,dea,invs", [((12, 12, 12), "pad", ("ke", "helicity", "enstrophy")),
                                         ((16, 16, 16), "23", ("ke", "helicity")),
                                         ((24, 20), "none", ("enstrophy", "ke"))])
def test_several_invariants_one_density_pass(fit, oracle, ns, dea, invs):
    """fit_spectral_fluxes_with_nhat: several invariants' T, Π from one N̂ in one pass equal the
    oracle's per-invariant fluxes (and the one-invariant entry point)."""
    say("case multi-invariant density", ns, dea, invs)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    u = oracle.synthetic_field(ns, nd, seed=21) if nd == 3 else rand_c(np.random.default_rng(5), ns + (2,))
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy()}
    from fit_amd.grid import shell_setup
    c, e, sidx = shell_setup(ks, fit.LinearBinning(1.0), None)
    for prec, tol in ((np.float64, TOL64), (np.float32, 1e-5)):
        with fit.Plan(ns, nd, nd, ks, deal, prec, sidx, len(c)) as plan:
            N = plan.nonlinear_term(u)
            multi = plan.spectral_fluxes_with_nhat([I[i] for i in invs], u, N)
            for inv, (T, Pi) in zip(invs, multi):
                _, To, Po = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), dea, inv)
                Ts, Ps = plan.spectral_flux_with_nhat(I[inv], u, N)
                e1, e2 = rel(T, To), rel(Pi, Po)
                say("  ", prec.__name__, inv, e1, e2)
                assert e1 < tol and e2 < tol
                assert rel(T, Ts) < 1e-12 and rel(Pi, Ps) < 1e-12


# last axis at a register FFT length: the fused last pass (k_lastdens_r) computes the densities
FUSED_CASES = [((12, 10, 128), "23", ("ke", "helicity")), ((10, 8, 192), "pad", ("ke", "helicity")),
               ((12, 12, 128), "none", ("enstrophy",)), ((10, 128), "23", ("ke", "enstrophy")),
               ((12, 192), "pad", ("ke",)), ((16, 256), "none", ("ke",))]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,dea,invs", FUSED_CASES)
def test_fused_flux_last_pass_matches_oracle(fit, oracle, ns, dea, invs, prec):
    """fit_spectral_fluxes / fit_spectral_flux with the density in the last forward pass
    (FIT_OPT_FUSED_FLUX: N̂ never stored; 3/2-padded Nyquist placement from the StNhat rule) against
    the oracle and the stored-N̂ form, for several invariants at once, non-Hermitian inputs."""
    say("case fused flux", ns, dea, invs, prec)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    rng = np.random.default_rng(len(ns) * 100 + ns[-1])
    u = rand_c(rng, ns + (nd,))
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy()}
    dt = np.complex128 if prec == "f64" else np.complex64
    tol = TOL64 if prec == "f64" else 1e-5
    from fit_amd.grid import shell_setup
    c, _, sidx = shell_setup(ks, fit.LinearBinning(1.0), None)
    from fit_amd import _lib
    with fit.Plan(ns, nd, nd, ks, deal, np.float64 if prec == "f64" else np.float32, sidx, len(c)) as plan:
        plan.set_option(_lib.OPT_FUSED_FLUX, 1)
        multi = plan.spectral_fluxes([I[i] for i in invs], u.astype(dt))
        single = plan.spectral_flux(I[invs[0]], u.astype(dt))
        plan.set_option(_lib.OPT_FUSED_FLUX, 0)
        stored = plan.spectral_fluxes([I[i] for i in invs], u.astype(dt))
    for (T, Pi), (Ts, Ps) in zip(multi, stored):
        assert rel(T, Ts) < tol and rel(Pi, Ps) < tol
    for inv, (T, Pi) in zip(invs, multi):
        _, To, Po = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), dea, inv)
        e1, e2 = rel(T, To), rel(Pi, Po)
        say("  ", inv, e1, e2)
        assert e1 < tol and e2 < tol
    _, To, _ = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), dea, invs[0])
    assert rel(single[0], To) < tol


@pytest.mark.parametrize("ns,dea,inv", [((12, 10, 128), "23", "ke"), ((10, 12, 192), "none", "helicity"),
                                        ((12, 192), "23", "enstrophy"), ((10, 256), "none", "ke")])
def test_batched_giver_loop_register_last_pass(fit, oracle, ns, dea, inv):
    """The batched giver loop with its last pass on the register kernel (pair density against the
    precomputed Hermitian carrier) against the oracle."""
    say("case batched register last pass", ns, dea, inv)
    nd = len(ns)
    ks = ks_of(oracle, ns)
    u = oracle.synthetic_field(ns, nd, seed=17)
    u = u / np.max(np.abs(u))
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds()}[dea]
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy()}[inv]
    r = fit.calculate_shell_to_shell_transfer(u, ks, binning=fit.CustomBinning(range(0, 6)), dealiasing=deal,
                                              invariant=I)
    _, _, To, _, _ = oracle.calculate_shell_to_shell_transfer(u, ks, oracle.CustomBinning(tuple(range(0, 6))), dea,
                                                               invariant=inv)
    assert np.max(np.abs(To)) > 1e-12
    e = rel(r.transfer_matrix, To)
    say("  T(n,m) err", e)
    assert e < TOL64


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_toroidal_poloidal_decomposition_and_partial_fluxes(fit, oracle, prec):
    """ToroidalPoloidalDecomposition (Decomposition.jl:107-141) on the device against the oracle:
    the two parts, the 8 partial-flux channels (SpectralFlux.jl:240-285) and the per-part fluxes
    of calculate_spectral_flux (SpectralFlux.jl:96-122)."""
    say("case toroidal/poloidal", prec)
    u, ks = oracle.curl_velocity_3d(16, 95)
    u = np.concatenate([u, 0.3 * u[..., :1]], axis=3)        # a fourth component: zero in both parts
    dt = np.complex128 if prec == "f64" else np.complex64
    tol = TOL64 if prec == "f64" else TOL32
    dec_o = oracle.toroidal_poloidal_decompose(u, ks)
    dec = fit.decompose_field(fit.ToroidalPoloidalDecomposition(), u.astype(dt), ks)
    assert list(dec) == ["toroidal", "poloidal"]
    for name in dec:
        assert rel(dec[name], dec_o[name]) < tol
    u3 = u[..., :3]
    b = fit.LinearBinning(1.0)
    ch_o, tot_o, _ = oracle.calculate_partial_fluxes(u3, ks, oracle.LinearBinning(1.0), "23",
                                                     decomposition="ToroidalPoloidalDecomposition")
    pf = fit.calculate_partial_fluxes(u3.astype(dt), ks, decomposition=fit.ToroidalPoloidalDecomposition(), binning=b)
    assert len(pf["channels"]) == 8
    scale = max(math.sqrt(np.sum(v[1] ** 2)) for v in ch_o.values())
    for key, (c, T, P) in ch_o.items():
        e = np.max(np.abs(pf["channels"][key].transfer_spectrum - T)) / scale
        assert e < (1e-10 if prec == "f64" else 2e-5), (key, e)
    per_o = oracle.calculate_spectral_flux(u3, ks, oracle.LinearBinning(1.0), "23",
                                           decomposition="ToroidalPoloidalDecomposition")
    per = fit.calculate_spectral_flux(u3.astype(dt), ks, binning=b, decomposition=fit.ToroidalPoloidalDecomposition())
    for name in ("toroidal", "poloidal"):
        assert rel(per[name].flux, per_o[name][2]) < (TOL64 if prec == "f64" else 1e-5)
