"""GPU: the physical -> spectral front-end (û = fft(u)/Np per component, the coefficient convention
of the reference's distributed front-end, ext/FlowInvariantTransferPencilFFTsExt.jl:83-88) against
numpy.fft, and the fluxes of physical inputs (pencil_spectral_flux, physical ensembles) against the
oracle on the same fields.

Tolerances (normwise): FP64 1e-12, FP32 1e-6 for the transform; flux tolerances as elsewhere."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def to_dev(a):
    import torch
    rev = tuple(reversed(range(a.ndim)))
    return torch.from_numpy(np.ascontiguousarray(a.transpose(rev))).to("cuda").permute(rev)


@pytest.mark.parametrize("ns", [(16,), (12,), (30,), (16, 12), (10, 24), (16, 12, 8), (9, 10, 12), (128, 24, 20)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_physical_to_spectral_matches_numpy(fit, oracle, ns, prec):
    say("case front-end", ns, prec)
    rng = np.random.default_rng(sum(ns))
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    u = rng.standard_normal(ns + (3,))
    ref = np.stack([np.fft.fftn(u[..., c]) / np.prod(ns) for c in range(3)], axis=-1)
    dt = np.float64 if prec == "f64" else np.float32
    tol = 1e-12 if prec == "f64" else 1e-6
    got = fit.physical_to_spectral(u.astype(dt), ks)                      # host in, host out
    assert got.shape == ref.shape and got.dtype == (np.complex128 if prec == "f64" else np.complex64)
    assert rel(got, ref) < tol
    gd = fit.physical_to_spectral(to_dev(u.astype(dt)), ks)               # device in, device out
    assert gd.is_cuda and rel(gd.cpu().numpy(), ref) < tol
    one = fit.physical_to_spectral(u[..., 0].astype(dt), ks)              # a single (ns...) field
    assert one.shape == ns and rel(one, ref[..., 0]) < tol
    z = u[..., 0] + 1j * u[..., 1]                                        # complex physical input
    zz = fit.physical_to_spectral(z.astype(np.complex128 if prec == "f64" else np.complex64), ks)
    assert rel(zz, np.fft.fftn(z) / np.prod(ns)) < tol


def test_pencil_spectral_flux_matches_oracle(fit, oracle):
    say("case pencil flux")
    for ns in ((32, 32), (16, 16, 16)):
        nd = len(ns)
        ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
        uh = oracle.synthetic_field(ns, nd, seed=99)
        u = np.stack([np.real(np.fft.ifftn(uh[..., c]) * np.prod(ns)) for c in range(nd)], axis=-1)
        _, T, Pi = oracle.calculate_spectral_flux(uh, ks, oracle.LinearBinning(1.0), "23")
        r = fit.pencil_spectral_flux(tuple(u[..., c] for c in range(nd)), ks, binning=fit.LinearBinning(1.0))
        say("  ", ns, rel(r.transfer_spectrum, T), rel(r.flux, Pi))
        assert rel(r.transfer_spectrum, T) < 1e-11 and rel(r.flux, Pi) < 1e-11
    ks = oracle.wavenumber_grid((8, 8), (2 * math.pi,) * 2)
    with pytest.raises(ValueError, match="KineticEnergy only"):
        fit.pencil_spectral_flux(np.zeros((8, 8, 2)), ks, binning=fit.LinearBinning(1.0), invariant=fit.Enstrophy())
    with pytest.raises(ValueError, match="D == nd"):
        fit.pencil_spectral_flux(np.zeros((8, 8, 3)), ks, binning=fit.LinearBinning(1.0))
    with pytest.raises(ValueError, match="isotropic"):
        fit.pencil_spectral_flux(np.zeros((8, 8, 2)), ks, binning=fit.LinearBinning(1.0),
                                 geometry=fit.PerpendicularShells())


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_physical_ensemble_matches_spectral_ensemble(fit, oracle, prec):
    import torch
    say("case physical ensemble", prec)
    shapes = [(32, 32), (64, 64), (32, 32), (128, 128)]
    dt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-11 if prec == "f64" else 1e-5
    uhs = [oracle.synthetic_field(s, 2, seed=120 + j, spectrum="k-3") for j, s in enumerate(shapes)]
    ths = [oracle.synthetic_field(s, 1, seed=220 + j, spectrum="k-1", solenoidal=False)[..., 0] for j, s in enumerate(shapes)]
    us = [np.stack([np.real(np.fft.ifftn(u[..., c]) * np.prod(s)) for c in range(2)], -1) for u, s in zip(uhs, shapes)]
    ts = [np.real(np.fft.ifftn(t) * np.prod(s)) for t, s in zip(ths, shapes)]
    kss = [oracle.wavenumber_grid(s, (2 * math.pi,) * 2) for s in shapes]
    out = fit.calculate_ensemble_fluxes([to_dev(u).to(dt) for u in us], kss, [to_dev(t).to(dt) for t in ts],
                                        binning=fit.LinearBinning(1.0), physical=True)
    for j, (ke, sc) in enumerate(out):
        _, T, Pi = oracle.calculate_spectral_flux(uhs[j], kss[j], oracle.LinearBinning(1.0), "23")
        _, Ts, Ps = oracle.calculate_scalar_flux(uhs[j], ths[j], kss[j], oracle.LinearBinning(1.0), "23")
        assert rel(ke.flux, Pi) < tol and rel(sc.flux, Ps) < tol
