"""Smooth band-to-band transfer T(K,Q) on the device (fit_band_to_band through the C ABI) against
the FP64 oracle (itself pinned by the DirectSum twin and the bilinear identity T = W S Wᵀ with the
mode-to-mode closed form), plus the reference test's properties (runtests.jl:661-681).

Tolerances (normwise): FP64 compute 1e-11, FP32 compute 1e-5.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


CASES = [((16, 16), "23", "ke", (1.0, 2.0, 3.0, 4.0), 0.5, None, False),
         ((16, 16), "none", "enstrophy", (1.0, 2.0, 4.0), 0.6, None, False),
         ((16, 16), "23", "scalar", (1.0, 2.0, 3.0, 4.0), 0.5, None, True),
         ((12, 12, 12), "23", "helicity", (1.0, 2.0, 3.0), 0.6, None, False),
         ((12, 10, 8), "pad", "ke", (1.0, 1.5, 2.5, 4.0), 0.4, "perp", False),
         ((128, 128), "23", "ke", (2.0, 4.0, 8.0, 16.0, 32.0), 0.6, None, False)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,dea,inv,centers,lw,geom,scalar", CASES)
def test_band_to_band_matches_oracle(fit, oracle, ns, dea, inv, centers, lw, geom, scalar, prec):
    print("case band", ns, dea, inv, len(centers), geom, scalar, prec, flush=True)
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    if nd == 2 and ns[0] == ns[1]:
        u, _ = oracle.streamfunction_velocity_2d(ns[0], 71)
    else:
        u = oracle.synthetic_field(ns, nd, seed=13)
    th = oracle.synthetic_field(ns, 1, seed=14, spectrum="k-1", solenoidal=False) if scalar else None
    g_o = oracle.PerpendicularShells() if geom == "perp" else None
    g_f = fit.PerpendicularShells() if geom == "perp" else None
    f, adv = (th, u) if scalar else (u, None)
    _, To, neto, asymo = oracle.calculate_band_to_band_transfer(
        f, ks, oracle.SmoothBands(centers, lw), dea, inv, advecting_hat=adv, geometry=g_o)
    assert np.max(np.abs(To)) > 1e-12
    dt = np.complex128 if prec == "f64" else np.complex64
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    I = {"ke": fit.KineticEnergy(), "helicity": fit.Helicity(), "enstrophy": fit.Enstrophy(),
         "scalar": fit.PassiveScalar()}[inv]
    r = fit.calculate_band_to_band_transfer(
        f.astype(dt), ks, bands=fit.SmoothBands(centers, lw), dealiasing=deal, invariant=I,
        advecting_hat=None if adv is None else adv.astype(dt), geometry=g_f)
    e = rel(r.transfer_matrix, To)
    print("  T err", e, flush=True)
    tol = 1e-11 if prec == "f64" else 1e-5
    assert e < tol and rel(r.net_transfer, neto) < tol
    assert r.transfer_matrix.dtype == (np.float64 if prec == "f64" else np.float32)
    if dea == "23" and inv in ("ke", "scalar"):
        T = r.transfer_matrix.astype(np.float64)
        Tn = np.sqrt(np.sum(T ** 2))
        assert r.max_antisymmetry_error < (1e-10 if prec == "f64" else 1e-5) * Tn
        assert abs(T.sum()) < (1e-10 if prec == "f64" else 1e-5) * Tn


def test_band_to_band_device_arrays_and_errors(fit, oracle):
    import torch
    ns = (16, 16)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 2)
    u, _ = oracle.streamfunction_velocity_2d(16, 71)
    b = fit.SmoothBands((1.0, 2.0, 3.0, 4.0), 0.5)
    host = fit.calculate_band_to_band_transfer(u, ks, bands=b)
    ud = torch.from_numpy(np.ascontiguousarray(u.transpose(2, 1, 0))).to("cuda").permute(2, 1, 0)
    dev = fit.calculate_band_to_band_transfer(ud, ks, bands=b)
    assert rel(dev.transfer_matrix, host.transfer_matrix) < 1e-13
    with pytest.raises(ValueError):
        fit.calculate_band_to_band_transfer(u, ks, bands=fit.SmoothBands((2.0, 1.0), 0.5))
    with pytest.raises(ValueError):
        fit.calculate_band_to_band_transfer(u, ks, bands=b, invariant=fit.Helicity())
