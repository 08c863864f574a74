"""The HIP path, called through the C ABI by the host mirror of the reference's API, against the
committed golden vectors (tests/golden/*.npz: DirectSum / brute-force / one-mode-loop values).

Tolerances (normwise, max|x - x_ref| / max|x_ref|): FP64 compute 1e-11, FP32 compute 2e-5 for
per-mode arrays and 1e-5 for shell sums.
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))
NAMES = sorted(MANIFEST)


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64 if not np.iscomplexobj(a) else np.complex128)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_golden(fit, name, prec):
    print("case golden", name, prec, flush=True)
    meta, z = MANIFEST[name], load(name)
    ns = tuple(meta["ns"])
    nd = len(ns)
    ks = fit.wavenumber_grid(ns, (2 * math.pi,) * nd)
    dt = np.complex128 if prec == "f64" else np.complex64
    tol_mode, tol_shell = (1e-11, 1e-11) if prec == "f64" else (2e-5, 1e-5)
    u = z["u_hat"].astype(dt)
    th = z["theta_hat"].astype(dt) if "theta_hat" in z else None
    if "u_plus" in z:
        h = fit.decompose_field(fit.HelicalDecomposition(), u, ks)
        assert rel(h["positive"], z["u_plus"]) < tol_mode
        assert rel(h["negative"], z["u_minus"]) < tol_mode
        return
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(),
            "pad": fit.PaddedThreeHalves()}[meta["dealias"]]
    inv = {"ke": fit.KineticEnergy(), "scalar": fit.PassiveScalar(), "helicity": fit.Helicity(),
           "enstrophy": fit.Enstrophy()}[meta["invariant"]]
    if "S" in z:
        if th is not None:
            r = fit.calculate_scalar_mode_to_mode_transfer(u, th, ks, dealiasing=deal)
        else:
            r = fit.calculate_mode_to_mode_transfer(u, ks, invariant=inv, dealiasing=deal)
        e_s, e_n = rel(r.transfer, z["S"]), rel(r.net_transfer, z["net"])
        print("  S err", e_s, "net err", e_n, flush=True)
        assert e_s < tol_mode and e_n < tol_mode
        return
    b = fit.LinearBinning(1.0)
    if th is not None:
        N = fit.compute_nonlinear_term(th.reshape(ns + (1,)), ks, dealiasing=deal, advecting_hat=u)
        r = fit.calculate_scalar_flux(u, th, ks, binning=b, dealiasing=deal)
    else:
        N = fit.compute_nonlinear_term(u, ks, dealiasing=deal)
        r = fit.calculate_spectral_flux(u, ks, binning=b, dealiasing=deal, invariant=inv)
    e_n, e_t, e_p = rel(N, z["N_hat"]), rel(r.transfer_spectrum, z["T"]), rel(r.flux, z["Pi"])
    print("  N err", e_n, "T err", e_t, "Pi err", e_p, flush=True)
    assert e_n < tol_mode and e_t < tol_shell and e_p < tol_shell
    assert len(r.k_shells) == len(z["centers"]) and rel(r.k_shells, z["centers"]) < 1e-6
    for extra, I in (("helicity", fit.Helicity()), ("enstrophy", fit.Enstrophy())):
        if "T_" + extra in z:
            r2 = fit.calculate_spectral_flux(u, ks, binning=b, dealiasing=deal, invariant=I)
            e = rel(r2.flux, z["Pi_" + extra])
            print(" ", extra, "Pi err", e, flush=True)
            assert e < tol_shell
    if "Tnm" in z:
        if th is not None:
            s = fit.calculate_scalar_shell_to_shell_transfer(u, th, ks, binning=b, dealiasing=deal)
        else:
            s = fit.calculate_shell_to_shell_transfer(u, ks, binning=b, dealiasing=deal, invariant=inv)
        e = rel(s.transfer_matrix, z["Tnm"])
        print("  T(n,m) err", e, flush=True)
        assert e < tol_shell
        assert rel(s.net_transfer, z["Tnm"].sum(axis=1)) < tol_shell
