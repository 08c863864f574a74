"""Committed golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU checks:
  * the fixtures hold the identities the mathematics fixes, read from the stored arrays alone:
    Π = cumsum(T); Σ_k T(k) = 0 where the reference's tests expect conservation; T(n,m)
    antisymmetric with Σ_m T(n,m) = T(k); S(k|p) antisymmetric with Σ_p S(k|p) = net(k);
    u+ + u- = û with u± divergence-free;
  * the FP64 FFT restatement (oracle) reproduces every stored expected array to 1e-12.
The HIP path is checked against the same fixtures in test_gpu_golden.py.
"""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))
NAMES = sorted(MANIFEST)


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def ks_of(oracle, ns):
    return oracle.wavenumber_grid(tuple(ns), (2 * math.pi,) * len(ns))


def test_manifest_lists_every_fixture():
    files = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert files == NAMES
    for name in NAMES:
        assert sorted(load(name)) == MANIFEST[name]["arrays"]


@pytest.mark.parametrize("name", NAMES)
def test_fixture_identities(name):
    meta, z = MANIFEST[name], load(name)
    for tk, pk in (("T", "Pi"), ("T_helicity", "Pi_helicity"), ("T_enstrophy", "Pi_enstrophy")):
        if tk in z:
            assert rel(np.cumsum(z[tk]), z[pk]) < 1e-13
            # conservation (runtests.jl:175-200): dealiased or alias-free fields with no
            # Nyquist content; the NoDealiasing / Nyquist-populated cases carry aliasing error,
            # and 3D enstrophy is not an invariant (vortex stretching)
            conserved = meta["dealias"] == "23" or name == "ke3d_8_padded"
            if conserved and not (tk == "T_enstrophy" and len(meta["ns"]) == 3):
                assert abs(z[tk].sum()) < 1e-10 * np.abs(z[tk]).sum()
    if "Tnm" in z:
        Tm = z["Tnm"]
        nrm = np.sqrt(np.sum(Tm ** 2))
        if meta["dealias"] == "23":
            # runtests.jl:371-438; without dealiasing the corner modes at |k| = k_max sit in no
            # shell (ShellBinning.jl:126-140), so they advect in T(k) but give no column
            assert np.max(np.abs(Tm + Tm.T)) < 1e-12 * nrm
            assert rel(Tm.sum(axis=1), z["T"]) < 1e-12
    if "S" in z:
        S = z["S"]
        Np = int(np.prod(meta["ns"]))
        S2 = S.reshape(Np, Np)
        assert rel(S2.sum(axis=1), z["net"].reshape(-1)) < 1e-12
        if meta["invariant"] in ("ke", "scalar") and meta["dealias"] == "23":
            assert np.max(np.abs(S2 + S2.T)) < 1e-12 * np.max(np.abs(S2))   # runtests.jl:1072-1111
    if "u_plus" in z:
        assert rel(z["u_plus"] + z["u_minus"], z["u_hat"]) < 1e-13
        ks = [np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n).astype(float)
              for n in meta["ns"]]
        mesh = np.meshgrid(*ks, indexing="ij")
        for key in ("u_plus", "u_minus"):
            div = sum(mesh[j] * z[key][..., j] for j in range(3))
            assert np.max(np.abs(div)) < 1e-13 * np.max(np.abs(z[key]))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(oracle, name):
    meta, z = MANIFEST[name], load(name)
    ns = tuple(meta["ns"])
    ks = ks_of(oracle, ns)
    u = z["u_hat"]
    th = z.get("theta_hat")
    if "u_plus" in z:
        h = oracle.helical_decompose(u, ks)
        assert rel(h["positive"], z["u_plus"]) < 1e-12 and rel(h["negative"], z["u_minus"]) < 1e-12
        return
    dea, inv = meta["dealias"], meta["invariant"]
    f, adv = (th, u) if th is not None else (u, None)
    if "S" in z:
        net, S = oracle.calculate_mode_to_mode_transfer(f, ks, inv, dea, advecting_hat=adv)
        assert rel(S, z["S"]) < 1e-12 and rel(net, z["net"]) < 1e-12
        return
    N = oracle.compute_nonlinear_term(f, ks, dea, advecting_hat=adv)
    assert rel(N, z["N_hat"]) < 1e-12
    b = oracle.LinearBinning(1.0)
    _, T, Pi = oracle.calculate_spectral_flux(f, ks, b, dea, inv, advecting_hat=adv)
    assert rel(T, z["T"]) < 1e-12 and rel(Pi, z["Pi"]) < 1e-12
    for extra in ("helicity", "enstrophy"):
        if "T_" + extra in z:
            _, T, _ = oracle.calculate_spectral_flux(u, ks, b, dea, extra)
            assert rel(T, z["T_" + extra]) < 1e-12
    if "Tnm" in z:
        Tm = oracle.calculate_shell_to_shell_transfer(f, ks, b, dea, invariant=inv, advecting_hat=adv)[2]
        assert rel(Tm, z["Tnm"]) < 1e-12


def _padded_brute_force(u, ks, nyq_freq_sign):
    """Exact 3/2-padded nonlinear term of a 2D field by integer-wavevector convolution, with no
    FFT: every original index i is a frequency f(i) in [-n/2, n/2) (the fftshift embedding,
    FFTWExt.jl:134-141, puts index n/2 at -n/2; nyq_freq_sign=+1 plants the mistake of putting it
    at +n/2).  The derivative factor is i·ks[j][i] at the original index (Utils.jl:59: +n/2 at
    the Nyquist index).  real() of the padded inverse transform is the Hermitian part
    H(q) = (E(q) + conj E(-q))/2 of the embedded spectrum E; the product's transform, truncated
    back to [-n/2, n/2) and scaled by Mtot/Ntot² (FFTWExt.jl:144-149, :167), is
    N̂(k) = Ntot^-2 Σ_{p+q≡k (mod Mp)} Σ_j Hu_j(q) Hg_jc(p) on the padded length Mp = 12: with
    Nyquist content |p+q| reaches n, which wraps onto the kept bin -n/2."""
    n = u.shape[0]
    M = u.shape[2]
    h = n // 2
    Mp = 12 if n == 8 else None      # FFTWExt.jl:128-131 for n = 8 (3n/2, even excess)
    assert Mp is not None

    def freq(i):
        f = i if i < h else i - n
        return -f if (i == h and nyq_freq_sign > 0) else f

    def herm(E):            # dict freq -> value  ->  Hermitian part on [-n/2, n/2]
        H = {}
        for q, v in E.items():
            H[q] = H.get(q, 0) + 0.5 * v
            mq = (-q[0], -q[1])
            H[mq] = H.get(mq, 0) + 0.5 * np.conj(v)
        return H

    Hu = [herm({(freq(a), freq(b)): u[a, b, j] for a in range(n) for b in range(n)}) for j in range(2)]
    out = np.zeros_like(u)
    for c in range(M):
        for j in range(2):
            kj = ks[j]
            Hg = herm({(freq(a), freq(b)): 1j * (kj[a] if j == 0 else kj[b]) * u[a, b, c]
                       for a in range(n) for b in range(n)})
            for q, hu in Hu[j].items():
                for p, hg in Hg.items():
                    k = tuple((p[d] + q[d] + Mp // 2) % Mp - Mp // 2 for d in range(2))
                    if -h <= k[0] < h and -h <= k[1] < h:
                        out[k[0] % n, k[1] % n, c] += hu * hg
    return out / float(n * n) ** 2


def test_padded_nyquist_fixture_by_brute_force(oracle):
    """ke2d_8_padded_nyquist (written by the oracle's FFT restatement) equals the FFT-free
    integer convolution with the reference's Nyquist placement, and the planted mistake (Nyquist
    bin at +n/2) is told apart."""
    g = load("ke2d_8_padded_nyquist")
    u = g["u_hat"]
    ks = ks_of(oracle, u.shape[:2])
    assert np.max(np.abs(u[4, :, :])) > 1e-3 and np.max(np.abs(u[:, 4, :])) > 1e-3  # Nyquist populated
    bf = _padded_brute_force(u, ks, -1)
    assert rel(g["N_hat"], bf) < 1e-12
    assert rel(oracle.compute_nonlinear_term(u, ks, "pad"), bf) < 1e-12
    wrong = _padded_brute_force(u, ks, +1)
    assert rel(wrong, bf) > 1e-3
