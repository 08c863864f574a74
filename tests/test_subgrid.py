"""The band-limited sub-grid identity the giver loop relies on, checked on the oracle (CPU).

When every shell lies in |k_d| <= R_d with 4 R_d < n_d, T(n,m) only sums triads k = p + q with
|p_d|, |k_d| <= R_d and |q_d| <= 2 R_d.  Copying that box of the input onto a grid of
L_d >= |q|+|p|+|k|+1 points (no dealiasing there), running the same giver loop and rescaling by
(Np_sub/Np)^2 (N̂ carries 1/Np^2 of the grid it is taken on) must reproduce the full-grid T(n,m)
of every dealiasing mode.  The library's plan builds exactly this child grid; these checks pin the
construction independently of any GPU code.
"""
import math

import numpy as np
import pytest


def sub_len(need):
    cands = [8 * 2 ** a for a in range(20)] + [12 * 2 ** a for a in range(20)]
    return min(c for c in cands if c >= need)


def fold(i, n):
    return i if 2 * i <= n else i - n


def subgrid_T(oracle, u, ks, sidx, nsh, dealiasing, invariant="ke", adv=None):
    nd = len(ks)
    ns = u.shape[:nd]
    band_in = [n // 3 - 1 if dealiasing == "23" else n // 2 for n in ns]
    nz = np.nonzero(sidx > 0)
    R = [min(max(min(int(i), n - int(i)) for i in nz[d]), band_in[d]) for d, n in enumerate(ns)]
    assert all(4 * r < n for r, n in zip(R, ns))
    Q = [min(2 * r, band_in[d]) if dealiasing == "23" else 2 * r for d, r in enumerate(R)]
    Ls = [min(sub_len(q + 2 * r + 1), n) for q, r, n in zip(Q, R, ns)]
    kfold = [np.array([fold(i, L) for i in range(L)]) for L in Ls]
    src = [np.where(kf >= 0, kf, kf + n) for kf, n in zip(kfold, ns)]
    inq = np.ones(Ls, dtype=bool)
    inr = np.ones(Ls, dtype=bool)
    for d in range(nd):
        shp = [1] * nd
        shp[d] = Ls[d]
        inq &= (np.abs(kfold[d]) <= Q[d]).reshape(shp)
        inr &= (np.abs(kfold[d]) <= R[d]).reshape(shp)
    grid = np.ix_(*src)

    def gather(a):
        out = a[grid + (slice(None),)].copy()
        out[~inq] = 0
        return out

    ks_c = [np.asarray(k)[s] for k, s in zip(ks, src)]
    sidx_c = np.where(inr, sidx[grid], 0)
    u_c = gather(u)
    a_c = gather(adv) if adv is not None else None
    T = oracle.shell_to_shell_columns(u_c, ks_c, sidx_c, nsh, list(range(1, nsh + 1)), "none",
                                      invariant, a_c)
    return T * (np.prod(Ls) / np.prod(ns)) ** 2, Ls


CASES = [
    ((48, 48, 48), "23", 7, "ke"),
    ((32, 32, 32), "none", 5, "ke"),
    ((32, 32, 32), "pad", 5, "ke"),
    ((40, 40, 40), "23", 6, "helicity"),
    ((96, 96), "23", 10, "ke"),
    ((64, 64), "pad", 8, "enstrophy"),
]


@pytest.mark.parametrize("ns,dealias,kcut,inv", CASES)
def test_subgrid_equals_full_grid(oracle, ns, dealias, kcut, inv):
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    # T carries 1/Np^2: scale the field so that T(n,m) is O(1)
    u = oracle.synthetic_field(ns, nd, seed=40 + kcut, spectrum="k-5/3") * np.prod(ns) ** (2 / 3)
    _, _, sidx = oracle.shell_setup(ks, oracle.CustomBinning(tuple(range(0, kcut + 1))))
    nsh = kcut
    full = oracle.shell_to_shell_columns(u, ks, sidx, nsh, list(range(1, nsh + 1)), dealias, inv)
    sub, Ls = subgrid_T(oracle, u, ks, sidx, nsh, dealias, inv)
    assert np.prod(Ls) * 2 <= np.prod([oracle.padded_len(n) if dealias == "pad" else n for n in ns])
    assert np.max(np.abs(full)) > 1e-8
    assert np.max(np.abs(sub - full)) <= 1e-12 * np.max(np.abs(full))


def test_subgrid_scalar_with_advecting_field(oracle):
    ns = (48, 48, 48)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 3)
    u = oracle.synthetic_field(ns, 3, seed=71) * np.prod(ns) ** (2 / 3)
    th = oracle.synthetic_field(ns, 1, seed=72, solenoidal=False) * np.prod(ns) ** (2 / 3)
    _, _, sidx = oracle.shell_setup(ks, oracle.CustomBinning(tuple(range(0, 8))))
    full = oracle.shell_to_shell_columns(th, ks, sidx, 7, list(range(1, 8)), "23", "scalar", u)
    sub, _ = subgrid_T(oracle, th, ks, sidx, 7, "23", "scalar", adv=u)
    assert np.max(np.abs(full)) > 1e-8
    assert np.max(np.abs(sub - full)) <= 1e-12 * np.max(np.abs(full))


def test_subgrid_too_small_fails(oracle):
    # planted mistake in the construction: one point too few per axis aliases the receivers
    ns = (48, 48, 48)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 3)
    u = oracle.synthetic_field(ns, 3, seed=47) * np.prod(ns) ** (2 / 3)
    _, _, sidx = oracle.shell_setup(ks, oracle.CustomBinning(tuple(range(0, 8))))
    full = oracle.shell_to_shell_columns(u, ks, sidx, 7, list(range(1, 8)), "23", "ke")
    # R = 6, Q = 12 -> the alias-free bound is 25 points; 24 lets q + p wrap onto receivers
    L = 24
    kf = np.array([fold(i, L) for i in range(L)])
    src = np.where(kf >= 0, kf, kf + 48)
    grid = np.ix_(src, src, src)
    inq = np.ones((L,) * 3, bool)
    inr = np.ones((L,) * 3, bool)
    for d in range(3):
        shp = [1, 1, 1]
        shp[d] = L
        inq &= (np.abs(kf) <= 12).reshape(shp)
        inr &= (np.abs(kf) <= 6).reshape(shp)
    u_c = u[grid + (slice(None),)].copy()
    u_c[~inq] = 0
    T = oracle.shell_to_shell_columns(u_c, [np.asarray(k)[src] for k in ks], np.where(inr, sidx[grid], 0),
                                      7, list(range(1, 8)), "none", "ke")
    T *= (L ** 3 / 48 ** 3) ** 2
    assert np.max(np.abs(T - full)) > 1e-6 * np.max(np.abs(full))
