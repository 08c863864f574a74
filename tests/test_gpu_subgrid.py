"""GPU: the giver loop on the band-limited sub-grid against the full-grid oracle.

Shell sets confined to |k_d| <= R_d with 4 R_d < n_d make the plan attach a grid of
L_d >= 4 R_d + 1 points (2^a or 3·2^a); T(n,m) from it must match the oracle's full-grid giver
loop (FFTWExt.jl:209-295) to the same tolerances as every other parity test, and the plan's own
grid (sub-grid switched off) to FP64 roundoff.
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


DEAL = {"none": "NoDealiasing", "23": "OrszagTwoThirds", "pad": "PaddedThreeHalves"}
INV = {"ke": "KineticEnergy", "helicity": "Helicity", "enstrophy": "Enstrophy", "scalar": "PassiveScalar"}

# (grid, dealiasing, shells CustomBinning(0:kcut), invariant, expected sub-grid)
CASES = [
    ((48, 48, 48), "23", 7, "ke", (32, 32, 32)),
    ((32, 32, 32), "none", 5, "ke", (24, 24, 24)),
    ((32, 32, 32), "pad", 5, "ke", (24, 24, 24)),
    ((48, 48, 48), "23", 6, "helicity", (24, 24, 24)),
    ((64, 48, 40), "23", 5, "enstrophy", (24, 24, 24)),
    ((96, 96), "23", 10, "ke", (48, 48)),
    ((64, 64), "pad", 8, "enstrophy", (32, 32)),
    ((256, 256), "23", 32, "ke", (128, 128)),
    ((512, 512), "23", 48, "ke", (192, 192)),
]


def field(oracle, ns, seed):
    if len(ns) == 2:
        if ns[0] == ns[1]:
            u, _ = oracle.streamfunction_velocity_2d(ns[0], seed)
            return u
        return oracle.synthetic_field(ns, 2, seed=seed)
    return oracle.synthetic_field(ns, 3, seed=seed)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ns,dea,kcut,inv,sub", CASES)
def test_subgrid_shell_to_shell(fit, oracle, ns, dea, kcut, inv, sub, prec):
    say("case subgrid s2s", ns, dea, kcut, inv, prec)
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    u = field(oracle, ns, 60 + kcut)
    c, e, sidx = fit.grid.shell_setup(ks, fit.CustomBinning(range(0, kcut + 1)), fit.IsotropicShells())
    assert len(c) == kcut
    T = oracle.shell_to_shell_columns(u, ks, sidx, kcut, list(range(1, kcut + 1)), dea, inv, workers=8)
    assert np.max(np.abs(T)) > 0
    if inv == "ke":   # conservative triads: a non-degenerate T is antisymmetric (runtests.jl:371-409)
        assert np.max(np.abs(T + T.T)) < 1e-9 * np.max(np.abs(T))
    dt = np.float64 if prec == "f64" else np.float32
    cdt = np.complex128 if prec == "f64" else np.complex64
    I = getattr(fit, INV[inv])()
    with fit.Plan(ns, nd, nd, ks, getattr(fit, DEAL[dea])(), dt, sidx, kcut) as plan:
        assert plan.subgrid == sub
        Tg, net, asym = plan.shell_to_shell(I, u.astype(cdt))
        e_sub = rel(Tg, T)
        say("  sub-grid", plan.subgrid, "T(n,m) err", e_sub)
        assert e_sub < (TOL64 if prec == "f64" else TOL32)
        assert rel(net, T.sum(axis=1)) < (TOL64 if prec == "f64" else TOL32) * 10
        if prec == "f64":
            plan.set_subgrid(False)
            Tf, _, _ = plan.shell_to_shell(I, u.astype(cdt))
            say("  own grid vs sub-grid", rel(Tg, Tf))
            assert rel(Tg, Tf) < 1e-12


def test_subgrid_scalar_with_advecting_field(fit, oracle):
    say("case subgrid scalar s2s")
    ns = (48, 48, 48)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 3)
    u = oracle.synthetic_field(ns, 3, seed=71)
    th = oracle.synthetic_field(ns, 1, seed=72, solenoidal=False)
    c, e, sidx = fit.grid.shell_setup(ks, fit.CustomBinning(range(0, 8)), fit.IsotropicShells())
    T = oracle.shell_to_shell_columns(th, ks, sidx, 7, list(range(1, 8)), "23", "scalar", u)
    with fit.Plan(ns, 1, 3, ks, fit.OrszagTwoThirds(), np.float64, sidx, 7) as plan:
        assert plan.subgrid == (32, 32, 32)
        Tg, _, _ = plan.shell_to_shell(fit.PassiveScalar(), th, adv=u)
    say("  T(n,m) err", rel(Tg, T))
    assert np.max(np.abs(T)) > 0 and rel(Tg, T) < TOL64


def test_no_subgrid_when_shells_reach_quarter_grid(fit, oracle):
    # default LinearBinning covers the whole spectrum: the plan keeps its own grid
    ns = (16, 16, 16)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 3)
    c, e, sidx = fit.grid.shell_setup(ks, fit.LinearBinning(1.0), fit.IsotropicShells())
    with fit.Plan(ns, 3, 3, ks, fit.OrszagTwoThirds(), np.float64, sidx, len(c)) as plan:
        assert plan.subgrid is None


def test_giver_classes_on_smaller_grids(fit, oracle):
    """Givers of small support run on smaller grids (L_d >= 2 R_d + 2 P_d + 1): the 64³ plan with
    shells 0:9 (R = 8) splits its givers into classes on 24³, 32³ and 48³; every class's columns
    equal the full-grid oracle and the plan's own grid."""
    from fit_amd.grid import shell_setup
    ns = (64, 64, 64)
    say("case giver classes", ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 3)
    u = oracle.synthetic_field(ns, 3, seed=23)
    c, _, sidx = shell_setup(ks, fit.CustomBinning(range(0, 10)), None)
    for prec, tol in ((np.float64, TOL64), (np.float32, TOL32)):
        with fit.Plan(ns, 3, 3, ks, fit.OrszagTwoThirds(), prec, sidx, len(c)) as plan:
            cls = plan.subgrid_classes
            say("  classes", cls)
            assert [k[0] for k in cls] == [(24,) * 3, (32,) * 3, (48,) * 3]
            assert cls[0][1][0] == 1 and cls[-1][1][1] == len(c)
            assert all(cls[i][1][1] + 1 == cls[i + 1][1][0] for i in range(len(cls) - 1))
            assert cls[0][0][0] < cls[-1][0][0] and plan.subgrid == cls[-1][0]
            Tm, _, _ = plan.shell_to_shell(fit.KineticEnergy(), u.astype(np.complex128 if prec == np.float64 else np.complex64))
            plan.set_subgrid(False)
            To, _, _ = plan.shell_to_shell(fit.KineticEnergy(), u.astype(np.complex128 if prec == np.float64 else np.complex64))
        _, _, Tr, _, _ = oracle.calculate_shell_to_shell_transfer(u, ks, oracle.CustomBinning(tuple(range(0, 10))), "23")
        e1, e2 = rel(Tm, Tr), rel(Tm, To)
        say("  ", prec.__name__, "vs oracle", e1, "vs own grid", e2)
        assert e1 < tol and e2 < tol
