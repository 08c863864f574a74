"""CPU-side checks: the C-ABI library loads and exports every declared entry point, and the
host logic (grid, shells, argument validation) matches the reference's rules."""
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fit_hip.h")).read()
    return sorted(set(re.findall(r"\b(fit_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(fit):
    lib = fit.load_library()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    assert b"gfx950" in lib.fit_version()
    # every declared symbol has a ctypes signature in the binding
    from fit_amd import _lib
    assert set(syms) == set(_lib.SIGNATURES)


def test_library_is_gfx950_code_object(fit):
    data = open(fit.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_grid_matches_oracle(fit, oracle):
    for ns in ((8,), (12, 10), (6, 8, 10)):
        a = fit.wavenumber_grid(ns, (2 * math.pi,) * len(ns))
        b = oracle.wavenumber_grid(ns, (2 * math.pi,) * len(ns))
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    ks = fit.wavenumber_grid((8, 8), (np.float32(2 * math.pi),) * 2, dtype=np.float32)
    assert ks[0].dtype == np.float32                       # Float32 propagation (runtests:779-804)


@pytest.mark.parametrize("binning", ["lin", "log", "dyadic", "custom", "lin_frac"])
@pytest.mark.parametrize("geom", [None, "perp", "par"])
def test_shell_setup_matches_oracle(fit, oracle, binning, geom):
    ns = (12, 10, 8)
    ks = fit.wavenumber_grid(ns, (2 * math.pi,) * 3)
    B = {"lin": ("LinearBinning", (1.0,)), "log": ("LogarithmicBinning", (1.0, 1.5)),
         "dyadic": ("DyadicBinning", (1.0,)), "custom": ("CustomBinning", ((0.0, 1.0, 2.5, 4.0),)),
         "lin_frac": ("LinearBinning", (2 * math.pi / 8,))}[binning]
    G = {None: (None, None), "perp": ("PerpendicularShells", ()), "par": ("ParallelShells", ())}[geom]
    fb = getattr(fit, B[0])(*B[1])
    ob = getattr(oracle, B[0])(*B[1])
    fg = getattr(fit, G[0])() if G[0] else None
    og = getattr(oracle, G[0])() if G[0] else None
    kc_f = fit.shell_coordinate(fg, ks)
    kc_o = oracle.shell_coordinate(og, ks)
    assert np.array_equal(kc_f, kc_o)
    ef = fit.shell_edges(fb, kc_f.max())
    eo = oracle.shell_edges(ob, kc_o.max())
    assert np.array_equal(ef, eo)
    assert np.allclose(fit.shell_centers(fb, kc_f.max()), oracle.shell_centers(ob, kc_o.max()))
    assert np.array_equal(fit.assign_shells(kc_f, ef), oracle.assign_shells(kc_o, eo))


def test_argument_errors_before_device(fit):
    # Helicity / enstrophy dimension guards (Invariants.jl:63, :99-101), mode guard
    # (ScaleToScaleTransfer.jl:64-67), backend checks -- all raised on the host.
    ks2 = fit.wavenumber_grid((4, 4), (2 * math.pi,) * 2)
    u2 = np.zeros((4, 4, 2), complex)
    with pytest.raises(ValueError):
        fit.calculate_spectral_flux(u2, ks2, invariant=fit.Helicity())
    with pytest.raises(ValueError):
        fit.calculate_mode_to_mode_transfer(u2, ks2, invariant=fit.Helicity())
    ks1 = fit.wavenumber_grid((8,), (2 * math.pi,))
    with pytest.raises(ValueError):
        fit.calculate_spectral_flux(np.zeros((8, 1), complex), ks1, invariant=fit.Enstrophy())
    ks3 = fit.wavenumber_grid((12, 12, 12), (2 * math.pi,) * 3)
    with pytest.raises(ValueError):
        fit.calculate_mode_to_mode_transfer(np.zeros((12, 12, 12, 3), complex), ks3)
    with pytest.raises(ValueError):
        fit.compute_nonlinear_term(u2, ks2, spectral="FFTBackend")
    with pytest.raises(ValueError):
        fit.as_component_field(np.zeros((4, 4, 1, 1)), 2)
    assert fit.as_component_field(np.zeros((4, 4)), 2).shape == (4, 4, 1)
    with pytest.raises(fit.DimensionMismatch):
        fit.calculate_spectral_flux(np.zeros((4, 4), complex), ks2)


def test_missing_library_fails_loudly(fit, tmp_path):
    from fit_amd import _lib
    with pytest.raises(RuntimeError):
        _lib.load_library(str(tmp_path / "nope.so"))


def test_batch_map_reductions(fit):
    # MPIExt.jl:25-37: :gather keeps input order; :sum / :mean fold with +; a binary function
    # folds with it; anything else is an ArgumentError (ValueError here).
    items = [np.arange(3.0) + i for i in range(4)]
    assert [list(x) for x in fit.batch_map(lambda a: 2 * a, items)] == [list(2 * a) for a in items]
    assert [list(x) for x in fit.batch_map(lambda a: a, items, reduce=":gather")] == [list(a) for a in items]
    assert np.array_equal(fit.batch_map(lambda a: a, items, reduce=":sum"), np.array([6.0, 10.0, 14.0]))
    assert np.array_equal(fit.batch_map(lambda a: a, items, reduce="mean"), np.array([1.5, 2.5, 3.5]))
    assert np.array_equal(fit.batch_map(lambda a: a, items, reduce=np.maximum), items[-1])
    assert fit.batch_map(lambda i: i, range(5), reduce=lambda a, b: a * 10 + b) == 1234
    with pytest.raises(ValueError, match="reduce must be"):
        fit.batch_map(lambda a: a, items, reduce=":max")
