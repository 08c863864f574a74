"""GPU parity of the compile-time-length FFT kernels (lengths 256..3072, radix 8/4/2/3) in each
axis position: x (fused contiguous-axis pass), y and z (strided column passes), FP64 and FP32,
both forms of the fused x pass.  Inputs are random complex fields (non-degenerate N̂)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(float(np.max(np.abs(b))), 1e-300))


def say(*a):
    print(*a, flush=True)


SHAPES = [(128, 6), (192, 6), (6, 128), (6, 192), (4, 4, 128), (4, 4, 192), (256, 6), (384, 6), (512, 4), (768, 4), (6, 256), (6, 384), (4, 512), (4, 768),
          (4, 4, 256), (4, 4, 384), (4, 4, 512), (4, 4, 768), (1024, 4), (4, 1024), (1536, 2), (2, 1536)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("xblock", [0, 1])
@pytest.mark.parametrize("ns", SHAPES)
def test_register_lengths_nonlinear(fit, oracle, ns, prec, xblock):
    from fit_amd import _lib
    say("case lengths", ns, prec, xblock)
    rng = np.random.default_rng(sum(ns))
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    u = rng.standard_normal(ns + (nd,)) + 1j * rng.standard_normal(ns + (nd,))
    ref = oracle.compute_nonlinear_term(u, ks, "none")
    dt = np.float64 if prec == "f64" else np.float32
    with fit.Plan(ns, nd, nd, ks, fit.NoDealiasing(), dt) as plan:
        plan.set_option(_lib.OPT_XPASS_BLOCK, xblock)
        got = plan.nonlinear_term(u)
    e = rel(got, ref)
    say("  err", e)
    assert e < (1e-11 if prec == "f64" else 2e-5)


# every axis a register length: the shared-transform gradient passes and the self-advection
# fused x pass (u formed in registers) are used; xself = 0 stores u to device memory instead
SELF_CASES = [((128, 192), "23"), ((192, 128), "23"), ((256, 384), "23"), ((384, 256), "23"),
              ((512, 128), "23"), ((768, 128), "23"), ((1024, 128), "23"), ((1536, 128), "23"),
              ((2048, 128), "23"), ((3072, 128), "23"), ((128, 3072), "23"), ((128, 128, 128), "23"),
              ((128, 128), "pad"), ((256, 128), "pad"), ((512, 128), "pad"), ((1024, 128), "pad"),
              ((2048, 128), "pad"), ((128, 128, 128), "pad")]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("xself", [1, 0])
@pytest.mark.parametrize("ns,dea", SELF_CASES)
def test_self_advection_register_lengths(fit, oracle, ns, dea, xself, prec):
    from fit_amd import _lib
    say("case self-advection lengths", ns, dea, xself, prec)
    rng = np.random.default_rng(sum(ns) + 7)
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    u = rng.standard_normal(ns + (nd,)) + 1j * rng.standard_normal(ns + (nd,))
    ref = oracle.compute_nonlinear_term(u, ks, dea)
    deal = {"23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    dt = np.float64 if prec == "f64" else np.float32
    with fit.Plan(ns, nd, nd, ks, deal, dt) as plan:
        plan.set_option(_lib.OPT_XPASS_SELF, xself)
        got = plan.nonlinear_term(u)
    e = rel(got, ref)
    say("  err", e)
    assert e < (1e-11 if prec == "f64" else 2e-5)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("xself", [1, 0])
@pytest.mark.parametrize("dea", ["23", "pad"])
def test_self_advection_three_components_on_a_plane(fit, oracle, dea, xself, prec):
    """A 2D grid carrying three velocity components (M = 3 > nd = 2): the first two advect all
    three, (u·∇)f_c with u = (f_1, f_2), as the oracle's self-advection does."""
    from fit_amd import _lib
    say("case self-advection 2D3C", dea, xself, prec)
    ns = (256, 128)
    rng = np.random.default_rng(91)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 2)
    u = rng.standard_normal(ns + (3,)) + 1j * rng.standard_normal(ns + (3,))
    ref = oracle.compute_nonlinear_term(u, ks, dea)
    deal = {"23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    dt = np.float64 if prec == "f64" else np.float32
    with fit.Plan(ns, 3, 2, ks, deal, dt) as plan:
        plan.set_option(_lib.OPT_XPASS_SELF, xself)
        got = plan.nonlinear_term(u)
    e = rel(got, ref)
    say("  err", e)
    assert e < (1e-11 if prec == "f64" else 2e-5)


# LDS-staged FP32 self-advection pass (x lines of 512 and 768 points, one row pair per wave):
# odd row counts leave the last pair half empty and row-pair counts that are not a multiple of 4
# leave waves of the last workgroup without rows
STAGE_CASES = [((512, 42), 2, "pad"), ((512, 128), 2, "23"), ((768, 10), 2, "23"), ((512, 42), 3, "pad"),
               ((512, 20, 6), 3, "pad"), ((512, 24, 8), 3, "23"), ((768, 6, 5), 3, "23")]


@pytest.mark.parametrize("stage", [1, 0])
@pytest.mark.parametrize("ns,M,dea", STAGE_CASES)
def test_staged_self_advection_pass(fit, oracle, ns, M, dea, stage):
    from fit_amd import _lib
    say("case staged self-advection", ns, M, dea, stage)
    rng = np.random.default_rng(sum(ns) + M)
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    u = rng.standard_normal(ns + (M,)) + 1j * rng.standard_normal(ns + (M,))
    ref = oracle.compute_nonlinear_term(u, ks, dea)
    deal = {"23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    with fit.Plan(ns, M, nd, ks, deal, np.float32) as plan:
        plan.set_option(_lib.OPT_XPASS_STAGE, stage)
        got = plan.nonlinear_term(u)
        again = plan.nonlinear_term(u)
    e = rel(got, ref)
    say("  err", e)
    assert e < 2e-5
    assert np.array_equal(got, again)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("split", [1, 0])
@pytest.mark.parametrize("ns,dea", [((128, 96, 64), "pad"), ((96, 128, 64), "23"), ((64, 64, 128), "23")])
def test_gradient_pass_split(fit, oracle, ns, dea, split, prec):
    """The y transform of the last-axis derivative as its own column pass (split = 1) or inside
    the two-line gradient kernel (split = 0): both against the oracle."""
    from fit_amd import _lib
    say("case gradient split", ns, dea, split, prec)
    rng = np.random.default_rng(sum(ns) + split)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 3)
    u = rng.standard_normal(ns + (3,)) + 1j * rng.standard_normal(ns + (3,))
    ref = oracle.compute_nonlinear_term(u, ks, dea)
    deal = {"23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    dt = np.float64 if prec == "f64" else np.float32
    with fit.Plan(ns, 3, 3, ks, deal, dt) as plan:
        plan.set_option(_lib.OPT_GRAD_SPLIT, split)
        got = plan.nonlinear_term(u)
    e = rel(got, ref)
    say("  err", e)
    assert e < (1e-11 if prec == "f64" else 2e-5)


COLF64_SHAPES = [((6, 768), "23"), ((4, 4, 768), "none"), ((4, 512, 4), "23"), ((4, 192, 128), "none"),
                 ((8, 6, 8), "pad"), ((4, 6, 256), "pad"), ((2, 3072), "none")]


@pytest.mark.parametrize("mode", [0, 1, 2, 4, 6, 7])
@pytest.mark.parametrize("ns,dea", COLF64_SHAPES)
def test_col_f64_arithmetic(fit, oracle, ns, dea, mode):
    """FP32 plans with the strided register passes computing in FP64 (FIT_OPT_COL_F64: inverse,
    forward y, forward z and their combinations, or none) against the oracle, at FP32 tolerance,
    and bitwise equal on a rerun."""
    from fit_amd import _lib
    say("case col_f64", ns, dea, mode)
    rng = np.random.default_rng(sum(ns) + mode)
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    u = rng.standard_normal(ns + (nd,)) + 1j * rng.standard_normal(ns + (nd,))
    ref = oracle.compute_nonlinear_term(u, ks, dea)
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    with fit.Plan(ns, nd, nd, ks, deal, np.float32) as plan:
        plan.set_option(_lib.OPT_COL_F64, mode)
        assert plan.get_option(_lib.OPT_COL_F64) == mode
        got = plan.nonlinear_term(u)
        again = plan.nonlinear_term(u)
    e = rel(got, ref)
    say("  err", e)
    assert e < 2e-5
    assert np.array_equal(got, again)
