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
