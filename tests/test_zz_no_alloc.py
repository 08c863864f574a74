"""fit_spectral_flux_with_nhat_noalloc (the allocation-free entry point) through the C ABI against
fit_spectral_flux_with_nhat on the same plan and inputs: two calls in a row sharing the library's
static device scratch, the second on a smaller grid with fewer shells, FP64 and FP32; T and Π come
back in device memory.  Also: the documented rejections (host arrays, another precision, more
shells than the scratch holds)."""
import ctypes
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


def _dev(torch, a):
    rev = tuple(reversed(range(a.ndim)))
    return torch.from_numpy(np.ascontiguousarray(a.transpose(rev))).to("cuda").permute(rev)


def _both(fit, torch, plan, inv, u, N):
    """(T, Π) from the existing entry point (host outputs) and from the allocation-free one
    (device outputs), both called directly through the C ABI."""
    from fit_amd import _lib
    lib = _lib.load_library()
    n = plan.n_shells
    T0, P0 = np.zeros(n), np.zeros(n)
    dp = ctypes.POINTER(ctypes.c_double)
    _lib.check(lib.fit_spectral_flux_with_nhat(plan.handle, inv.code, ctypes.c_void_p(u.data_ptr()),
                                               ctypes.c_void_p(N.data_ptr()), T0.ctypes.data_as(dp),
                                               P0.ctypes.data_as(dp), plan.io, _lib.PTR_DEVICE))
    T1 = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    P1 = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(lib.fit_spectral_flux_with_nhat_noalloc(plan.handle, inv.code, ctypes.c_void_p(u.data_ptr()),
                                                       ctypes.c_void_p(N.data_ptr()),
                                                       ctypes.c_void_p(T1.data_ptr()),
                                                       ctypes.c_void_p(P1.data_ptr()), plan.io, _lib.PTR_DEVICE))
    return (T0, P0), (T1.cpu().numpy(), P1.cpu().numpy())


@pytest.mark.parametrize("prec", [np.float64, np.float32])
def test_noalloc_flux_tail_equals_existing_entry(fit, oracle, prec):
    import torch
    from fit_amd.grid import shell_setup
    cdt = torch.complex128 if prec == np.float64 else torch.complex64
    tol = 1e-12 if prec == np.float64 else 1e-6
    cases = [((32, 32, 32), 3, fit.KineticEnergy(), "ke"),      # first call: 28 shells
             ((16, 16), 2, fit.Enstrophy(), "enstrophy")]        # second call: smaller grid, 12 shells
    for ns, nd, inv, tag in cases:
        say("case noalloc", ns, tag, prec.__name__)
        u = oracle.synthetic_field(ns, nd, seed=90 + nd)
        ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
        c, _, sidx = shell_setup(ks, fit.LinearBinning(1.0), None)
        ud = _dev(torch, u).to(cdt)
        with fit.Plan(ns, nd, nd, ks, fit.OrszagTwoThirds(), prec, sidx, len(c)) as plan:
            N = plan.nonlinear_term(ud)
            (T0, P0), (T1, P1) = _both(fit, torch, plan, inv, ud, N)
        _, To, Po = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23", tag)
        e_same = max(rel(T1, T0), rel(P1, P0))
        e_orc = max(rel(T1, To), rel(P1, Po))
        say("  noalloc vs existing", e_same, "vs oracle", e_orc)
        assert len(T1) == len(c) and np.all(np.isfinite(T1)) and np.all(np.isfinite(P1))
        assert e_same < tol
        assert e_orc < (1e-11 if prec == np.float64 else 1e-5)


def test_noalloc_rejections(fit, oracle):
    import torch
    from fit_amd import _lib
    from fit_amd.grid import shell_setup
    lib = _lib.load_library()
    ns = (64, 64)
    u = oracle.synthetic_field(ns, 2, seed=97)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 2)
    ud = _dev(torch, u)
    T = torch.zeros(8192, dtype=torch.float64, device="cuda")
    P = torch.zeros(8192, dtype=torch.float64, device="cuda")
    args = lambda plan, io, kind: (plan.handle, 0, ctypes.c_void_p(ud.data_ptr()), ctypes.c_void_p(ud.data_ptr()),
                                   ctypes.c_void_p(T.data_ptr()), ctypes.c_void_p(P.data_ptr()), io, kind)
    c, _, sidx = shell_setup(ks, fit.LinearBinning(1.0), None)
    with fit.Plan(ns, 2, 2, ks, fit.OrszagTwoThirds(), np.float64, sidx, len(c)) as plan:
        assert lib.fit_spectral_flux_with_nhat_noalloc(*args(plan, _lib.F64, _lib.PTR_HOST)) == 1
        assert lib.fit_spectral_flux_with_nhat_noalloc(*args(plan, _lib.F32, _lib.PTR_DEVICE)) == 1
        assert lib.fit_spectral_flux_with_nhat_noalloc(*args(plan, _lib.F64, _lib.PTR_DEVICE)) == 0
    c, _, sidx = shell_setup(ks, fit.LinearBinning(0.01), None)      # about 4500 shells
    assert len(c) > 4096
    with fit.Plan(ns, 2, 2, ks, fit.OrszagTwoThirds(), np.float64, sidx, len(c)) as plan:
        rc = lib.fit_spectral_flux_with_nhat_noalloc(*args(plan, _lib.F64, _lib.PTR_DEVICE))
        assert rc == 3 and b"4096" in lib.fit_last_error()
