"""The call sequence of the device-resident Julia binding, replayed through the C ABI with ctypes
(every call below is one `ccall` of that binding):

  * workspace construction: one plan per workspace -- `Workspaces._make_fft_plans(û, ks)`
    (Workspaces.jl:22, called once by each NonlinearTermWorkspace constructor, :63-75) returns it;
  * many nonlinear-term calls on fresh device arrays (hook `_compute_nonlinear_term!`,
    NonlinearTerm.jl:102-116) whose N̂ goes through the reference's own density and host shell sum
    (`_calculate_spectral_flux_with_N̂!`, SpectralFlux.jl:148-180), then the same tail on the
    device (fit_spectral_flux_with_nhat) and the whole in-place flux (`calculate_spectral_flux!`,
    :131-146 -> fit_spectral_flux) on the same plan;
  * the decomposed flux (`_calculate_spectral_flux_decomposed(::NamedTuple, ...)`, :96-122): one
    N̂, the helical parts from fit_decompose, one device tail per part;
  * shell-to-shell on its workspace's plan (`_calculate_shell_to_shell!`,
    ShellToShellTransfer.jl:152-159) and mode-to-mode on a cached plan
    (`calculate_mode_to_mode_transfer`, ScaleToScaleTransfer.jl:49-91), each called repeatedly.

Every result is checked against the FP64 oracle for each new field, and fit_plan_live_count shows
that the sequence holds exactly the plans its workspaces made: none is created per call."""
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


def _colmajor_dev(torch, a, cdt):
    """A device array with the memory layout of a Julia (column-major) array of a's shape."""
    rev = tuple(reversed(range(a.ndim)))
    return torch.from_numpy(np.ascontiguousarray(a.transpose(rev))).to("cuda").to(cdt).permute(rev)


def _colmajor_host(t):
    rev = tuple(reversed(range(t.dim())))
    return t.permute(rev).contiguous().cpu().numpy().transpose(rev)


class JuliaLikeBinding:
    """The ccalls of the binding, nothing else: plans are made only by make_plan (the
    `_make_fft_plans` hook and the mode-to-mode plan cache)."""

    def __init__(self, lib, _lib, io):
        self.lib, self._lib, self.io = lib, _lib, io
        self.keep = []

    def make_plan(self, ns, M, D, ks, dealias, shell_idx=None, n_shells=0):
        nd = len(ns)
        ksd = [np.ascontiguousarray(np.asarray(k, dtype=np.float64)) for k in ks]
        dp = ctypes.POINTER(ctypes.c_double)
        ksp = (dp * nd)(*[k.ctypes.data_as(dp) for k in ksd])
        sh = None
        if n_shells:
            # the converted map is kept alive across the call (GC.@preserve in the binding)
            shm = np.asfortranarray(np.asarray(shell_idx, dtype=np.int64))
            self.keep.append(shm)
            sh = shm.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        h = ctypes.c_void_p()
        self._lib.check(self.lib.fit_plan_create(ctypes.byref(h), nd, (ctypes.c_int64 * nd)(*ns), M, D, self.io,
                                                 dealias, ksp, sh, n_shells, 0))
        return h

    def wait(self, plan, torch):
        self._lib.check(self.lib.fit_plan_wait_stream(plan, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def nonlinear_term(self, plan, u, N):
        self._lib.check(self.lib.fit_nonlinear_term(plan, ctypes.c_void_p(u.data_ptr()), None,
                                                    ctypes.c_void_p(N.data_ptr()), self.io, self._lib.PTR_DEVICE))

    def flux_with_nhat(self, plan, inv, c, N, nsh):
        T, P = np.zeros(nsh), np.zeros(nsh)
        dp = ctypes.POINTER(ctypes.c_double)
        self._lib.check(self.lib.fit_spectral_flux_with_nhat(plan, inv, ctypes.c_void_p(c.data_ptr()),
                                                             ctypes.c_void_p(N.data_ptr()), T.ctypes.data_as(dp),
                                                             P.ctypes.data_as(dp), self.io, self._lib.PTR_DEVICE))
        return T, P

    def flux(self, plan, inv, u, nsh):
        T, P = np.zeros(nsh), np.zeros(nsh)
        dp = ctypes.POINTER(ctypes.c_double)
        self._lib.check(self.lib.fit_spectral_flux(plan, inv, ctypes.c_void_p(u.data_ptr()), None,
                                                   T.ctypes.data_as(dp), P.ctypes.data_as(dp), self.io,
                                                   self._lib.PTR_DEVICE))
        return T, P

    def decompose(self, plan, code, u, a, b):
        self._lib.check(self.lib.fit_decompose(plan, code, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                                               ctypes.c_void_p(b.data_ptr()), self.io, self._lib.PTR_DEVICE))

    def shell_to_shell(self, plan, inv, u, nsh):
        T = np.zeros((nsh, nsh), order="F")
        net, asym = np.zeros(nsh), np.zeros(1)
        dp = ctypes.POINTER(ctypes.c_double)
        self._lib.check(self.lib.fit_shell_to_shell(plan, inv, ctypes.c_void_p(u.data_ptr()), None,
                                                    T.ctypes.data_as(dp), net.ctypes.data_as(dp),
                                                    asym.ctypes.data_as(dp), 1, self.io, self._lib.PTR_DEVICE))
        return T, net, float(asym[0])

    def mode_to_mode(self, plan, inv, u, S, net):
        self._lib.check(self.lib.fit_mode_to_mode(plan, inv, ctypes.c_void_p(u.data_ptr()), None,
                                                  ctypes.c_void_p(S.data_ptr()), ctypes.c_void_p(net.data_ptr()),
                                                  self.io, self._lib.PTR_DEVICE))


@pytest.mark.parametrize("prec", [np.float32, np.float64])
def test_binding_call_sequence_reuses_one_plan_per_workspace(fit, oracle, prec):
    import torch
    from fit_amd import _lib
    lib = _lib.load_library()
    io = _lib.F64 if prec == np.float64 else _lib.F32
    cdt = torch.complex128 if prec == np.float64 else torch.complex64
    rdt = torch.float64 if prec == np.float64 else torch.float32
    tol = 1e-11 if prec == np.float64 else 2e-5
    B = JuliaLikeBinding(lib, _lib, io)
    base = lib.fit_plan_live_count()
    say("case binding sequence", prec.__name__, "live plans before:", base)

    # --- SpectralFluxWorkspace(û, ks, LinearBinning(1.0)) -> one plan --------------------------
    ns = (32, 32, 32)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * 3)
    centers, _, sidx = oracle.shell_setup(ks, oracle.LinearBinning(1.0))
    nsh = len(centers)
    pflux = B.make_plan(ns, 3, 3, ks, fit.OrszagTwoThirds().code, sidx, nsh)
    # --- ShellToShellWorkspace(û, ks, CustomBinning(0:6)) -> its own plan ----------------------
    c2, _, sidx2 = oracle.shell_setup(ks, oracle.CustomBinning(tuple(range(0, 7))))
    ptnm = B.make_plan(ns, 3, 3, ks, fit.OrszagTwoThirds().code, sidx2, len(c2))
    # --- mode-to-mode plan cache entry for (6³, M=3) -------------------------------------------
    nm = (6, 6, 6)
    ksm = oracle.wavenumber_grid(nm, (2 * math.pi,) * 3)
    pm2m = B.make_plan(nm, 3, 3, ksm, fit.OrszagTwoThirds().code)
    made = 3
    assert lib.fit_plan_live_count() == base + made

    N = torch.empty((3,) + ns[::-1], dtype=cdt, device="cuda").permute(3, 2, 1, 0)    # ws.nonlinear.N̂
    up = torch.empty_like(N)
    um = torch.empty_like(N)
    Np = int(np.prod(nm))
    S = torch.empty((Np * Np,), dtype=rdt, device="cuda")
    net = torch.empty((Np,), dtype=rdt, device="cuda")
    KE, HEL = fit.KineticEnergy().code, fit.Helicity().code
    for step in range(4):                                      # time steps of a simulation
        u = oracle.synthetic_field(ns, 3, seed=300 + step)
        ud = _colmajor_dev(torch, u, cdt)
        _, To, Po = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23", "ke")
        # hook 2 + the reference's own density and host shell sum
        B.wait(pflux, torch)
        B.nonlinear_term(pflux, ud, N)
        Nh = _colmajor_host(N).astype(np.complex128)
        Th = oracle.shell_sum(oracle.transfer_density("ke", u, Nh, ks), sidx, nsh)
        e_host = max(rel(Th, To), rel(np.cumsum(Th), Po))
        # the same tail on the device, then the whole in-place flux
        Td, Pd = B.flux_with_nhat(pflux, KE, ud, N, nsh)
        Tf, Pf = B.flux(pflux, KE, ud, nsh)
        e_dev = max(rel(Td, To), rel(Pd, Po))
        e_full = max(rel(Tf, To), rel(Pf, Po))
        # decomposed flux: one N̂, the helical parts, one tail each
        hel = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23", "ke",
                                             decomposition="HelicalDecomposition")
        B.decompose(pflux, fit.HelicalDecomposition().code, ud, up, um)
        e_dec = 0.0
        for name, part in (("positive", up), ("negative", um)):
            Tp, Pp = B.flux_with_nhat(pflux, KE, part, N, nsh)
            e_dec = max(e_dec, rel(Tp, hel[name][1]), rel(Pp, hel[name][2]))
        # helicity flux on the same N̂
        _, Tho, Pho = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), "23", "helicity")
        Thd, Phd = B.flux_with_nhat(pflux, HEL, ud, N, nsh)
        e_hel = max(rel(Thd, Tho), rel(Phd, Pho))
        say(f"  step {step}: host tail {e_host:.2e} device tail {e_dev:.2e} flux {e_full:.2e} "
            f"helical parts {e_dec:.2e} helicity {e_hel:.2e}")
        assert max(e_host, e_dev, e_full, e_dec, e_hel) < tol
        assert lib.fit_plan_live_count() == base + made
        if step < 2:
            _, _, Tm, netm, _ = oracle.calculate_shell_to_shell_transfer(
                u, ks, oracle.CustomBinning(tuple(range(0, 7))), "23")
            B.wait(ptnm, torch)
            Tg, netg, asym = B.shell_to_shell(ptnm, KE, ud, len(c2))
            e_s2s = max(rel(Tg, Tm), rel(netg, netm))
            um6 = oracle.synthetic_field(nm, 3, seed=310 + step)
            net_o, S_o = oracle.calculate_mode_to_mode_transfer(um6, ksm, "ke", "23", force=True)
            u6 = _colmajor_dev(torch, um6, cdt)
            B.wait(pm2m, torch)
            B.mode_to_mode(pm2m, KE, u6, S, net)
            Sg = S.cpu().numpy().reshape(Np, Np, order="F")     # S[k, p], receiver first
            e_m2m = max(rel(Sg, S_o.reshape(Np, Np, order="F")), rel(net.cpu().numpy(), net_o.reshape(-1, order="F")))
            say(f"  step {step}: T(n,m) {e_s2s:.2e} (asym {asym:.1e}) S(k|p) {e_m2m:.2e}")
            assert e_s2s < tol and e_m2m < (tol if prec == np.float64 else 1e-4)
            assert lib.fit_plan_live_count() == base + made
    for h in (pflux, ptnm, pm2m):
        _lib.check(lib.fit_plan_destroy(h))
    assert lib.fit_plan_live_count() == base
