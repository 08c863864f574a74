"""Host-side mirror of the reference's public API for the pseudospectral transfer path, with every
computation routed to libfit_hip.so (MI355X).

Same names, argument meaning, defaults and error behaviour as the reference (src/SpectralFlux.jl,
src/ShellToShell/ShellToShellTransfer.jl, src/ScaleToScale/ScaleToScaleTransfer.jl,
src/NonlinearTerm.jl, src/Invariants.jl), with the backend singletons replaced by
``HIPSpectral()`` / ``HIPExecution()``.  Arrays use the reference's layout: shape (ns..., C) with
the component axis last, column-major in memory.  Host inputs are numpy arrays; device inputs are
torch tensors on a HIP device (results then stay on the device).
"""
import ctypes
import functools
import operator
import os
import threading
from collections import OrderedDict

import numpy as np

from . import _lib
from .grid import as_component_field, shell_coordinate, shell_setup
from .types import (AbstractDealiasing, AbstractInvariant, BandToBandResult, HelicalDecomposition,
                    HIPExecution, HIPSpectral, KineticEnergy, ModeToModeTriadResult, NoDealiasing,
                    NoDecomposition, OrszagTwoThirds, PaddedThreeHalves, PassiveScalar,
                    ShellMagnitude, ShellToShellResult, SmoothBands, SpectralFluxResult,
                    ToroidalPoloidalDecomposition)

_SPLITS = (HelicalDecomposition, ToroidalPoloidalDecomposition)

try:  # torch is only plumbing for device arrays
    import torch
except Exception:  # pragma: no cover
    torch = None


# ---------------------------------------------------------------------------------------------
# array marshalling
# ---------------------------------------------------------------------------------------------

def _is_device(a):
    return torch is not None and isinstance(a, torch.Tensor) and a.device.type == "cuda"


def _real_dtype(a):
    if _is_device(a):
        return np.float64 if a.dtype in (torch.complex128, torch.float64) else np.float32
    dt = np.asarray(a).dtype
    return np.float64 if dt in (np.complex128, np.float64) else np.float32


def _fortran_order(t):
    """True when tensor t is column-major contiguous (read from its strides; no view is made)."""
    acc = 1
    for n, st in zip(t.shape, t.stride()):
        if n != 1 and st != acc:
            return False
        acc *= n
    return True


def _marshal(a, io):
    """-> (pointer, keep-alive object, ptr_kind) for a complex array in io precision (F32/F64),
    column-major."""
    if _is_device(a):
        want = torch.complex64 if io == _lib.F32 else torch.complex128
        t = a if a.dtype == want else a.to(want)
        if not _fortran_order(t):
            rev = tuple(reversed(range(t.dim())))
            t = t.permute(rev).contiguous().permute(rev)
        return t.data_ptr(), t, _lib.PTR_DEVICE
    want = np.complex64 if io == _lib.F32 else np.complex128
    arr = np.asfortranarray(np.asarray(a, dtype=want))
    return arr.ctypes.data, arr, _lib.PTR_HOST


def _caller_stream(t):
    """hipStream_t of torch's current stream on the device of tensor t."""
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _marshal_real(a, io):
    """-> (pointer, keep-alive object, ptr_kind, io) for a real array, column-major; io follows the
    array's own precision (F32 / F64)."""
    if _is_device(a):
        t = a if a.dtype in (torch.float32, torch.float64) else a.to(torch.float64)
        if not _fortran_order(t):
            rev = tuple(reversed(range(t.dim())))
            t = t.permute(rev).contiguous().permute(rev)
        return t.data_ptr(), t, _lib.PTR_DEVICE, (_lib.F64 if t.dtype == torch.float64 else _lib.F32)
    arr = np.asarray(a)
    if arr.dtype not in (np.float32, np.float64):
        arr = arr.astype(np.float64)
    arr = np.asfortranarray(arr)
    return arr.ctypes.data, arr, _lib.PTR_HOST, (_lib.F64 if arr.dtype == np.float64 else _lib.F32)


def _alloc(like_device, shape, io, complex_=True):
    if like_device is not None:
        if complex_:
            dt = torch.complex64 if io == _lib.F32 else torch.complex128
        else:
            dt = torch.float32 if io == _lib.F32 else torch.float64
        rev = tuple(reversed(range(len(shape))))
        t = torch.empty(tuple(reversed(shape)), dtype=dt, device=like_device.device).permute(rev)
        return t, t.data_ptr()
    if complex_:
        dt = np.complex64 if io == _lib.F32 else np.complex128
    else:
        dt = np.float32 if io == _lib.F32 else np.float64
    arr = np.empty(shape, dtype=dt, order="F")
    return arr, arr.ctypes.data


def _dptr(arr):
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check_invariant(invariant, nd, M):
    """Invariants.jl:63 and :99-101 -- raised before any device work."""
    from .types import Enstrophy, Helicity
    if isinstance(invariant, Helicity) and nd != 3:
        raise ValueError(f"Helicity transfer is defined in 3D only (got nd={nd}).")
    if isinstance(invariant, Enstrophy) and nd not in (2, 3):
        raise ValueError("Enstrophy transfer is defined in 2D (scalar vorticity, conserved) or 3D "
                         f"(vector vorticity, non-conservative via stretching); got nd={nd}.")
    if isinstance(invariant, (Helicity, Enstrophy)) and M < nd:
        raise _lib.DimensionMismatch("vorticity-weighted invariants need nd velocity components")


def _check_spectral(spectral):
    if not isinstance(spectral, HIPSpectral):
        raise ValueError(f"spectral={spectral!r}: this package provides the HIP backend only "
                         "(pass spectral=HIPSpectral()).")


# ---------------------------------------------------------------------------------------------
# plans (device workspaces)
# ---------------------------------------------------------------------------------------------

class Plan:
    """Device-resident workspace for one grid (the reference's NonlinearTermWorkspace /
    SpectralFluxWorkspace / ShellToShellWorkspace, Workspaces.jl:46-170, plus FFT plans)."""

    def __init__(self, ns, M, D, ks, dealiasing=OrszagTwoThirds(), precision=np.float32,
                 shell_idx=None, n_shells=0, device=0):
        self.lib = _lib.load_library()
        self.ns = tuple(int(n) for n in ns)
        self.nd = len(self.ns)
        self.M, self.D = int(M), int(D)
        self.io = _lib.F64 if np.dtype(precision) == np.float64 else _lib.F32
        self.n_shells = int(n_shells)
        nsarr = (ctypes.c_int64 * self.nd)(*self.ns)
        self._ks = [np.ascontiguousarray(np.asarray(k, dtype=np.float64)) for k in ks]
        if len(self._ks) != self.nd or any(len(k) != n for k, n in zip(self._ks, self.ns)):
            raise _lib.DimensionMismatch("ks must hold nd vectors of lengths ns")
        ksp = (ctypes.POINTER(ctypes.c_double) * self.nd)(*[_dptr(k) for k in self._ks])
        sp = None
        if self.n_shells > 0:
            self._shell = np.asfortranarray(np.asarray(shell_idx, dtype=np.int64))
            if self._shell.shape != self.ns:
                raise _lib.DimensionMismatch("shell_idx must have shape ns")
            sp = self._shell.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        self.handle = ctypes.c_void_p()
        rc = self.lib.fit_plan_create(ctypes.byref(self.handle), self.nd, nsarr, self.M, self.D,
                                      self.io, dealiasing.code, ksp, sp, self.n_shells, device)
        _lib.check(rc)

    def close(self):
        if getattr(self, "handle", None) and self.handle.value:
            self.lib.fit_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, option, value):
        """FIT_OPT_* knobs: plain=True runs the unfused reference dataflow (comparison baseline);
        giver_batch caps the givers per batched launch."""
        _lib.check(self.lib.fit_plan_set_option(self.handle, option, int(value)))

    def get_option(self, option):
        v = ctypes.c_int()
        _lib.check(self.lib.fit_plan_get_option(self.handle, option, ctypes.byref(v)))
        return int(v.value)

    def options(self):
        """{name: value} of every execution knob."""
        return {name: self.get_option(code) for code, name in _lib.OPT_NAMES.items()}

    def autotune(self, workload, invariant, f, adv=None, reps=2):
        """fit_plan_autotune on this plan with the given input (workload "flux" or
        "shell_to_shell"); returns {"report": ..., "options": {...}}."""
        code = {"flux": _lib.TUNE_FLUX, "shell_to_shell": _lib.TUNE_SHELL_TO_SHELL}[workload]
        fp, ap, keep, kind = self._fa(f, adv)
        self._order(kind, keep[0])
        _lib.check(self.lib.fit_plan_autotune(self.handle, code, invariant.code, fp, ap, self.io, kind, int(reps)))
        buf = ctypes.create_string_buffer(2048)
        self.lib.fit_plan_tuning(self.handle, buf, 2048)
        self.tuned = getattr(self, "tuned", set()) | {workload}
        return {"report": buf.value.decode(), "options": self.options()}

    def set_plain(self, on=True):
        self.set_option(_lib.OPT_PLAIN_PIPELINE, int(bool(on)))

    def set_giver_batch(self, n):
        self.set_option(_lib.OPT_GIVER_BATCH, int(n))

    def set_subgrid(self, on=True):
        """Use (default) or bypass the band-limited sub-grid of the giver loop."""
        self.set_option(_lib.OPT_SUBGRID, int(bool(on)))

    @property
    def subgrid(self):
        """Grid of the band-limited giver loop, or None when the plan has none attached."""
        dims = (ctypes.c_int64 * 3)()
        rc = self.lib.fit_plan_subgrid(self.handle, dims)
        if rc < 0:
            _lib.check(1)
        return tuple(int(d) for d in dims[:self.nd]) if rc == 1 else None

    @property
    def subgrid_classes(self):
        """[(grid dims, (first giver, last giver))] of the band-limited giver loop's classes."""
        dims = (ctypes.c_int64 * 96)()
        giv = (ctypes.c_int * 64)()
        n = self.lib.fit_plan_subgrid_classes(self.handle, dims, giv, 32)
        if n < 0:
            _lib.check(1)
        return [(tuple(int(dims[3 * i + d]) for d in range(self.nd)), (int(giv[2 * i]), int(giv[2 * i + 1])))
                for i in range(min(n, 32))]

    @property
    def device_bytes(self):
        return int(self.lib.fit_plan_device_bytes(self.handle))

    @property
    def stream(self):
        return self.lib.fit_plan_stream(self.handle)

    # -- profiling hooks -----------------------------------------------------------------------
    def profile(self, on=True):
        _lib.check(self.lib.fit_profile_enable(self.handle, int(on)))

    def profile_stats(self):
        out = {}
        for i, name in enumerate(_lib.K_CLASSES):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            _lib.check(self.lib.fit_profile_query(self.handle, i, ctypes.byref(n), ctypes.byref(ms),
                                                  ctypes.byref(by)))
            out[name] = {"launches": n.value, "ms": ms.value, "alg_bytes": by.value}
        return out

    def kernel_stats(self):
        """Per-kernel-label totals of the profiled launches (launches, ms, algorithmic bytes)."""
        return _lib.kernel_stats(self.lib.fit_profile_kernel_count, self.lib.fit_profile_kernel, self.handle)

    # -- entry points --------------------------------------------------------------------------
    def _order(self, kind, t):
        """Device arrays: the plan's stream waits for torch's current stream, where the inputs
        (and the copies _marshal made of them) were produced and where recycled output blocks
        may still be in use."""
        if kind == _lib.PTR_DEVICE:
            _lib.check(self.lib.fit_plan_wait_stream(self.handle, _caller_stream(t)))

    def _fa(self, f, adv):
        fp, fk, kind = _marshal(f, self.io)
        if adv is None:
            return fp, None, (fk,), kind
        ap, ak, kind2 = _marshal(adv, self.io)
        if kind2 != kind:
            raise ValueError("advected and advecting fields must both be host or both device arrays")
        return fp, ap, (fk, ak), kind

    def nonlinear_term(self, f, adv=None):
        fp, ap, keep, kind = self._fa(f, adv)
        out, op = _alloc(keep[0] if kind == _lib.PTR_DEVICE else None, self.ns + (self.M,), self.io)
        self._order(kind, keep[0])
        _lib.check(self.lib.fit_nonlinear_term(self.handle, fp, ap, op, self.io, kind))
        return out

    def transfer_density(self, invariant, c_hat, N_hat):
        cp, ck, kind = _marshal(c_hat, self.io)
        npp, nk, _ = _marshal(N_hat, self.io)
        out, op = _alloc(ck if kind == _lib.PTR_DEVICE else None, self.ns, self.io, complex_=False)
        self._order(kind, ck)
        _lib.check(self.lib.fit_transfer_density(self.handle, invariant.code, cp, npp, op, self.io,
                                                 kind))
        return out

    def spectral_flux(self, invariant, f, adv=None):
        fp, ap, keep, kind = self._fa(f, adv)
        T = np.zeros(self.n_shells)
        Pi = np.zeros(self.n_shells)
        self._order(kind, keep[0])
        _lib.check(self.lib.fit_spectral_flux(self.handle, invariant.code, fp, ap, _dptr(T),
                                              _dptr(Pi), self.io, kind))
        return T, Pi

    def spectral_fluxes(self, invariants, f, adv=None):
        """[(T, Π)] for several invariants of f advected by adv (or f) with one shared N̂, whose
        densities the last forward pass computes (fit_spectral_fluxes)."""
        fp, ap, keep, kind = self._fa(f, adv)
        q = len(invariants)
        T = np.zeros(q * self.n_shells)
        Pi = np.zeros(q * self.n_shells)
        codes = (ctypes.c_int * q)(*[inv.code for inv in invariants])
        self._order(kind, keep[0])
        _lib.check(self.lib.fit_spectral_fluxes(self.handle, q, codes, fp, ap, _dptr(T), _dptr(Pi), self.io, kind))
        n = self.n_shells
        return [(T[i * n:(i + 1) * n], Pi[i * n:(i + 1) * n]) for i in range(q)]

    def spectral_flux_with_nhat(self, invariant, c_hat, N_hat):
        cp, ck, kind = _marshal(c_hat, self.io)
        npp, nk, _ = _marshal(N_hat, self.io)
        T = np.zeros(self.n_shells)
        Pi = np.zeros(self.n_shells)
        self._order(kind, ck)
        _lib.check(self.lib.fit_spectral_flux_with_nhat(self.handle, invariant.code, cp, npp,
                                                        _dptr(T), _dptr(Pi), self.io, kind))
        return T, Pi

    def spectral_fluxes_with_nhat(self, invariants, c_hat, N_hat):
        """[(T, Π)] for several invariants from one carrier and N̂ in one density pass."""
        cp, ck, kind = _marshal(c_hat, self.io)
        npp, nk, _ = _marshal(N_hat, self.io)
        q = len(invariants)
        T = np.zeros(q * self.n_shells)
        Pi = np.zeros(q * self.n_shells)
        codes = (ctypes.c_int * q)(*[inv.code for inv in invariants])
        self._order(kind, ck)
        _lib.check(self.lib.fit_spectral_fluxes_with_nhat(self.handle, q, codes, cp, npp, _dptr(T),
                                                          _dptr(Pi), self.io, kind))
        n = self.n_shells
        return [(T[i * n:(i + 1) * n], Pi[i * n:(i + 1) * n]) for i in range(q)]

    def shell_to_shell(self, invariant, f, adv=None, verify=True):
        fp, ap, keep, kind = self._fa(f, adv)
        n = self.n_shells
        T = np.zeros((n, n), order="F")
        net = np.zeros(n)
        asym = np.zeros(1)
        self._order(kind, keep[0])
        _lib.check(self.lib.fit_shell_to_shell(self.handle, invariant.code, fp, ap, _dptr(T),
                                               _dptr(net), _dptr(asym), int(bool(verify)), self.io,
                                               kind))
        return T, net, float(asym[0])

    def mode_to_mode(self, invariant, f, adv=None):
        fp, ap, keep, kind = self._fa(f, adv)
        like = keep[0] if kind == _lib.PTR_DEVICE else None
        S, sp = _alloc(like, self.ns + self.ns, self.io, complex_=False)
        net, npp = _alloc(like, self.ns, self.io, complex_=False)
        self._order(kind, keep[0])
        _lib.check(self.lib.fit_mode_to_mode(self.handle, invariant.code, fp, ap, sp, npp, self.io,
                                             kind))
        return net, S

    def decompose(self, decomposition, u):
        """The two named parts of u (fit_decompose): helical (positive, negative) or
        toroidal/poloidal (toroidal, poloidal)."""
        up, uk, kind = _marshal(u, self.io)
        like = uk if kind == _lib.PTR_DEVICE else None
        a, ap = _alloc(like, self.ns + (self.M,), self.io)
        b, bp = _alloc(like, self.ns + (self.M,), self.io)
        self._order(kind, uk)
        _lib.check(self.lib.fit_decompose(self.handle, decomposition.code, up, ap, bp, self.io, kind))
        return a, b

    def physical_to_spectral(self, u_phys):
        """û = fft(u)/Np per component of a real physical field (ns..., C) (fit_physical_to_spectral),
        in the field's own precision; device in, device out."""
        up, uk, kind, io = _marshal_real(u_phys, self.io)
        C = int(u_phys.shape[self.nd]) if len(u_phys.shape) == self.nd + 1 else 1
        out, op = _alloc(uk if kind == _lib.PTR_DEVICE else None, self.ns + (C,), io)
        self._order(kind, uk)
        _lib.check(self.lib.fit_physical_to_spectral(self.handle, up, C, op, io, kind))
        return out

    def helical_decompose(self, u):
        return self.decompose(HelicalDecomposition(), u)


# A plan is not safe for concurrent calls (its workspaces are shared by its launches), so cached
# plans are per host thread, as the reference keeps one workspace per task
# (OhMyThreadsExt.jl:46).  The caches themselves are guarded by one lock.
_PLANS = OrderedDict()
_MAX_PLANS = 12
_CACHE_LOCK = threading.Lock()


def _plan_for(ns, M, D, ks, dealiasing, precision, device, binning=None, geometry=None,
              shells=False):
    key = (threading.get_ident(), tuple(ns), M, D,
           tuple(np.asarray(k, dtype=np.float64).tobytes() for k in ks),
           type(dealiasing).__name__, np.dtype(precision).name, device,
           repr(binning) if shells else None, repr(geometry) if shells else None)
    with _CACHE_LOCK:
        if key in _PLANS:
            _PLANS.move_to_end(key)
            return _PLANS[key]
    centers = edges = None
    sidx = None
    nsh = 0
    if shells:
        centers, edges, sidx = shell_setup(ks, binning, geometry)
        nsh = len(centers)
    plan = Plan(ns, M, D, ks, dealiasing, precision, sidx, nsh, device)
    plan.centers, plan.edges = centers, edges
    with _CACHE_LOCK:
        _PLANS[key] = plan
        while len(_PLANS) > _MAX_PLANS:
            # not closed here: another thread may be inside a call on it; the plan is destroyed
            # when its last reference goes away (Plan.__del__)
            _PLANS.popitem(last=False)
    return plan


def _maybe_tune(plan, workload, invariant, f, adv):
    """With FIT_AUTOTUNE=1 in the environment, a cached plan is autotuned on its first call of a
    workload, with that call's input (the choices then hold for the plan and its key)."""
    if os.environ.get("FIT_AUTOTUNE") != "1" or workload in getattr(plan, "tuned", set()):
        return
    plan.autotune(workload, invariant, f, adv, reps=1)


def clear_plan_cache():
    with _CACHE_LOCK:
        plans = list(_PLANS.values())
        ens = list(_ENSEMBLES.values())
        _PLANS.clear()
        _ENSEMBLES.clear()
    for p in plans + ens:
        p.close()


def _shape_info(f_hat, ks, advecting_hat):
    nd = len(ks)
    ns = tuple(int(s) for s in f_hat.shape[:nd])
    if len(f_hat.shape) != nd + 1:
        raise _lib.DimensionMismatch(f"expected an (ns..., M) array with nd={nd}")
    M = int(f_hat.shape[nd])
    D = int(advecting_hat.shape[nd]) if advecting_hat is not None else M
    if advecting_hat is not None and tuple(advecting_hat.shape[:nd]) != ns:
        raise _lib.DimensionMismatch("advecting_hat grid differs from the advected field's")
    return nd, ns, M, D


def _ft_cast(x, ft):
    return np.asarray(x).astype(ft)


# ---------------------------------------------------------------------------------------------
# public API (same names / keywords as the reference)
# ---------------------------------------------------------------------------------------------

def compute_nonlinear_term(velocity_hat, ks, *, dealiasing=OrszagTwoThirds(),
                           spectral=HIPSpectral(), advecting_hat=None):
    """NonlinearTerm.jl:65-76 -- N̂ = FFT[(u_adv·∇) f] with the reference's normalisation."""
    _check_spectral(spectral)
    nd, ns, M, D = _shape_info(velocity_hat, ks, advecting_hat)
    ft = _real_dtype(velocity_hat)
    plan = _plan_for(ns, M, D, ks, dealiasing, spectral.compute_dtype(ft), spectral.device)
    plan_io = plan.io
    out = plan.nonlinear_term(velocity_hat, advecting_hat)
    if not _is_device(out):
        out = out.astype(np.complex64 if ft == np.float32 else np.complex128)
    return out


def transfer_density(invariant, velocity_hat, N_hat, ks, *, spectral=HIPSpectral()):
    """Invariants.jl:112-118 (allocating form of transfer_density!)."""
    _check_spectral(spectral)
    nd = len(ks)
    ns = tuple(int(s) for s in velocity_hat.shape[:nd])
    M = int(velocity_hat.shape[nd])
    _check_invariant(invariant, nd, M)
    ft = _real_dtype(velocity_hat)
    plan = _plan_for(ns, M, max(M, nd), ks, OrszagTwoThirds(), spectral.compute_dtype(ft),
                     spectral.device)
    t = plan.transfer_density(invariant, velocity_hat, N_hat)
    return t if _is_device(t) else t.astype(ft)


def calculate_spectral_flux(velocity_hat, ks, *, binning=None, dealiasing=OrszagTwoThirds(),
                            invariant=KineticEnergy(), decomposition=NoDecomposition(),
                            spectral=HIPSpectral(), advecting_hat=None, geometry=None):
    """SpectralFlux.jl:49-94: T(k_n) = Σ_{shell n} t(k), Π = +cumsum(T)."""
    _check_spectral(spectral)
    nd, ns, M, D = _shape_info(velocity_hat, ks, advecting_hat)
    _check_invariant(invariant, nd, M)
    ft = _real_dtype(velocity_hat)
    geometry = geometry if geometry is not None else ShellMagnitude(None)
    plan = _plan_for(ns, M, D, ks, dealiasing, spectral.compute_dtype(ft), spectral.device,
                     binning, geometry, shells=True)
    centers = _ft_cast(plan.centers, ft)
    if isinstance(decomposition, _SPLITS):
        # SpectralFlux.jl:96-122: one N̂ of the full field, the density of each named part
        N = plan.nonlinear_term(velocity_hat, advecting_hat)
        parts = plan.decompose(decomposition, velocity_hat)
        res = {}
        for name, comp in zip(decomposition.names, parts):
            T, Pi = plan.spectral_flux_with_nhat(invariant, comp, N)
            res[name] = SpectralFluxResult(centers, _ft_cast(T, ft), _ft_cast(Pi, ft))
        return res
    if not isinstance(decomposition, NoDecomposition):
        raise ValueError(f"decomposition {decomposition!r} is not on the HIP path")
    _maybe_tune(plan, "flux", invariant, velocity_hat, advecting_hat)
    T, Pi = plan.spectral_flux(invariant, velocity_hat, advecting_hat)
    return SpectralFluxResult(centers, _ft_cast(T, ft), _ft_cast(Pi, ft))


def calculate_scalar_flux(velocity_hat, scalar_hat, ks, *, binning=None,
                          dealiasing=OrszagTwoThirds(), spectral=HIPSpectral(), geometry=None):
    """SpectralFlux.jl:203-215."""
    th = as_component_field(scalar_hat, len(ks))
    return calculate_spectral_flux(th, ks, binning=binning, dealiasing=dealiasing,
                                   invariant=PassiveScalar(), advecting_hat=velocity_hat,
                                   spectral=spectral, geometry=geometry)


def calculate_shell_to_shell_transfer(velocity_hat, ks, *, binning=None,
                                      dealiasing=OrszagTwoThirds(), verify_antisymmetry=True,
                                      invariant=KineticEnergy(), spectral=HIPSpectral(),
                                      execution=HIPExecution(), advecting_hat=None, geometry=None):
    """ShellToShellTransfer.jl:78-104 with the giver loop on the device."""
    _check_spectral(spectral)
    if not isinstance(execution, HIPExecution):
        raise ValueError(f"execution={execution!r}: pass execution=HIPExecution()")
    nd, ns, M, D = _shape_info(velocity_hat, ks, advecting_hat)
    _check_invariant(invariant, nd, M)
    ft = _real_dtype(velocity_hat)
    geometry = geometry if geometry is not None else ShellMagnitude(None)
    plan = _plan_for(ns, M, D, ks, dealiasing, spectral.compute_dtype(ft), spectral.device,
                     binning, geometry, shells=True)
    _maybe_tune(plan, "shell_to_shell", invariant, velocity_hat, advecting_hat)
    T, net, asym = plan.shell_to_shell(invariant, velocity_hat, advecting_hat,
                                       verify_antisymmetry)
    return ShellToShellResult(_ft_cast(plan.centers, ft), _ft_cast(plan.edges, ft),
                              np.asfortranarray(T.astype(ft)), _ft_cast(net, ft), ft(asym))


def calculate_scalar_shell_to_shell_transfer(velocity_hat, scalar_hat, ks, **kw):
    """ShellToShellTransfer.jl:143-147."""
    th = as_component_field(scalar_hat, len(ks))
    return calculate_shell_to_shell_transfer(th, ks, invariant=PassiveScalar(),
                                             advecting_hat=velocity_hat, **kw)


def calculate_mode_to_mode_transfer(velocity_hat, ks, *, invariant=KineticEnergy(),
                                    dealiasing=OrszagTwoThirds(), spectral=HIPSpectral(),
                                    max_modes=1024, force=False, advecting_hat=None):
    """ScaleToScaleTransfer.jl:49-91 (guard at :64-67)."""
    _check_spectral(spectral)
    nd, ns, M, D = _shape_info(velocity_hat, ks, advecting_hat)
    _check_invariant(invariant, nd, M)
    Nm = int(np.prod(ns))
    if not (force or Nm <= max_modes):
        raise ValueError(
            f"calculate_mode_to_mode_transfer resolves S(k|p) over all {Nm}×{Nm} mode pairs "
            f"(O(N^{{2D}}) time/memory); N_modes={Nm} exceeds max_modes={max_modes}. Use "
            "calculate_shell_to_shell_transfer / calculate_spectral_flux for the aggregates, or "
            "pass force=True.")
    ft = _real_dtype(velocity_hat)
    plan = _plan_for(ns, M, D, ks, dealiasing, spectral.compute_dtype(ft), spectral.device)
    net, S = plan.mode_to_mode(invariant, velocity_hat, advecting_hat)
    if not _is_device(S):
        net, S = net.astype(ft), S.astype(ft)
    return ModeToModeTriadResult(invariant, ks, net, S)


def calculate_band_to_band_transfer(velocity_hat, ks, *, bands, dealiasing=OrszagTwoThirds(),
                                    invariant=KineticEnergy(), spectral=HIPSpectral(),
                                    advecting_hat=None, geometry=None):
    """BandTransfer.jl:66-113 -- smooth band-to-band transfer T(n,m) between the graded bands of
    `bands` (SmoothBands): one nonlinear term of w_m ⊙ û per band, band-weighted density sums on
    the device.  Returns BandToBandResult(centers, transfer_matrix, net_transfer,
    max_antisymmetry_error)."""
    _check_spectral(spectral)
    if not isinstance(bands, SmoothBands):
        raise ValueError(f"bands must be SmoothBands, got {bands!r}")
    cen = np.asarray(bands.centers, dtype=np.float64)
    if cen.size == 0 or np.any(cen <= 0) or np.any(np.diff(cen) <= 0):
        raise ValueError("SmoothBands centers must be positive and increasing")
    nd, ns, M, D = _shape_info(velocity_hat, ks, advecting_hat)
    _check_invariant(invariant, nd, M)
    ft = _real_dtype(velocity_hat)
    geometry = geometry if geometry is not None else ShellMagnitude(None)
    plan = _plan_for(ns, M, D, ks, dealiasing, spectral.compute_dtype(ft), spectral.device)
    kc = np.asfortranarray(shell_coordinate(geometry, ks).astype(np.float64))
    fp, ap, keep, kind = plan._fa(velocity_hat, advecting_hat)
    nb = cen.size
    T = np.zeros((nb, nb), order="F")
    net = np.zeros(nb)
    asym = np.zeros(1)
    plan._order(kind, keep[0])
    _lib.check(plan.lib.fit_band_to_band(plan.handle, invariant.code, _dptr(cen), nb,
                                         float(bands.logwidth), _dptr(kc), fp, ap, _dptr(T),
                                         _dptr(net), _dptr(asym), plan.io, kind))
    return BandToBandResult(_ft_cast(cen, ft), np.asfortranarray(T.astype(ft)), _ft_cast(net, ft),
                            ft(asym[0]))


def calculate_scalar_mode_to_mode_transfer(velocity_hat, scalar_hat, ks, **kw):
    """ScaleToScaleTransfer.jl:101-105."""
    th = as_component_field(scalar_hat, len(ks))
    return calculate_mode_to_mode_transfer(th, ks, invariant=PassiveScalar(),
                                           advecting_hat=velocity_hat, **kw)


def decompose_field(decomposition, velocity_hat, ks, *, spectral=HIPSpectral()):
    """Decomposition.jl:42-141 (NoDecomposition -> identity; the helical and toroidal/poloidal
    splits on the device, returned as named parts in the reference's order)."""
    if isinstance(decomposition, NoDecomposition):
        return velocity_hat
    if not isinstance(decomposition, _SPLITS):
        raise ValueError(f"decomposition {decomposition!r} is not on the HIP path")
    nd = len(ks)
    name = type(decomposition).__name__
    if nd != 3:
        raise ValueError(f"{name} is defined in 3D only (got nd={nd}).")
    ns = tuple(int(s) for s in velocity_hat.shape[:nd])
    M = int(velocity_hat.shape[nd])
    if M < 3:
        raise ValueError(f"{name} needs ≥3 velocity components (got D={M}).")
    ft = _real_dtype(velocity_hat)
    plan = _plan_for(ns, M, max(M, nd), ks, OrszagTwoThirds(), spectral.compute_dtype(ft),
                     spectral.device)
    a, b = plan.decompose(decomposition, velocity_hat)
    return dict(zip(decomposition.names, (a, b)))


def physical_to_spectral(u_phys, ks, *, spectral=HIPSpectral()):
    """Spectral coefficients of a physical field in the package's convention, û = fft(u)/Np per
    component (the front-end of PencilFFTsExt.jl:83-88), on the device.  u_phys: (ns...,) or
    (ns..., C); a complex field is transformed as its real and imaginary parts,
    fft(a + i b) = fft(a) + i fft(b)."""
    _check_spectral(spectral)
    nd = len(ks)
    shape = tuple(int(n) for n in u_phys.shape)
    ns = shape[:nd]
    single = len(shape) == nd
    if single:
        u_phys = u_phys[..., None]
    elif len(shape) != nd + 1:
        raise _lib.DimensionMismatch(f"expected an (ns...) or (ns..., C) array with nd={nd}")
    C = int(u_phys.shape[nd])
    is_complex = (u_phys.is_complex() if _is_device(u_phys) else np.iscomplexobj(u_phys))
    if is_complex:
        re = physical_to_spectral(u_phys.real, ks, spectral=spectral)
        im = physical_to_spectral(u_phys.imag, ks, spectral=spectral)
        out = re + 1j * im
        return out[..., 0] if single else out
    if _is_device(u_phys):
        ft = np.float64 if u_phys.dtype == torch.float64 else np.float32
    else:
        u_phys = np.asarray(u_phys)
        ft = np.float32 if u_phys.dtype == np.float32 else np.float64
    plan = _plan_for(ns, max(C, 1), max(C, nd), ks, NoDealiasing(), spectral.compute_dtype(ft), spectral.device)
    out = plan.physical_to_spectral(u_phys)
    return out[..., 0] if single else out


def pencil_spectral_flux(u_phys, ks, *, binning, dealiasing=OrszagTwoThirds(), invariant=KineticEnergy(),
                         geometry=None, spectral=HIPSpectral()):
    """PencilFFTsExt.jl:40-137 on one GPU: the kinetic-energy flux of a PHYSICAL velocity (a tuple of
    nd component arrays, or one (ns..., nd) array), with û = fft(u)/Np formed on the device.  Same
    argument checks and result fields (k_shells, transfer_spectrum, flux) as the reference."""
    if not isinstance(invariant, KineticEnergy):
        raise ValueError("pencil_spectral_flux currently supports KineticEnergy only "
                         f"(got {type(invariant).__name__}).")
    if geometry is not None and not (isinstance(geometry, ShellMagnitude) and geometry.dims is None):
        raise ValueError("pencil_spectral_flux currently supports isotropic |k| shells (IsotropicShells()) only.")
    nd = len(ks)
    if isinstance(u_phys, (tuple, list)):
        D = len(u_phys)
        if _is_device(u_phys[0]):
            u_phys = torch.stack(list(u_phys), dim=-1)
        else:
            u_phys = np.stack([np.asarray(c) for c in u_phys], axis=-1)
    else:
        D = int(u_phys.shape[nd]) if len(u_phys.shape) == nd + 1 else 1
    if D != nd:
        raise ValueError(f"got {D} velocity components for an {nd}-D grid; pencil flux needs D == nd.")
    uh = physical_to_spectral(u_phys, ks, spectral=spectral)
    return calculate_spectral_flux(uh, ks, binning=binning, dealiasing=dealiasing, invariant=invariant,
                                   spectral=spectral)


def calculate_partial_fluxes(velocity_hat, ks, *, decomposition=HelicalDecomposition(),
                             binning=None, dealiasing=OrszagTwoThirds(), spectral=HIPSpectral(),
                             geometry=None):
    """SpectralFlux.jl:240-285: channels[(s_k, s_p, s_q)] from (u_{s_p}·∇)u_{s_q} carried by
    u_{s_k}; returns {"channels": ..., "total": ..., "k_shells": ...}."""
    _check_spectral(spectral)
    comps = decompose_field(decomposition, velocity_hat, ks, spectral=spectral)
    if not isinstance(comps, dict):
        raise ValueError("calculate_partial_fluxes needs a decomposition that splits u into ≥2 "
                         f"named components; got {decomposition!r}.")
    nd, ns, M, D = _shape_info(velocity_hat, ks, None)
    ft = _real_dtype(velocity_hat)
    geometry = geometry if geometry is not None else ShellMagnitude(None)
    plan = _plan_for(ns, M, M, ks, dealiasing, spectral.compute_dtype(ft), spectral.device,
                     binning, geometry, shells=True)
    centers = _ft_cast(plan.centers, ft)
    channels = {}
    for sp in comps:
        for sq in comps:
            N = plan.nonlinear_term(comps[sq], comps[sp])
            for sk in comps:
                T, Pi = plan.spectral_flux_with_nhat(KineticEnergy(), comps[sk], N)
                channels[(sk, sp, sq)] = SpectralFluxResult(centers, _ft_cast(T, ft),
                                                            _ft_cast(Pi, ft))
    totT = sum(c.transfer_spectrum.astype(np.float64) for c in channels.values())
    totP = sum(c.flux.astype(np.float64) for c in channels.values())
    total = SpectralFluxResult(centers, totT.astype(ft), totP.astype(ft))
    return {"channels": channels, "total": total, "k_shells": centers}


def calculate_helical_partial_fluxes(velocity_hat, ks, **kw):
    """SpectralFlux.jl:295-296."""
    return calculate_partial_fluxes(velocity_hat, ks, decomposition=HelicalDecomposition(), **kw)


# ---------------------------------------------------------------------------------------------
# batch axis (the reference's mpi_batch_map semantics on one GPU)
# ---------------------------------------------------------------------------------------------

def batch_map(f, items, *, reduce="gather"):
    """Apply f to every item and return the results in input order -- the collation contract of
    the reference's batch axis (MPIExt.jl:14-23, :39-68), executed on one device -- then combine
    them as `reduce` asks (MPIExt.jl:25-37): "gather" (or ":gather") returns the ordered list,
    "sum" the left fold with +, "mean" that sum divided by the item count, and a binary callable
    the left fold with it; anything else raises ValueError (the reference's ArgumentError)."""
    tag = reduce.lstrip(":") if isinstance(reduce, str) else None
    if tag not in ("gather", "sum", "mean") and not callable(reduce):
        raise ValueError(f"reduce must be :gather, :sum, :mean, or a binary function (got {reduce!r}).")
    flat = [f(x) for x in items]
    if tag == "gather":
        return flat
    if tag in ("sum", "mean"):
        tot = functools.reduce(operator.add, flat)
        return tot / len(flat) if tag == "mean" else tot
    return functools.reduce(reduce, flat)


class Ensemble:
    """Device-resident ensemble of snapshots (fit_ensemble): the reference's batch axis
    (mpi_batch_map, MPIExt.jl:39-68) with the kinetic-energy flux and, optionally, the scalar
    flux of every snapshot.  Snapshots on the same grid run batched in the same launches (on a
    3/2-padded grid: one after another through the padded pipeline) and the whole ensemble
    replays as one captured HIP graph.

    shapes: per-snapshot grid sizes; ks_list: per-snapshot wavenumber vectors."""

    def __init__(self, shapes, ks_list, *, with_scalar=False, binning=None,
                 dealiasing=OrszagTwoThirds(), precision=np.float32, geometry=None, device=0):
        self.lib = _lib.load_library()
        self.shapes = [tuple(int(n) for n in s) for s in shapes]
        if len(self.shapes) != len(ks_list) or not self.shapes:
            raise ValueError("ks_list must match shapes one to one (and be non-empty)")
        self.nd = len(self.shapes[0])
        if any(len(s) != self.nd for s in self.shapes):
            raise _lib.DimensionMismatch("every snapshot of an ensemble has the same dimension")
        self.io = _lib.F64 if np.dtype(precision) == np.float64 else _lib.F32
        self.with_scalar = bool(with_scalar)
        geometry = geometry if geometry is not None else ShellMagnitude(None)
        n = len(self.shapes)
        self._ks = []
        self._shells = []
        self.centers = []
        cache = {}
        for s, ks in zip(self.shapes, ks_list):
            kk = [np.ascontiguousarray(np.asarray(k, dtype=np.float64)) for k in ks]
            if len(kk) != self.nd or any(len(k) != m for k, m in zip(kk, s)):
                raise _lib.DimensionMismatch("ks must hold nd vectors of lengths ns")
            key = (s, tuple(k.tobytes() for k in kk))
            if key not in cache:
                c, _, sidx = shell_setup(kk, binning, geometry)
                cache[key] = (c, np.asfortranarray(np.asarray(sidx, dtype=np.int64)))
            c, sidx = cache[key]
            self._ks.extend(kk)
            self._shells.append(sidx)
            self.centers.append(c)
        self.n_shells = [len(c) for c in self.centers]
        nsarr = (ctypes.c_int64 * (n * self.nd))(*[m for s in self.shapes for m in s])
        ksp = (ctypes.POINTER(ctypes.c_double) * len(self._ks))(*[_dptr(k) for k in self._ks])
        shp = (ctypes.POINTER(ctypes.c_int64) * n)(
            *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)) for a in self._shells])
        nsh = (ctypes.c_int * n)(*self.n_shells)
        self.handle = ctypes.c_void_p()
        rc = self.lib.fit_ensemble_create(ctypes.byref(self.handle), n, self.nd, nsarr, self.io,
                                          dealiasing.code, ksp, shp, nsh, int(self.with_scalar),
                                          device)
        _lib.check(rc)
        self.offsets = np.concatenate([[0], np.cumsum(self.n_shells)]).astype(np.int64)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.fit_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def groups(self):
        return int(self.lib.fit_ensemble_groups(self.handle))

    def set_option(self, option, value):
        _lib.check(self.lib.fit_ensemble_set_option(self.handle, int(option), int(value)))

    def set_graph(self, on=True):
        self.set_option(_lib.ENS_OPT_GRAPH, int(bool(on)))

    def set_batch(self, n):
        self.set_option(_lib.ENS_OPT_BATCH, int(n))

    def set_plan_option(self, option, value):
        """Plan option `option` (the plans' OPT_* codes) on every same-grid group's plan."""
        self.set_option(_lib.ENS_OPT_PLAN + int(option), int(value))

    def stream(self):
        return self.lib.fit_ensemble_stream(self.handle)

    def profile(self, on=True):
        _lib.check(self.lib.fit_ensemble_profile_enable(self.handle, int(on)))

    def profile_stats(self):
        out = {}
        for i, name in enumerate(_lib.K_CLASSES):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            _lib.check(self.lib.fit_ensemble_profile_query(self.handle, i, ctypes.byref(n),
                                                           ctypes.byref(ms), ctypes.byref(by)))
            out[name] = {"launches": n.value, "ms": ms.value, "alg_bytes": by.value}
        return out

    def kernel_stats(self):
        return _lib.kernel_stats(self.lib.fit_ensemble_profile_kernel_count, self.lib.fit_ensemble_profile_kernel,
                                 self.handle)

    def raw_fluxes(self, velocities, scalars=None):
        """-> (T_ke, Pi_ke, T_sc, Pi_sc) concatenated over snapshots (offsets in self.offsets)."""
        n = len(self.shapes)
        if len(velocities) != n or (self.with_scalar and (scalars is None or len(scalars) != n)):
            raise ValueError("one velocity (and one scalar) per snapshot")
        keep = []
        up = (ctypes.c_void_p * n)()
        tp = (ctypes.c_void_p * n)()
        kind = None
        for j in range(n):
            s = self.shapes[j]
            if tuple(int(m) for m in velocities[j].shape) != s + (self.nd,):
                raise _lib.DimensionMismatch(f"snapshot {j}: velocity must be (ns..., {self.nd})")
            p, k, kd = _marshal(velocities[j], self.io)
            keep.append(k)
            up[j] = p
            kinds = {kd}
            if self.with_scalar:
                th = scalars[j]
                if tuple(int(m) for m in th.shape[:self.nd]) != s:
                    raise _lib.DimensionMismatch(f"snapshot {j}: scalar grid differs from the velocity's")
                p, k, kd2 = _marshal(th, self.io)
                keep.append(k)
                tp[j] = p
                kinds.add(kd2)
            if kind is None:
                kind = kd
            kinds.add(kind)
            if len(kinds) != 1:
                raise ValueError("all ensemble arrays must be host arrays or all device arrays")
        tot = int(self.offsets[-1])
        T, Pi = np.zeros(tot), np.zeros(tot)
        Ts, Ps = np.zeros(tot), np.zeros(tot)
        if kind == _lib.PTR_DEVICE:
            _lib.check(self.lib.fit_ensemble_wait_stream(self.handle, _caller_stream(keep[0])))
        _lib.check(self.lib.fit_ensemble_flux(self.handle, up, tp if self.with_scalar else None,
                                              _dptr(T), _dptr(Pi), _dptr(Ts), _dptr(Ps), self.io,
                                              kind))
        return T, Pi, Ts, Ps

    def fluxes(self, velocities, scalars=None):
        """Per snapshot, in input order: (kinetic-energy SpectralFluxResult, scalar-variance
        SpectralFluxResult or None)."""
        T, Pi, Ts, Ps = self.raw_fluxes(velocities, scalars)
        out = []
        for j in range(len(self.shapes)):
            ft = _real_dtype(velocities[j])
            a, b = self.offsets[j], self.offsets[j + 1]
            c = _ft_cast(self.centers[j], ft)
            ke = SpectralFluxResult(c, _ft_cast(T[a:b], ft), _ft_cast(Pi[a:b], ft))
            sc = SpectralFluxResult(c, _ft_cast(Ts[a:b], ft), _ft_cast(Ps[a:b], ft)) \
                if self.with_scalar else None
            out.append((ke, sc))
        return out


def calculate_ensemble_fluxes(velocities, ks_list, scalars=None, *, binning=None,
                              dealiasing=OrszagTwoThirds(), spectral=HIPSpectral(),
                              invariant=KineticEnergy(), geometry=None, physical=False):
    """Spectral fluxes of an ensemble of snapshots of possibly different sizes, in input order
    (the collation contract of mpi_batch_map, MPIExt.jl:14-23, :39-68).

    velocities[j]: (ns_j..., D) array; ks_list[j]: its wavenumber vectors; scalars[j] (optional):
    a scalar advected by velocities[j]; physical=True: the arrays are physical fields, transformed
    to û = fft(u)/Np on the device first.  Returns a list of (flux of `invariant`, scalar-variance
    flux or None).  Kinetic-energy ensembles of 2D / 3D snapshots (D = nd) run as one
    graph-captured fit_ensemble; anything else runs snapshot by snapshot on cached plans."""
    _check_spectral(spectral)
    if physical:
        # physical snapshots: û = fft(u)/Np on the device first (PencilFFTsExt.jl:83-88)
        velocities = [physical_to_spectral(v, k, spectral=spectral) for v, k in zip(velocities, ks_list)]
        if scalars is not None:
            scalars = [physical_to_spectral(t, k, spectral=spectral) for t, k in zip(scalars, ks_list)]
    if scalars is not None and len(scalars) != len(velocities):
        raise ValueError("scalars must match velocities one to one")
    if len(ks_list) != len(velocities):
        raise ValueError("ks_list must match velocities one to one")
    if not velocities:
        return []
    nd = len(ks_list[0])
    batched = (isinstance(invariant, KineticEnergy) and nd in (2, 3)
               and all(len(k) == nd and int(v.shape[nd]) == nd and len(v.shape) == nd + 1
                       for v, k in zip(velocities, ks_list))
               and len({_real_dtype(v) for v in velocities}) == 1
               and len({_is_device(v) for v in velocities}) == 1)
    if batched:
        ft = _real_dtype(velocities[0])
        shapes = [tuple(int(m) for m in v.shape[:nd]) for v in velocities]
        key = (tuple(shapes), tuple(tuple(np.asarray(k, dtype=np.float64).tobytes() for k in ks)
                                    for ks in ks_list),
               type(dealiasing).__name__, np.dtype(spectral.compute_dtype(ft)).name, spectral.device,
               repr(binning), repr(geometry), scalars is not None)
        key = (threading.get_ident(),) + key
        with _CACHE_LOCK:
            ens = _ENSEMBLES.get(key)
        if ens is None:
            ens = Ensemble(shapes, ks_list, with_scalar=scalars is not None, binning=binning,
                           dealiasing=dealiasing, precision=spectral.compute_dtype(ft),
                           geometry=geometry, device=spectral.device)
            with _CACHE_LOCK:
                for k in [k for k in _ENSEMBLES if k[0] == key[0]]:
                    del _ENSEMBLES[k]   # one cached ensemble per thread
                _ENSEMBLES[key] = ens
        sc = None
        if scalars is not None:
            sc = [as_component_field(s, nd) for s in scalars]
        return ens.fluxes(velocities, sc)

    def one(j):
        ke = calculate_spectral_flux(velocities[j], ks_list[j], binning=binning, dealiasing=dealiasing,
                                     invariant=invariant, spectral=spectral, geometry=geometry)
        sc = None
        if scalars is not None:
            sc = calculate_scalar_flux(velocities[j], scalars[j], ks_list[j], binning=binning,
                                       dealiasing=dealiasing, spectral=spectral, geometry=geometry)
        return ke, sc

    return batch_map(one, range(len(velocities)))


_ENSEMBLES = {}
