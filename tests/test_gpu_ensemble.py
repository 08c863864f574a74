"""Ensembles (fit_ensemble through the C ABI): uneven snapshot sizes, batched per grid and replayed
as one HIP graph, against the FP64 oracle snapshot by snapshot (the reference's mpi_batch_map
collation: results in input order, MPIExt.jl:14-23, :39-68).

Tolerances (normwise per snapshot): FP64 compute 1e-11, FP32 compute 1e-5 (shell sums).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def nerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def make_ensemble(oracle, shapes, seed0, with_scalar=True):
    us, ths, kss = [], [], []
    for j, s in enumerate(shapes):
        nd = len(s)
        us.append(oracle.synthetic_field(s, nd, seed=seed0 + j, spectrum="k-3" if nd == 2 else "k-5/3"))
        if with_scalar:
            ths.append(oracle.synthetic_field(s, 1, seed=seed0 + 100 + j, spectrum="k-1",
                                              solenoidal=False)[..., 0])
        kss.append(oracle.wavenumber_grid(s, (2 * math.pi,) * nd))
    return us, (ths if with_scalar else None), kss


def to_dev(a, dtype):
    import torch
    rev = tuple(reversed(range(a.ndim)))
    return torch.from_numpy(np.ascontiguousarray(a.transpose(rev))).to("cuda").permute(rev).to(dtype)


def check(fit, oracle, res, us, ths, kss, tol, dealias="23"):
    for j, (ke, sc) in enumerate(res):
        _, T, Pi = oracle.calculate_spectral_flux(us[j], kss[j], oracle.LinearBinning(1.0), dealias)
        e = max(nerr(ke.transfer_spectrum, T), nerr(ke.flux, Pi))
        assert e < tol, (j, "kinetic energy", e)
        if ths is not None:
            _, Ts, Ps = oracle.calculate_scalar_flux(us[j], ths[j], kss[j], oracle.LinearBinning(1.0), dealias)
            e = max(nerr(sc.transfer_spectrum, Ts), nerr(sc.flux, Ps))
            assert e < tol, (j, "scalar", e)
        else:
            assert sc is None


SHAPES_2D = [(16, 16), (32, 32), (16, 16), (24, 24), (32, 32), (16, 16), (16, 24), (24, 24)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ensemble_2d_uneven_matches_oracle(fit, oracle, prec):
    import torch
    print("case ensemble 2D", prec, flush=True)
    dt = torch.complex128 if prec == "f64" else torch.complex64
    tol = 1e-11 if prec == "f64" else 1e-5
    us, ths, kss = make_ensemble(oracle, SHAPES_2D, 4000)
    ud = [to_dev(u, dt) for u in us]
    td = [to_dev(t, dt) for t in ths]
    ens = fit.Ensemble(SHAPES_2D, kss, with_scalar=True, binning=fit.LinearBinning(1.0),
                       precision=np.float64 if prec == "f64" else np.float32)
    assert ens.groups == 4
    res = ens.fluxes(ud, td)                # eager run + graph capture
    check(fit, oracle, res, us, ths, kss, tol)
    # replay of the captured graph on new arrays (only the pointer tables change)
    us2, ths2, _ = make_ensemble(oracle, SHAPES_2D, 5000)
    res2 = ens.fluxes([to_dev(u, dt) for u in us2], [to_dev(t, dt) for t in ths2])
    check(fit, oracle, res2, us2, ths2, kss, tol)
    # and once more on the first arrays: a replay reproduces the eager run bit for bit (the shell
    # bins add in a fixed order inside and across workgroups)
    res3 = ens.fluxes(ud, td)
    for (a, b), (c, d) in zip(res, res3):
        assert np.array_equal(c.flux, a.flux) and np.array_equal(d.flux, b.flux)
        assert np.array_equal(c.transfer_spectrum, a.transfer_spectrum)
    ens.close()


def test_ensemble_batches_graph_and_host_inputs_agree(fit, oracle):
    import torch
    print("case ensemble options", flush=True)
    us, ths, kss = make_ensemble(oracle, SHAPES_2D, 4100)
    ud = [to_dev(u, torch.complex128) for u in us]
    td = [to_dev(t, torch.complex128) for t in ths]
    ref = fit.Ensemble(SHAPES_2D, kss, with_scalar=True, precision=np.float64)
    base = ref.raw_fluxes(ud, td)
    # batches of one snapshot, eager (no graph)
    ens = fit.Ensemble(SHAPES_2D, kss, with_scalar=True, precision=np.float64)
    ens.set_batch(1)
    ens.set_graph(False)
    for a, b in zip(ens.raw_fluxes(ud, td), base):
        assert nerr(a, b) < 1e-13
    # host numpy inputs (staged through the device) and batches of two, replayed twice
    ens.set_batch(2)
    ens.set_graph(True)
    for _ in range(2):
        for a, b in zip(ens.raw_fluxes(us, ths), base):
            assert nerr(a, b) < 1e-13
    # FP64 inputs into an FP32 ensemble (converted on the device)
    e32 = fit.Ensemble(SHAPES_2D, kss, with_scalar=True, precision=np.float32)
    for a, b in zip(e32.raw_fluxes(ud, td), base):
        assert nerr(a, b) < 1e-5
    for e in (ref, ens, e32):
        e.close()


def test_ensemble_3d_kinetic_energy_only(fit, oracle):
    import torch
    print("case ensemble 3D", flush=True)
    shapes = [(12, 12, 12), (16, 16, 16), (12, 12, 12), (8, 12, 16)]
    us, _, kss = make_ensemble(oracle, shapes, 4200, with_scalar=False)
    ud = [to_dev(u, torch.complex64) for u in us]
    res = fit.calculate_ensemble_fluxes(ud, kss, binning=fit.LinearBinning(1.0))
    check(fit, oracle, res, us, None, kss, 1e-5)
    res2 = fit.calculate_ensemble_fluxes(ud, kss, binning=fit.LinearBinning(1.0))   # graph replay
    for (a, _), (b, _) in zip(res, res2):
        assert nerr(b.flux, a.flux) < 1e-6


def test_ensemble_api_uses_batched_path(fit, oracle):
    import torch
    us, ths, kss = make_ensemble(oracle, SHAPES_2D[:3], 4300)
    res = fit.calculate_ensemble_fluxes([to_dev(u, torch.complex64) for u in us], kss,
                                        [to_dev(t, torch.complex64) for t in ths])
    check(fit, oracle, res, us, ths, kss, 1e-5)


def test_ensemble_plan_options_reach_every_group(fit, oracle):
    """FIT_ENS_OPT_PLAN + o sets plan option o on every group's plan: the workgroup-form x pass and
    a 1 KiB bin budget (ordered commits) give the same fluxes; bad values are rejected."""
    import torch
    from fit_amd import _lib
    us, ths, kss = make_ensemble(oracle, SHAPES_2D[:4], 4600)
    ud = [to_dev(u, torch.complex64) for u in us]
    td = [to_dev(t, torch.complex64) for t in ths]
    ens = fit.Ensemble(SHAPES_2D[:4], kss, with_scalar=True, binning=fit.LinearBinning(1.0))
    base = ens.raw_fluxes(ud, td)
    ens.set_plan_option(_lib.OPT_XPASS_BLOCK, 1)
    ens.set_plan_option(_lib.OPT_LAST_BINS, 1)
    for a, b in zip(ens.raw_fluxes(ud, td), base):
        assert nerr(a, b) < 1e-5
    with pytest.raises(ValueError):
        ens.set_plan_option(_lib.OPT_LAST_BINS, 0)
    with pytest.raises(ValueError):
        ens.set_plan_option(999, 1)
    ens.close()


PADDED_SHAPES = {2: [(16, 16), (32, 32), (16, 16), (24, 24), (16, 24)],
                 3: [(8, 8, 8), (12, 12, 12), (8, 8, 8)]}


@pytest.mark.parametrize("nd,prec", [(2, "f64"), (2, "f32"), (3, "f32")])
def test_ensemble_padded_matches_oracle(fit, oracle, nd, prec):
    """PaddedThreeHalves ensembles: every snapshot through the padded pipeline, against the
    oracle's padded flux; new input arrays capture the graph again, and a replay on the first
    arrays reproduces the first call bit for bit."""
    import torch
    print("case ensemble padded", nd, prec, flush=True)
    dt = torch.complex128 if prec == "f64" else torch.complex64
    tol = 1e-11 if prec == "f64" else 1e-5
    shapes = PADDED_SHAPES[nd]
    us, ths, kss = make_ensemble(oracle, shapes, 4400 + nd)
    ud = [to_dev(u, dt) for u in us]
    td = [to_dev(t, dt) for t in ths]
    ens = fit.Ensemble(shapes, kss, with_scalar=True, binning=fit.LinearBinning(1.0),
                       dealiasing=fit.PaddedThreeHalves(),
                       precision=np.float64 if prec == "f64" else np.float32)
    res = ens.fluxes(ud, td)
    check(fit, oracle, res, us, ths, kss, tol, "pad")
    us2, ths2, _ = make_ensemble(oracle, shapes, 4500 + nd)
    res2 = ens.fluxes([to_dev(u, dt) for u in us2], [to_dev(t, dt) for t in ths2])
    check(fit, oracle, res2, us2, ths2, kss, tol, "pad")
    res3 = ens.fluxes(ud, td)
    res4 = ens.fluxes(ud, td)               # graph replay on the same pointers
    for r in (res3, res4):
        for (a, b), (c, d) in zip(res, r):
            assert np.array_equal(c.flux, a.flux) and np.array_equal(d.flux, b.flux)
    ens.close()
    # the functional API routes padded ensembles through fit_ensemble as well
    res5 = fit.calculate_ensemble_fluxes(ud, kss, td, binning=fit.LinearBinning(1.0),
                                         dealiasing=fit.PaddedThreeHalves())
    check(fit, oracle, res5, us, ths, kss, tol, "pad")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ensemble_register_lengths_carrier_in_last_pass(fit, oracle, prec):
    """Snapshots whose last axis has a register FFT (128, 192, 256): the last pass forms each
    snapshot's Hermitian carrier from its own array (no carrier pass) -- against the oracle."""
    import torch
    print("case ensemble register lengths", prec, flush=True)
    dt = torch.complex128 if prec == "f64" else torch.complex64
    tol = 1e-11 if prec == "f64" else 1e-5
    shapes = [(128, 128), (256, 256), (128, 128), (192, 192), (256, 256)]
    us, ths, kss = make_ensemble(oracle, shapes, 700)
    res = fit.calculate_ensemble_fluxes([to_dev(u, dt) for u in us], kss, [to_dev(t, dt) for t in ths],
                                        binning=fit.LinearBinning(1.0))
    res = fit.calculate_ensemble_fluxes([to_dev(u, dt) for u in us], kss, [to_dev(t, dt) for t in ths],
                                        binning=fit.LinearBinning(1.0))          # graph replay
    check(fit, oracle, res, us, ths, kss, tol)
