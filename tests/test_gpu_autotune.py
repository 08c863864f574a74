"""GPU: the autotuner (fit_plan_autotune).  Tuned plans give the oracle's results (to the same
tolerances as the default plans), the report names the workload and its times, the choice is
cached per plan key (a new plan with the same key starts from it) and appended to the file
FIT_TUNE_CACHE names.  The whole parity suite also runs under tuned plans when FIT_AUTOTUNE=1."""
import math
import os
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


@pytest.mark.parametrize("ns,dea,workload", [((32, 32, 32), "pad", "flux"), ((48, 40, 32), "23", "flux"),
                                             ((48, 48, 48), "23", "shell_to_shell"),
                                             ((40, 36), "none", "shell_to_shell")])
def test_autotuned_plan_matches_oracle(fit, oracle, ns, dea, workload, tmp_path):
    from fit_amd.grid import shell_setup
    say("case autotune", ns, dea, workload)
    nd = len(ns)
    ks = oracle.wavenumber_grid(ns, (2 * math.pi,) * nd)
    u = oracle.synthetic_field(ns, nd, seed=130)
    deal = {"none": fit.NoDealiasing(), "23": fit.OrszagTwoThirds(), "pad": fit.PaddedThreeHalves()}[dea]
    binning = fit.LinearBinning(1.0) if workload == "flux" else fit.CustomBinning(range(0, 8))
    c, _, sidx = shell_setup(ks, binning, None)
    cache = tmp_path / "tune.txt"
    os.environ["FIT_TUNE_CACHE"] = str(cache)
    try:
        with fit.Plan(ns, nd, nd, ks, deal, np.float64, sidx, len(c)) as plan:
            default = plan.options()
            rep = plan.autotune(workload, fit.KineticEnergy(), u, reps=2)
            say("  ", rep["report"])
            assert rep["report"].startswith("workload=" + workload) and "tuned_ms=" in rep["report"]
            tuned = plan.options()
            if workload == "flux":
                _, To, _ = oracle.calculate_spectral_flux(u, ks, oracle.LinearBinning(1.0), dea)
                T, _ = plan.spectral_flux(fit.KineticEnergy(), u)
            else:
                _, _, To, _, _ = oracle.calculate_shell_to_shell_transfer(
                    u, ks, oracle.CustomBinning(tuple(range(0, 8))), dea)
                T, _, _ = plan.shell_to_shell(fit.KineticEnergy(), u)
            assert rel(T, To) < 1e-11
        with fit.Plan(ns, nd, nd, ks, deal, np.float64, sidx, len(c)) as again:
            assert again.options() == tuned                     # the cached choice is applied
        lines = cache.read_text().splitlines()
        assert lines and lines[-1].split("|")[0].endswith(":flux" if workload == "flux" else ":s2s")
        say("  default", default, "tuned", tuned)
    finally:
        os.environ.pop("FIT_TUNE_CACHE", None)
