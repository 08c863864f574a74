"""MI355X (gfx950) backend for FlowInvariantTransfer.jl's pseudospectral transfer path.

Public names mirror the reference (src/FlowInvariantTransfer.jl exports for this path); every
computation runs in libfit_hip.so (hand-written HIP kernels), loaded through ctypes.
"""
from ._lib import DimensionMismatch, LIB_PATH, load_library
from .api import (Ensemble, Plan, batch_map, calculate_band_to_band_transfer, calculate_ensemble_fluxes, calculate_helical_partial_fluxes, calculate_mode_to_mode_transfer,
                  calculate_partial_fluxes, calculate_scalar_flux,
                  calculate_scalar_mode_to_mode_transfer, calculate_scalar_shell_to_shell_transfer,
                  calculate_shell_to_shell_transfer, calculate_spectral_flux, clear_plan_cache,
                  compute_nonlinear_term, decompose_field, pencil_spectral_flux, physical_to_spectral,
                  transfer_density)
from .grid import (as_component_field, assign_shells, default_binning, n_shells, shell_centers,
                   shell_coordinate, shell_edges, wavenumber_grid, wavenumber_magnitude_grid)
from .types import (BandToBandResult, CustomBinning, DyadicBinning, Enstrophy, Helicity, HelicalDecomposition,
                    HIPExecution, HIPSpectral, IsotropicShells, KineticEnergy, LinearBinning,
                    LogarithmicBinning, ModeToModeTriadResult, NoDealiasing, NoDecomposition,
                    OrszagTwoThirds, PaddedThreeHalves, ParallelShells, PassiveScalar,
                    PerpendicularShells, ShellMagnitude, ShellToShellResult, SmoothBands,
                    SpectralFluxResult, ToroidalPoloidalDecomposition)

__all__ = [n for n in dir() if not n.startswith("_")]
