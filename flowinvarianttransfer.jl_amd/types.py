"""Vocabulary mirroring the reference's singleton types (src/Types.jl) plus the HIP backend's
own backend singletons.  Keyword values select behaviour exactly as in the reference."""
from dataclasses import dataclass
from typing import Any, Optional, Tuple

import numpy as np


class _Singleton:
    """Value-less tag type (the reference's singleton structs)."""
    code = -1

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)

    def __repr__(self):
        return f"{type(self).__name__}()"


# --- invariants (Types.jl:69-132) -------------------------------------------------------------
class AbstractInvariant(_Singleton):
    pass


class KineticEnergy(AbstractInvariant):
    code = 0


class Helicity(AbstractInvariant):
    code = 1


class Enstrophy(AbstractInvariant):
    code = 2


class PassiveScalar(AbstractInvariant):
    code = 3


# --- dealiasing (Types.jl:451-476) ------------------------------------------------------------
class AbstractDealiasing(_Singleton):
    pass


class NoDealiasing(AbstractDealiasing):
    code = 0


class OrszagTwoThirds(AbstractDealiasing):
    code = 1


class PaddedThreeHalves(AbstractDealiasing):
    code = 2


# --- decompositions (Types.jl:136-310; helical only on this path) ----------------------------
class AbstractFieldDecomposition(_Singleton):
    pass


class NoDecomposition(AbstractFieldDecomposition):
    pass


class HelicalDecomposition(AbstractFieldDecomposition):
    """Types.jl: positive / negative helical parts (Decomposition.jl:63-98)."""
    names = ("positive", "negative")
    code = 1


class ToroidalPoloidalDecomposition(AbstractFieldDecomposition):
    """Types.jl:195-209: toroidal / poloidal Craya-Herring parts (Decomposition.jl:107-141)."""
    names = ("toroidal", "poloidal")
    code = 2


# --- binning (Types.jl:370-421) ---------------------------------------------------------------
@dataclass(frozen=True)
class LinearBinning:
    dk: float


@dataclass(frozen=True)
class LogarithmicBinning:
    k0: float
    lam: float


@dataclass(frozen=True)
class DyadicBinning:
    k0: float


@dataclass(frozen=True)
class CustomBinning:
    edges: Tuple[float, ...]

    def __init__(self, edges):
        object.__setattr__(self, "edges", tuple(float(e) for e in edges))


# --- geometry (Types.jl:487-531); dims are 1-based like the reference ------------------------
@dataclass(frozen=True)
class ShellMagnitude:
    dims: Optional[Tuple[int, ...]] = None


def IsotropicShells():
    return ShellMagnitude(None)


def PerpendicularShells(dims=(1, 2)):
    return ShellMagnitude(tuple(dims))


def ParallelShells(dims=(3,)):
    return ShellMagnitude(tuple(dims))


# --- backends (Types.jl:544-635) --------------------------------------------------------------
class AbstractSpectralBackend(_Singleton):
    pass


class AbstractExecutionBackend(_Singleton):
    pass


@dataclass(frozen=True)
class HIPSpectral(AbstractSpectralBackend):
    """MI355X pseudospectral transform backend (hand-written Stockham FFT kernels).

    precision: numpy real dtype the device computes in (np.float32 / np.float64), or None to
    follow the input's precision (the reference's FT = real(eltype(û)) rule)."""
    precision: Any = None
    device: int = 0

    def compute_dtype(self, input_real_dtype):
        p = self.precision if self.precision is not None else input_real_dtype
        return np.dtype(p)


@dataclass(frozen=True)
class HIPExecution(AbstractExecutionBackend):
    """Giver-shell loop executed on the device (replaces GPUBackend{B} of the reference)."""
    device: int = 0


# --- results (Types.jl:654-751) ---------------------------------------------------------------
@dataclass
class SpectralFluxResult:
    k_shells: np.ndarray
    transfer_spectrum: np.ndarray
    flux: np.ndarray


@dataclass
class ShellToShellResult:
    shell_centers: np.ndarray
    shell_edges: np.ndarray
    transfer_matrix: np.ndarray
    net_transfer: np.ndarray
    max_antisymmetry_error: float


@dataclass
class BandToBandResult:
    """The NamedTuple calculate_band_to_band_transfer returns (BandTransfer.jl:110-112)."""
    centers: np.ndarray
    transfer_matrix: np.ndarray
    net_transfer: np.ndarray
    max_antisymmetry_error: float


@dataclass(frozen=True)
class SmoothBands:
    """Graded spectral bands for band-to-band transfer (Types.jl:414-430): log-Gaussian weights
    about `centers` (increasing, > 0) of width `logwidth` in ln κ, renormalised to a partition of
    unity."""
    centers: Tuple[float, ...]
    logwidth: float = 0.6

    def __post_init__(self):
        object.__setattr__(self, "centers", tuple(float(c) for c in self.centers))
        object.__setattr__(self, "logwidth", float(self.logwidth))


@dataclass
class ModeToModeTriadResult:
    invariant: AbstractInvariant
    ks: tuple
    net_transfer: Any
    transfer: Any
