/*
 * fit_hip.h -- C ABI of libfit_hip.so, the MI355X (gfx950) backend for FlowInvariantTransfer.jl's
 * pseudospectral transfer path.
 *
 * Every entry point replaces one override point of the reference (file:line in the reference
 * repository).  Arrays follow the reference's Julia layout: column-major (first index fastest),
 * shape (ns..., C) with the component axis last, complex values interleaved (re, im).
 *   - complex inputs/outputs: FIT_F32 -> ComplexF32 (8 B), FIT_F64 -> ComplexF64 (16 B)
 *   - ks: nd arrays of ns[d] doubles in FFTW index order (wavenumber_grid, Utils.jl:51-63)
 *   - shell_idx: Np Int64 values, 1-based shell number, 0 = outside every shell
 *     (assign_shells, ShellBinning.jl:126-140)
 *   - T, Pi, Tmat, net are host arrays of doubles; Tmat is column-major n_sh x n_sh,
 *     Tmat[n + n_sh*m] = transfer INTO shell n FROM giver shell m.
 * Return value: 0 on success, otherwise a FIT_E* code; fit_last_error() gives the message
 * (thread-local).  A plan is not safe for concurrent calls; use one plan per host thread.
 */
#ifndef FIT_HIP_H
#define FIT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fit_plan fit_plan;

/* compute / io precision */
enum { FIT_F32 = 0, FIT_F64 = 1 };
/* dealiasing: NoDealiasing, OrszagTwoThirds, PaddedThreeHalves (Types.jl:451-476) */
enum { FIT_DEALIAS_NONE = 0, FIT_DEALIAS_TWOTHIRDS = 1, FIT_DEALIAS_PADDED = 2 };
/* invariants (Types.jl:69-132) */
enum { FIT_KINETIC_ENERGY = 0, FIT_HELICITY = 1, FIT_ENSTROPHY = 2, FIT_PASSIVE_SCALAR = 3 };
/* where a data pointer lives */
enum { FIT_PTR_HOST = 0, FIT_PTR_DEVICE = 1 };

/* error codes: ARGUMENT ~ Julia ArgumentError, DIMENSION ~ DimensionMismatch */
enum {
  FIT_OK = 0,
  FIT_E_ARGUMENT = 1,
  FIT_E_DIMENSION = 2,
  FIT_E_UNSUPPORTED = 3,
  FIT_E_OUT_OF_MEMORY = 4,
  FIT_E_DEVICE = 5
};

/* Message of the last failing call on this thread ("" if none). */
const char* fit_last_error(void);

/* Library version string. */
const char* fit_version(void);

/*
 * Build a plan: device-resident workspaces, twiddle tables, wavenumbers and shell map for one
 * grid.  Replaces Workspaces._make_fft_plans (Workspaces.jl:22, FFTWExt.jl:37-49) together with
 * the NonlinearTermWorkspace / SpectralFluxWorkspace / ShellToShellWorkspace constructors
 * (Workspaces.jl:63-75, 104-117, 154-170).
 *   nd in 1..3; ns[d] of the form 2^a 3^b 5^c; M = advected components; D = advecting velocity
 *   components (>= nd); shell_idx may be NULL when n_shells == 0 (nonlinear term only).
 */
int fit_plan_create(fit_plan** plan, int nd, const int64_t* ns, int M, int D, int compute_dtype,
                    int dealias, const double* const* ks, const int64_t* shell_idx, int n_shells,
                    int device);

int fit_plan_destroy(fit_plan* plan);

/* Device bytes the plan holds. */
int64_t fit_plan_device_bytes(const fit_plan* plan);

/* Number of plans made by fit_plan_create and not yet destroyed in this process (all threads,
 * all devices).  No reference counterpart: a check for bindings that must build one plan per
 * workspace (Workspaces.jl:22 is called once per NonlinearTermWorkspace) and reuse it, and for
 * plan leaks. */
int64_t fit_plan_live_count(void);

/* Plan options (no reference counterpart: execution knobs of this backend).
 *   FIT_OPT_PLAIN_PIPELINE = 1: value 1 runs the plain dataflow (unfused, unpruned passes, one
 *     giver at a time, N̂ through HBM) -- the comparison baseline; 0 (default) the fused one.
 *   FIT_OPT_GIVER_BATCH = 2: cap on giver shells per batched launch (0 = device memory budget).
 *   FIT_OPT_XPASS_GIVERS = 3: givers per workgroup in the fused contiguous-axis pass (0 = 4).
 *   FIT_OPT_XPASS_BLOCK = 4: value 1 selects the workgroup-barrier form of that pass.
 *   FIT_OPT_SUBGRID = 5: value 0 keeps the giver loop on the plan's own grid even when a
 *     band-limited sub-grid is attached (1, default: use it).
 *   FIT_OPT_COL_TILE = 6: columns per workgroup of the strided-axis passes (power of two; 0 =
 *     automatic: 16 within 64 KB of LDS).
 *   FIT_OPT_XPASS_SELF = 7: 1 (default) forms the advecting velocity of a self-advected field in
 *     registers inside the fused contiguous-axis pass; 0 stores it to device memory first (one
 *     more pass: 2.6 ms more on the 29 ms flux of a 512³ field with 3/2 padding).
 *   FIT_OPT_XPASS_STAGE = 8: 1 (default) lets the FP32 self-advection pass at 512- and 768-point
 *     lines stage the next gradient line into LDS by DMA while the current one is transformed;
 *     0 loads each line into registers just before its transform.
 *   FIT_OPT_GRAD_SPLIT = 9: 1 runs the y transform of the last-axis derivative (3D shared-
 *     transform gradients) as its own one-line column pass; 0 keeps it in the two-line gradient
 *     kernel after the other two; -1 (default) splits for y lines shorter than 512 points
 *     (measured: 0.17 ms faster at 192, 0.11 ms slower at 768 on the 512³ snapshot).
 *   FIT_OPT_FUSED_FLUX = 10: 1 makes fit_spectral_flux / fit_spectral_fluxes compute the densities
 *     in the last forward pass of the nonlinear term (N̂ never stored; M <= 3, up to 2 invariants,
 *     register FFT length on the last axis); 0 (default) stores N̂ and reads it in one density
 *     pass (measured faster on the 512³ flux with 3/2 padding).
 *   FIT_OPT_RADIX_ORDER = 11: radix order of the LDS-engine FFT passes (lengths without a
 *     register FFT): 0 (default) radix 8 first, then 4, 2, 3, 5; 1 radix 4 first; 2 the odd
 *     radices first.
 *   FIT_OPT_COL_F64 = 12: FP32 plans only; which strided-axis register-FFT passes (y and z
 *     transforms) run their arithmetic in FP64, with FP64 twiddles and LDS exchanges, loading and
 *     storing the fields in FP32, so that each value is rounded to FP32 once per pass instead of at
 *     every butterfly stage: bit 0 the inverse passes (gradient passes included), bit 1 the forward
 *     y pass, bit 2 the forward z pass.  6 (default): both forward passes.  FP32 error of the
 *     kinetic-energy Π(K) at 512³ with 3/2 padding against FP64 (eight seeded fields, rms /
 *     worst): 6: 3.0e-7 / 4.9e-7 for 1.2 ms more per flux than 0; 4: 3.5e-7 / 6.8e-7 (+0.6 ms);
 *     2: 4.2e-7 / 6.0e-7 (+0.7 ms); 0 (all FP32): 4.5e-7 / 6.7e-7; 7 also the inverse passes.
 *   FIT_OPT_COL_PIPE = 13: 0 (default) one tile of the other index per workgroup in the strided
 *     register-FFT passes; n > 0 the software-pipelined form, each workgroup walking n tiles and
 *     loading the next tile's lines while it transforms the current one.
 *   FIT_OPT_LAST_BINS = 14: LDS budget in KiB (1 to 128, default 64) for per-wave private FP64
 *     shell bins in the fused last pass; above it the waves share one copy and add their points
 *     in order through one wave.  When the private copies fit only with half the columns per
 *     workgroup, the pass runs half the columns (2D ensemble, 2048² and 1024² last passes: 1.52
 *     -> 1.39 and 0.32 -> 0.26 ms against a 16 KiB budget).
 */
enum { FIT_OPT_PLAIN_PIPELINE = 1, FIT_OPT_GIVER_BATCH = 2, FIT_OPT_XPASS_GIVERS = 3,
       FIT_OPT_XPASS_BLOCK = 4, FIT_OPT_SUBGRID = 5, FIT_OPT_COL_TILE = 6, FIT_OPT_XPASS_SELF = 7,
       FIT_OPT_XPASS_STAGE = 8, FIT_OPT_GRAD_SPLIT = 9, FIT_OPT_FUSED_FLUX = 10, FIT_OPT_RADIX_ORDER = 11,
       FIT_OPT_COL_F64 = 12, FIT_OPT_COL_PIPE = 13, FIT_OPT_LAST_BINS = 14 };
int fit_plan_set_option(fit_plan* plan, int option, int value);
int fit_plan_get_option(const fit_plan* plan, int option, int* value);

/* Autotuner (no reference counterpart; the reference's FFT planning, FFTWExt.jl:42-45, is the
 * nearest equivalent): times the plan's execution knobs on the caller's own input by coordinate
 * descent and keeps the fastest -- column tile, x-pass form and givers per workgroup, gradient
 * split, LDS staging, self-advection pass, fused density, giver batch depth (against the device
 * memory the batches may use), band-limited sub-grid, radix order.  workload FIT_TUNE_FLUX times
 * fit_spectral_flux of `invariant`, FIT_TUNE_SHELL_TO_SHELL times fit_shell_to_shell; each timing
 * is the best of `reps` calls after one untimed call, and a change must win by 2%.  The choice is
 * cached per (grid, components, precision, dealiasing, shells, workload): plans created later with
 * the same key start from it, also in other processes when the environment variable
 * FIT_TUNE_CACHE names a file (one line per choice, appended).  fit_plan_tuning writes the last
 * report ("workload=... start_ms=... optN=V ... tuned_ms=...") and returns its length. */
enum { FIT_TUNE_FLUX = 1, FIT_TUNE_SHELL_TO_SHELL = 2 };
int fit_plan_autotune(fit_plan* plan, int workload, int invariant, const void* f_hat, const void* adv_hat,
                      int io_dtype, int ptr_kind, int reps);
int fit_plan_tuning(const fit_plan* plan, char* buf, int len);

/* Band-limited sub-grid of the giver loop (no reference counterpart).  When every shell lies in
 * |k_d| <= R_d with 4 R_d < n_d, fit_plan_create attaches a grid of L_d >= 4 R_d + 1 points per
 * axis on which fit_shell_to_shell evaluates the same triad sums alias-free (ShellToShell
 * results equal the full-grid ones in exact arithmetic).  Returns 1 and writes L_d to dims[0..2]
 * when such a grid is attached, 0 otherwise; negative on a NULL plan. */
int fit_plan_subgrid(const fit_plan* plan, int64_t* dims);

/* Giver classes of the band-limited giver loop: givers of smaller spectral support need fewer
 * advecting modes (|q_d| <= R_d + P_d for support P_d), so consecutive giver shells are grouped by
 * the smallest alias-free grid (L_d >= 2 R_d + 2 P_d + 1) and each class runs on its own grid.
 * Writes up to max_classes entries: dims[3i..3i+2] = L_d of class i, givers[2i], givers[2i+1] =
 * its first and last giver shell (1-based).  Returns the number of classes (0: no sub-grid),
 * negative on a NULL plan. */
int fit_plan_subgrid_classes(const fit_plan* plan, int64_t* dims, int* givers, int max_classes);

/*
 * N̂ = keep·FFT[(u_adv·∇) f]/Np into N_hat (ns..., M).  Replaces
 * NonlinearTerm._compute_nonlinear_term! for the FFT path (NonlinearTerm.jl:102-116) i.e.
 * FFTWExt._nonlinear_term_fft! (FFTWExt.jl:61-121) and _nonlinear_term_padded_fft! (:160-183).
 *   f_hat (ns..., M); adv_hat (ns..., D) or NULL for self-advection (M == D required).
 */
int fit_nonlinear_term(fit_plan* plan, const void* f_hat, const void* adv_hat, void* N_hat,
                       int io_dtype, int ptr_kind);

/*
 * Per-mode density t(k) (ns...) in io_dtype real.  Replaces Invariants.transfer_density!
 * (Invariants.jl:55-104).  c_hat: carrier (ns..., M); N_hat: (ns..., M).
 */
int fit_transfer_density(fit_plan* plan, int invariant, const void* c_hat, const void* N_hat,
                         void* t, int io_dtype, int ptr_kind);

/*
 * Spectral transfer T(k_n) and flux Pi(K_n) = +cumsum(T) over the plan's n_shells shells.
 * Replaces SpectralFlux.calculate_spectral_flux! (SpectralFlux.jl:131-180): nonlinear term,
 * density and shell scatter-sum all on the device.  T and Pi: host arrays of n_shells doubles.
 */
int fit_spectral_flux(fit_plan* plan, int invariant, const void* f_hat, const void* adv_hat,
                      double* T, double* Pi, int io_dtype, int ptr_kind);

/*
 * T and Pi of n_inv (1..4) invariants of one field: calculate_spectral_flux (SpectralFlux.jl:49-94)
 * for each invariant with the single N̂ they share (as a decomposition shares it, :96-122).  For up
 * to 2 invariants and M <= 3 the last forward pass of the nonlinear term computes the densities
 * and shell sums itself (N̂ is never stored).  T, Pi: host arrays of n_inv*n_shells doubles,
 * invariant q's shells at offset q*n_shells.
 */
int fit_spectral_fluxes(fit_plan* plan, int n_inv, const int* invariants, const void* f_hat, const void* adv_hat,
                        double* T, double* Pi, int io_dtype, int ptr_kind);

/*
 * T and Pi from a given carrier and N̂ (both (ns..., M)): the density + shell sum + cumsum tail
 * of the flux, SpectralFlux._calculate_spectral_flux_with_N̂! (SpectralFlux.jl:148-180).  Used for
 * decomposed carriers (helical components, SpectralFlux.jl:96-122) and partial fluxes (:240-296).
 */
int fit_spectral_flux_with_nhat(fit_plan* plan, int invariant, const void* c_hat, const void* N_hat,
                                double* T, double* Pi, int io_dtype, int ptr_kind);

/*
 * T and Pi of n_inv (1..4) invariants from one carrier and one N̂ in a single density pass over
 * them.  The reference evaluates _calculate_spectral_flux_with_N̂! (SpectralFlux.jl:148-180) once
 * per invariant; this is that tail for several invariants sharing N̂ (as calculate_spectral_flux
 * with a decomposition shares it, SpectralFlux.jl:96-122).  invariants: n_inv codes; T, Pi: host
 * arrays of n_inv*n_shells doubles, invariant q's shells at offset q*n_shells.
 */
int fit_spectral_fluxes_with_nhat(fit_plan* plan, int n_inv, const int* invariants, const void* c_hat,
                                  const void* N_hat, double* T, double* Pi, int io_dtype, int ptr_kind);

/*
 * Allocation-free variant of fit_spectral_flux_with_nhat.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no allocating routine: c_hat and N_hat must
 * be DEVICE arrays in the plan's compute precision (io_dtype == compute dtype, ptr_kind ==
 * FIT_PTR_DEVICE; anything else is rejected, since staging would allocate), and T and Pi are
 * DEVICE arrays of n_shells doubles.  The shell partial sums live in a static device buffer of
 * FIT_NOALLOC_SCRATCH_BYTES that the library holds on each GPU until the process ends; plans
 * with more than FIT_NOALLOC_MAX_SHELLS shells are rejected (FIT_E_UNSUPPORTED).  Every call on
 * a GPU shares that buffer: calls that overlap on one GPU (several host threads, several plans)
 * race.  Returns after the plan's stream has finished.
 */
#define FIT_NOALLOC_SCRATCH_BYTES 4194304
#define FIT_NOALLOC_MAX_SHELLS 4096
int fit_spectral_flux_with_nhat_noalloc(fit_plan* plan, int invariant, const void* c_hat, const void* N_hat,
                                        double* T, double* Pi, int io_dtype, int ptr_kind);

/*
 * Shell-to-shell matrix T(n,m) over the plan's shells (every shell is a giver).  Replaces
 * ShellToShellTransfer._calculate_shell_to_shell! (ShellToShellTransfer.jl:152-159), i.e. the
 * FFT giver loop FFTWExt._shell_to_shell_fft! (FFTWExt.jl:209-295) and the KernelAbstractions
 * GPUBackend loop (KernelAbstractionsExt.jl:68-138).  Tmat: n_sh*n_sh doubles (column-major),
 * net: n_sh doubles (Σ_m T(n,m)), *max_asym = max|T+Tᵀ| or NaN when verify == 0.
 */
int fit_shell_to_shell(fit_plan* plan, int invariant, const void* f_hat, const void* adv_hat,
                       double* Tmat, double* net, double* max_asym, int verify, int io_dtype,
                       int ptr_kind);

/*
 * Mode-resolved triad transfer S(k|p) (ns..., ns...) in io_dtype real (receiver k first) and
 * net(k) = Σ_p S(k|p) (ns...) in io_dtype real.  Replaces calculate_mode_to_mode_transfer's giver
 * loop (ScaleToScaleTransfer.jl:75-88).  S and net must be DEVICE pointers when Np^2 values do
 * not fit host memory; ptr_kind applies to all four pointers.  Every dealiasing: with
 * PADDED the triad sum runs on the padded grid exactly as the reference's padded nonlinear term
 * of each one-mode giver field (FFTWExt.jl:160-183).  The plan's transform lengths must be
 * 2^a 3^b 5^c (no transform runs, but the plan is created like any other).
 */
int fit_mode_to_mode(fit_plan* plan, int invariant, const void* f_hat, const void* adv_hat,
                     void* S, void* net, int io_dtype, int ptr_kind);

/*
 * Smooth band-to-band transfer T(n,m) between graded spectral bands (SmoothBands: log-Gaussian
 * weights w_n(κ) about centers[n] of width logwidth in ln κ, renormalised to a partition of unity,
 * 0 at κ <= 0).  Replaces BandTransfer.calculate_band_to_band_transfer (BandTransfer.jl:66-113):
 * for every band m the nonlinear term of w_m ⊙ f̂ advected by adv_hat (or f̂), and
 * T(n,m) = Σ_k w_n(k) t_m(k).  k_coord: host array of Np doubles, the band coordinate of every
 * mode (shell_coordinate, ShellBinning.jl:19-36).  Tmat: n_bands² doubles (column-major,
 * Tmat[n + n_bands*m] into band n from band m), net: n_bands doubles, *max_asym = max|T+Tᵀ|.
 */
int fit_band_to_band(fit_plan* plan, int invariant, const double* centers, int n_bands, double logwidth,
                     const double* k_coord, const void* f_hat, const void* adv_hat, double* Tmat, double* net,
                     double* max_asym, int io_dtype, int ptr_kind);

/*
 * Helical (Craya-Herring) split of a 3D velocity into u+ and u- (each (ns..., M), components past
 * the third zero).  Replaces Decomposition._decompose_field_spectral(::HelicalDecomposition)
 * (Decomposition.jl:63-98).  Same as fit_decompose(plan, FIT_DECOMP_HELICAL, ...).
 */
int fit_helical_decompose(fit_plan* plan, const void* u_hat, void* u_plus, void* u_minus,
                          int io_dtype, int ptr_kind);

/*
 * Physical -> spectral front-end: u_hat[..., c] = fft(u_phys[..., c]) / Np for ncomp REAL physical
 * components (ns..., ncomp) in io_dtype, into the full complex spectral array (ns..., ncomp) in
 * io_dtype -- the package's coefficient convention, as the reference's distributed front-end
 * forms it (ext/FlowInvariantTransferPencilFFTsExt.jl:83-88).  The plan must not be 3/2-padded;
 * 1 <= ncomp <= max(M, 3).
 */
int fit_physical_to_spectral(fit_plan* plan, const void* u_phys, int ncomp, void* u_hat, int io_dtype,
                             int ptr_kind);

/* spectral decompositions of a 3D velocity into two named components */
enum { FIT_DECOMP_HELICAL = 1, FIT_DECOMP_TOROIDAL_POLOIDAL = 2 };

/*
 * Split a 3D velocity (ns..., M >= 3) into two components (each (ns..., M), components past the
 * third zero): FIT_DECOMP_HELICAL -> (positive, negative) helical parts (Decomposition.jl:63-98);
 * FIT_DECOMP_TOROIDAL_POLOIDAL -> (toroidal, poloidal) Craya-Herring parts (Decomposition.jl:107-141,
 * e1 = k × ẑ / |k × ẑ|, e2 = k̂ × e1, (x̂, ŷ) for k ∥ ẑ).  Replaces
 * Decomposition._decompose_field_spectral for both.
 */
int fit_decompose(fit_plan* plan, int decomposition, const void* u_hat, void* out_a, void* out_b, int io_dtype,
                  int ptr_kind);

/* ---- ensembles of snapshots ---------------------------------------------------------------- */

typedef struct fit_ensemble fit_ensemble;

/*
 * Ensemble of n_snap snapshots, possibly of different sizes, for the kinetic-energy flux and
 * (with_scalar != 0) the passive-scalar flux of every snapshot.  Replaces the reference's batch
 * axis mpi_batch_map(f, snapshots) (ext/FlowInvariantTransferMPIExt.jl:39-68) with f =
 * calculate_spectral_flux (SpectralFlux.jl:49-94) + calculate_scalar_flux (SpectralFlux.jl:203-215)
 * per snapshot, results collated in input order.  Snapshots on the same grid (sizes, wavenumbers
 * and shell map equal) are batched into the same launches; the whole ensemble is captured into
 * one HIP graph on the first call and replayed afterwards.
 *   nd in {2, 3}; ns: n_snap*nd sizes (snapshot j: ns[j*nd .. j*nd+nd-1]); velocities have D = nd
 *   components; ks: n_snap*nd arrays (snapshot j, axis d at ks[j*nd + d]); shell_idx[j]: Np_j
 *   Int64 shell numbers (as fit_plan_create); n_shells[j] >= 1; any dealiasing.  With PADDED
 *   each snapshot runs the plan's padded flux pipeline in turn (the batched launches serve
 *   NONE and TWOTHIRDS), and the graph is captured again whenever the snapshot pointers change.
 */
int fit_ensemble_create(fit_ensemble** ens, int n_snap, int nd, const int64_t* ns, int compute_dtype,
                        int dealias, const double* const* ks, const int64_t* const* shell_idx,
                        const int* n_shells, int with_scalar, int device);

int fit_ensemble_destroy(fit_ensemble* ens);

/* Number of same-grid groups the snapshots were sorted into. */
int fit_ensemble_groups(const fit_ensemble* ens);

/* Ensemble options (no reference counterpart):
 *   FIT_ENS_OPT_GRAPH = 1: 1 (default) capture the ensemble into a HIP graph, 0 run eagerly.
 *   FIT_ENS_OPT_BATCH = 2: cap on snapshots per batched launch (0 = device memory budget).
 *   FIT_ENS_OPT_PLAN + o: plan option o (FIT_OPT_*) on every group's plan. */
enum { FIT_ENS_OPT_GRAPH = 1, FIT_ENS_OPT_BATCH = 2, FIT_ENS_OPT_PLAN = 1000 };
int fit_ensemble_set_option(fit_ensemble* ens, int option, int value);

/*
 * Fluxes of every snapshot.  u_hat[j]: (ns_j..., nd) velocity; theta_hat[j]: (ns_j...) scalar
 * (ignored, may be NULL, without a scalar).  Outputs are host arrays of Σ_j n_shells[j] doubles,
 * snapshot j's shells at offset Σ_{i<j} n_shells[i]: T_ke / Pi_ke (kinetic energy, Π = cumsum T),
 * T_sc / Pi_sc (scalar variance); any output may be NULL.  ptr_kind applies to every snapshot.
 */
int fit_ensemble_flux(fit_ensemble* ens, const void* const* u_hat, const void* const* theta_hat, double* T_ke,
                      double* Pi_ke, double* T_sc, double* Pi_sc, int io_dtype, int ptr_kind);

/* Per-kernel-class timing of the ensemble (as fit_profile_*, summed over its groups); while
 * enabled, calls run eagerly instead of replaying the graph. */
int fit_ensemble_profile_enable(fit_ensemble* ens, int on);
int fit_ensemble_profile_query(const fit_ensemble* ens, int kernel_class, int64_t* launches, double* total_ms,
                               double* alg_bytes);

/* The HIP stream the ensemble launches on (as void*: hipStream_t). */
void* fit_ensemble_stream(const fit_ensemble* ens);

/* ---- instrumentation (no reference counterpart: measurement hooks for the benchmark) ---- */

/* Enable (1) / disable (0) per-kernel HIP-event timing on the plan's stream; resets counters. */
int fit_profile_enable(fit_plan* plan, int on);

/* Kernel classes timed by the profiler. */
enum {
  FIT_K_FFT_COLUMN = 0,      /* strided-axis FFT pass */
  FIT_K_FFT_ROW_INVERSE = 1, /* contiguous-axis c2r pass (+ fused product) */
  FIT_K_FFT_ROW_FORWARD = 2, /* contiguous-axis r2c pass */
  FIT_K_DENSITY_SHELL = 3,   /* density + shell partial sums */
  FIT_K_REDUCE = 4,          /* fixed-order FP64 partial reduction */
  FIT_K_OTHER = 5,
  FIT_K_FUSED_X = 6,         /* fused contiguous-axis pass: c2r of gradients, product, r2c */
  FIT_K_COUNT = 7
};

/* Launch count, summed kernel time (ms) and summed algorithmic HBM bytes for one class. */
int fit_profile_query(const fit_plan* plan, int kernel_class, int64_t* launches, double* total_ms,
                      double* alg_bytes);

/* Per-kernel totals of the profiled launches (labels such as "k_grad2_r<f32,768>": kernel,
 * compute precision, transform length, loader / storer): the number of distinct labels, and for
 * label `index` its name (NUL-terminated, truncated to name_len), kernel class, launch count,
 * summed time (ms) and summed algorithmic HBM bytes. */
int fit_profile_kernel_count(const fit_plan* plan);
int fit_profile_kernel(const fit_plan* plan, int index, char* name, int name_len, int* kernel_class,
                       int64_t* launches, double* total_ms, double* alg_bytes);
int fit_ensemble_profile_kernel_count(const fit_ensemble* ens);
int fit_ensemble_profile_kernel(const fit_ensemble* ens, int index, char* name, int name_len, int* kernel_class,
                                int64_t* launches, double* total_ms, double* alg_bytes);

/* The HIP stream the plan launches on (as void*: hipStream_t). */
void* fit_plan_stream(const fit_plan* plan);

/* Stream ordering for device arrays (no reference counterpart: the reference's CPU and
 * KernelAbstractions paths run on the caller's own task or queue).  A plan launches on its own
 * stream; before a call that reads device arrays produced on another stream (the caller's
 * ROCArray / torch stream), call this with that stream (a hipStream_t, NULL = the null stream):
 * everything already queued on it finishes before the plan's next launch.  Calls return after
 * the plan's stream has finished, so results may be used on any stream afterwards. */
int fit_plan_wait_stream(fit_plan* plan, void* stream);
int fit_ensemble_wait_stream(fit_ensemble* ens, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FIT_HIP_H */
