// fit_internal.h -- plan / ensemble state shared by the translation units of libfit_hip.so.
//
//   fit_hip.hip  : plan construction, staging, profiling, the C ABI (host code, small kernels)
//   fit_f32.hip  : the pipeline templates of fit_pipeline.h instantiated for FP32
//   fit_f64.hip  : the same for FP64
// The two pipeline units hold nearly all device code and compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fit_hip.h"
#include "fit_kernels.h"

namespace fitx {
using namespace fit;

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);

#define FIT_HIP(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      return ::fitx::fail(e_ == hipErrorOutOfMemory ? FIT_E_OUT_OF_MEMORY : FIT_E_DEVICE,      \
                          std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    }                                                                                          \
  } while (0)

bool make_desc(int L, FftDesc& d, int order);
int padded_len(int n);

// one class of givers [m0, m1] of the band-limited sub-grid giver loop
struct SubGrid {
  fit_plan* c = nullptr;         // child plan on the class's grid (no dealiasing, same shells)
  int q[3] = {0, 0, 0};          // folded box of input modes copied into the child
  double tscale = 1.0;           // (Np_sub / Np)^2: N̂ carries 1/Np^2 of the grid it is taken on
  int m0 = 1, m1 = 0;
};

struct ProfRec {
  hipEvent_t a = nullptr, b = nullptr;
  int cls = 0;
  double bytes = 0;
  std::string name;  // kernel label, e.g. "k_grad2_r<f32,768>"
};

// per-kernel-label totals of the profiled launches
struct KernAgg {
  std::string name;
  int cls = 0;
  int64_t launches = 0;
  double ms = 0, bytes = 0;
};

}  // namespace fitx

struct fit_plan {
  int nd = 0, M = 0, D = 0, dtype = 0, dealias = 0, device = 0;
  int n[3] = {1, 1, 1}, G[3] = {1, 1, 1};
  int h1 = 1;
  int64_t Np = 1, Gp = 1, Hp = 1;
  int nsh = 0;
  bool shell_symmetric = true;
  size_t rsz = 4;  // sizeof(R)
  hipStream_t stream = nullptr;
  fit::FftDesc fd[3];
  void* tw[3] = {nullptr, nullptr, nullptr};
  void* ks[3] = {nullptr, nullptr, nullptr};
  int* shell = nullptr;
  void* work = nullptr;    // one half spectrum on the transform grid
  void* uphys = nullptr;   // nd physical advecting components
  void* nphys = nullptr;   // one physical product
  void* nhat = nullptr;    // M full original-grid N̂
  void* nhat_dst = nullptr;  // caller's device N̂ array when written in place of nhat
  bool nhat_half_req = false;   // flux: the forward pass may store N̂ on the original half grid
  bool nhat_half_done = false;  // ... and did (nhat then holds M half-grid components)
  void* in_f = nullptr;    // staged primary field (M comps)
  void* in_a = nullptr;    // staged advecting field (D comps)
  void* in_c = nullptr;    // staged carrier / extra input (M comps)
  void* out_tmp = nullptr; // staged output
  size_t out_tmp_bytes = 0;
  double* partials = nullptr;
  int nblk_density = 0;
  double* dcol = nullptr;  // nsh doubles
  int64_t bytes = 0;
  // fused / pruned pipeline state
  int plain = 0;                 // 1: unfused, unpruned, one giver per launch (reference dataflow)
  int band_in[3] = {0, 0, 0};    // folded support of spectral inputs (keep / embedded block)
  int band_rx[3] = {0, 0, 0};    // folded support of the receiving shells (∩ output mask)
  std::vector<int4> shell_band;  // per shell (index s-1): folded support ∩ band_in (-1: empty)
  int4* d_shell_band = nullptr;
  int* shh = nullptr;            // shell map on the original half grid (x <= n1/2)
  void* eh = nullptr;            // Hermitian part of the masked carrier, M comps, half grid
  void* work2 = nullptr; size_t work2_bytes = 0;
  void* nxbuf = nullptr; size_t nxbuf_bytes = 0;
  double* partials2 = nullptr; size_t partials2_bytes = 0;
  double* dcolG = nullptr; size_t dcolG_bytes = 0;
  size_t eh_bytes = 0;
  int giver_batch = 0;           // cap on givers per batch (0 = memory budget only)
  int xfused_gc = 0;             // givers per workgroup in the wave-per-line x pass (0 = 4)
  int xfused_block = 0;          // 1: use the workgroup-barrier x pass instead
  int col_tx = 0;                // FIT_OPT_COL_TILE (0: automatic)
  int xself = 1;                 // FIT_OPT_XPASS_SELF: self-advection x pass forms u in registers
  int xstage = 1;                // FIT_OPT_XPASS_STAGE: FP32 self pass stages the next line by LDS DMA
  int grad_split = -1;           // FIT_OPT_GRAD_SPLIT: slot 2 of the y gradient pass as a column pass
                                 // (-1: for y lines shorter than 512 points)
  // shared-transform gradients: wavenumbers odd under k -> -k on every populated bin
  bool commute = false;
  void* kt[2] = {nullptr, nullptr};  // transform-grid wavenumbers of x and y (compute precision)
  bool grads_commuted = false;       // the last gradient passes produced slot 0 = h
  // band-limited sub-grid for the giver loop: when every shell lies in |k_d| <= R_d with
  // 4 R_d < n_d, T(n,m) only involves advecting modes |q_d| <= 2 R_d and the exact convolution
  // is evaluated alias-free on a grid of at least 4 R_d + 1 points per axis
  // (givers of smaller support need fewer advecting modes, so the givers are split into classes
  // of consecutive shells, each on the smallest grid alias-free for it)
  std::vector<fitx::SubGrid> subs;  // giver classes, each with its child plan
  fit_plan* sub = nullptr;       // the child with the largest grid (reported by fit_plan_subgrid)
  int giver_lo = 1, giver_hi = 0;  // giver shells this plan's loop runs (giver_hi = 0: all)
  void* sub_f = nullptr; size_t sub_f_bytes = 0;
  void* sub_a = nullptr; size_t sub_a_bytes = 0;
  bool own_stream = true;
  bool user_plan = false;     // made by fit_plan_create (counted by fit_plan_live_count)
  int use_sub = 1;                // FIT_OPT_SUBGRID
  // fused last pass + density of the flux (flux_multi): requested invariants and whether the
  // pipeline took the fused form (densities in dcolG) or stored N̂
  int fused_flux = 0;            // FIT_OPT_FUSED_FLUX: 1 lets the flux compute its densities in the last pass
  int fused_ninv = 0;
  int fused_inv[2] = {0, 0};
  const void* fused_car = nullptr;
  bool fused_done = false;
  // ordering after a caller's stream (fit_plan_wait_stream)
  hipEvent_t order_ev = nullptr;
  int radix_order = 0;            // FIT_OPT_RADIX_ORDER of the LDS-engine passes
  int col_pipe = 0;               // FIT_OPT_COL_PIPE: tiles per workgroup of the pipelined column passes (0: off)
  int last_bins_kb = 64;          // FIT_OPT_LAST_BINS: LDS budget of private shell bins in the fused last pass
  int col_f64 = 6;                // FIT_OPT_COL_F64: FP32 plans run the strided register passes in FP64
                                  // (bit 0: inverse passes, bit 1: forward y, bit 2: forward z)
  void* tw64[3] = {nullptr, nullptr, nullptr};  // FP64 twiddles of an FP32 plan (FIT_OPT_COL_F64)
  std::string tune_report;        // the last fit_plan_autotune: workload, times, choices
  // profiling
  bool prof = false;
  std::vector<fitx::ProfRec> recs;
  size_t nrec = 0;
  int64_t launches[FIT_K_COUNT] = {0};
  double ms[FIT_K_COUNT] = {0};
  double abytes[FIT_K_COUNT] = {0};
  std::vector<fitx::KernAgg> kagg;
};

namespace fitx {

struct EnsGroup {
  fit_plan* p = nullptr;          // the group's grid, shells and workspaces (M = D = nd)
  std::vector<int> snaps;         // snapshot indices, input order
  void** d_tab = nullptr;         // 2B device pointers: velocities, then scalars
  std::vector<void*> h_tab;       // host copy of the table (outlives the async upload)
  void* ub = nullptr; size_t ub_bytes = 0;    // physical velocities of one batch (ng, nd, Gp)
  void* ehk = nullptr; size_t ehk_bytes = 0;  // kinetic-energy carrier, Hermitian part (ng, nd, Hn)
  void* ehs = nullptr; size_t ehs_bytes = 0;  // scalar carrier (ng, Hn)
  double* rows = nullptr;         // (2, B, nsh): kinetic-energy rows, then scalar rows
  std::vector<double> h_rows;
  void* stage = nullptr; size_t stage_bytes = 0;  // inputs staged from the host / other dtype
  int ng = 0;                     // snapshots per batch (device memory budget)
};

}  // namespace fitx

struct fit_ensemble {
  int n = 0, nd = 0, dtype = 0, dealias = 0, device = 0, with_scalar = 0;
  hipStream_t stream = nullptr;
  std::vector<fitx::EnsGroup> groups;
  std::vector<int> nsh;           // per snapshot
  std::vector<int64_t> off;       // per snapshot: first row of its shells in the outputs
  int use_graph = 1;
  int batch_cap = 0;
  bool prof = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_io = -1;              // (ptr_kind, io_dtype) the graph was captured for
  std::vector<void*> graph_ptrs;  // padded groups' input pointers the graph's launches hold
  hipEvent_t order_ev = nullptr;  // ordering after a caller's stream (fit_ensemble_wait_stream)
};

namespace fitx {

// ---- host helpers (fit_hip.hip) -------------------------------------------------------------
int dalloc(fit_plan* p, void** ptr, size_t bytes);
void dfree(void*& ptr);
// kname<prec,L extra>: the label a launch is aggregated under (fit_profile_kernel)
void prof_begin(fit_plan* p, int cls, double bytes, const char* kname, int L = 0, const char* extra = "");
void prof_end(fit_plan* p);
void prof_collect(fit_plan* p);
int grow(fit_plan* p, void** buf, size_t* have, size_t need);
// Fixed-order reduction of ng groups of nblk partial rows of nsh bins (k_reduce_rows).
void launch_reduce(fit_plan* p, const double* partials, int nblk, int nsh, int ng, double* out);
int stage_in(fit_plan* p, const void* src, int ncomp, int io_dtype, int ptr_kind, void** staging,
             const void** out);
int stage_out_real(fit_plan* p, const void* dsrc, void* dst, int64_t nvals, int io_dtype, int ptr_kind);
int check_plan(fit_plan* p);
int check_invariant(fit_plan* p, int inv, int ncomp);
int stage_fa(fit_plan* p, const void* f_hat, const void* adv_hat, int io_dtype, int ptr_kind, const void** f,
             const void** a);
int finish(fit_plan* p);
int spectral_flux_tail(fit_plan* p, double* T, double* Pi);
int rows_tail(fit_plan* p, int n_inv, double* T, double* Pi);

extern size_t g_max_lds;  // per-workgroup LDS the launches may use
constexpr int kThreads = 256;
// LDS the deterministic shell bins may take as private per-wave copies (above it the waves commit
// in turn): streaming density passes, and the FFT + density passes (which keep their exchange LDS)
constexpr size_t kDensBinBudget = 64 * 1024, kFftBinBudget = 16 * 1024;

// ---- small shared helpers ---------------------------------------------------------------------
// twiddle table of axis d in the compute precision C of a pass over fields stored in R
template <typename R, typename C> inline const cpx<C>* twc(const fit_plan* p, int d) {
  return (const cpx<C>*)(sizeof(C) != sizeof(R) ? p->tw64[d] : p->tw[d]);
}

// largest power-of-two tile with 2*L*T*sizeof(cpx) <= 64 KiB (ping-pong LDS), capped at cap
inline int tile_for(int L, size_t csz, int cap) {
  int t = cap;
  while (t > 1 && (size_t)2 * L * t * csz > 65536) t /= 2;
  return t;
}

inline Band full_band(int G) { return Band{G, G}; }

// lengths with a compile-time register FFT (others use the generic LDS engine)
#define FIT_REG_LENGTHS(X) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) X(3072)
constexpr int reg_T(int L) { return L / (((L & (L - 1)) == 0) ? 8 : 24); }
constexpr int two_T(int L) { return L / gE(L); }   // threads per line of the two-line kernels
constexpr int xblk_T(int L) { return L / xE(L); }  // threads per line of the workgroup x pass
inline bool reg_supported(int L) {
  switch (L) {
#define FIT_CASE(LL) case LL:
    FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
    return true;
    default: return false;
  }
}

// x-extent of compact storage for a folded band on the half grid
inline int xext(const fit_plan* p, int a) { return std::min(a, p->G[0] / 2) + 1; }
// row pitch of compact storage: the extent rounded up to 16 elements, so that every row starts
// on a 128-byte line and column tiles never straddle two lines
inline int xpitch(int xe) { return (xe + 15) / 16 * 16; }

// shared-transform gradient passes apply: exact for this plan, register lengths on every axis
inline bool commute_ok(const fit_plan* p) {
  if (!p->commute || p->plain || p->nd < 2) return false;
  for (int d = 0; d < p->nd; ++d)
    if (!reg_supported(p->G[d])) return false;
  return true;
}

// Columns per workgroup of the two-line gradient passes (512 threads at most): 16 within 64 KB
// of LDS, or the plan's column tile within 160 KB.
inline int grad_tile(const fit_plan* p, int L, size_t csz) {
  const int T = two_T(L);
  int TX = std::min(p->col_tx > 0 ? p->col_tx : 16, 512 / T);
  const size_t cap = p->col_tx > 0 ? (size_t)160 * 1024 : g_max_lds;
  while (TX > 1 && (size_t)L * TX * csz > cap) TX /= 2;
  return TX;
}

template <typename R> Grid grid_of(const fit_plan* p) {
  Grid g;
  g.nd = p->nd;
  g.n1 = p->n[0]; g.n2 = p->n[1]; g.n3 = p->n[2];
  g.G1 = p->G[0]; g.G2 = p->G[1]; g.G3 = p->G[2];
  g.h1 = p->h1;
  g.Np = p->Np;
  return g;
}

// loader / storer tags in kernel labels
template <class T> struct KTag { static constexpr const char* v = ""; };
template <typename R> struct KTag<LdSpectral<R>> { static constexpr const char* v = ",LdSpectral"; };
template <typename R> struct KTag<LdSpectralB<R>> { static constexpr const char* v = ",LdSpectralB"; };
template <typename R> struct KTag<LdHalf<R>> { static constexpr const char* v = ",LdHalf"; };
template <typename R> struct KTag<LdReal<R>> { static constexpr const char* v = ",LdReal"; };
template <typename R> struct KTag<StHalf<R>> { static constexpr const char* v = ",StHalf"; };
template <typename R> struct KTag<StReal<R>> { static constexpr const char* v = ",StReal"; };
template <typename R> struct KTag<StProduct<R>> { static constexpr const char* v = ",StProduct"; };
template <typename R> struct KTag<StNhat<R>> { static constexpr const char* v = ",StNhat"; };
template <typename R> struct KTag<StNhatHalf<R>> { static constexpr const char* v = ",StNhatHalf"; };

inline size_t csize(int dtype) { return dtype == FIT_F64 ? 16 : 8; }

// restores a plan's component count when a batch runs the scalar part with M = 1
struct MGuard {
  fit_plan* p; int m;
  MGuard(fit_plan* q, int mm) : p(q), m(q->M) { q->M = mm; }
  ~MGuard() { p->M = m; }
};

// ---- pipeline entry points (fit_pipeline.h; instantiated in fit_f32.hip and fit_f64.hip) ----
template <typename R> int nonlinear_term_impl(fit_plan* p, const void* f, const void* a);
template <typename R> int density_launch(fit_plan* p, int inv, const void* car, const void* nh, void* t_out,
                                         bool shells);
template <typename R> int density_multi(fit_plan* p, int ninv, const int* invs, const void* car, const void* nh);
template <typename R> int flux_multi(fit_plan* p, int ninv, const int* invs, const void* f, const void* a);
template <typename R> int phys_to_spec(fit_plan* p, const void* u, int phys_dtype, int ncomp, void* out);
template <typename R> int density_shells_into(fit_plan* p, int inv, const void* car, const void* nh, double* partials,
                                              int nblk, double* out);
template <typename R> int s2s_entry(fit_plan* p, int inv, const void* f, const void* a, std::vector<double>& Tm);
template <typename R> int b2b_impl(fit_plan* p, int inv, const void* f, const void* a, const R* kc,
                                   const double* cen, int nb, double inv2s2, std::vector<double>& Tm);
template <typename R> int m2m_impl(fit_plan* p, int inv, const void* f, const void* a, void* S, void* net,
                                   int io_dtype, int ptr_kind);
template <typename R> int ens_run(fit_ensemble* e);

#define FIT_PIPELINE_EXTERN(R)                                                                                  \
  extern template int nonlinear_term_impl<R>(fit_plan*, const void*, const void*);                              \
  extern template int density_launch<R>(fit_plan*, int, const void*, const void*, void*, bool);                 \
  extern template int density_multi<R>(fit_plan*, int, const int*, const void*, const void*);                   \
  extern template int flux_multi<R>(fit_plan*, int, const int*, const void*, const void*);                      \
  extern template int phys_to_spec<R>(fit_plan*, const void*, int, int, void*);                                \
  extern template int density_shells_into<R>(fit_plan*, int, const void*, const void*, double*, int, double*);   \
  extern template int s2s_entry<R>(fit_plan*, int, const void*, const void*, std::vector<double>&);            \
  extern template int b2b_impl<R>(fit_plan*, int, const void*, const void*, const R*, const double*, int, double, \
                                  std::vector<double>&);                                                        \
  extern template int m2m_impl<R>(fit_plan*, int, const void*, const void*, void*, void*, int, int);            \
  extern template int ens_run<R>(fit_ensemble*);

}  // namespace fitx
