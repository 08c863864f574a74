// fit_hip.hip -- plan management, pipeline orchestration and the C ABI of libfit_hip.so.
//
// One plan = one grid on one GPU: device-resident twiddles, wavenumbers, shell map and
// workspaces, one HIP stream.  Every call runs the whole hot path on the device and returns
// FP64 shell sums (or device/host arrays) to the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fit_hip.h"
#include "fit_kernels.h"

using namespace fit;

namespace {

thread_local std::string g_last_error;
std::mutex g_create_mutex;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define FIT_HIP(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      return fail(e_ == hipErrorOutOfMemory ? FIT_E_OUT_OF_MEMORY : FIT_E_DEVICE,              \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                          \
    }                                                                                          \
  } while (0)

// radix plan: as many 8s as possible, then 4 / 2, then 3s and 5s (any order is valid Stockham)
bool make_desc(int L, FftDesc& d) {
  d.L = L;
  d.nst = 0;
  int r = L;
  while (r % 8 == 0 && d.nst < FIT_MAX_STAGES) { d.rad[d.nst++] = 8; r /= 8; }
  while (r % 4 == 0 && d.nst < FIT_MAX_STAGES) { d.rad[d.nst++] = 4; r /= 4; }
  while (r % 2 == 0 && d.nst < FIT_MAX_STAGES) { d.rad[d.nst++] = 2; r /= 2; }
  while (r % 3 == 0 && d.nst < FIT_MAX_STAGES) { d.rad[d.nst++] = 3; r /= 3; }
  while (r % 5 == 0 && d.nst < FIT_MAX_STAGES) { d.rad[d.nst++] = 5; r /= 5; }
  return r == 1;
}

int padded_len(int n) {  // FFTWExt.jl:128-131
  int m = (3 * n + 1) / 2;
  return ((m - n) % 2 == 0) ? m : m + 1;
}

struct ProfRec {
  hipEvent_t a = nullptr, b = nullptr;
  int cls = 0;
  double bytes = 0;
};

}  // namespace

struct fit_plan {
  int nd = 0, M = 0, D = 0, dtype = 0, dealias = 0, device = 0;
  int n[3] = {1, 1, 1}, G[3] = {1, 1, 1};
  int h1 = 1;
  int64_t Np = 1, Gp = 1, Hp = 1;
  int nsh = 0;
  bool shell_symmetric = true;
  size_t rsz = 4;  // sizeof(R)
  hipStream_t stream = nullptr;
  FftDesc fd[3];
  void* tw[3] = {nullptr, nullptr, nullptr};
  void* ks[3] = {nullptr, nullptr, nullptr};
  int* shell = nullptr;
  void* work = nullptr;    // one half spectrum on the transform grid
  void* uphys = nullptr;   // nd physical advecting components
  void* nphys = nullptr;   // one physical product
  void* nhat = nullptr;    // M full original-grid N̂
  void* nhat_dst = nullptr;  // caller's device N̂ array when written in place of nhat
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
  fit_plan* sub = nullptr;       // child plan on that grid (no dealiasing, same shells)
  int sub_q[3] = {0, 0, 0};      // folded box of input modes copied into the child
  double sub_tscale = 1.0;       // (Np_sub / Np)^2: N̂ carries 1/Np^2 of the grid it is taken on
  void* sub_f = nullptr; size_t sub_f_bytes = 0;
  void* sub_a = nullptr; size_t sub_a_bytes = 0;
  bool own_stream = true;
  int use_sub = 1;                // FIT_OPT_SUBGRID
  // profiling
  bool prof = false;
  std::vector<ProfRec> recs;
  size_t nrec = 0;
  int64_t launches[FIT_K_COUNT] = {0};
  double ms[FIT_K_COUNT] = {0};
  double abytes[FIT_K_COUNT] = {0};
};

namespace {

int dalloc(fit_plan* p, void** ptr, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(e == hipErrorOutOfMemory ? FIT_E_OUT_OF_MEMORY : FIT_E_DEVICE,
                "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
  }
  p->bytes += (int64_t)bytes;
  return FIT_OK;
}

void dfree(void*& ptr) {
  if (ptr) hipFree(ptr);
  ptr = nullptr;
}

// profiling brackets
void prof_begin(fit_plan* p, int cls, double bytes) {
  if (!p->prof) return;
  if (p->nrec >= p->recs.size()) {
    ProfRec r;
    hipEventCreate(&r.a);
    hipEventCreate(&r.b);
    p->recs.push_back(r);
  }
  ProfRec& r = p->recs[p->nrec];
  r.cls = cls;
  r.bytes = bytes;
  hipEventRecord(r.a, p->stream);
}

void prof_end(fit_plan* p) {
  if (!p->prof) return;
  hipEventRecord(p->recs[p->nrec].b, p->stream);
  p->nrec++;
}

// launches recorded by the plan and by its sub-grid child are reported together
void prof_collect(fit_plan* p) {
  if (!p->prof) return;
  for (fit_plan* s : {p, p->sub}) {
    if (!s || !s->prof) continue;
    for (size_t i = 0; i < s->nrec; ++i) {
      float t = 0.f;
      hipEventElapsedTime(&t, s->recs[i].a, s->recs[i].b);
      p->launches[s->recs[i].cls] += 1;
      p->ms[s->recs[i].cls] += t;
      p->abytes[s->recs[i].cls] += s->recs[i].bytes;
    }
    s->nrec = 0;
  }
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

// largest power-of-two tile with 2*L*T*sizeof(cpx) <= 64 KiB (ping-pong LDS), capped at cap
int tile_for(int L, size_t csz, int cap) {
  int t = cap;
  while (t > 1 && (size_t)2 * L * t * csz > 65536) t /= 2;
  return t;
}

constexpr int kThreads = 256;

// ------------------------------ pass launchers ------------------------------------------------

Band full_band(int G) { return Band{G, G}; }

// lengths with a compile-time register FFT (others use the generic LDS engine)
#define FIT_REG_LENGTHS(X) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) X(3072)
constexpr int reg_T(int L) { return L / (((L & (L - 1)) == 0) ? 8 : 24); }
constexpr int two_T(int L) { return L / gE(L); }  // threads per line of the two-line kernels
constexpr int xblk_T(int L) { return L / xE(L); }  // threads per line of the workgroup x pass
bool reg_supported(int L) {
  switch (L) {
#define FIT_CASE(LL) case LL:
    FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
    return true;
    default: return false;
  }
}
size_t g_max_lds = 65536;  // per-workgroup LDS the launches may use

// Fixed-order reduction of ng groups of nblk partial rows of nsh bins (k_reduce_rows).
void launch_reduce(fit_plan* p, const double* partials, int nblk, int nsh, int ng, double* out) {
  if (nsh <= 0 || ng <= 0) return;
  hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)((nsh + kRedBins - 1) / kRedBins), (unsigned)ng),
                     dim3(kRedBins * kRedRows), 0, p->stream, partials, nblk, nsh, out);
}

// Column pass over fields [0, nfields): x < nx, other-index band ob, input band ib, output band wb.
template <typename R, int SIGN, class LD, class ST>
int launch_col(fit_plan* p, LD ld, ST st, int axis, int nfields, int nx, Band ob, Band ib, Band wb,
               const int4* fband = nullptr, int fdiv = 1) {
  const int L = p->G[axis];
  const size_t csz = sizeof(cpx<R>);
  const int nob = ob.count();
  if (nx <= 0 || nob <= 0 || nfields <= 0) return FIT_OK;
  const double rows_in = ib.count(), rows_out = wb.count();
  const int gother = axis == 2 ? p->G[1] : p->G[2];
  prof_begin(p, FIT_K_FFT_COLUMN, (double)nx * nob * (rows_in + rows_out) * csz * nfields);
  if (!p->plain && reg_supported(L)) {
    const int T = reg_T(L);
    // columns per workgroup: 16 (128-byte rows in FP32) within 64 KB of LDS, or the plan's cap
    // within the 160 KB a workgroup may hold
    int TX = std::min(p->col_tx > 0 ? p->col_tx : 16, col_threads<R>() / T);
    const size_t lds_cap = p->col_tx > 0 ? (size_t)160 * 1024 : g_max_lds;
    while (TX > 1 && (size_t)L * TX * csz > lds_cap) TX /= 2;
    dim3 grid((nx + TX - 1) / TX, nob, nfields);
    const size_t lds = (size_t)L * TX * csz;
    switch (L) {
#define FIT_CASE(LL)                                                                                       \
  case LL:                                                                                                 \
    hipLaunchKernelGGL((k_col_r<R, LL, SIGN, LD, ST>), grid, dim3(reg_T(LL) * TX), lds, p->stream, ld, st, \
                       (const cpx<R>*)p->tw[axis], nx, TX, axis, ob, ib, wb, fband, fdiv, gother);           \
    break;
      FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
      default: break;
    }
  } else {
    const int TX = tile_for(L, csz, 16);
    dim3 grid((nx + TX - 1) / TX, nob, nfields);
    const size_t lds = (size_t)2 * L * TX * csz;
    hipLaunchKernelGGL((k_col<R, SIGN, LD, ST>), grid, dim3(kThreads), lds, p->stream, ld, st, p->fd[axis],
                       (const cpx<R>*)p->tw[axis], nx, TX, axis, ob, ib, wb, fband, fdiv, gother);
  }
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

template <typename R, int SIGN, class LD, class ST>
int launch_col_full(fit_plan* p, LD ld, ST st, int axis, int nfields) {
  const int other = axis == 1 ? p->G[2] : p->G[1];
  return launch_col<R, SIGN>(p, ld, st, axis, nfields, p->h1, full_band(other), full_band(p->G[axis]),
                             full_band(p->G[axis]));
}

template <typename R, class LD, class ST>
int launch_row_inv(fit_plan* p, LD ld, ST st, int nfields, double extra_bytes_per_row, int kxa) {
  const int L = p->G[0];
  const size_t csz = sizeof(cpx<R>);
  const int P = tile_for(L, csz, 8);
  const int nrows = p->G[1] * p->G[2];
  const int npairs = (nrows + 1) / 2;
  dim3 grid((npairs + P - 1) / P, 1, nfields);
  const size_t lds = (size_t)2 * L * P * csz;
  const int nin = std::min(kxa + 1, p->h1);
  prof_begin(p, FIT_K_FFT_ROW_INVERSE,
             ((double)nin * csz + (double)L * sizeof(R) + extra_bytes_per_row) * nrows * nfields);
  if (!p->plain && reg_supported(L)) {
    const int T = reg_T(L);
    const size_t LL = (size_t)L + L / 16;
    int B = std::max(1, kThreads / T);
    while (B > 1 && B * LL * csz > g_max_lds) --B;
    dim3 rgrid((npairs + B - 1) / B, 1, nfields);
    switch (L) {
#define FIT_CASE(LLEN)                                                                                      \
  case LLEN:                                                                                                \
    hipLaunchKernelGGL((k_row_inv_r<R, LLEN, LD, ST>), rgrid, dim3(B * reg_T(LLEN)), B * LL * csz, p->stream, \
                       ld, st, (const cpx<R>*)p->tw[0], p->G[1], nrows, B, kxa);                             \
    break;
      FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
      default: break;
    }
  } else {
    hipLaunchKernelGGL((k_row_inv<R, LD, ST>), grid, dim3(kThreads), lds, p->stream, ld, st, p->fd[0],
                       (const cpx<R>*)p->tw[0], p->G[1], nrows, P, kxa);
  }
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

template <typename R, class LD, class ST>
int launch_row_fwd(fit_plan* p, LD ld, ST st, int nfields, int nout) {
  const int L = p->G[0];
  const size_t csz = sizeof(cpx<R>);
  const int P = tile_for(L, csz, 8);
  const int nrows = p->G[1] * p->G[2];
  const int npairs = (nrows + 1) / 2;
  dim3 grid((npairs + P - 1) / P, 1, nfields);
  const size_t lds = (size_t)2 * L * P * csz;
  prof_begin(p, FIT_K_FFT_ROW_FORWARD, ((double)nout * csz + (double)L * sizeof(R)) * nrows * nfields);
  hipLaunchKernelGGL((k_row_fwd<R, LD, ST>), grid, dim3(kThreads), lds, p->stream, ld, st, p->fd[0],
                     (const cpx<R>*)p->tw[0], p->G[1], nrows, P, nout);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

template <typename R> LdHalf<R> ld_work(fit_plan* p) {
  return LdHalf<R>{(const cpx<R>*)p->work, (int64_t)p->h1, (int64_t)p->h1 * p->G[1], p->Hp};
}
template <typename R> StHalf<R> st_work(fit_plan* p) {
  return StHalf<R>{(cpx<R>*)p->work, (int64_t)p->h1, (int64_t)p->h1 * p->G[1], p->Hp};
}

// Inverse transform of one spectral-source field into a physical row storer RS.  `bi` is the
// folded support of the source (pruning); the plain pipeline passes full bands.
template <typename R, class RS>
int inverse_field(fit_plan* p, const LdSpectral<R>& ls, RS rs, double extra_row_bytes, const int* bi) {
  int rc;
  const int xa = std::min(bi[0], p->G[0] / 2) + 1;
  if (p->nd == 3) {
    if ((rc = launch_col<R, 1>(p, ls, st_work<R>(p), 2, 1, xa, Band{bi[1], p->G[1]}, Band{bi[2], p->G[2]},
                               full_band(p->G[2]))))
      return rc;
    if ((rc = launch_col<R, 1>(p, ld_work<R>(p), st_work<R>(p), 1, 1, xa, full_band(p->G[2]),
                               Band{bi[1], p->G[1]}, full_band(p->G[1]))))
      return rc;
    return launch_row_inv<R>(p, ld_work<R>(p), rs, 1, extra_row_bytes, xa - 1);
  }
  if (p->nd == 2) {
    if ((rc = launch_col<R, 1>(p, ls, st_work<R>(p), 1, 1, xa, full_band(1), Band{bi[1], p->G[1]},
                               full_band(p->G[1]))))
      return rc;
    return launch_row_inv<R>(p, ld_work<R>(p), rs, 1, extra_row_bytes, xa - 1);
  }
  return launch_row_inv<R>(p, ls, rs, 1, extra_row_bytes, xa - 1);
}

// Forward transform of the physical product (nphys) into full original-grid N̂ component c.
template <typename R>
int forward_field(fit_plan* p, int c, double scale) {
  Grid g = grid_of<R>(p);
  LdReal<R> lr{(const R*)p->nphys, (int64_t)p->G[0], (int64_t)p->G[0] * p->G[1], 0};
  StNhat<R> sn{g, (cpx<R>*)(p->nhat_dst ? p->nhat_dst : p->nhat) + (int64_t)c * p->Np, 0, p->dealias == FIT_DEALIAS_TWOTHIRDS,
               (R)scale};
  int rc;
  if (p->nd == 1) return launch_row_fwd<R>(p, lr, sn, 1, p->h1);
  if ((rc = launch_row_fwd<R>(p, lr, st_work<R>(p), 1, p->h1))) return rc;
  if (p->nd == 2) return launch_col_full<R, -1>(p, ld_work<R>(p), sn, 1, 1);
  if ((rc = launch_col_full<R, -1>(p, ld_work<R>(p), st_work<R>(p), 1, 1))) return rc;
  return launch_col_full<R, -1>(p, ld_work<R>(p), sn, 2, 1);
}

template <typename R> LdSpectral<R> spectral_loader(fit_plan* p, const void* src, int gradj, int giver) {
  LdSpectral<R> ls;
  ls.g = grid_of<R>(p);
  ls.src = (const cpx<R>*)src;
  ls.sf = p->Np;
  ls.trunc = p->dealias == FIT_DEALIAS_TWOTHIRDS;
  ls.gradj = gradj;
  ls.gradj_per_field = 0;
  ls.src_per_field_div = 1;
  ls.ks0 = (const R*)p->ks[0];
  ls.ks1 = (const R*)p->ks[1];
  ls.ks2 = (const R*)p->ks[2];
  ls.shell = p->shell;
  ls.giver = giver;
  ls.srcs = nullptr;
  ls.cps = 1;
  ls.gstride = 0;
  return ls;
}

const int* full_bi(fit_plan* p) {
  static thread_local int b[3];
  b[0] = p->G[0]; b[1] = p->G[1]; b[2] = p->G[2];
  return b;
}

// Advecting velocity u_j on the transform grid (FFTWExt.jl:80-86 / :171): u_j = Re ifft(keep â_j)
template <typename R>
int compute_u(fit_plan* p, const void* adv) {
  const double inv_mtot = 1.0 / (double)p->Gp;
  const int* bi = p->plain ? full_bi(p) : p->band_in;
  for (int j = 0; j < p->nd; ++j) {
    LdSpectral<R> ls = spectral_loader<R>(p, (const cpx<R>*)adv + (int64_t)j * p->Np, -1, 0);
    StReal<R> sr{(R*)p->uphys + (int64_t)j * p->Gp, (int64_t)p->G[0], (int64_t)p->G[0] * p->G[1], 0,
                 (R)inv_mtot};
    int rc = inverse_field<R>(p, ls, sr, 0.0, bi);
    if (rc) return rc;
  }
  return FIT_OK;
}

// Plain pipeline: N̂ for the primary field f (optionally restricted to giver shell m), u computed.
template <typename R>
int nonlinear_from_u(fit_plan* p, const void* f, int giver) {
  const double inv_mtot = 1.0 / (double)p->Gp;
  const double fscale = (double)p->Gp / ((double)p->Np * (double)p->Np);
  const int64_t sy = p->G[0], sz = (int64_t)p->G[0] * p->G[1];
  for (int c = 0; c < p->M; ++c) {
    for (int j = 0; j < p->nd; ++j) {
      LdSpectral<R> ls = spectral_loader<R>(p, (const cpx<R>*)f + (int64_t)c * p->Np, j, giver);
      StProduct<R> sp{(R*)p->nphys, (const R*)p->uphys + (int64_t)j * p->Gp, sy, sz, (R)inv_mtot, j == 0};
      int rc = inverse_field<R>(p, ls, sp, (double)p->G[0] * sizeof(R) * (j == 0 ? 2.0 : 3.0), full_bi(p));
      if (rc) return rc;
    }
    int rc = forward_field<R>(p, c, fscale);
    if (rc) return rc;
  }
  return FIT_OK;
}

// ------------------------------ fused / pruned pipeline ---------------------------------------

int grow(fit_plan* p, void** buf, size_t* have, size_t need) {
  if (*have >= need) return FIT_OK;
  if (*buf) { (void)hipStreamSynchronize(p->stream); (void)hipFree(*buf); p->bytes -= (int64_t)*have; *buf = nullptr; *have = 0; }
  int rc = dalloc(p, buf, need);
  if (rc) return rc;
  *have = need;
  return FIT_OK;
}

// x-extent of compact storage for a folded band on the half grid
int xext(fit_plan* p, int a) { return std::min(a, p->G[0] / 2) + 1; }
// row pitch of compact storage: the extent rounded up to 16 elements, so that every row starts
// on a 128-byte line and column tiles never straddle two lines
int xpitch(int xe) { return (xe + 15) / 16 * 16; }

// shared-transform gradient passes apply: exact for this plan, register lengths on every axis
bool commute_ok(const fit_plan* p) {
  if (!p->commute || p->plain || p->nd < 2) return false;
  for (int d = 0; d < p->nd; ++d)
    if (!reg_supported(p->G[d])) return false;
  return true;
}

// Columns per workgroup of the two-line gradient passes (512 threads at most): 16 within 64 KB
// of LDS, or the plan's column tile within 160 KB.
int grad_tile(const fit_plan* p, int L, size_t csz) {
  const int T = two_T(L);
  int TX = std::min(p->col_tx > 0 ? p->col_tx : 16, 512 / T);
  const size_t cap = p->col_tx > 0 ? (size_t)160 * 1024 : g_max_lds;
  while (TX > 1 && (size_t)L * TX * csz > cap) TX /= 2;
  return TX;
}

// First and (3D) y passes of the shared-transform gradients of ng*M fields into work2.
template <typename R>
int grads_shared(fit_plan* p, const void* f, int giver0, int ng, const int* bi, const int4* fband, int xa,
                 StHalf<R> st, const void* const* srcs) {
  const size_t csz = sizeof(cpx<R>);
  const int nd = p->nd, ngc = ng * p->M;
  const int ax = nd - 1;
  LdSpectral<R> ls = spectral_loader<R>(p, f, 0, 0);
  if (srcs) {
    ls.srcs = (const cpx<R>* const*)srcs;
    ls.cps = p->M;
    ls.gstride = p->M;
  }
  {
    const int L = p->G[ax];
    const int TX = grad_tile(p, L, csz);
    const Band ob = nd == 3 ? Band{bi[1], p->G[1]} : full_band(1);
    const Band ib = Band{bi[ax], L};
    const Band wb = full_band(L);
    const int nob = ob.count();
    dim3 grid((xa + TX - 1) / TX, nob, ngc);
    prof_begin(p, FIT_K_FFT_COLUMN, (double)ngc * xa * nob * (2.0 * ib.count() + 2.0 * wb.count()) * csz);
    switch (L) {
#define FIT_CASE(LL)                                                                                         \
  case LL:                                                                                                   \
    hipLaunchKernelGGL((k_grad1_r<R, LL>), grid, dim3(two_T(LL) * TX), (size_t)L * TX * csz, p->stream, ls,  \
                       p->M, giver0, st, (const cpx<R>*)p->tw[ax], xa, TX, ax, ob, ib, wb, fband, p->G[1]);  \
    break;
      FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
      default: break;
    }
    prof_end(p);
    FIT_HIP(hipGetLastError());
  }
  if (nd == 3) {
    const int L = p->G[1];
    const int TX = grad_tile(p, L, csz);
    const Band ob = full_band(p->G[2]);
    const Band ib = Band{bi[1], L};
    const Band wb = full_band(L);
    const bool split = p->grad_split < 0 ? L < 512 : p->grad_split != 0;
    dim3 grid((xa + TX - 1) / TX, ob.count(), ngc);
    prof_begin(p, FIT_K_FFT_COLUMN,
               (double)ngc * xa * ob.count() * ((split ? 1.0 : 2.0) * ib.count() + (split ? 2.0 : 3.0) * wb.count()) * csz);
    switch (L) {
#define FIT_CASE(LL)                                                                                         \
  case LL:                                                                                                   \
    hipLaunchKernelGGL((k_grad2_r<R, LL>), grid, dim3(two_T(LL) * TX), (size_t)L * TX * csz, p->stream,      \
                       st.p, st.sy, st.sz, st.sf, (const R*)p->kt[1], (const cpx<R>*)p->tw[1], xa, TX, ob,   \
                       ib, wb, fband, p->M, split ? 0 : 1);                                                  \
    break;
      FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
      default: break;
    }
    prof_end(p);
    FIT_HIP(hipGetLastError());
    if (split) {
      // slot 2 (the last-axis derivative) needs only its own y transform: the one-line column
      // pass, which holds twice the waves of the two-line kernel
      LdHalf<R> l2{st.p + 2 * st.sf, st.sy, st.sz, 3 * st.sf};
      StHalf<R> s2{st.p + 2 * st.sf, st.sy, st.sz, 3 * st.sf};
      int rc = launch_col<R, 1>(p, l2, s2, 1, ngc, xa, ob, ib, wb, fband, p->M);
      if (rc) return rc;
    }
  }
  return FIT_OK;
}

// Inverse passes (all but x) of ng*M*nd gradient fields (giver g, component c, direction j) of
// f̂ into work2 (compact, pitch xpitch(xa)).  giver0 = 0: no shell filter.  fband: per-giver bounds.
// srcs (ensembles): device table of ng snapshot arrays (M components each); group g reads srcs[g].
template <typename R>
int v2_inverse_grads(fit_plan* p, const void* f, int giver0, int ng, const int* bi, const int4* fband,
                     const void* const* srcs = nullptr) {
  const int nf = ng * p->M * p->nd;
  const int xa = xext(p, bi[0]), pa = xpitch(xa);
  const size_t csz = sizeof(cpx<R>);
  int rc = grow(p, &p->work2, &p->work2_bytes, (size_t)nf * pa * p->G[1] * p->G[2] * csz);
  if (rc) return rc;
  const int64_t sy = pa, sz = (int64_t)pa * p->G[1], sf = sz * p->G[2];
  StHalf<R> st{(cpx<R>*)p->work2, sy, sz, sf};
  p->grads_commuted = false;
  if (commute_ok(p)) {
    if ((rc = grads_shared<R>(p, f, giver0, ng, bi, fband, xa, st, srcs))) return rc;
    p->grads_commuted = true;
    return FIT_OK;
  }
  LdSpectralB<R> ls{spectral_loader<R>(p, f, 0, 0), p->M, giver0};
  if (srcs) {
    ls.base.srcs = (const cpx<R>* const*)srcs;
    ls.base.cps = p->M;
    ls.base.gstride = p->M;
  }
  LdHalf<R> ld{(const cpx<R>*)p->work2, sy, sz, sf};
  const int fdiv = p->M * p->nd;
  if (p->nd == 3) {
    if ((rc = launch_col<R, 1>(p, ls, st, 2, nf, xa, Band{bi[1], p->G[1]}, Band{bi[2], p->G[2]},
                               full_band(p->G[2]), fband, fdiv)))
      return rc;
    return launch_col<R, 1>(p, ld, st, 1, nf, xa, full_band(p->G[2]), Band{bi[1], p->G[1]},
                            full_band(p->G[1]), fband, fdiv);
  }
  return launch_col<R, 1>(p, ls, st, 1, nf, xa, full_band(1), Band{bi[1], p->G[1]}, full_band(p->G[1]),
                          fband, fdiv);
}

// Fused x pass: ng*M field groups -> forward-x half rows kx <= bo[0] in nxbuf (pitch xpitch(xb)).
// u: physical advecting velocity (nd components, pitch Gp) of group 0; ugs: its stride per group
// (0: one velocity for every group, as for the giver shells; ensembles: nd*Gp).
template <typename R>
int v2_xfused(fit_plan* p, int ng, const int* bi, const int4* fband, const int* bo, const void* u = nullptr,
              int64_t ugs = 0) {
  if (!u) u = p->uphys;
  const int nfg = ng * p->M;
  const int xa = xext(p, bi[0]), pa = xpitch(xa);
  const int xb = xext(p, bo[0]), pb = xpitch(xb);
  const size_t csz = sizeof(cpx<R>);
  int rc = grow(p, &p->nxbuf, &p->nxbuf_bytes, (size_t)nfg * pb * p->G[1] * p->G[2] * csz);
  if (rc) return rc;
  const int L = p->G[0];
  const int nrows = p->G[1] * p->G[2];
  const int npairs = (nrows + 1) / 2;
  const int64_t gsy = pa, gsz = (int64_t)pa * p->G[1], gsf = gsz * p->G[2];
  const int64_t osy = pb, osz = (int64_t)pb * p->G[1], osf = osz * p->G[2];
  const int64_t usy = L, usz = (int64_t)L * p->G[1];
  const double bytes = (double)nfg * nrows * (p->nd * (xa * csz) + xb * csz) + (double)p->nd * nrows * L * sizeof(R);
  const R* kxt = p->grads_commuted ? (const R*)p->kt[0] : nullptr;
  if (p->grads_commuted && !reg_supported(L)) return fail(FIT_E_UNSUPPORTED, "shared-transform gradients need a register x pass");
  prof_begin(p, FIT_K_FUSED_X, bytes);
  // wave form for short register footprints (E <= 16); the workgroup form otherwise
  const bool wave_ok = reg_supported(L) && !p->xfused_block && wE(L) <= 16 &&
                       (size_t)4 * (64 / (L / wE(L))) * ((size_t)L + L / 16) * csz <= g_max_lds;
  if (wave_ok) {
    // wave-per-line kernel: 4 waves per workgroup, 64/T row pairs per wave
    const int T = L / wE(L);
    const int pairs_per_block = 4 * (64 / T);
    const size_t LL = (size_t)L + L / 16;
    const size_t lds = (size_t)pairs_per_block * LL * csz;
    const int nrg = (npairs + pairs_per_block - 1) / pairs_per_block;
    const int ngiv = nfg / p->M;
    const int GC = ugs ? 1 : std::max(1, std::min(ngiv, p->xfused_gc > 0 ? p->xfused_gc : 4));
    const int nchunk = (ngiv + GC - 1) / GC;
    // only lengths with at most 16 points per lane are instantiated (the others never qualify)
    switch (L) {
#define FIT_CASE(LLEN)                                                                                            \
  case LLEN:                                                                                                      \
    if constexpr (wE(LLEN) <= 16) {                                                                               \
      hipLaunchKernelGGL((k_xfused_w<R, LLEN>), dim3((unsigned)((int64_t)nrg * nchunk)), dim3(256), lds,          \
                         p->stream, (const cpx<R>*)p->tw[0], (const cpx<R>*)p->work2, gsy, gsz, gsf, (const R*)u, \
                         usy, usz, p->Gp, (cpx<R>*)p->nxbuf, osy, osz, osf, p->nd, p->M, p->G[1], nrows,          \
                         (R)(1.0 / (double)p->Gp), xa - 1, fband, xb - 1, ngiv, GC, kxt, ugs);                     \
    } else {                                                                                                      \
      prof_end(p);                                                                                                \
      return fail(FIT_E_UNSUPPORTED, "wave-form x pass needs at most 16 points per lane");                        \
    }                                                                                                             \
    break;
      FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
      default: break;
    }
  } else if (reg_supported(L)) {
    // workgroup form: B row pairs per workgroup, single LDS exchange buffer
    const int T = xblk_T(L);
    const size_t LL = (size_t)L + L / 16;
    int B = std::max(1, 256 / T);
    while (B > 1 && B * LL * csz > g_max_lds) --B;
    const int nrg = (npairs + B - 1) / B;
    const size_t lds = B * LL * csz;
    // field groups per workgroup: every component of up to xfused_gc givers (or one snapshot)
    const int ngiv = nfg / p->M;
    const int GCM = p->M * (ugs ? 1 : std::max(1, std::min(ngiv, p->xfused_gc > 0 ? p->xfused_gc : 4)));
    const int nchunk = (nfg + GCM - 1) / GCM;
    switch (L) {
#define FIT_CASE(LLEN)                                                                                         \
  case LLEN:                                                                                                   \
    hipLaunchKernelGGL((k_xfused_r<R, LLEN>), dim3((unsigned)((int64_t)nrg * nchunk)), dim3(B * xblk_T(LLEN)), \
                       lds, p->stream, (const cpx<R>*)p->tw[0], (const cpx<R>*)p->work2, gsy, gsz, gsf,        \
                       (const R*)u, usy, usz, p->Gp, (cpx<R>*)p->nxbuf, osy, osz, osf, p->nd, p->M,            \
                       p->G[1], nrows, B, (R)(1.0 / (double)p->Gp), xa - 1, fband, xb - 1, nfg, kxt, ugs, GCM); \
    break;
      FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
      default: break;
    }
  } else {
    int P = 8;
    while (P > 1 && (size_t)3 * L * P * csz > 65536) --P;
    const int nrg = (npairs + P - 1) / P;
    hipLaunchKernelGGL(k_xfused<R>, dim3((unsigned)((int64_t)nrg * nfg)), dim3(kThreads), (size_t)3 * L * P * csz,
                       p->stream, p->fd[0], (const cpx<R>*)p->tw[0], (const cpx<R>*)p->work2, gsy, gsz, gsf,
                       (const R*)u, usy, usz, p->Gp, (cpx<R>*)p->nxbuf, osy, osz, osf, p->nd, p->M, p->G[1],
                       nrows, P, (R)(1.0 / (double)p->Gp), xa - 1, fband, xb - 1, nfg, ugs);
  }
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

// Fused x pass of a self-advected field whose gradients went through the shared-transform
// passes: u_j is formed in registers from slot 0 (no u in HBM), then N_c for every component.
template <typename R>
int v2_xfused_self(fit_plan* p, const int* bi, const int* bo) {
  const int nfg = p->M;
  const int xa = xext(p, bi[0]), pa = xpitch(xa);
  const int xb = xext(p, bo[0]), pb = xpitch(xb);
  const size_t csz = sizeof(cpx<R>);
  int rc = grow(p, &p->nxbuf, &p->nxbuf_bytes, (size_t)nfg * pb * p->G[1] * p->G[2] * csz);
  if (rc) return rc;
  const int L = p->G[0];
  const int nrows = p->G[1] * p->G[2];
  const int npairs = (nrows + 1) / 2;
  const int64_t gsy = pa, gsz = (int64_t)pa * p->G[1], gsf = gsz * p->G[2];
  const int64_t osy = pb, osz = (int64_t)pb * p->G[1], osf = osz * p->G[2];
  const double bytes = (double)nrows * (((double)p->nd + (double)nfg * p->nd) * xa * csz + (double)nfg * xb * csz);
  prof_begin(p, FIT_K_FUSED_X, bytes);
  if constexpr (std::is_same<R, float>::value) {
    // one line pair per wave (T = 64): the LDS-staged form, 4 pairs per workgroup
    if (p->xstage && (L == 512 || L == 768)) {
      const int nrg = (npairs + 3) / 4;
#define FIT_STG(LLEN)                                                                                         \
  hipLaunchKernelGGL((k_xfused_self_stg<LLEN, sE(LLEN)>), dim3((unsigned)nrg), dim3(256), 0, p->stream,        \
                     (const cpx<float>*)p->tw[0], (const cpx<float>*)p->work2, gsy, gsz, gsf,                 \
                     (cpx<float>*)p->nxbuf, osy, osz, osf, p->nd, p->M, p->G[1], nrows,                       \
                     (float)(1.0 / (double)p->Gp), xa - 1, xb - 1, (const float*)p->kt[0])
      if (L == 512) FIT_STG(512);
      else FIT_STG(768);
#undef FIT_STG
      prof_end(p);
      FIT_HIP(hipGetLastError());
      return FIT_OK;
    }
  }
  switch (L) {
#define FIT_CASE(LLEN)                                                                                          \
  case LLEN: {                                                                                                  \
    constexpr int T = LLEN / sE(LLEN);                                                                          \
    const int B = std::max(1, kThreads / T);                                                                    \
    const size_t lds = (size_t)B * ldsline(LLEN) * csz;                                                         \
    const int nrg = (npairs + B - 1) / B;                                                                       \
    hipLaunchKernelGGL((k_xfused_self<R, LLEN, sE(LLEN)>), dim3((unsigned)nrg), dim3(B * T), lds, p->stream,   \
                       (const cpx<R>*)p->tw[0], (const cpx<R>*)p->work2, gsy, gsz, gsf, (cpx<R>*)p->nxbuf, osy, \
                       osz, osf, p->nd, p->M, p->G[1], nrows, B, (R)(1.0 / (double)p->Gp), xa - 1, xb - 1,     \
                       (const R*)p->kt[0]);                                                                     \
    break;                                                                                                      \
  }
    FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
    default:
      return fail(FIT_E_UNSUPPORTED, "self-advection x pass needs a register FFT length");
  }
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

// Remaining forward passes of nxbuf (M fields, one field group per component) into full N̂.
template <typename R>
int v2_forward_nhat(fit_plan* p, const int* bo) {
  const int xb = xext(p, bo[0]), pb = xpitch(xb);
  const int64_t sy = pb, sz = (int64_t)pb * p->G[1], sf = sz * p->G[2];
  LdHalf<R> ld{(const cpx<R>*)p->nxbuf, sy, sz, sf};
  StHalf<R> st{(cpx<R>*)p->nxbuf, sy, sz, sf};
  Grid g = grid_of<R>(p);
  const bool trunc = p->dealias == FIT_DEALIAS_TWOTHIRDS;
  const double scale = (double)p->Gp / ((double)p->Np * (double)p->Np);
  void* dst = p->nhat_dst ? p->nhat_dst : p->nhat;
  StNhat<R> sn{g, (cpx<R>*)dst, p->Np, trunc, (R)scale};
  int rc;
  if (trunc) FIT_HIP(hipMemsetAsync(dst, 0, (size_t)p->M * p->Np * sizeof(cpx<R>), p->stream));
  if (p->nd == 3) {
    if ((rc = launch_col<R, -1>(p, ld, st, 1, p->M, xb, full_band(p->G[2]), full_band(p->G[1]),
                                Band{bo[1], p->G[1]})))
      return rc;
    return launch_col<R, -1>(p, ld, sn, 2, p->M, xb, Band{bo[1], p->G[1]}, full_band(p->G[2]),
                             Band{bo[2], p->G[2]});
  }
  return launch_col<R, -1>(p, ld, sn, 1, p->M, xb, full_band(1), full_band(p->G[1]), Band{bo[1], p->G[1]});
}

// N̂ of f advected by u (already in uphys) through the fused pipeline.
template <typename R>
int v2_nonlinear(fit_plan* p, const void* f) {
  const int* bo = p->band_in;  // every mode the input band feeds (2/3: kept box; padded: block)
  int rc;
  if ((rc = v2_inverse_grads<R>(p, f, 0, 1, p->band_in, nullptr))) return rc;
  if ((rc = v2_xfused<R>(p, 1, p->band_in, nullptr, bo))) return rc;
  return v2_forward_nhat<R>(p, bo);
}

// density (+ shell partial sums, reduced into p->dcol) from carrier and N̂ (both full grid)
template <typename R>
int density_launch(fit_plan* p, int inv, const void* car, const void* nh, void* t_out, bool shells) {
  Grid g = grid_of<R>(p);
  const int nsh = shells ? p->nsh : 0;
  const int nblk = p->nblk_density;
  const size_t lds = (size_t)std::max(nsh, 1) * sizeof(double);
  const size_t csz = sizeof(cpx<R>);
  prof_begin(p, FIT_K_DENSITY_SHELL,
             (double)p->Np * (2.0 * p->M * csz + (shells ? 4.0 : 0.0) + (t_out ? sizeof(R) : 0.0)));
  hipLaunchKernelGGL(k_density<R>, dim3(nblk), dim3(kThreads), lds, p->stream, g, int4{inv, 0, 0, 0}, 1, p->M,
                     (const cpx<R>*)car, p->Np, (const cpx<R>*)nh, p->Np, (const R*)p->ks[0],
                     (const R*)p->ks[1], (const R*)p->ks[2], shells ? p->shell : nullptr, nsh,
                     p->partials, (R*)t_out);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  if (shells) {
    prof_begin(p, FIT_K_REDUCE, (double)nblk * nsh * 8.0 + nsh * 8.0);
    launch_reduce(p, p->partials, nblk, nsh, 1, p->dcol);
    prof_end(p);
    FIT_HIP(hipGetLastError());
  }
  return FIT_OK;
}

// shell sums of ninv (<= 4) invariants in one density pass over carrier and N̂: rows of nsh
// doubles, invariant q at dcolG + q*nsh
template <typename R>
int density_multi(fit_plan* p, int ninv, const int* invs, const void* car, const void* nh) {
  Grid g = grid_of<R>(p);
  const int nsh = p->nsh, nb = ninv * nsh;
  const int nblk = p->nblk_density;
  const size_t csz = sizeof(cpx<R>);
  int rc;
  if ((rc = grow(p, (void**)&p->partials2, &p->partials2_bytes, (size_t)nblk * nb * sizeof(double)))) return rc;
  if ((rc = grow(p, (void**)&p->dcolG, &p->dcolG_bytes, (size_t)nb * sizeof(double)))) return rc;
  const int4 iv{invs[0], ninv > 1 ? invs[1] : 0, ninv > 2 ? invs[2] : 0, ninv > 3 ? invs[3] : 0};
  prof_begin(p, FIT_K_DENSITY_SHELL, (double)p->Np * (2.0 * p->M * csz + 4.0));
  hipLaunchKernelGGL(k_density<R>, dim3(nblk), dim3(kThreads), (size_t)nb * sizeof(double), p->stream, g, iv, ninv,
                     p->M, (const cpx<R>*)car, p->Np, (const cpx<R>*)nh, p->Np, (const R*)p->ks[0],
                     (const R*)p->ks[1], (const R*)p->ks[2], p->shell, nsh, p->partials2, (R*)nullptr);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  prof_begin(p, FIT_K_REDUCE, (double)nblk * nb * 8.0 + nb * 8.0);
  launch_reduce(p, p->partials2, nblk, nb, 1, p->dcolG);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

// ------------------------------ input / output staging ----------------------------------------

size_t csize(int dtype) { return dtype == FIT_F64 ? 16 : 8; }

// Make a device array of `ncomp` full-grid complex components in compute precision available.
// Returns the device pointer to use (user pointer if usable as is, else the staging buffer).
int stage_in(fit_plan* p, const void* src, int ncomp, int io_dtype, int ptr_kind, void** staging,
             const void** out) {
  const int64_t n = (int64_t)ncomp * p->Np;
  const size_t cb = csize(p->dtype);
  if (ptr_kind == FIT_PTR_DEVICE && io_dtype == p->dtype) {
    *out = src;
    return FIT_OK;
  }
  if (!*staging) {
    int rc = dalloc(p, staging, (size_t)std::max(p->M, p->D) * p->Np * cb);
    if (rc) return rc;
  }
  if (io_dtype == p->dtype) {
    FIT_HIP(hipMemcpyAsync(*staging, src, n * cb, hipMemcpyHostToDevice, p->stream));
  } else {
    const void* dsrc = src;
    void* tmp = nullptr;
    if (ptr_kind == FIT_PTR_HOST) {
      FIT_HIP(hipMalloc(&tmp, n * csize(io_dtype)));
      FIT_HIP(hipMemcpyAsync(tmp, src, n * csize(io_dtype), hipMemcpyHostToDevice, p->stream));
      dsrc = tmp;
    }
    const int64_t nr = 2 * n;
    if (io_dtype == FIT_F64)
      hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, p->stream, (const double*)dsrc,
                         (float*)*staging, nr);
    else
      hipLaunchKernelGGL((k_convert<float, double>), dim3(1024), dim3(256), 0, p->stream, (const float*)dsrc,
                         (double*)*staging, nr);
    FIT_HIP(hipGetLastError());
    if (tmp) {
      FIT_HIP(hipStreamSynchronize(p->stream));
      hipFree(tmp);
    }
  }
  *out = *staging;
  return FIT_OK;
}

// copy nvals reals (compute precision) from device src to dst (io_dtype, host or device)
int stage_out_real(fit_plan* p, const void* dsrc, void* dst, int64_t nvals, int io_dtype, int ptr_kind) {
  const size_t rs = p->rsz;
  const size_t os = io_dtype == FIT_F64 ? 8 : 4;
  if (io_dtype == p->dtype) {
    FIT_HIP(hipMemcpyAsync(dst, dsrc, nvals * rs,
                           ptr_kind == FIT_PTR_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                           p->stream));
    return FIT_OK;
  }
  void* conv = dst;
  if (ptr_kind == FIT_PTR_HOST) {
    if (p->out_tmp_bytes < nvals * os) {
      dfree(p->out_tmp);
      int rc = dalloc(p, &p->out_tmp, nvals * os);
      if (rc) return rc;
      p->out_tmp_bytes = nvals * os;
    }
    conv = p->out_tmp;
  }
  if (io_dtype == FIT_F64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(1024), dim3(256), 0, p->stream, (const float*)dsrc,
                       (double*)conv, nvals);
  else
    hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, p->stream, (const double*)dsrc,
                       (float*)conv, nvals);
  FIT_HIP(hipGetLastError());
  if (ptr_kind == FIT_PTR_HOST)
    FIT_HIP(hipMemcpyAsync(dst, conv, nvals * os, hipMemcpyDeviceToHost, p->stream));
  return FIT_OK;
}

int check_plan(fit_plan* p) {
  if (!p) return fail(FIT_E_ARGUMENT, "null plan");
  FIT_HIP(hipSetDevice(p->device));
  return FIT_OK;
}

int check_invariant(fit_plan* p, int inv, int ncomp) {
  if (inv < 0 || inv > 3) return fail(FIT_E_ARGUMENT, "unknown invariant code");
  if (inv == FIT_HELICITY && p->nd != 3)
    return fail(FIT_E_ARGUMENT, "Helicity transfer is defined in 3D only (got nd=" + std::to_string(p->nd) + ").");
  if (inv == FIT_ENSTROPHY && p->nd != 2 && p->nd != 3)
    return fail(FIT_E_ARGUMENT, "Enstrophy transfer is defined in 2D or 3D; got nd=" + std::to_string(p->nd) + ".");
  if ((inv == FIT_HELICITY || inv == FIT_ENSTROPHY) && ncomp < p->nd)
    return fail(FIT_E_DIMENSION, "vorticity-weighted invariants need M >= nd components");
  return FIT_OK;
}

// stage f / adv for a nonlinear-term call
int stage_fa(fit_plan* p, const void* f_hat, const void* adv_hat, int io_dtype, int ptr_kind, const void** f,
             const void** a) {
  if (!f_hat) return fail(FIT_E_ARGUMENT, "f_hat is NULL");
  if (io_dtype != FIT_F32 && io_dtype != FIT_F64) return fail(FIT_E_ARGUMENT, "io_dtype must be FIT_F32 or FIT_F64");
  if (ptr_kind != FIT_PTR_HOST && ptr_kind != FIT_PTR_DEVICE) return fail(FIT_E_ARGUMENT, "bad ptr_kind");
  int rc = stage_in(p, f_hat, p->M, io_dtype, ptr_kind, &p->in_f, f);
  if (rc) return rc;
  if (adv_hat) {
    rc = stage_in(p, adv_hat, p->D, io_dtype, ptr_kind, &p->in_a, a);
    if (rc) return rc;
  } else {
    if (p->M < p->nd)
      return fail(FIT_E_DIMENSION, "self-advection needs M >= nd velocity components; pass adv_hat");
    *a = *f;
  }
  return FIT_OK;
}

template <typename R>
int nonlinear_term_impl(fit_plan* p, const void* f, const void* a) {
  int rc;
  if (a == f && commute_ok(p)) {
    // self-advection: u_j = Re ifft(keep f̂_j) is the slot-0 transform of component j, so only
    // its x pass remains (the y and z passes are shared with the gradients); the fused x pass
    // forms it in registers
    if ((rc = v2_inverse_grads<R>(p, f, 0, 1, p->band_in, nullptr))) return rc;
    if (p->xself) {
      if ((rc = v2_xfused_self<R>(p, p->band_in, p->band_in))) return rc;
      return v2_forward_nhat<R>(p, p->band_in);
    }
    const int xa = xext(p, p->band_in[0]), pa = xpitch(xa);
    const int64_t gsf = (int64_t)pa * p->G[1] * p->G[2];
    LdHalf<R> ld{(const cpx<R>*)p->work2, (int64_t)pa, (int64_t)pa * p->G[1], (int64_t)p->nd * gsf};
    StReal<R> sr{(R*)p->uphys, (int64_t)p->G[0], (int64_t)p->G[0] * p->G[1], p->Gp, (R)(1.0 / (double)p->Gp)};
    if ((rc = launch_row_inv<R>(p, ld, sr, p->nd, 0.0, xa - 1))) return rc;
    if ((rc = v2_xfused<R>(p, 1, p->band_in, nullptr, p->band_in))) return rc;
    return v2_forward_nhat<R>(p, p->band_in);
  }
  rc = compute_u<R>(p, a);
  if (rc) return rc;
  if (p->plain || p->nd == 1) return nonlinear_from_u<R>(p, f, 0);
  return v2_nonlinear<R>(p, f);
}

int finish(fit_plan* p) {
  FIT_HIP(hipStreamSynchronize(p->stream));
  prof_collect(p);
  return FIT_OK;
}

int spectral_flux_tail(fit_plan* p, double* T, double* Pi) {
  std::vector<double> h(p->nsh);
  FIT_HIP(hipMemcpyAsync(h.data(), p->dcol, p->nsh * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  int rc = finish(p);
  if (rc) return rc;
  double acc = 0.0;
  for (int s = 0; s < p->nsh; ++s) {
    if (T) T[s] = h[s];
    acc += h[s];
    if (Pi) Pi[s] = acc;  // Π = +cumsum(T) (SpectralFlux.jl:172-177)
  }
  return FIT_OK;
}

// Plain giver loop: one giver per pipeline pass, full transforms, N̂ through HBM.
template <typename R>
int s2s_plain(fit_plan* p, int inv, const void* f, std::vector<double>& Tm) {
  int rc;
  const int nsh = p->nsh;
  for (int m = 1; m <= nsh; ++m) {
    if ((rc = nonlinear_from_u<R>(p, f, m))) return rc;
    if ((rc = density_launch<R>(p, inv, f, p->nhat, nullptr, true))) return rc;
    FIT_HIP(hipMemcpyAsync(Tm.data() + (size_t)(m - 1) * nsh, p->dcol, nsh * sizeof(double), hipMemcpyDeviceToHost,
                           p->stream));
    // dcol is reused by the next giver: order the copy before the next reduction
    FIT_HIP(hipStreamSynchronize(p->stream));
  }
  return FIT_OK;
}

// Tail of a batched pass for gb field groups already through the fused x pass (nxbuf, pitch
// xpitch(xext(band_rx))): the remaining forward pass(es), the pair density against eh (+ g*ehgs for group
// g) and FP64 shell bins, reduced in fixed order into out (gb rows of nsh doubles, device).
template <typename R>
int batched_lastdens(fit_plan* p, int gb, const void* eh, int64_t ehgs, double* out) {
  const int nsh = p->nsh;
  const int* bo = p->band_rx;
  const size_t csz = sizeof(cpx<R>);
  const int n1h = p->n[0] / 2 + 1;
  const int64_t Hn = (int64_t)n1h * p->n[1] * p->n[2];
  const int lax = p->nd == 3 ? 2 : 1;
  const int L = p->G[lax];
  // columns per workgroup: up to 16 (128-byte rows) within the 160 KB a workgroup may hold; long
  // columns (2048-point y lines of 2D ensembles) keep at least two lines per workgroup this way
  int TX = 16;
  // (TX*L <= 16*256 keeps the per-thread accumulators at 16: about 160 VGPRs, three waves per SIMD)
  while (TX > 1 && ((size_t)2 * L * TX * csz + (size_t)nsh * 8 > (size_t)160 * 1024 || TX * L > 16 * kThreads))
    TX /= 2;
  const int xb = std::max(1, xext(p, bo[0]));
  const Band ob = p->nd == 3 ? Band{bo[1], p->G[1]} : full_band(1);
  const Band wb = p->nd == 3 ? Band{bo[2], p->G[2]} : Band{bo[1], p->G[1]};
  const int gx = (xb + TX - 1) / TX, gy = ob.count();
  const int nblk = gx * gy;
  int rc;
  if ((rc = grow(p, (void**)&p->partials2, &p->partials2_bytes, (size_t)gb * nblk * nsh * sizeof(double)))) return rc;
  const double scale = 1.0 / (double)p->Np;
  const int pb = xpitch(xb);
  const int64_t sy = pb, sz = (int64_t)pb * p->G[1], sf = sz * p->G[2];
  if (p->nd == 3) {
    LdHalf<R> ld{(const cpx<R>*)p->nxbuf, sy, sz, sf};
    StHalf<R> st{(cpx<R>*)p->nxbuf, sy, sz, sf};
    if ((rc = launch_col<R, -1>(p, ld, st, 1, gb * p->M, xb, full_band(p->G[2]), full_band(p->G[1]),
                                Band{bo[1], p->G[1]})))
      return rc;
  }
  prof_begin(p, FIT_K_DENSITY_SHELL,
             (double)gb * p->M * xb * gy * L * csz + (double)gb * p->M * xb * gy * wb.count() * csz);
  // per-thread accumulators: the smallest register array that covers TX*L points per workgroup
  const int qneed = (L * TX + kThreads - 1) / kThreads;
#define FIT_LASTDENS(QQ)                                                                                         \
  hipLaunchKernelGGL((k_lastdens<R, QQ>), dim3(gx, gy, gb), dim3(kThreads), (size_t)2 * L * TX * csz + (size_t)nsh * 8, \
                     p->stream, p->fd[lax], (const cpx<R>*)p->tw[lax], (const cpx<R>*)p->nxbuf, sy, sz, sf, xb, TX, \
                     lax, ob, wb, p->M, (R)scale, (const cpx<R>*)eh, (int64_t)n1h, (int64_t)n1h * p->n[1], Hn,     \
                     p->shh, nsh, p->n[0], p->partials2, ehgs)
  if (qneed <= 4) FIT_LASTDENS(4);
  else if (qneed <= 8) FIT_LASTDENS(8);
  else if (qneed <= 16) FIT_LASTDENS(16);
  else FIT_LASTDENS(32);
#undef FIT_LASTDENS
  prof_end(p);
  FIT_HIP(hipGetLastError());
  prof_begin(p, FIT_K_REDUCE, (double)gb * nblk * nsh * 8.0);
  launch_reduce(p, p->partials2, nblk, nsh, gb, out);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

// Hermitian part of the masked effective carrier on the receiving band, for nb snapshots
// (srcs: device table of their arrays, or nullptr for the single array f) at eh + b*M*Hn.
template <typename R>
int carrier_half(fit_plan* p, int inv, const void* f, const void* const* srcs, int nb, void* eh) {
  const int* bo = p->band_rx;
  const size_t csz = sizeof(cpx<R>);
  const int n1h = p->n[0] / 2 + 1;
  const int64_t Hn = (int64_t)n1h * p->n[1] * p->n[2];
  Grid g = grid_of<R>(p);
  prof_begin(p, FIT_K_OTHER, (double)nb * Hn * p->M * csz * 3.0);
  hipLaunchKernelGGL(k_carrier_half<R>, dim3(srcs ? 256 : 2048, nb), dim3(256), 0, p->stream, g, inv, p->M,
                     p->dealias == FIT_DEALIAS_TWOTHIRDS ? 1 : 0, (const cpx<R>*)f, p->Np, (const R*)p->ks[0],
                     (const R*)p->ks[1], (const R*)p->ks[2], Band{bo[1], p->n[1]}, Band{bo[2], p->n[2]}, bo[0],
                     (cpx<R>*)eh, Hn, (const cpx<R>* const*)srcs, (int64_t)p->M * Hn);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

// Bytes of batched workspace one field group (giver or snapshot) needs in the pruned pipeline.
template <typename R>
double batch_bytes_per_group(fit_plan* p, int xa_max) {
  const int xb = std::max(1, xext(p, p->band_rx[0]));
  return ((double)p->M * p->nd * xpitch(xa_max) + (double)p->M * xpitch(xb)) * p->G[1] * p->G[2] * sizeof(cpx<R>);
}

// Batched giver loop: every giver of a batch goes through each pruned pass in one launch; the
// last forward pass computes the pair density and the shell sums in place (no N̂ in HBM).
template <typename R>
int s2s_batched(fit_plan* p, int inv, const void* f, std::vector<double>& Tm) {
  const int nsh = p->nsh;
  const int* bo = p->band_rx;
  const size_t csz = sizeof(cpx<R>);
  const int n1h = p->n[0] / 2 + 1;
  const int64_t Hn = (int64_t)n1h * p->n[1] * p->n[2];
  int rc;
  // carrier Hermitian part on the receiving band
  if ((rc = grow(p, &p->eh, &p->eh_bytes, (size_t)p->M * Hn * csz))) return rc;
  if ((rc = carrier_half<R>(p, inv, f, nullptr, 1, p->eh))) return rc;
  // batch size from the device memory budget
  int bmax[3] = {-1, -1, -1};
  for (int s = 0; s < nsh; ++s) {
    bmax[0] = std::max(bmax[0], p->shell_band[s].x);
    bmax[1] = std::max(bmax[1], p->shell_band[s].y);
    bmax[2] = std::max(bmax[2], p->shell_band[s].z);
  }
  const int xa_max = std::max(1, xext(p, bmax[0]));
  const double per_giver = batch_bytes_per_group<R>(p, xa_max);
  size_t freeb = 0, totalb = 0;
  FIT_HIP(hipMemGetInfo(&freeb, &totalb));
  const double budget = 0.6 * ((double)freeb + (double)p->work2_bytes + (double)p->nxbuf_bytes);
  int ng = (int)std::max(1.0, std::min((double)nsh, budget / per_giver));
  if (p->giver_batch > 0) ng = std::min(ng, p->giver_batch);
  if ((rc = grow(p, (void**)&p->dcolG, &p->dcolG_bytes, (size_t)ng * nsh * sizeof(double)))) return rc;
  for (int m0 = 1; m0 <= nsh; m0 += ng) {
    const int gb = std::min(ng, nsh - m0 + 1);
    int bi[3] = {-1, -1, -1};
    for (int s = m0 - 1; s < m0 - 1 + gb; ++s) {
      bi[0] = std::max(bi[0], p->shell_band[s].x);
      bi[1] = std::max(bi[1], p->shell_band[s].y);
      bi[2] = std::max(bi[2], p->shell_band[s].z);
    }
    if (bi[0] < 0) continue;  // empty givers: zero columns
    const int4* fb = p->d_shell_band + (m0 - 1);
    if ((rc = v2_inverse_grads<R>(p, f, m0, gb, bi, fb))) return rc;
    if ((rc = v2_xfused<R>(p, gb, bi, fb, bo))) return rc;
    if ((rc = batched_lastdens<R>(p, gb, p->eh, 0, p->dcolG))) return rc;
    FIT_HIP(hipMemcpyAsync(Tm.data() + (size_t)(m0 - 1) * nsh, p->dcolG, (size_t)gb * nsh * sizeof(double),
                           hipMemcpyDeviceToHost, p->stream));
    FIT_HIP(hipStreamSynchronize(p->stream));
  }
  return FIT_OK;
}

template <typename R>
int s2s_impl(fit_plan* p, int inv, const void* f, const void* a, std::vector<double>& Tm) {
  int rc = compute_u<R>(p, a);
  if (rc) return rc;
  if (p->plain || p->nd == 1 || p->dealias == FIT_DEALIAS_PADDED || !p->shell_symmetric)
    return s2s_plain<R>(p, inv, f, Tm);
  return s2s_batched<R>(p, inv, f, Tm);
}

// Giver loop on the band-limited sub-grid.  With receivers and givers in |k_d| <= R_d, every
// triad k = p + q that T(n,m) sums has |q_d| <= 2 R_d, and on a grid of L_d > |q|+|p|+|k| points
// the product of those modes is alias-free at the receivers, so the child grid yields the same
// convolution the full grid does (FFTWExt.jl:234-274).  N̂ = conv/Np² on either grid, hence the
// (Np_sub/Np)² rescale of the shell sums.
template <typename R>
int s2s_sub(fit_plan* p, int inv, const void* f, const void* a, std::vector<double>& Tm) {
  fit_plan* c = p->sub;
  const size_t csz = sizeof(cpx<R>);
  auto gather = [&](const void* src, int ncomp, void** dst, size_t* have) -> int {
    int r = grow(p, dst, have, (size_t)ncomp * c->Np * csz);
    if (r) return r;
    prof_begin(p, FIT_K_OTHER, (double)ncomp * c->Np * csz * 2.0);
    hipLaunchKernelGGL(k_subgrid_gather<R>, dim3(2048), dim3(256), 0, p->stream, (const cpx<R>*)src, p->n[0],
                       p->n[1], p->n[2], p->Np, (cpx<R>*)*dst, c->n[0], c->n[1], c->n[2], c->Np, p->sub_q[0],
                       p->sub_q[1], p->sub_q[2], ncomp);
    prof_end(p);
    FIT_HIP(hipGetLastError());
    return FIT_OK;
  };
  int rc = gather(f, p->M, &p->sub_f, &p->sub_f_bytes);
  if (rc) return rc;
  const void* ca = p->sub_f;
  if (a != f) {
    if ((rc = gather(a, p->D, &p->sub_a, &p->sub_a_bytes))) return rc;
    ca = p->sub_a;
  }
  c->prof = p->prof;
  if ((rc = s2s_impl<R>(c, inv, p->sub_f, ca, Tm))) return rc;
  FIT_HIP(hipStreamSynchronize(p->stream));
  for (double& v : Tm) v *= p->sub_tscale;
  return FIT_OK;
}

template <typename R>
int s2s_entry(fit_plan* p, int inv, const void* f, const void* a, std::vector<double>& Tm) {
  if (p->sub && p->use_sub && !p->plain) return s2s_sub<R>(p, inv, f, a, Tm);
  return s2s_impl<R>(p, inv, f, a, Tm);
}

// Smooth band-to-band transfer (BandTransfer.jl:66-113): for every band m, the nonlinear term of
// w_m ⊙ f̂ advected by the full velocity (fused pipeline, u computed once), then the band sums
// Σ_k w_n(k) t(k) of its density for every n in one pass.  kc: per-mode band coordinate.
template <typename R>
int b2b_impl(fit_plan* p, int inv, const void* f, const void* a, const R* kc, const double* cen, int nb,
             double inv2s2, std::vector<double>& Tm) {
  int rc = compute_u<R>(p, a);
  if (rc) return rc;
  const size_t cs = sizeof(cpx<R>);
  if (!p->in_c && (rc = dalloc(p, &p->in_c, (size_t)std::max(p->M, p->D) * p->Np * cs))) return rc;
  const int nblk = p->nblk_density;
  if ((rc = grow(p, (void**)&p->partials2, &p->partials2_bytes, (size_t)nblk * nb * sizeof(double)))) return rc;
  if ((rc = grow(p, (void**)&p->dcolG, &p->dcolG_bytes, (size_t)nb * sizeof(double)))) return rc;
  Grid g = grid_of<R>(p);
  p->nhat_dst = nullptr;
  for (int m = 0; m < nb; ++m) {
    prof_begin(p, FIT_K_OTHER, (double)p->Np * (2.0 * p->M * cs + sizeof(R)));
    hipLaunchKernelGGL(k_band_field<R>, dim3(1024), dim3(256), 0, p->stream, p->Np, p->M, (const cpx<R>*)f, p->Np,
                       kc, cen, nb, inv2s2, m, (cpx<R>*)p->in_c);
    prof_end(p);
    FIT_HIP(hipGetLastError());
    rc = (p->plain || p->nd == 1) ? nonlinear_from_u<R>(p, p->in_c, 0) : v2_nonlinear<R>(p, p->in_c);
    if (rc) return rc;
    prof_begin(p, FIT_K_DENSITY_SHELL, (double)p->Np * (2.0 * p->M * cs + sizeof(R)));
    hipLaunchKernelGGL(k_band_density<R>, dim3(nblk), dim3(kThreads), (size_t)nb * sizeof(double), p->stream, g, inv,
                       p->M, (const cpx<R>*)f, p->Np, (const cpx<R>*)p->nhat, p->Np, (const R*)p->ks[0],
                       (const R*)p->ks[1], (const R*)p->ks[2], kc, cen, nb, inv2s2, p->partials2);
    prof_end(p);
    FIT_HIP(hipGetLastError());
    prof_begin(p, FIT_K_REDUCE, (double)nblk * nb * 8.0);
    launch_reduce(p, p->partials2, nblk, nb, 1, p->dcolG);
    prof_end(p);
    FIT_HIP(hipGetLastError());
    FIT_HIP(hipMemcpyAsync(Tm.data() + (size_t)m * nb, p->dcolG, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost,
                           p->stream));
    FIT_HIP(hipStreamSynchronize(p->stream));
  }
  return FIT_OK;
}

template <typename R>
int m2m_impl(fit_plan* p, int inv, const void* f, const void* a, void* S, void* net, int io_dtype,
                    int ptr_kind) {
  Grid g = grid_of<R>(p);
  const size_t cs = 2 * sizeof(R);
  void *A = nullptr, *E = nullptr, *Sd = nullptr, *netpart = nullptr, *netd = nullptr;
  int rc = FIT_OK;
  const int64_t Np = p->Np;
  const int64_t pchunk = 256;
  const int nchunk = (int)((Np + pchunk - 1) / pchunk);
  const size_t os = io_dtype == FIT_F64 ? 8 : 4;
  auto cleanup = [&]() { if (A) (void)hipFree(A); if (E) (void)hipFree(E); if (Sd && Sd != S) (void)hipFree(Sd); if (netpart) (void)hipFree(netpart); if (netd) (void)hipFree(netd); };
  if (hipMalloc(&A, (size_t)p->nd * Np * cs) != hipSuccess || hipMalloc(&E, (size_t)p->M * Np * cs) != hipSuccess ||
      hipMalloc(&netpart, (size_t)nchunk * Np * sizeof(double)) != hipSuccess ||
      hipMalloc(&netd, (size_t)Np * sizeof(double)) != hipSuccess) {
    cleanup();
    return fail(FIT_E_OUT_OF_MEMORY, "mode-to-mode scratch allocation failed");
  }
  if (ptr_kind == FIT_PTR_DEVICE) Sd = S;
  else if (hipMalloc(&Sd, (size_t)Np * Np * os) != hipSuccess) { cleanup(); return fail(FIT_E_OUT_OF_MEMORY, "S staging"); }
  const int trunc = p->dealias == FIT_DEALIAS_TWOTHIRDS;
  hipLaunchKernelGGL(k_m2m_prep<R>, dim3(512), dim3(256), 0, p->stream, g, trunc, inv, p->M, (const cpx<R>*)f, Np,
                     (const cpx<R>*)a, Np, (const R*)p->ks[0], (const R*)p->ks[1], (const R*)p->ks[2], (cpx<R>*)A,
                     (cpx<R>*)E);
  dim3 grid((unsigned)((Np + 255) / 256), nchunk);
  prof_begin(p, FIT_K_OTHER, (double)Np * Np * os);
  if (io_dtype == FIT_F64)
    hipLaunchKernelGGL((k_m2m<R, double>), grid, dim3(256), 0, p->stream, g, trunc, p->M, (const cpx<R>*)f, Np,
                       (const cpx<R>*)A, (const cpx<R>*)E, (const R*)p->ks[0], (const R*)p->ks[1], (const R*)p->ks[2],
                       pchunk, (double*)Sd, (double*)netpart);
  else
    hipLaunchKernelGGL((k_m2m<R, float>), grid, dim3(256), 0, p->stream, g, trunc, p->M, (const cpx<R>*)f, Np,
                       (const cpx<R>*)A, (const cpx<R>*)E, (const R*)p->ks[0], (const R*)p->ks[1], (const R*)p->ks[2],
                       pchunk, (float*)Sd, (double*)netpart);
  prof_end(p);
  launch_reduce(p, (const double*)netpart, nchunk, (int)Np, 1, (double*)netd);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { cleanup(); return fail(FIT_E_DEVICE, hipGetErrorString(e)); }
  // net in io precision
  if (io_dtype == FIT_F64) {
    e = hipMemcpyAsync(net, netd, Np * 8, ptr_kind == FIT_PTR_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                       p->stream);
  } else {
    void* nf = nullptr;
    if (hipMalloc(&nf, Np * 4) != hipSuccess) { cleanup(); return fail(FIT_E_OUT_OF_MEMORY, "net staging"); }
    hipLaunchKernelGGL((k_convert<double, float>), dim3(256), dim3(256), 0, p->stream, (const double*)netd, (float*)nf, Np);
    e = hipMemcpyAsync(net, nf, Np * 4, ptr_kind == FIT_PTR_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                       p->stream);
    (void)hipStreamSynchronize(p->stream);
    (void)hipFree(nf);
  }
  if (e == hipSuccess && ptr_kind == FIT_PTR_HOST)
    e = hipMemcpyAsync(S, Sd, (size_t)Np * Np * os, hipMemcpyDeviceToHost, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  cleanup();
  if (e != hipSuccess) return fail(FIT_E_DEVICE, hipGetErrorString(e));
  (void)rc;
  return FIT_OK;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
// Shared-transform gradients are exact when ks is odd (ks[0] = 0, ks[n-i] = -ks[i] off the
// Nyquist index) and no self-conjugate Nyquist bin is populated: Orszag 2/3 masks it, the padded
// embedding leaves its mirror empty, odd lengths have none, and a sub-grid copies no mode that far.
static int setup_commute(fit_plan* p, const double* const* ks, bool nyquist_zero) {
  p->commute = false;
  if (p->nd < 2) return FIT_OK;
  bool odd_all = true;
  for (int d = 0; d < p->nd; ++d) {
    const int n = p->n[d];
    if (ks[d][0] != 0.0) return FIT_OK;
    for (int i = 1; 2 * i < n; ++i)
      if (ks[d][n - i] != -ks[d][i]) return FIT_OK;
    odd_all = odd_all && (n % 2 == 1);
  }
  if (!(nyquist_zero || odd_all || p->dealias == FIT_DEALIAS_PADDED || p->dealias == FIT_DEALIAS_TWOTHIRDS))
    return FIT_OK;
  for (int d = 0; d < 2; ++d) {
    const int G = p->G[d], n = p->n[d];
    std::vector<double> kt((size_t)G, 0.0);
    for (int i = 0; i < G; ++i) {
      int o = 0, om = 0;
      const int mi = i == 0 ? 0 : G - i;
      if (pad_pre(i, n, G, o)) kt[i] = ks[d][o];
      else if (pad_pre(mi, n, G, om)) kt[i] = -ks[d][om];
    }
    int rc = dalloc(p, &p->kt[d], (size_t)G * p->rsz);
    if (rc) return rc;
    if (p->dtype == FIT_F64) {
      FIT_HIP(hipMemcpy(p->kt[d], kt.data(), (size_t)G * 8, hipMemcpyHostToDevice));
    } else {
      std::vector<float> kf(kt.begin(), kt.end());
      FIT_HIP(hipMemcpy(p->kt[d], kf.data(), (size_t)G * 4, hipMemcpyHostToDevice));
    }
  }
  p->commute = true;
  return FIT_OK;
}

// Plan construction (the caller holds the creation lock).  allow_sub: may attach a band-limited
// sub-grid child for the giver loop (children never nest).
static int create_plan(fit_plan** out, int nd, const int64_t* ns, int M, int D, int compute_dtype, int dealias,
                       const double* const* ks, const int64_t* shell_idx, int n_shells, int device,
                       bool allow_sub);

// smallest transform length of the form 2^a or 3*2^a (at least 8) that is >= need
static int subgrid_len(int need) {
  int best = 1 << 30;
  for (int b = 8; b < (1 << 24); b *= 2)
    if (b >= need) { best = std::min(best, b); break; }
  for (int b = 12; b < (1 << 24); b *= 2)
    if (b >= need) { best = std::min(best, b); break; }
  return best;
}

// Attach the band-limited sub-grid child when every shell lies in |k_d| <= R_d with 4 R_d < n_d
// and the child grid is at most half the transform grid.  Child: no dealiasing, the API grid's
// wavenumbers at the same folded indices, the shell map restricted to the box |k_d| <= R_d.
static int setup_subgrid(fit_plan* p, const double* const* ks, const std::vector<int>& sh) {
  if (p->nd < 2 || p->nsh <= 0 || !p->shell_symmetric) return FIT_OK;
  int R[3] = {0, 0, 0}, Q[3] = {0, 0, 0}, Ls[3] = {1, 1, 1};
  double ratio = 1.0;
  for (int d = 0; d < p->nd; ++d) {
    R[d] = std::max(p->band_rx[d], 0);
    if (4 * R[d] >= p->n[d]) return FIT_OK;
    Q[d] = 2 * R[d];
    if (p->dealias == FIT_DEALIAS_TWOTHIRDS) Q[d] = std::min(Q[d], p->band_in[d]);
    const int need = Q[d] + 2 * R[d] + 1;  // <= 4 R_d + 1 <= n_d
    Ls[d] = subgrid_len(need);
    if (Ls[d] > p->n[d]) Ls[d] = p->n[d];
    ratio *= (double)Ls[d] / (double)p->G[d];
  }
  if (ratio > 0.5) return FIT_OK;
  std::vector<double> kc[3];
  const double* kp[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < p->nd; ++d) {
    kc[d].resize(Ls[d]);
    for (int i = 0; i < Ls[d]; ++i) {
      const int k = 2 * i <= Ls[d] ? i : i - Ls[d];
      kc[d][i] = ks[d][k >= 0 ? k : k + p->n[d]];
    }
    kp[d] = kc[d].data();
  }
  const int64_t npc = (int64_t)Ls[0] * Ls[1] * Ls[2];
  std::vector<int64_t> shc((size_t)npc, 0);
  for (int64_t o = 0; o < npc; ++o) {
    const int i = (int)(o % Ls[0]), j = (int)((o / Ls[0]) % Ls[1]), l = (int)(o / ((int64_t)Ls[0] * Ls[1]));
    const int ki = 2 * i <= Ls[0] ? i : i - Ls[0];
    const int kj = 2 * j <= Ls[1] ? j : j - Ls[1];
    const int kl = 2 * l <= Ls[2] ? l : l - Ls[2];
    if (std::abs(ki) > R[0] || std::abs(kj) > R[1] || std::abs(kl) > R[2]) continue;
    const int si = ki >= 0 ? ki : ki + p->n[0], sj = kj >= 0 ? kj : kj + p->n[1], sl = kl >= 0 ? kl : kl + p->n[2];
    shc[(size_t)o] = sh[(size_t)(si + (int64_t)p->n[0] * (sj + (int64_t)p->n[1] * sl))];
  }
  const int64_t nsc[3] = {Ls[0], Ls[1], Ls[2]};
  fit_plan* c = nullptr;
  int rc = create_plan(&c, p->nd, nsc, p->M, p->D, p->dtype, FIT_DEALIAS_NONE, kp, shc.data(), p->nsh, p->device,
                       false);
  if (rc) return rc;
  // inputs beyond the copied box are zero: prune the child's input band to it
  for (int d = 0; d < p->nd; ++d) c->band_in[d] = std::min(c->band_in[d], Q[d]);
  dfree(c->kt[0]); dfree(c->kt[1]);
  if ((rc = setup_commute(c, kp, true))) { fit_plan_destroy(c); return rc; }
  (void)hipStreamDestroy(c->stream);
  c->stream = p->stream;
  c->own_stream = false;
  p->sub = c;
  for (int d = 0; d < 3; ++d) p->sub_q[d] = Q[d];
  const double r = (double)c->Np / (double)p->Np;
  p->sub_tscale = r * r;
  return FIT_OK;
}

static int create_plan(fit_plan** out, int nd, const int64_t* ns, int M, int D, int compute_dtype, int dealias,
                       const double* const* ks, const int64_t* shell_idx, int n_shells, int device,
                       bool allow_sub) {
  if (!out) return fail(FIT_E_ARGUMENT, "plan out-pointer is NULL");
  *out = nullptr;
  if (nd < 1 || nd > 3) return fail(FIT_E_ARGUMENT, "nd must be 1, 2 or 3");
  if (!ns || !ks) return fail(FIT_E_ARGUMENT, "ns / ks is NULL");
  if (M < 1) return fail(FIT_E_DIMENSION, "M must be >= 1");
  if (D < nd) return fail(FIT_E_DIMENSION, "the advecting velocity needs D >= nd components");
  if (compute_dtype != FIT_F32 && compute_dtype != FIT_F64) return fail(FIT_E_ARGUMENT, "bad compute dtype");
  if (dealias < 0 || dealias > 2) return fail(FIT_E_ARGUMENT, "bad dealiasing code");
  if (n_shells < 0 || (n_shells > 0 && !shell_idx)) return fail(FIT_E_ARGUMENT, "shell_idx missing");
  if (n_shells > 8000) return fail(FIT_E_UNSUPPORTED, "at most 8000 shells per plan");
  fit_plan* p = new fit_plan();
  p->nd = nd; p->M = M; p->D = D; p->dtype = compute_dtype; p->dealias = dealias; p->device = device;
  p->rsz = compute_dtype == FIT_F64 ? 8 : 4;
  for (int d = 0; d < nd; ++d) {
    if (ns[d] < 1 || ns[d] > (1 << 20)) { delete p; return fail(FIT_E_DIMENSION, "bad grid size"); }
    p->n[d] = (int)ns[d];
    p->G[d] = dealias == FIT_DEALIAS_PADDED ? padded_len((int)ns[d]) : (int)ns[d];
  }
  for (int d = 0; d < 3; ++d) {
    if (!make_desc(p->G[d], p->fd[d])) {
      int bad = p->G[d];
      delete p;
      return fail(FIT_E_UNSUPPORTED, "transform length " + std::to_string(bad) + " is not of the form 2^a 3^b 5^c");
    }
    if ((size_t)2 * p->G[d] * (compute_dtype == FIT_F64 ? 16 : 8) > 65536) {
      int bad = p->G[d];
      delete p;
      return fail(FIT_E_UNSUPPORTED, "transform length " + std::to_string(bad) + " exceeds the LDS line limit");
    }
  }
  p->h1 = p->G[0] / 2 + 1;
  p->Np = (int64_t)p->n[0] * p->n[1] * p->n[2];
  p->Gp = (int64_t)p->G[0] * p->G[1] * p->G[2];
  p->Hp = (int64_t)p->h1 * p->G[1] * p->G[2];
  p->nsh = n_shells;
  int rc = FIT_OK;
  auto bail = [&](int code) { fit_plan_destroy(p); return code; };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { delete p; return fail(FIT_E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e)); }
  e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete p; return fail(FIT_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
  const size_t rs = p->rsz, cs = 2 * rs;
  // twiddles exp(-2πi k/L) (FP64 host arithmetic, rounded once) and wavenumbers
  for (int d = 0; d < 3; ++d) {
    const int L = p->G[d];
    std::vector<double> tw(2 * (size_t)L);
    for (int k = 0; k < L; ++k) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)L;
      tw[2 * k] = (double)cosl(ang);
      tw[2 * k + 1] = (double)sinl(ang);
    }
    if ((rc = dalloc(p, &p->tw[d], (size_t)L * cs))) return bail(rc);
    const int nk = d < nd ? p->n[d] : 1;
    std::vector<double> kv(nk, 0.0);
    if (d < nd) for (int i = 0; i < nk; ++i) kv[i] = ks[d][i];
    if ((rc = dalloc(p, &p->ks[d], (size_t)nk * rs))) return bail(rc);
    if (compute_dtype == FIT_F64) {
      hipMemcpy(p->tw[d], tw.data(), (size_t)L * cs, hipMemcpyHostToDevice);
      hipMemcpy(p->ks[d], kv.data(), (size_t)nk * rs, hipMemcpyHostToDevice);
    } else {
      std::vector<float> twf(2 * (size_t)L), kf(nk);
      for (size_t i = 0; i < twf.size(); ++i) twf[i] = (float)tw[i];
      for (int i = 0; i < nk; ++i) kf[i] = (float)kv[i];
      hipMemcpy(p->tw[d], twf.data(), (size_t)L * cs, hipMemcpyHostToDevice);
      hipMemcpy(p->ks[d], kf.data(), (size_t)nk * rs, hipMemcpyHostToDevice);
    }
  }
  if ((rc = setup_commute(p, ks, false))) return bail(rc);
  // folded support of the spectral inputs: kept box (2/3), all modes otherwise (the embedded
  // block on the padded grid has the same folded extent n/2)
  for (int d = 0; d < 3; ++d) {
    if (d >= nd) { p->band_in[d] = 0; continue; }
    p->band_in[d] = dealias == FIT_DEALIAS_TWOTHIRDS ? p->n[d] / 3 - 1 : p->n[d] / 2;
  }
  // shell map (int32 on device); symmetry under k -> -k is recorded
  if (n_shells > 0) {
    std::vector<int> sh((size_t)p->Np);
    for (int64_t i = 0; i < p->Np; ++i) {
      const int64_t s = shell_idx[i];
      if (s < 0 || s > n_shells) { fit_plan_destroy(p); return fail(FIT_E_ARGUMENT, "shell index out of range"); }
      sh[i] = (int)s;
    }
    for (int64_t o = 0; o < p->Np && p->shell_symmetric; ++o) {
      const int ix = (int)(o % p->n[0]), iy = (int)((o / p->n[0]) % p->n[1]), iz = (int)(o / ((int64_t)p->n[0] * p->n[1]));
      const int mx = ix ? p->n[0] - ix : 0, my = iy ? p->n[1] - iy : 0, mz = iz ? p->n[2] - iz : 0;
      if (sh[o] != sh[mx + (int64_t)p->n[0] * (my + (int64_t)p->n[1] * mz)]) p->shell_symmetric = false;
    }
    if ((rc = dalloc(p, (void**)&p->shell, (size_t)p->Np * sizeof(int)))) return bail(rc);
    hipMemcpy(p->shell, sh.data(), (size_t)p->Np * sizeof(int), hipMemcpyHostToDevice);
    // per-shell folded support (∩ input band) and the receivers' support (∩ output mask)
    p->shell_band.assign(n_shells, int4{-1, -1, -1, 0});
    for (int64_t o = 0; o < p->Np; ++o) {
      const int s = sh[o];
      if (s <= 0) continue;
      const int ix = (int)(o % p->n[0]), iy = (int)((o / p->n[0]) % p->n[1]), iz = (int)(o / ((int64_t)p->n[0] * p->n[1]));
      const int fx = std::min(ix, p->n[0] - ix), fy = std::min(iy, p->n[1] - iy), fz = std::min(iz, p->n[2] - iz);
      int4& b = p->shell_band[s - 1];
      b.x = std::max(b.x, fx); b.y = std::max(b.y, fy); b.z = std::max(b.z, fz);
    }
    p->band_rx[0] = p->band_rx[1] = p->band_rx[2] = -1;
    for (auto& b : p->shell_band) {
      if (b.x < 0) continue;
      b.x = std::min(b.x, p->band_in[0]); b.y = std::min(b.y, p->band_in[1]); b.z = std::min(b.z, p->band_in[2]);
      if (b.x < 0 || b.y < 0 || b.z < 0) { b = int4{-1, -1, -1, 0}; continue; }
      p->band_rx[0] = std::max(p->band_rx[0], b.x);
      p->band_rx[1] = std::max(p->band_rx[1], b.y);
      p->band_rx[2] = std::max(p->band_rx[2], b.z);
    }
    for (int d = 0; d < 3; ++d) p->band_rx[d] = std::max(p->band_rx[d], 0);
    if ((rc = dalloc(p, (void**)&p->d_shell_band, (size_t)n_shells * sizeof(int4)))) return bail(rc);
    hipMemcpy(p->d_shell_band, p->shell_band.data(), (size_t)n_shells * sizeof(int4), hipMemcpyHostToDevice);
    // shell map on the original half grid (x <= n1/2) for the fused density pass
    if (dealias != FIT_DEALIAS_PADDED) {
      const int n1h = p->n[0] / 2 + 1;
      const int64_t Hn = (int64_t)n1h * p->n[1] * p->n[2];
      std::vector<int> shh((size_t)Hn);
      for (int64_t h = 0; h < Hn; ++h) {
        const int64_t ix = h % n1h, rest = h / n1h;
        shh[h] = sh[ix + (int64_t)p->n[0] * rest];
      }
      if ((rc = dalloc(p, (void**)&p->shh, (size_t)Hn * sizeof(int)))) return bail(rc);
      hipMemcpy(p->shh, shh.data(), (size_t)Hn * sizeof(int), hipMemcpyHostToDevice);
    }
    if (allow_sub && (rc = setup_subgrid(p, ks, sh))) return bail(rc);
  }
  if ((rc = dalloc(p, &p->work, (size_t)p->Hp * cs))) return bail(rc);
  if ((rc = dalloc(p, &p->uphys, (size_t)nd * p->Gp * rs))) return bail(rc);
  if ((rc = dalloc(p, &p->nphys, (size_t)p->Gp * rs))) return bail(rc);
  if ((rc = dalloc(p, &p->nhat, (size_t)std::max(M, 3) * p->Np * cs))) return bail(rc);
  p->nblk_density = 1024;
  if ((rc = dalloc(p, (void**)&p->partials, (size_t)p->nblk_density * std::max(n_shells, 1) * sizeof(double))))
    return bail(rc);
  if ((rc = dalloc(p, (void**)&p->dcol, (size_t)std::max(n_shells, 1) * sizeof(double)))) return bail(rc);
  hipError_t le = hipDeviceSynchronize();
  if (le != hipSuccess) { fit_plan_destroy(p); return fail(FIT_E_DEVICE, hipGetErrorString(le)); }
  *out = p;
  return FIT_OK;
}

// =============================================================================================
// Ensembles of snapshots: the batch axis of the reference (mpi_batch_map, MPIExt.jl:39-68) with
// f = calculate_spectral_flux + calculate_scalar_flux (SpectralFlux.jl:49-94, :203-215) per
// snapshot, on one GPU.  Snapshots on the same grid (size, wavenumbers and shell map) form a
// group served by one plan; every pass of a group runs a batch of its snapshots in one launch
// (batch index = the field-group index the giver loop uses) and reads each snapshot through a
// device table of array pointers, so the caller's arrays are used in place.  Per batch: the
// advecting velocities (shared by both fluxes), the kinetic-energy N̂ fused with its density and
// shell sums, then the same for the scalar.  The sequence for all groups is captured once into
// a HIP graph and replayed by later calls, which only rewrite the pointer tables.
// =============================================================================================
namespace {

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

}  // namespace

struct fit_ensemble {
  int n = 0, nd = 0, dtype = 0, dealias = 0, device = 0, with_scalar = 0;
  hipStream_t stream = nullptr;
  std::vector<EnsGroup> groups;
  std::vector<int> nsh;           // per snapshot
  std::vector<int64_t> off;       // per snapshot: first row of its shells in the outputs
  int use_graph = 1;
  int batch_cap = 0;
  bool prof = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_io = -1;              // (ptr_kind, io_dtype) the graph was captured for
};

namespace {

// restores a plan's component count when a batch runs the scalar part with M = 1
struct MGuard {
  fit_plan* p; int m;
  MGuard(fit_plan* q, int mm) : p(q), m(q->M) { q->M = mm; }
  ~MGuard() { p->M = m; }
};

// u_j = Re ifft(keep û_j) for the gb snapshots tabled at tu: fields (b, j) -> ub (b*nd + j)*Gp
template <typename R>
int ens_velocity(fit_plan* p, const void* const* tu, int gb, void* ub) {
  const int nd = p->nd, nf = gb * nd;
  const int* bi = p->band_in;
  const int xa = xext(p, bi[0]), pa = xpitch(xa);
  const size_t csz = sizeof(cpx<R>);
  int rc = grow(p, &p->work2, &p->work2_bytes, (size_t)nf * pa * p->G[1] * p->G[2] * csz);
  if (rc) return rc;
  const int64_t sy = pa, sz = (int64_t)pa * p->G[1], sf = sz * p->G[2];
  StHalf<R> st{(cpx<R>*)p->work2, sy, sz, sf};
  LdHalf<R> ld{(const cpx<R>*)p->work2, sy, sz, sf};
  LdSpectral<R> ls = spectral_loader<R>(p, nullptr, -1, 0);
  ls.srcs = (const cpx<R>* const*)tu;
  ls.cps = nd;
  if (nd == 3) {
    if ((rc = launch_col<R, 1>(p, ls, st, 2, nf, xa, Band{bi[1], p->G[1]}, Band{bi[2], p->G[2]}, full_band(p->G[2]))))
      return rc;
    if ((rc = launch_col<R, 1>(p, ld, st, 1, nf, xa, full_band(p->G[2]), Band{bi[1], p->G[1]}, full_band(p->G[1]))))
      return rc;
  } else {
    if ((rc = launch_col<R, 1>(p, ls, st, 1, nf, xa, full_band(1), Band{bi[1], p->G[1]}, full_band(p->G[1]))))
      return rc;
  }
  StReal<R> sr{(R*)ub, (int64_t)p->G[0], (int64_t)p->G[0] * p->G[1], p->Gp, (R)(1.0 / (double)p->Gp)};
  return launch_row_inv<R>(p, ld, sr, nf, 0.0, xa - 1);
}

// one batch [b0, b0 + gb) of a group: velocities, kinetic-energy rows, scalar rows
template <typename R>
int ens_batch(fit_ensemble* e, EnsGroup& G, int b0, int gb) {
  fit_plan* p = G.p;
  const int nd = p->nd, nsh = p->nsh;
  const int B = (int)G.snaps.size();
  const int64_t Hn = (int64_t)(p->n[0] / 2 + 1) * p->n[1] * p->n[2];
  const void* const* tu = (const void* const*)(G.d_tab + b0);
  const void* const* tt = (const void* const*)(G.d_tab + B + b0);
  int rc;
  if ((rc = ens_velocity<R>(p, tu, gb, G.ub))) return rc;
  const int64_t ugs = (int64_t)nd * p->Gp;
  {
    MGuard mg(p, nd);
    if ((rc = carrier_half<R>(p, FIT_KINETIC_ENERGY, nullptr, tu, gb, G.ehk))) return rc;
    if ((rc = v2_inverse_grads<R>(p, nullptr, 0, gb, p->band_in, nullptr, tu))) return rc;
    if ((rc = v2_xfused<R>(p, gb, p->band_in, nullptr, p->band_rx, G.ub, ugs))) return rc;
    if ((rc = batched_lastdens<R>(p, gb, G.ehk, (int64_t)nd * Hn, G.rows + (size_t)b0 * nsh))) return rc;
  }
  if (e->with_scalar) {
    MGuard mg(p, 1);
    if ((rc = carrier_half<R>(p, FIT_PASSIVE_SCALAR, nullptr, tt, gb, G.ehs))) return rc;
    if ((rc = v2_inverse_grads<R>(p, nullptr, 0, gb, p->band_in, nullptr, tt))) return rc;
    if ((rc = v2_xfused<R>(p, gb, p->band_in, nullptr, p->band_rx, G.ub, ugs))) return rc;
    if ((rc = batched_lastdens<R>(p, gb, G.ehs, Hn, G.rows + (size_t)(B + b0) * nsh))) return rc;
  }
  return FIT_OK;
}

// batch size of a group from the device memory budget (first run only: the graph reuses it)
template <typename R>
int ens_batch_size(fit_ensemble* e, EnsGroup& G) {
  if (G.ng > 0) return FIT_OK;
  fit_plan* p = G.p;
  const int nd = p->nd;
  const int B = (int)G.snaps.size();
  const size_t csz = sizeof(cpx<R>);
  const int64_t Hn = (int64_t)(p->n[0] / 2 + 1) * p->n[1] * p->n[2];
  const int xa = xext(p, p->band_in[0]);
  double per = 0.0;
  {
    MGuard mg(p, nd);
    per = batch_bytes_per_group<R>(p, xa);
  }
  per += (double)nd * p->Gp * sizeof(R) + (double)(nd + 1) * Hn * csz;
  size_t freeb = 0, totalb = 0;
  FIT_HIP(hipMemGetInfo(&freeb, &totalb));
  int ng = (int)std::max(1.0, std::min((double)B, 0.5 * (double)freeb / per));
  if (e->batch_cap > 0) ng = std::min(ng, e->batch_cap);
  G.ng = ng;
  int rc;
  if ((rc = grow(p, &G.ub, &G.ub_bytes, (size_t)ng * nd * p->Gp * sizeof(R)))) return rc;
  if ((rc = grow(p, &G.ehk, &G.ehk_bytes, (size_t)ng * nd * Hn * csz))) return rc;
  if (e->with_scalar && (rc = grow(p, &G.ehs, &G.ehs_bytes, (size_t)ng * Hn * csz))) return rc;
  return FIT_OK;
}

template <typename R>
int ens_run(fit_ensemble* e) {
  for (EnsGroup& G : e->groups) {
    int rc = ens_batch_size<R>(e, G);
    if (rc) return rc;
    const int B = (int)G.snaps.size();
    for (int b0 = 0; b0 < B; b0 += G.ng)
      if ((rc = ens_batch<R>(e, G, b0, std::min(G.ng, B - b0)))) return rc;
  }
  return FIT_OK;
}

int ens_run_any(fit_ensemble* e) { return e->dtype == FIT_F64 ? ens_run<double>(e) : ens_run<float>(e); }

void ens_drop_graph(fit_ensemble* e) {
  if (e->exec) hipGraphExecDestroy(e->exec);
  if (e->graph) hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->graph_io = -1;
}

}  // namespace

extern "C" {

const char* fit_last_error(void) { return g_last_error.c_str(); }

const char* fit_version(void) { return "fit_hip 0.1 (gfx950)"; }

int fit_plan_create(fit_plan** out, int nd, const int64_t* ns, int M, int D, int compute_dtype, int dealias,
                    const double* const* ks, const int64_t* shell_idx, int n_shells, int device) {
  std::lock_guard<std::mutex> lock(g_create_mutex);
  g_last_error.clear();
  return create_plan(out, nd, ns, M, D, compute_dtype, dealias, ks, shell_idx, n_shells, device, true);
}

int fit_plan_destroy(fit_plan* p) {
  if (!p) return FIT_OK;
  hipSetDevice(p->device);
  if (p->stream) hipStreamSynchronize(p->stream);
  if (p->sub) { fit_plan_destroy(p->sub); p->sub = nullptr; }
  dfree(p->sub_f); dfree(p->sub_a);
  for (int d = 0; d < 3; ++d) { dfree(p->tw[d]); dfree(p->ks[d]); }
  dfree(p->kt[0]); dfree(p->kt[1]);
  void* sh = p->shell; dfree(sh);
  dfree(p->work); dfree(p->uphys); dfree(p->nphys); dfree(p->nhat);
  dfree(p->in_f); dfree(p->in_a); dfree(p->in_c); dfree(p->out_tmp);
  void* pa = p->partials; dfree(pa);
  { void* q = p->d_shell_band; dfree(q); }
  { void* q = p->shh; dfree(q); }
  dfree(p->eh); dfree(p->work2); dfree(p->nxbuf);
  { void* q = p->partials2; dfree(q); }
  { void* q = p->dcolG; dfree(q); }
  void* dc = p->dcol; dfree(dc);
  for (auto& r : p->recs) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
  if (p->stream && p->own_stream) hipStreamDestroy(p->stream);
  delete p;
  return FIT_OK;
}


int fit_plan_subgrid(const fit_plan* p, int64_t* dims) {
  if (!p) { fail(FIT_E_ARGUMENT, "null plan"); return -1; }
  if (!p->sub) return 0;
  if (dims)
    for (int d = 0; d < 3; ++d) dims[d] = p->sub->n[d];
  return 1;
}

int64_t fit_plan_device_bytes(const fit_plan* p) { return p ? p->bytes + (p->sub ? p->sub->bytes : 0) : 0; }

int fit_plan_set_option(fit_plan* p, int option, int value) {
  if (!p) return fail(FIT_E_ARGUMENT, "null plan");
  // pipeline-shape options apply to the sub-grid child as well
  if (p->sub && option != FIT_OPT_PLAIN_PIPELINE && option != FIT_OPT_SUBGRID) {
    int rc = fit_plan_set_option(p->sub, option, value);
    if (rc) return rc;
  }
  switch (option) {
    case FIT_OPT_PLAIN_PIPELINE: p->plain = value != 0; return FIT_OK;
    case FIT_OPT_GIVER_BATCH:
      if (value < 0) return fail(FIT_E_ARGUMENT, "giver batch must be >= 0");
      p->giver_batch = value;
      return FIT_OK;
    case FIT_OPT_XPASS_GIVERS:
      if (value < 0) return fail(FIT_E_ARGUMENT, "givers per workgroup must be >= 0");
      p->xfused_gc = value;
      return FIT_OK;
    case FIT_OPT_XPASS_BLOCK: p->xfused_block = value != 0; return FIT_OK;
    case FIT_OPT_SUBGRID: p->use_sub = value != 0; return FIT_OK;
    case FIT_OPT_COL_TILE:
      if (value < 0 || value > 64 || (value & (value - 1)) != 0)
        return fail(FIT_E_ARGUMENT, "column tile must be 0 or a power of two <= 64");
      p->col_tx = value;
      return FIT_OK;
    case FIT_OPT_XPASS_SELF: p->xself = value != 0; return FIT_OK;
    case FIT_OPT_XPASS_STAGE: p->xstage = value != 0; return FIT_OK;
    case FIT_OPT_GRAD_SPLIT: p->grad_split = value < 0 ? -1 : (value != 0); return FIT_OK;
    default: return fail(FIT_E_ARGUMENT, "unknown plan option");
  }
}

void* fit_plan_stream(const fit_plan* p) { return p ? (void*)p->stream : nullptr; }

int fit_profile_enable(fit_plan* p, int on) {
  if (!p) return fail(FIT_E_ARGUMENT, "null plan");
  p->prof = on != 0;
  p->nrec = 0;
  for (int i = 0; i < FIT_K_COUNT; ++i) { p->launches[i] = 0; p->ms[i] = 0; p->abytes[i] = 0; }
  if (p->sub) return fit_profile_enable(p->sub, on);
  return FIT_OK;
}

int fit_profile_query(const fit_plan* p, int cls, int64_t* launches, double* total_ms, double* alg_bytes) {
  if (!p || cls < 0 || cls >= FIT_K_COUNT) return fail(FIT_E_ARGUMENT, "bad plan / kernel class");
  if (launches) *launches = p->launches[cls];
  if (total_ms) *total_ms = p->ms[cls];
  if (alg_bytes) *alg_bytes = p->abytes[cls];
  return FIT_OK;
}

int fit_nonlinear_term(fit_plan* p, const void* f_hat, const void* adv_hat, void* N_hat, int io_dtype,
                       int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!N_hat) return fail(FIT_E_ARGUMENT, "N_hat is NULL");
  const void *f, *a;
  if ((rc = stage_fa(p, f_hat, adv_hat, io_dtype, ptr_kind, &f, &a))) return rc;
  // a device output of the compute precision that does not overlap the inputs is written by the
  // last pass directly (no staging copy)
  const size_t cb = csize(p->dtype);
  auto disjoint = [](const void* x, size_t xb, const void* y, size_t yb) {
    const uintptr_t x0 = (uintptr_t)x, y0 = (uintptr_t)y;
    return x0 + xb <= y0 || y0 + yb <= x0;
  };
  const size_t nb = (size_t)p->M * p->Np * cb;
  const bool direct = ptr_kind == FIT_PTR_DEVICE && io_dtype == p->dtype &&
                      disjoint(N_hat, nb, f, (size_t)p->M * p->Np * cb) &&
                      disjoint(N_hat, nb, a, (size_t)p->D * p->Np * cb);
  p->nhat_dst = direct ? N_hat : nullptr;
  rc = p->dtype == FIT_F64 ? nonlinear_term_impl<double>(p, f, a) : nonlinear_term_impl<float>(p, f, a);
  p->nhat_dst = nullptr;
  if (rc) return rc;
  if (!direct && (rc = stage_out_real(p, p->nhat, N_hat, 2 * (int64_t)p->M * p->Np, io_dtype, ptr_kind))) return rc;
  return finish(p);
}

int fit_transfer_density(fit_plan* p, int inv, const void* c_hat, const void* N_hat, void* t, int io_dtype,
                         int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!c_hat || !N_hat || !t) return fail(FIT_E_ARGUMENT, "NULL array");
  if ((rc = check_invariant(p, inv, p->M))) return rc;
  const void *c, *n;
  if ((rc = stage_in(p, c_hat, p->M, io_dtype, ptr_kind, &p->in_f, &c))) return rc;
  if ((rc = stage_in(p, N_hat, p->M, io_dtype, ptr_kind, &p->in_c, &n))) return rc;
  void* tdev = p->nphys;  // Np <= Gp reals
  rc = p->dtype == FIT_F64 ? density_launch<double>(p, inv, c, n, tdev, false)
                           : density_launch<float>(p, inv, c, n, tdev, false);
  if (rc) return rc;
  if ((rc = stage_out_real(p, tdev, t, p->Np, io_dtype, ptr_kind))) return rc;
  return finish(p);
}

int fit_spectral_flux(fit_plan* p, int inv, const void* f_hat, const void* adv_hat, double* T, double* Pi,
                      int io_dtype, int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (p->nsh <= 0) return fail(FIT_E_ARGUMENT, "plan has no shells");
  if ((rc = check_invariant(p, inv, p->M))) return rc;
  const void *f, *a;
  if ((rc = stage_fa(p, f_hat, adv_hat, io_dtype, ptr_kind, &f, &a))) return rc;
  if (p->dtype == FIT_F64) {
    if ((rc = nonlinear_term_impl<double>(p, f, a))) return rc;
    if ((rc = density_launch<double>(p, inv, f, p->nhat, nullptr, true))) return rc;
  } else {
    if ((rc = nonlinear_term_impl<float>(p, f, a))) return rc;
    if ((rc = density_launch<float>(p, inv, f, p->nhat, nullptr, true))) return rc;
  }
  return spectral_flux_tail(p, T, Pi);
}

int fit_spectral_flux_with_nhat(fit_plan* p, int inv, const void* c_hat, const void* N_hat, double* T,
                                double* Pi, int io_dtype, int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (p->nsh <= 0) return fail(FIT_E_ARGUMENT, "plan has no shells");
  if ((rc = check_invariant(p, inv, p->M))) return rc;
  const void *c, *n;
  if ((rc = stage_in(p, c_hat, p->M, io_dtype, ptr_kind, &p->in_f, &c))) return rc;
  if ((rc = stage_in(p, N_hat, p->M, io_dtype, ptr_kind, &p->in_c, &n))) return rc;
  rc = p->dtype == FIT_F64 ? density_launch<double>(p, inv, c, n, nullptr, true)
                           : density_launch<float>(p, inv, c, n, nullptr, true);
  if (rc) return rc;
  return spectral_flux_tail(p, T, Pi);
}

int fit_spectral_fluxes_with_nhat(fit_plan* p, int n_inv, const int* invariants, const void* c_hat,
                                  const void* N_hat, double* T, double* Pi, int io_dtype, int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (p->nsh <= 0) return fail(FIT_E_ARGUMENT, "plan has no shells");
  if (!invariants || n_inv < 1 || n_inv > 4) return fail(FIT_E_ARGUMENT, "1 to 4 invariants per call");
  for (int q = 0; q < n_inv; ++q)
    if ((rc = check_invariant(p, invariants[q], p->M))) return rc;
  if ((size_t)n_inv * p->nsh * sizeof(double) > (size_t)160 * 1024)
    return fail(FIT_E_UNSUPPORTED, "n_inv * n_shells exceeds the shell bins one workgroup holds");
  const void *c, *n;
  if ((rc = stage_in(p, c_hat, p->M, io_dtype, ptr_kind, &p->in_f, &c))) return rc;
  if ((rc = stage_in(p, N_hat, p->M, io_dtype, ptr_kind, &p->in_c, &n))) return rc;
  rc = p->dtype == FIT_F64 ? density_multi<double>(p, n_inv, invariants, c, n)
                           : density_multi<float>(p, n_inv, invariants, c, n);
  if (rc) return rc;
  const int nb = n_inv * p->nsh;
  std::vector<double> h((size_t)nb);
  FIT_HIP(hipMemcpyAsync(h.data(), p->dcolG, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  if ((rc = finish(p))) return rc;
  for (int q = 0; q < n_inv; ++q) {
    double acc = 0.0;
    for (int s = 0; s < p->nsh; ++s) {
      const double v = h[(size_t)q * p->nsh + s];
      acc += v;  // Π = +cumsum(T) (SpectralFlux.jl:172-177)
      if (T) T[(size_t)q * p->nsh + s] = v;
      if (Pi) Pi[(size_t)q * p->nsh + s] = acc;
    }
  }
  return FIT_OK;
}

int fit_shell_to_shell(fit_plan* p, int inv, const void* f_hat, const void* adv_hat, double* Tmat, double* net,
                       double* max_asym, int verify, int io_dtype, int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (p->nsh <= 0) return fail(FIT_E_ARGUMENT, "plan has no shells");
  if ((rc = check_invariant(p, inv, p->M))) return rc;
  const void *f, *a;
  if ((rc = stage_fa(p, f_hat, adv_hat, io_dtype, ptr_kind, &f, &a))) return rc;
  const int nsh = p->nsh;
  std::vector<double> Tm((size_t)nsh * nsh, 0.0);
  rc = p->dtype == FIT_F64 ? s2s_entry<double>(p, inv, f, a, Tm) : s2s_entry<float>(p, inv, f, a, Tm);
  if (rc) return rc;
  if ((rc = finish(p))) return rc;
  if (Tmat) std::memcpy(Tmat, Tm.data(), Tm.size() * sizeof(double));
  if (net) {
    for (int n = 0; n < nsh; ++n) {
      double s = 0.0;
      for (int m = 0; m < nsh; ++m) s += Tm[n + (size_t)nsh * m];
      net[n] = s;
    }
  }
  if (max_asym) {
    if (verify) {
      double v = 0.0;
      for (int n = 0; n < nsh; ++n)
        for (int m = 0; m < nsh; ++m) v = std::max(v, std::fabs(Tm[n + (size_t)nsh * m] + Tm[m + (size_t)nsh * n]));
      *max_asym = v;
    } else {
      *max_asym = std::nan("");
    }
  }
  return FIT_OK;
}

int fit_mode_to_mode(fit_plan* p, int inv, const void* f_hat, const void* adv_hat, void* S, void* net, int io_dtype,
                     int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!S || !net) return fail(FIT_E_ARGUMENT, "NULL output");
  if (p->dealias == FIT_DEALIAS_PADDED)
    return fail(FIT_E_UNSUPPORTED, "mode-to-mode supports NoDealiasing and OrszagTwoThirds");
  if ((rc = check_invariant(p, inv, p->M))) return rc;
  if (p->M > 3 && (inv == FIT_HELICITY || inv == FIT_ENSTROPHY)) return fail(FIT_E_DIMENSION, "M > 3");
  const void *f, *a;
  if ((rc = stage_fa(p, f_hat, adv_hat, io_dtype, ptr_kind, &f, &a))) return rc;
  rc = p->dtype == FIT_F64 ? m2m_impl<double>(p, inv, f, a, S, net, io_dtype, ptr_kind)
                           : m2m_impl<float>(p, inv, f, a, S, net, io_dtype, ptr_kind);
  if (rc) return rc;
  return finish(p);
}

int fit_helical_decompose(fit_plan* p, const void* u_hat, void* u_plus, void* u_minus, int io_dtype, int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (p->nd != 3) return fail(FIT_E_ARGUMENT, "HelicalDecomposition is defined in 3D only (got nd=" + std::to_string(p->nd) + ").");
  if (p->M < 3) return fail(FIT_E_ARGUMENT, "HelicalDecomposition needs >=3 velocity components");
  if (!u_hat || !u_plus || !u_minus) return fail(FIT_E_ARGUMENT, "NULL array");
  const void* u;
  if ((rc = stage_in(p, u_hat, p->M, io_dtype, ptr_kind, &p->in_f, &u))) return rc;
  const int64_t Np = p->Np;
  const size_t cs = 2 * p->rsz;
  // outputs into nhat (3 comps) and a second buffer (in_c)
  if (!p->in_c) { if ((rc = dalloc(p, &p->in_c, (size_t)std::max(p->M, p->D) * Np * cs))) return rc; }
  if (p->dtype == FIT_F64) {
    Grid g = grid_of<double>(p);
    hipLaunchKernelGGL(k_helical<double>, dim3(1024), dim3(256), 0, p->stream, g, (const double2*)u, Np,
                       (const double*)p->ks[0], (const double*)p->ks[1], (const double*)p->ks[2], (double2*)p->nhat,
                       (double2*)p->in_c, Np);
  } else {
    Grid g = grid_of<float>(p);
    hipLaunchKernelGGL(k_helical<float>, dim3(1024), dim3(256), 0, p->stream, g, (const float2*)u, Np,
                       (const float*)p->ks[0], (const float*)p->ks[1], (const float*)p->ks[2], (float2*)p->nhat,
                       (float2*)p->in_c, Np);
  }
  FIT_HIP(hipGetLastError());
  // components beyond 3 are zero in both outputs
  if (p->M > 3) {
    FIT_HIP(hipMemsetAsync((char*)p->nhat + 3 * Np * cs, 0, (size_t)(p->M - 3) * Np * cs, p->stream));
    FIT_HIP(hipMemsetAsync((char*)p->in_c + 3 * Np * cs, 0, (size_t)(p->M - 3) * Np * cs, p->stream));
  }
  if ((rc = stage_out_real(p, p->nhat, u_plus, 2 * (int64_t)p->M * Np, io_dtype, ptr_kind))) return rc;
  if ((rc = finish(p))) return rc;
  if ((rc = stage_out_real(p, p->in_c, u_minus, 2 * (int64_t)p->M * Np, io_dtype, ptr_kind))) return rc;
  return finish(p);
}



int fit_band_to_band(fit_plan* p, int inv, const double* centers, int n_bands, double logwidth,
                     const double* k_coord, const void* f_hat, const void* adv_hat, double* Tmat, double* net,
                     double* max_asym, int io_dtype, int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if ((rc = check_invariant(p, inv, p->M))) return rc;
  if (!centers || !k_coord) return fail(FIT_E_ARGUMENT, "centers / k_coord is NULL");
  if (n_bands < 1 || n_bands > 4096) return fail(FIT_E_ARGUMENT, "SmoothBands needs 1..4096 centers");
  if (!(logwidth > 0.0)) return fail(FIT_E_ARGUMENT, "SmoothBands: logwidth must be positive");
  for (int n = 0; n < n_bands; ++n)
    if (!(centers[n] > 0.0)) return fail(FIT_E_ARGUMENT, "SmoothBands: centers must be positive");
  const void *f, *a;
  if ((rc = stage_fa(p, f_hat, adv_hat, io_dtype, ptr_kind, &f, &a))) return rc;
  void *kc = nullptr, *cen = nullptr;
  if ((rc = dalloc(p, &kc, (size_t)p->Np * p->rsz))) return rc;
  if ((rc = dalloc(p, &cen, (size_t)n_bands * sizeof(double)))) { dfree(kc); return rc; }
  auto release = [&]() {
    hipStreamSynchronize(p->stream);
    p->bytes -= (int64_t)p->Np * p->rsz + (int64_t)n_bands * sizeof(double);
    dfree(kc);
    dfree(cen);
  };
  hipError_t he = hipMemcpy(cen, centers, (size_t)n_bands * sizeof(double), hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    if (p->dtype == FIT_F64) {
      he = hipMemcpy(kc, k_coord, (size_t)p->Np * sizeof(double), hipMemcpyHostToDevice);
    } else {
      std::vector<float> k32((size_t)p->Np);
      for (int64_t i = 0; i < p->Np; ++i) k32[i] = (float)k_coord[i];
      he = hipMemcpy(kc, k32.data(), (size_t)p->Np * sizeof(float), hipMemcpyHostToDevice);
    }
  }
  if (he != hipSuccess) { release(); return fail(FIT_E_DEVICE, std::string("band upload: ") + hipGetErrorString(he)); }
  std::vector<double> Tm((size_t)n_bands * n_bands, 0.0);
  const double inv2s2 = 1.0 / (2.0 * logwidth * logwidth);
  rc = p->dtype == FIT_F64
           ? b2b_impl<double>(p, inv, f, a, (const double*)kc, (const double*)cen, n_bands, inv2s2, Tm)
           : b2b_impl<float>(p, inv, f, a, (const float*)kc, (const double*)cen, n_bands, inv2s2, Tm);
  release();
  if (rc) return rc;
  if ((rc = finish(p))) return rc;
  const int nb = n_bands;
  double asym = 0.0;
  for (int n = 0; n < nb; ++n) {
    double acc = 0.0;
    for (int m = 0; m < nb; ++m) {
      const double v = Tm[n + (size_t)nb * m];   // T(n, m): receiver band n, giver band m
      if (Tmat) Tmat[n + (size_t)nb * m] = v;
      acc += v;
      asym = std::max(asym, std::fabs(v + Tm[m + (size_t)nb * n]));
    }
    if (net) net[n] = acc;
  }
  if (max_asym) *max_asym = asym;
  return FIT_OK;
}

// ---------------------------------------------------------------------------------------------
// ensembles
// ---------------------------------------------------------------------------------------------

int fit_ensemble_create(fit_ensemble** out, int n_snap, int nd, const int64_t* ns, int compute_dtype,
                        int dealias, const double* const* ks, const int64_t* const* shell_idx,
                        const int* n_shells, int with_scalar, int device) {
  std::lock_guard<std::mutex> lock(g_create_mutex);
  g_last_error.clear();
  if (!out) return fail(FIT_E_ARGUMENT, "ensemble out-pointer is NULL");
  *out = nullptr;
  if (n_snap < 1) return fail(FIT_E_ARGUMENT, "an ensemble needs at least one snapshot");
  if (nd != 2 && nd != 3) return fail(FIT_E_UNSUPPORTED, "ensembles run 2D or 3D snapshots");
  if (!ns || !ks || !shell_idx || !n_shells) return fail(FIT_E_ARGUMENT, "ns / ks / shell_idx / n_shells is NULL");
  if (dealias == FIT_DEALIAS_PADDED)
    return fail(FIT_E_UNSUPPORTED, "ensembles run NoDealiasing or OrszagTwoThirds (use a plan per snapshot for padding)");
  fit_ensemble* e = new fit_ensemble();
  e->n = n_snap; e->nd = nd; e->dtype = compute_dtype; e->dealias = dealias; e->device = device;
  e->with_scalar = with_scalar != 0;
  auto bail = [&](int code) {
    std::string msg = g_last_error;
    fit_ensemble_destroy(e);
    g_last_error = msg;
    return code;
  };
  hipError_t he = hipSetDevice(device);
  if (he != hipSuccess) { delete e; return fail(FIT_E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(he)); }
  he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he != hipSuccess) { delete e; return fail(FIT_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(he)); }
  e->nsh.resize(n_snap);
  e->off.resize(n_snap);
  int64_t acc = 0;
  for (int j = 0; j < n_snap; ++j) {
    if (n_shells[j] < 1 || !shell_idx[j]) return bail(fail(FIT_E_ARGUMENT, "every snapshot needs a shell map"));
    e->nsh[j] = n_shells[j];
    e->off[j] = acc;
    acc += n_shells[j];
  }
  // group snapshots with the same grid: sizes, wavenumbers and shell map all equal
  std::vector<int> rep;  // representative snapshot of each group
  for (int j = 0; j < n_snap; ++j) {
    int64_t np = 1;
    for (int d = 0; d < nd; ++d) np *= ns[(size_t)j * nd + d];
    int gi = -1;
    for (size_t g = 0; g < rep.size() && gi < 0; ++g) {
      const int r = rep[g];
      bool same = n_shells[r] == n_shells[j];
      for (int d = 0; d < nd && same; ++d) same = ns[(size_t)r * nd + d] == ns[(size_t)j * nd + d];
      for (int d = 0; d < nd && same; ++d)
        same = std::memcmp(ks[(size_t)r * nd + d], ks[(size_t)j * nd + d], ns[(size_t)j * nd + d] * sizeof(double)) == 0;
      if (same) same = std::memcmp(shell_idx[r], shell_idx[j], (size_t)np * sizeof(int64_t)) == 0;
      if (same) gi = (int)g;
    }
    if (gi < 0) {
      gi = (int)rep.size();
      rep.push_back(j);
      e->groups.emplace_back();
    }
    e->groups[gi].snaps.push_back(j);
  }
  for (size_t g = 0; g < e->groups.size(); ++g) {
    EnsGroup& G = e->groups[g];
    const int r = rep[g];
    int rc = create_plan(&G.p, nd, ns + (size_t)r * nd, nd, nd, compute_dtype, dealias, ks + (size_t)r * nd,
                         shell_idx[r], n_shells[r], device, false);
    if (rc) return bail(rc);
    if (!G.p->shell_symmetric) return bail(fail(FIT_E_UNSUPPORTED, "ensemble shells must be symmetric under k -> -k"));
    hipStreamDestroy(G.p->stream);
    G.p->stream = e->stream;
    G.p->own_stream = false;
    const int B = (int)G.snaps.size();
    if ((rc = dalloc(G.p, (void**)&G.d_tab, (size_t)2 * B * sizeof(void*)))) return bail(rc);
    if ((rc = dalloc(G.p, (void**)&G.rows, (size_t)2 * B * G.p->nsh * sizeof(double)))) return bail(rc);
    G.h_tab.assign((size_t)2 * B, nullptr);
    G.h_rows.assign((size_t)2 * B * G.p->nsh, 0.0);
  }
  *out = e;
  return FIT_OK;
}

int fit_ensemble_destroy(fit_ensemble* e) {
  if (!e) return FIT_OK;
  hipSetDevice(e->device);
  if (e->stream) hipStreamSynchronize(e->stream);
  ens_drop_graph(e);
  for (EnsGroup& G : e->groups) {
    if (!G.p) continue;
    dfree(G.ub); dfree(G.ehk); dfree(G.ehs); dfree(G.stage);
    { void* q = G.d_tab; dfree(q); }
    { void* q = G.rows; dfree(q); }
    fit_plan_destroy(G.p);
  }
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return FIT_OK;
}

int fit_ensemble_set_option(fit_ensemble* e, int option, int value) {
  if (!e) return fail(FIT_E_ARGUMENT, "null ensemble");
  switch (option) {
    case FIT_ENS_OPT_GRAPH:
      e->use_graph = value != 0;
      if (!e->use_graph) ens_drop_graph(e);
      return FIT_OK;
    case FIT_ENS_OPT_BATCH:
      if (value < 0) return fail(FIT_E_ARGUMENT, "batch cap must be >= 0");
      e->batch_cap = value;
      for (EnsGroup& G : e->groups) G.ng = 0;
      ens_drop_graph(e);
      return FIT_OK;
    default:
      return fail(FIT_E_ARGUMENT, "unknown ensemble option");
  }
}

int fit_ensemble_groups(const fit_ensemble* e) { return e ? (int)e->groups.size() : -1; }

int fit_ensemble_flux(fit_ensemble* e, const void* const* u_hat, const void* const* theta_hat, double* T_ke,
                      double* Pi_ke, double* T_sc, double* Pi_sc, int io_dtype, int ptr_kind) {
  if (!e) return fail(FIT_E_ARGUMENT, "null ensemble");
  FIT_HIP(hipSetDevice(e->device));
  if (!u_hat) return fail(FIT_E_ARGUMENT, "u_hat is NULL");
  if (e->with_scalar && !theta_hat) return fail(FIT_E_ARGUMENT, "this ensemble was created with a scalar: theta_hat is NULL");
  if (io_dtype != FIT_F32 && io_dtype != FIT_F64) return fail(FIT_E_ARGUMENT, "io_dtype must be FIT_F32 or FIT_F64");
  if (ptr_kind != FIT_PTR_HOST && ptr_kind != FIT_PTR_DEVICE) return fail(FIT_E_ARGUMENT, "bad ptr_kind");
  for (int j = 0; j < e->n; ++j)
    if (!u_hat[j] || (e->with_scalar && !theta_hat[j])) return fail(FIT_E_ARGUMENT, "NULL snapshot array");
  const size_t cb = csize(e->dtype), ib = csize(io_dtype);
  const bool direct = ptr_kind == FIT_PTR_DEVICE && io_dtype == e->dtype;
  // pointer tables (staging host / other-precision inputs into the group's stage buffer)
  for (EnsGroup& G : e->groups) {
    fit_plan* p = G.p;
    const int B = (int)G.snaps.size();
    const int nc = e->nd + (e->with_scalar ? 1 : 0);
    if (!direct) {
      int rc = grow(p, &G.stage, &G.stage_bytes, (size_t)B * nc * p->Np * cb);
      if (rc) return rc;
    }
    void* tmp = nullptr;
    for (int b = 0; b < B; ++b) {
      const int j = G.snaps[b];
      for (int part = 0; part < (e->with_scalar ? 2 : 1); ++part) {
        const void* src = part == 0 ? u_hat[j] : theta_hat[j];
        const int ncomp = part == 0 ? e->nd : 1;
        if (direct) {
          G.h_tab[(size_t)part * B + b] = const_cast<void*>(src);
          continue;
        }
        char* dst = (char*)G.stage + ((size_t)b * nc + (part == 0 ? 0 : e->nd)) * p->Np * cb;
        G.h_tab[(size_t)part * B + b] = dst;
        const int64_t nvals = (int64_t)ncomp * p->Np;
        if (io_dtype == e->dtype) {
          FIT_HIP(hipMemcpyAsync(dst, src, nvals * cb, hipMemcpyHostToDevice, e->stream));
          continue;
        }
        const void* dsrc = src;
        if (ptr_kind == FIT_PTR_HOST) {
          if (!tmp) FIT_HIP(hipMalloc(&tmp, (size_t)nc * p->Np * ib));
          FIT_HIP(hipMemcpyAsync(tmp, src, nvals * ib, hipMemcpyHostToDevice, e->stream));
          dsrc = tmp;
        }
        if (io_dtype == FIT_F64)
          hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, e->stream, (const double*)dsrc,
                             (float*)dst, 2 * nvals);
        else
          hipLaunchKernelGGL((k_convert<float, double>), dim3(1024), dim3(256), 0, e->stream, (const float*)dsrc,
                             (double*)dst, 2 * nvals);
        FIT_HIP(hipGetLastError());
        if (tmp) FIT_HIP(hipStreamSynchronize(e->stream));
      }
    }
    if (tmp) hipFree(tmp);
    FIT_HIP(hipMemcpyAsync(G.d_tab, G.h_tab.data(), G.h_tab.size() * sizeof(void*), hipMemcpyHostToDevice, e->stream));
  }
  const int io_key = ptr_kind * 2 + io_dtype;
  if (e->exec && e->graph_io == io_key && !e->prof) {
    FIT_HIP(hipGraphLaunch(e->exec, e->stream));
  } else {
    int rc = ens_run_any(e);  // allocates every workspace the graph will reuse
    if (rc) return rc;
    if (e->use_graph && !e->prof) {
      ens_drop_graph(e);
      FIT_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      rc = ens_run_any(e);
      hipGraph_t g = nullptr;
      hipError_t ce = hipStreamEndCapture(e->stream, &g);
      if (rc) { if (g) hipGraphDestroy(g); return rc; }
      FIT_HIP(ce);
      e->graph = g;
      FIT_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
      e->graph_io = io_key;
    }
  }
  for (EnsGroup& G : e->groups)
    FIT_HIP(hipMemcpyAsync(G.h_rows.data(), G.rows, G.h_rows.size() * sizeof(double), hipMemcpyDeviceToHost,
                           e->stream));
  FIT_HIP(hipStreamSynchronize(e->stream));
  for (EnsGroup& G : e->groups) {
    prof_collect(G.p);
    const int B = (int)G.snaps.size(), nsh = G.p->nsh;
    for (int b = 0; b < B; ++b) {
      const int j = G.snaps[b];
      for (int part = 0; part < (e->with_scalar ? 2 : 1); ++part) {
        double* T = part == 0 ? T_ke : T_sc;
        double* Pi = part == 0 ? Pi_ke : Pi_sc;
        const double* h = G.h_rows.data() + ((size_t)part * B + b) * nsh;
        double a = 0.0;
        for (int s = 0; s < nsh; ++s) {
          a += h[s];  // Π = +cumsum(T) (SpectralFlux.jl:172-177)
          if (T) T[e->off[j] + s] = h[s];
          if (Pi) Pi[e->off[j] + s] = a;
        }
      }
    }
  }
  return FIT_OK;
}

int fit_ensemble_profile_enable(fit_ensemble* e, int on) {
  if (!e) return fail(FIT_E_ARGUMENT, "null ensemble");
  e->prof = on != 0;
  for (EnsGroup& G : e->groups) {
    int rc = fit_profile_enable(G.p, on);
    if (rc) return rc;
  }
  return FIT_OK;
}

int fit_ensemble_profile_query(const fit_ensemble* e, int cls, int64_t* launches, double* total_ms,
                               double* alg_bytes) {
  if (!e) return fail(FIT_E_ARGUMENT, "null ensemble");
  if (cls < 0 || cls >= FIT_K_COUNT) return fail(FIT_E_ARGUMENT, "bad kernel class");
  int64_t l = 0;
  double ms = 0, by = 0;
  for (const EnsGroup& G : e->groups) {
    l += G.p->launches[cls];
    ms += G.p->ms[cls];
    by += G.p->abytes[cls];
  }
  if (launches) *launches = l;
  if (total_ms) *total_ms = ms;
  if (alg_bytes) *alg_bytes = by;
  return FIT_OK;
}

void* fit_ensemble_stream(const fit_ensemble* e) { return e ? (void*)e->stream : nullptr; }

}  // extern "C"
