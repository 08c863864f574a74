// fit_hip.hip -- plan management, staging, profiling and the C ABI of libfit_hip.so.
//
// One plan = one grid on one GPU: device-resident twiddles, wavenumbers, shell map and
// workspaces, one HIP stream.  Every call runs the whole hot path on the device and returns
// FP64 shell sums (or device/host arrays) to the caller.  The pipeline itself (pass launchers
// and their kernels) is compiled per precision in fit_f32.hip / fit_f64.hip.
#include "fit_internal.h"

#include <chrono>
#include <cstdlib>

// Device scratch of fit_spectral_flux_with_nhat_noalloc: the shell partial rows and their fixed-
// order sums.  One copy per GPU, held until the process ends; every call on a GPU shares it.  In
// host code the array's name is only a stand-in: kernels get the per-device address found by
// hipGetSymbolAddress (scratch_for below).
constexpr size_t kScratchDoubles = FIT_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace fitx {
FIT_PIPELINE_EXTERN(float)
FIT_PIPELINE_EXTERN(double)

thread_local std::string g_last_error;
std::mutex g_create_mutex;
int64_t g_live_plans = 0;   // plans made by fit_plan_create and not yet destroyed (under g_create_mutex)

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// Radix plan of the LDS-engine FFT (any order is valid Stockham).  order 0: as many 8s as
// possible, then 4s, 2s, 3s, 5s; 1: 4s first, then 2s, 8s, 3s, 5s; 2: the odd radices first.
bool make_desc(int L, FftDesc& d, int order) {
  d.L = L;
  d.nst = 0;
  int r = L;
  static const int orders[3][5] = {{8, 4, 2, 3, 5}, {4, 2, 8, 3, 5}, {3, 5, 8, 4, 2}};
  const int* o = orders[(order >= 0 && order < 3) ? order : 0];
  for (int i = 0; i < 5; ++i)
    while (r % o[i] == 0 && d.nst < FIT_MAX_STAGES) { d.rad[d.nst++] = o[i]; r /= o[i]; }
  return r == 1;
}

int padded_len(int n) {  // FFTWExt.jl:128-131
  int m = (3 * n + 1) / 2;
  return ((m - n) % 2 == 0) ? m : m + 1;
}


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
void prof_begin(fit_plan* p, int cls, double bytes, const char* kname, int L, const char* extra) {
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
  r.name = std::string(kname) + "<" + (p->rsz == 8 ? "f64" : "f32") + (L > 0 ? "," + std::to_string(L) : "") +
           extra + ">";
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
  std::vector<fit_plan*> all{p};
  for (const SubGrid& sg : p->subs) all.push_back(sg.c);
  for (fit_plan* s : all) {
    if (!s || !s->prof) continue;
    for (size_t i = 0; i < s->nrec; ++i) {
      float t = 0.f;
      hipEventElapsedTime(&t, s->recs[i].a, s->recs[i].b);
      p->launches[s->recs[i].cls] += 1;
      p->ms[s->recs[i].cls] += t;
      p->abytes[s->recs[i].cls] += s->recs[i].bytes;
      KernAgg* k = nullptr;
      for (KernAgg& a : p->kagg)
        if (a.name == s->recs[i].name) { k = &a; break; }
      if (!k) {
        p->kagg.push_back(KernAgg{s->recs[i].name, s->recs[i].cls});
        k = &p->kagg.back();
      }
      k->launches += 1;
      k->ms += t;
      k->bytes += s->recs[i].bytes;
    }
    s->nrec = 0;
  }
}


size_t g_max_lds = 65536;

// Fixed-order reduction of ng groups of nblk partial rows of nsh bins (k_reduce_rows).
void launch_reduce(fit_plan* p, const double* partials, int nblk, int nsh, int ng, double* out) {
  if (nsh <= 0 || ng <= 0) return;
  hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)((nsh + kRedBins - 1) / kRedBins), (unsigned)ng),
                     dim3(kRedBins * kRedRows), 0, p->stream, partials, nblk, nsh, out);
}



int grow(fit_plan* p, void** buf, size_t* have, size_t need) {
  if (*have >= need) return FIT_OK;
  if (*buf) { (void)hipStreamSynchronize(p->stream); (void)hipFree(*buf); p->bytes -= (int64_t)*have; *buf = nullptr; *have = 0; }
  int rc = dalloc(p, buf, need);
  if (rc) return rc;
  *have = need;
  return FIT_OK;
}

// ------------------------------ input / output staging ----------------------------------------


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



int finish(fit_plan* p) {
  FIT_HIP(hipStreamSynchronize(p->stream));
  prof_collect(p);
  return FIT_OK;
}

// T and Π = +cumsum(T) (SpectralFlux.jl:172-177) of n_inv invariants from the shell-sum rows in
// dcolG (invariant q at q*nsh), into host arrays at the same offsets
int rows_tail(fit_plan* p, int n_inv, double* T, double* Pi) {
  const int nb = n_inv * p->nsh;
  std::vector<double> h((size_t)nb);
  FIT_HIP(hipMemcpyAsync(h.data(), p->dcolG, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  int rc = finish(p);
  if (rc) return rc;
  for (int q = 0; q < n_inv; ++q) {
    double acc = 0.0;
    for (int s = 0; s < p->nsh; ++s) {
      const double v = h[(size_t)q * p->nsh + s];
      acc += v;
      if (T) T[(size_t)q * p->nsh + s] = v;
      if (Pi) Pi[(size_t)q * p->nsh + s] = acc;
    }
  }
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

// Child plan of the band-limited giver loop on a grid of Ls points per axis: no dealiasing, the
// API grid's wavenumbers at the same folded indices, the shell map restricted to the box
// |k_d| <= R_d, input band pruned to the copied box |q_d| <= Q_d.
static int make_subgrid_child(fit_plan* p, const double* const* ks, const std::vector<int>& sh, const int* R,
                              const int* Q, const int* Ls, fit_plan** out) {
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
  *out = c;
  return FIT_OK;
}

// Band-limited sub-grids of the giver loop.  With receivers in |k_d| <= R_d and a giver's modes
// in |p_d| <= P_d, every triad k = p + q it feeds has |q_d| <= R_d + P_d; copying the advecting
// field's box |q_d| <= Q_d = R_d + P_d onto a grid of L_d >= Q_d + P_d + R_d + 1 points keeps the
// products alias-free at the receivers, so the giver's column of T(n,m) equals the full-grid one.
// Consecutive givers with the same grid form a class (adjacent classes merge when that costs less);
// a sub-grid is used when the largest class's grid is at most half the transform grid (4 R_d < n_d).
static int setup_subgrid(fit_plan* p, const double* const* ks, const std::vector<int>& sh) {
  if (p->nd < 2 || p->nsh <= 0 || !p->shell_symmetric) return FIT_OK;
  int R[3] = {0, 0, 0};
  for (int d = 0; d < p->nd; ++d) {
    R[d] = std::max(p->band_rx[d], 0);
    if (4 * R[d] >= p->n[d]) return FIT_OK;
  }
  const bool trunc = p->dealias == FIT_DEALIAS_TWOTHIRDS;
  auto grid_for = [&](const int* P, int* Q, int* Ls) {
    for (int d = 0; d < 3; ++d) {
      if (d >= p->nd) { Q[d] = 0; Ls[d] = 1; continue; }
      Q[d] = R[d] + std::max(P[d], 0);
      if (trunc) Q[d] = std::min(Q[d], p->band_in[d]);
      Ls[d] = std::min(subgrid_len(Q[d] + std::max(P[d], 0) + R[d] + 1), p->n[d]);
    }
  };
  {
    int Q[3], Ls[3];
    grid_for(R, Q, Ls);  // the largest giver support is the receivers' R
    double ratio = 1.0;
    for (int d = 0; d < p->nd; ++d) ratio *= (double)Ls[d] / (double)p->G[d];
    if (ratio > 0.5) return FIT_OK;
  }
  struct Cls { int m0, m1, Q[3], L[3]; };
  std::vector<Cls> cls;
  for (int m = 1; m <= p->nsh; ++m) {
    const int4 b = p->shell_band[m - 1];
    const int P[3] = {std::max(b.x, 0), std::max(b.y, 0), std::max(b.z, 0)};
    Cls c{m, m, {0, 0, 0}, {1, 1, 1}};
    grid_for(P, c.Q, c.L);
    if (!cls.empty() && cls.back().L[0] == c.L[0] && cls.back().L[1] == c.L[1] && cls.back().L[2] == c.L[2]) {
      Cls& k = cls.back();
      k.m1 = m;
      for (int d = 0; d < 3; ++d) k.Q[d] = std::max(k.Q[d], c.Q[d]);
    } else {
      cls.push_back(c);
    }
  }
  auto vol = [](const Cls& c) { return (double)c.L[0] * c.L[1] * c.L[2]; };
  // cost of a class: its givers plus about half a giver for the class's own velocity transforms
  // and input gathers, each at the class's grid volume; adjacent classes merge while that lowers
  // the total
  auto cost = [&](const Cls& c) { return ((double)(c.m1 - c.m0 + 1) + 0.5) * vol(c); };
  for (;;) {
    double best = 0.0;
    size_t bi = cls.size();
    for (size_t i = 0; i + 1 < cls.size(); ++i) {
      Cls m = cls[i + 1];
      m.m0 = cls[i].m0;
      for (int d = 0; d < 3; ++d) { m.L[d] = std::max(cls[i].L[d], m.L[d]); m.Q[d] = std::max(cls[i].Q[d], m.Q[d]); }
      const double saving = cost(cls[i]) + cost(cls[i + 1]) - cost(m);
      if (saving > best) { best = saving; bi = i; }
    }
    if (bi == cls.size()) break;
    Cls& b = cls[bi + 1];
    b.m0 = cls[bi].m0;
    for (int d = 0; d < 3; ++d) { b.L[d] = std::max(cls[bi].L[d], b.L[d]); b.Q[d] = std::max(cls[bi].Q[d], b.Q[d]); }
    cls.erase(cls.begin() + (long)bi);
  }
  size_t big = 0;
  for (const Cls& c : cls) {
    SubGrid sg;
    sg.m0 = c.m0;
    sg.m1 = c.m1;
    for (int d = 0; d < 3; ++d) sg.q[d] = c.Q[d];
    int rc = make_subgrid_child(p, ks, sh, R, c.Q, c.L, &sg.c);
    if (rc) return rc;
    sg.c->giver_lo = c.m0;
    sg.c->giver_hi = c.m1;
    const double r = (double)sg.c->Np / (double)p->Np;
    sg.tscale = r * r;
    p->subs.push_back(sg);
    if (vol(c) >= vol(cls[big])) big = p->subs.size() - 1;
  }
  p->sub = p->subs[big].c;
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
    if (!make_desc(p->G[d], p->fd[d], 0)) {
      int bad = p->G[d];
      delete p;
      return fail(FIT_E_UNSUPPORTED, "transform length " + std::to_string(bad) + " is not of the form 2^a 3^b 5^c");
    }
    // the LDS-engine passes hold two lines per workgroup (up to the 160 KB of LDS a gfx950
    // workgroup may take: FP64 lines up to 5120 points, FP32 up to 10240)
    if ((size_t)2 * p->G[d] * (compute_dtype == FIT_F64 ? 16 : 8) > (size_t)160 * 1024) {
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
      if ((rc = dalloc(p, &p->tw64[d], (size_t)L * 16))) return bail(rc);
      hipMemcpy(p->tw64[d], tw.data(), (size_t)L * 16, hipMemcpyHostToDevice);
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
    // shell map on the original half grid (x <= n1/2) for the pair density passes
    {
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
int ens_run_any(fit_ensemble* e) { return e->dtype == FIT_F64 ? ens_run<double>(e) : ens_run<float>(e); }

void ens_drop_graph(fit_ensemble* e) {
  if (e->exec) hipGraphExecDestroy(e->exec);
  if (e->graph) hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->graph_io = -1;
}

}  // namespace

// ---- allocation-free spectral flux tail --------------------------------------------------------
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_dev[kMaxDevices] = {nullptr};

// address of `scratch` on the current device (looked up once per device, with it current)
int scratch_for(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (device < 0 || device >= kMaxDevices) return fail(FIT_E_ARGUMENT, "device index out of range");
  if (!g_scratch_dev[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(FIT_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = ptr;
  }
  *out = (double*)g_scratch_dev[device];
  return FIT_OK;
}

// T (copy) and Π = +cumsum(T) (SpectralFlux.jl:172-177) into device arrays, in shell order
static __global__ void k_cumsum_shells(const double* __restrict__ sums, int nsh, double* __restrict__ T,
                                       double* __restrict__ Pi) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double acc = 0.0;
  for (int s = 0; s < nsh; ++s) {
    const double v = sums[s];
    acc += v;
    T[s] = v;
    Pi[s] = acc;
  }
}

}  // namespace fitx

using namespace fitx;

// ---- autotuning ----------------------------------------------------------------------------
// Choices of the plan knobs found by fit_plan_autotune, per (grid, components, precision,
// dealiasing, shells, workload); applied to every plan created afterwards with the same key
// (within the process, and across processes through the file FIT_TUNE_CACHE names).
namespace fitx {

struct TuneEntry {
  std::string key;
  std::vector<std::pair<int, int>> opts;
};
std::vector<TuneEntry> g_tune;
bool g_tune_file_read = false;

std::string tune_key(const fit_plan* p) {
  return std::to_string(p->nd) + ":" + std::to_string(p->n[0]) + "x" + std::to_string(p->n[1]) + "x" +
         std::to_string(p->n[2]) + ":M" + std::to_string(p->M) + ":D" + std::to_string(p->D) + ":" +
         (p->dtype == FIT_F64 ? "f64" : "f32") + ":dealias" + std::to_string(p->dealias) + ":sh" +
         std::to_string(p->nsh);
}

void tune_store(const std::string& key, const std::vector<std::pair<int, int>>& opts) {
  for (TuneEntry& e : g_tune)
    if (e.key == key) { e.opts = opts; return; }
  g_tune.push_back(TuneEntry{key, opts});
}

// lines "key|opt=value,opt=value" (the caller holds g_create_mutex)
void tune_read_file() {
  if (g_tune_file_read) return;
  g_tune_file_read = true;
  const char* path = std::getenv("FIT_TUNE_CACHE");
  if (!path || !*path) return;
  FILE* fp = std::fopen(path, "r");
  if (!fp) return;
  char line[1024];
  while (std::fgets(line, sizeof line, fp)) {
    std::string l(line);
    const size_t bar = l.find('|');
    if (bar == std::string::npos) continue;
    std::vector<std::pair<int, int>> opts;
    size_t i = bar + 1;
    while (i < l.size()) {
      int o = 0, v = 0, used = 0;
      if (std::sscanf(l.c_str() + i, "%d=%d%n", &o, &v, &used) != 2) break;
      opts.emplace_back(o, v);
      i += (size_t)used;
      if (i < l.size() && l[i] == ',') ++i;
    }
    tune_store(l.substr(0, bar), opts);
  }
  std::fclose(fp);
}

void tune_append_file(const std::string& key, const std::vector<std::pair<int, int>>& opts) {
  const char* path = std::getenv("FIT_TUNE_CACHE");
  if (!path || !*path) return;
  FILE* fp = std::fopen(path, "a");
  if (!fp) return;
  std::fprintf(fp, "%s|", key.c_str());
  for (size_t i = 0; i < opts.size(); ++i) std::fprintf(fp, "%s%d=%d", i ? "," : "", opts[i].first, opts[i].second);
  std::fprintf(fp, "\n");
  std::fclose(fp);
}

// apply the cached choice of this plan's key, if any (the caller holds g_create_mutex)
void tune_apply(fit_plan* p) {
  tune_read_file();
  const std::string key = tune_key(p);
  for (const TuneEntry& e : g_tune) {
    if (e.key.compare(0, key.size(), key) != 0) continue;
    for (const auto& ov : e.opts) (void)fit_plan_set_option(p, ov.first, ov.second);
  }
}
}  // namespace fitx

extern "C" {

const char* fit_last_error(void) { return g_last_error.c_str(); }

const char* fit_version(void) { return "fit_hip 0.1 (gfx950)"; }

int fit_plan_create(fit_plan** out, int nd, const int64_t* ns, int M, int D, int compute_dtype, int dealias,
                    const double* const* ks, const int64_t* shell_idx, int n_shells, int device) {
  std::lock_guard<std::mutex> lock(g_create_mutex);
  g_last_error.clear();
  int rc = create_plan(out, nd, ns, M, D, compute_dtype, dealias, ks, shell_idx, n_shells, device, true);
  if (rc == FIT_OK) {
    tune_apply(*out);
    (*out)->user_plan = true;
    ++g_live_plans;
  }
  return rc;
}

int64_t fit_plan_live_count(void) {
  std::lock_guard<std::mutex> lock(g_create_mutex);
  return g_live_plans;
}

int fit_plan_destroy(fit_plan* p) {
  if (!p) return FIT_OK;
  if (p->user_plan) {
    std::lock_guard<std::mutex> lock(g_create_mutex);
    --g_live_plans;
  }
  hipSetDevice(p->device);
  if (p->stream) hipStreamSynchronize(p->stream);
  for (SubGrid& sg : p->subs) fit_plan_destroy(sg.c);
  p->subs.clear();
  p->sub = nullptr;
  dfree(p->sub_f); dfree(p->sub_a);
  for (int d = 0; d < 3; ++d) { dfree(p->tw[d]); dfree(p->tw64[d]); dfree(p->ks[d]); }
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
  if (p->order_ev) hipEventDestroy(p->order_ev);
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

int fit_plan_subgrid_classes(const fit_plan* p, int64_t* dims, int* givers, int max_classes) {
  if (!p) { fail(FIT_E_ARGUMENT, "null plan"); return -1; }
  const int n = (int)p->subs.size();
  for (int i = 0; i < n && i < max_classes; ++i) {
    if (dims)
      for (int d = 0; d < 3; ++d) dims[3 * i + d] = p->subs[(size_t)i].c->n[d];
    if (givers) {
      givers[2 * i] = p->subs[(size_t)i].m0;
      givers[2 * i + 1] = p->subs[(size_t)i].m1;
    }
  }
  return n;
}

int64_t fit_plan_device_bytes(const fit_plan* p) {
  if (!p) return 0;
  int64_t b = p->bytes;
  for (const SubGrid& sg : p->subs) b += sg.c->bytes;
  return b;
}

int fit_plan_set_option(fit_plan* p, int option, int value) {
  if (!p) return fail(FIT_E_ARGUMENT, "null plan");
  // pipeline-shape options apply to the sub-grid child as well
  if (option != FIT_OPT_PLAIN_PIPELINE && option != FIT_OPT_SUBGRID)
    for (SubGrid& sg : p->subs) {
      int rc = fit_plan_set_option(sg.c, option, value);
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
    case FIT_OPT_FUSED_FLUX: p->fused_flux = value != 0; return FIT_OK;
    case FIT_OPT_RADIX_ORDER:
      if (value < 0 || value > 2) return fail(FIT_E_ARGUMENT, "radix order must be 0, 1 or 2");
      for (int d = 0; d < 3; ++d) make_desc(p->G[d], p->fd[d], value);
      p->radix_order = value;
      return FIT_OK;
    case FIT_OPT_COL_F64:
      if (value < 0 || value > 7) return fail(FIT_E_ARGUMENT, "FIT_OPT_COL_F64 takes 0 to 7");
      p->col_f64 = value;
      return FIT_OK;
    case FIT_OPT_COL_PIPE:
      if (value < 0 || value > 64) return fail(FIT_E_ARGUMENT, "FIT_OPT_COL_PIPE takes 0 to 64");
      p->col_pipe = value;
      return FIT_OK;
    case FIT_OPT_LAST_BINS:
      if (value < 1 || value > 128) return fail(FIT_E_ARGUMENT, "FIT_OPT_LAST_BINS takes 1 to 128 (KiB)");
      p->last_bins_kb = value;
      return FIT_OK;
    default: return fail(FIT_E_ARGUMENT, "unknown plan option");
  }
}

int fit_plan_get_option(const fit_plan* p, int option, int* value) {
  if (!p || !value) return fail(FIT_E_ARGUMENT, "null plan / value");
  switch (option) {
    case FIT_OPT_PLAIN_PIPELINE: *value = p->plain; return FIT_OK;
    case FIT_OPT_GIVER_BATCH: *value = p->giver_batch; return FIT_OK;
    case FIT_OPT_XPASS_GIVERS: *value = p->xfused_gc; return FIT_OK;
    case FIT_OPT_XPASS_BLOCK: *value = p->xfused_block; return FIT_OK;
    case FIT_OPT_SUBGRID: *value = p->use_sub; return FIT_OK;
    case FIT_OPT_COL_TILE: *value = p->col_tx; return FIT_OK;
    case FIT_OPT_XPASS_SELF: *value = p->xself; return FIT_OK;
    case FIT_OPT_XPASS_STAGE: *value = p->xstage; return FIT_OK;
    case FIT_OPT_GRAD_SPLIT: *value = p->grad_split; return FIT_OK;
    case FIT_OPT_FUSED_FLUX: *value = p->fused_flux; return FIT_OK;
    case FIT_OPT_RADIX_ORDER: *value = p->radix_order; return FIT_OK;
    case FIT_OPT_COL_F64: *value = p->col_f64; return FIT_OK;
    case FIT_OPT_COL_PIPE: *value = p->col_pipe; return FIT_OK;
    case FIT_OPT_LAST_BINS: *value = p->last_bins_kb; return FIT_OK;
    default: return fail(FIT_E_ARGUMENT, "unknown plan option");
  }
}

int fit_plan_autotune(fit_plan* p, int workload, int invariant, const void* f_hat, const void* adv_hat,
                      int io_dtype, int ptr_kind, int reps) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (workload != FIT_TUNE_FLUX && workload != FIT_TUNE_SHELL_TO_SHELL)
    return fail(FIT_E_ARGUMENT, "workload must be FIT_TUNE_FLUX or FIT_TUNE_SHELL_TO_SHELL");
  if (p->nsh <= 0) return fail(FIT_E_ARGUMENT, "plan has no shells");
  if (reps < 1) reps = 1;
  const int nsh = p->nsh;
  std::vector<double> T((size_t)nsh), Pi((size_t)nsh), Tm, net((size_t)nsh);
  double asym = 0.0;
  if (workload == FIT_TUNE_SHELL_TO_SHELL) Tm.resize((size_t)nsh * nsh);
  auto run = [&]() -> int {
    return workload == FIT_TUNE_FLUX
               ? fit_spectral_flux(p, invariant, f_hat, adv_hat, T.data(), Pi.data(), io_dtype, ptr_kind)
               : fit_shell_to_shell(p, invariant, f_hat, adv_hat, Tm.data(), net.data(), &asym, 0, io_dtype,
                                    ptr_kind);
  };
  // best of reps timed calls after one untimed call (every call synchronises its stream)
  auto timed = [&](double* ms) -> int {
    int r = run();
    if (r) return r;
    double best = 1e300;
    for (int i = 0; i < reps; ++i) {
      const auto t0 = std::chrono::steady_clock::now();
      if ((r = run())) return r;
      best = std::min(best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    *ms = best;
    return FIT_OK;
  };
  struct Knob { int opt; std::vector<int> vals; };
  std::vector<Knob> knobs;
  if (workload == FIT_TUNE_FLUX) {
    knobs = {{FIT_OPT_XPASS_SELF, {1, 0}}, {FIT_OPT_XPASS_STAGE, {1, 0}}, {FIT_OPT_GRAD_SPLIT, {-1, 0, 1}},
             {FIT_OPT_COL_TILE, {0, 8, 32}}, {FIT_OPT_FUSED_FLUX, {0, 1}}, {FIT_OPT_XPASS_BLOCK, {0, 1}},
             {FIT_OPT_RADIX_ORDER, {0, 1, 2}}};
  } else {
    knobs = {{FIT_OPT_XPASS_GIVERS, {0, 2, 8}}, {FIT_OPT_XPASS_BLOCK, {0, 1}}, {FIT_OPT_GRAD_SPLIT, {-1, 0, 1}},
             {FIT_OPT_COL_TILE, {0, 8, 32}}, {FIT_OPT_GIVER_BATCH, {0, std::max(1, nsh / 2), std::max(1, nsh / 4)}},
             {FIT_OPT_SUBGRID, {1, 0}}, {FIT_OPT_RADIX_ORDER, {0, 1, 2}}};
  }
  double cur = 0.0;
  if ((rc = timed(&cur))) return rc;
  const double t_start = cur;
  std::string log = std::string("workload=") + (workload == FIT_TUNE_FLUX ? "flux" : "shell_to_shell") +
                    " start_ms=" + std::to_string(t_start);
  std::vector<std::pair<int, int>> chosen;
  for (const Knob& k : knobs) {
    int best_v = 0;
    if ((rc = fit_plan_get_option(p, k.opt, &best_v))) return rc;
    for (int v : k.vals) {
      if (v == best_v) continue;
      if (fit_plan_set_option(p, k.opt, v) != FIT_OK) continue;
      double ms = 0.0;
      const int r = timed(&ms);
      if (r == FIT_E_DEVICE || r == FIT_E_OUT_OF_MEMORY) {
        (void)fit_plan_set_option(p, k.opt, best_v);
        return r;
      }
      if (r == FIT_OK && ms < cur * 0.98) {  // a choice must win by more than timing noise
        cur = ms;
        best_v = v;
      }
    }
    if ((rc = fit_plan_set_option(p, k.opt, best_v))) return rc;
    chosen.emplace_back(k.opt, best_v);
    log += " opt" + std::to_string(k.opt) + "=" + std::to_string(best_v);
  }
  log += " tuned_ms=" + std::to_string(cur);
  p->tune_report = log;
  {
    std::lock_guard<std::mutex> lock(g_create_mutex);
    tune_read_file();
    const std::string key = tune_key(p) + (workload == FIT_TUNE_FLUX ? ":flux" : ":s2s");
    tune_store(key, chosen);
    tune_append_file(key, chosen);
  }
  return FIT_OK;
}

int fit_plan_tuning(const fit_plan* p, char* buf, int len) {
  if (!p) return fail(FIT_E_ARGUMENT, "null plan");
  if (buf && len > 0) {
    std::strncpy(buf, p->tune_report.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return (int)p->tune_report.size();
}

static int kagg_out(const std::vector<KernAgg>& v, int i, char* name, int name_len, int* kernel_class,
                    int64_t* launches, double* total_ms, double* alg_bytes) {
  if (i < 0 || i >= (int)v.size()) return fail(FIT_E_ARGUMENT, "kernel index out of range");
  const KernAgg& a = v[(size_t)i];
  if (name && name_len > 0) {
    std::strncpy(name, a.name.c_str(), (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (kernel_class) *kernel_class = a.cls;
  if (launches) *launches = a.launches;
  if (total_ms) *total_ms = a.ms;
  if (alg_bytes) *alg_bytes = a.bytes;
  return FIT_OK;
}

int fit_profile_kernel_count(const fit_plan* p) { return p ? (int)p->kagg.size() : -1; }

int fit_profile_kernel(const fit_plan* p, int i, char* name, int name_len, int* kernel_class, int64_t* launches,
                       double* total_ms, double* alg_bytes) {
  if (!p) return fail(FIT_E_ARGUMENT, "null plan");
  return kagg_out(p->kagg, i, name, name_len, kernel_class, launches, total_ms, alg_bytes);
}

void* fit_plan_stream(const fit_plan* p) { return p ? (void*)p->stream : nullptr; }

// The plan's stream waits for an event recorded on the caller's stream: everything the caller
// queued there (the kernels that produced the inputs, pending users of recycled output memory)
// completes before the plan's next launch.
static int order_after(int device, hipEvent_t* ev, hipStream_t own, void* stream) {
  FIT_HIP(hipSetDevice(device));
  if (!*ev) FIT_HIP(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  FIT_HIP(hipEventRecord(*ev, (hipStream_t)stream));
  FIT_HIP(hipStreamWaitEvent(own, *ev, 0));
  return FIT_OK;
}

int fit_plan_wait_stream(fit_plan* p, void* stream) {
  if (!p) return fail(FIT_E_ARGUMENT, "null plan");
  return order_after(p->device, &p->order_ev, p->stream, stream);
}

int fit_profile_enable(fit_plan* p, int on) {
  if (!p) return fail(FIT_E_ARGUMENT, "null plan");
  p->prof = on != 0;
  p->nrec = 0;
  for (int i = 0; i < FIT_K_COUNT; ++i) { p->launches[i] = 0; p->ms[i] = 0; p->abytes[i] = 0; }
  p->kagg.clear();
  for (SubGrid& sg : p->subs) {
    int rc = fit_profile_enable(sg.c, on);
    if (rc) return rc;
  }
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
  rc = p->dtype == FIT_F64 ? flux_multi<double>(p, 1, &inv, f, a) : flux_multi<float>(p, 1, &inv, f, a);
  if (rc) return rc;
  return rows_tail(p, 1, T, Pi);
}

int fit_spectral_fluxes(fit_plan* p, int n_inv, const int* invariants, const void* f_hat, const void* adv_hat,
                        double* T, double* Pi, int io_dtype, int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (p->nsh <= 0) return fail(FIT_E_ARGUMENT, "plan has no shells");
  if (!invariants || n_inv < 1 || n_inv > 4) return fail(FIT_E_ARGUMENT, "1 to 4 invariants per call");
  for (int q = 0; q < n_inv; ++q)
    if ((rc = check_invariant(p, invariants[q], p->M))) return rc;
  if ((size_t)n_inv * p->nsh * sizeof(double) > (size_t)160 * 1024)
    return fail(FIT_E_UNSUPPORTED, "n_inv * n_shells exceeds the shell bins one workgroup holds");
  const void *f, *a;
  if ((rc = stage_fa(p, f_hat, adv_hat, io_dtype, ptr_kind, &f, &a))) return rc;
  rc = p->dtype == FIT_F64 ? flux_multi<double>(p, n_inv, invariants, f, a)
                           : flux_multi<float>(p, n_inv, invariants, f, a);
  if (rc) return rc;
  return rows_tail(p, n_inv, T, Pi);
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
  return rows_tail(p, n_inv, T, Pi);
}

int fit_spectral_flux_with_nhat_noalloc(fit_plan* p, int inv, const void* c_hat, const void* N_hat, double* T,
                                        double* Pi, int io_dtype, int ptr_kind) {
  int rc = check_plan(p);  // makes the plan's device current
  if (rc) return rc;
  if (p->nsh <= 0) return fail(FIT_E_ARGUMENT, "plan has no shells");
  if (p->nsh > FIT_NOALLOC_MAX_SHELLS)
    return fail(FIT_E_UNSUPPORTED, "fit_spectral_flux_with_nhat_noalloc: " + std::to_string(p->nsh) +
                                       " shells exceed its maximum of " + std::to_string(FIT_NOALLOC_MAX_SHELLS));
  if ((rc = check_invariant(p, inv, p->M))) return rc;
  if (!c_hat || !N_hat || !T || !Pi) return fail(FIT_E_ARGUMENT, "NULL array");
  if (ptr_kind != FIT_PTR_DEVICE || io_dtype != p->dtype)
    return fail(FIT_E_ARGUMENT, "fit_spectral_flux_with_nhat_noalloc takes device arrays in the plan's compute "
                                "precision (staging them would allocate)");
  double* scr = nullptr;
  if ((rc = scratch_for(p->device, &scr))) return rc;
  const int nsh = p->nsh;
  // partial rows fill the scratch up to the space the sums need; every row is written in full
  // by each call, so nothing has to start at zero
  const int nblk = (int)std::min<size_t>(1024, (kScratchDoubles - (size_t)nsh) / (size_t)nsh);
  double* partials = scr;
  double* sums = scr + (size_t)nblk * nsh;
  rc = p->dtype == FIT_F64 ? density_shells_into<double>(p, inv, c_hat, N_hat, partials, nblk, sums)
                           : density_shells_into<float>(p, inv, c_hat, N_hat, partials, nblk, sums);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cumsum_shells, dim3(1), dim3(64), 0, p->stream, (const double*)sums, nsh, T, Pi);
  FIT_HIP(hipGetLastError());
  FIT_HIP(hipStreamSynchronize(p->stream));
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
  if ((rc = check_invariant(p, inv, p->M))) return rc;
  const void *f, *a;
  if ((rc = stage_fa(p, f_hat, adv_hat, io_dtype, ptr_kind, &f, &a))) return rc;
  rc = p->dtype == FIT_F64 ? m2m_impl<double>(p, inv, f, a, S, net, io_dtype, ptr_kind)
                           : m2m_impl<float>(p, inv, f, a, S, net, io_dtype, ptr_kind);
  if (rc) return rc;
  return finish(p);
}

int fit_decompose(fit_plan* p, int decomposition, const void* u_hat, void* out_a, void* out_b, int io_dtype,
                  int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  const char* name = decomposition == FIT_DECOMP_HELICAL ? "HelicalDecomposition"
                     : decomposition == FIT_DECOMP_TOROIDAL_POLOIDAL ? "ToroidalPoloidalDecomposition"
                                                                     : nullptr;
  if (!name) return fail(FIT_E_ARGUMENT, "unknown decomposition code");
  if (p->nd != 3)
    return fail(FIT_E_ARGUMENT, std::string(name) + " is defined in 3D only (got nd=" + std::to_string(p->nd) + ").");
  if (p->M < 3) return fail(FIT_E_ARGUMENT, std::string(name) + " needs >=3 velocity components");
  if (!u_hat || !out_a || !out_b) return fail(FIT_E_ARGUMENT, "NULL array");
  const void* u;
  if ((rc = stage_in(p, u_hat, p->M, io_dtype, ptr_kind, &p->in_f, &u))) return rc;
  const int64_t Np = p->Np;
  const size_t cs = 2 * p->rsz;
  // outputs into nhat (3 comps) and a second buffer (in_c)
  if (!p->in_c) { if ((rc = dalloc(p, &p->in_c, (size_t)std::max(p->M, p->D) * Np * cs))) return rc; }
  if (p->dtype == FIT_F64) {
    Grid g = grid_of<double>(p);
    if (decomposition == FIT_DECOMP_HELICAL)
      hipLaunchKernelGGL(k_helical<double>, dim3(1024), dim3(256), 0, p->stream, g, (const double2*)u, Np,
                         (const double*)p->ks[0], (const double*)p->ks[1], (const double*)p->ks[2], (double2*)p->nhat,
                         (double2*)p->in_c, Np);
    else
      hipLaunchKernelGGL(k_torpol<double>, dim3(1024), dim3(256), 0, p->stream, g, (const double2*)u, Np,
                         (const double*)p->ks[0], (const double*)p->ks[1], (const double*)p->ks[2], (double2*)p->nhat,
                         (double2*)p->in_c, Np);
  } else {
    Grid g = grid_of<float>(p);
    if (decomposition == FIT_DECOMP_HELICAL)
      hipLaunchKernelGGL(k_helical<float>, dim3(1024), dim3(256), 0, p->stream, g, (const float2*)u, Np,
                         (const float*)p->ks[0], (const float*)p->ks[1], (const float*)p->ks[2], (float2*)p->nhat,
                         (float2*)p->in_c, Np);
    else
      hipLaunchKernelGGL(k_torpol<float>, dim3(1024), dim3(256), 0, p->stream, g, (const float2*)u, Np,
                         (const float*)p->ks[0], (const float*)p->ks[1], (const float*)p->ks[2], (float2*)p->nhat,
                         (float2*)p->in_c, Np);
  }
  FIT_HIP(hipGetLastError());
  // components beyond 3 are zero in both outputs
  if (p->M > 3) {
    FIT_HIP(hipMemsetAsync((char*)p->nhat + 3 * Np * cs, 0, (size_t)(p->M - 3) * Np * cs, p->stream));
    FIT_HIP(hipMemsetAsync((char*)p->in_c + 3 * Np * cs, 0, (size_t)(p->M - 3) * Np * cs, p->stream));
  }
  if ((rc = stage_out_real(p, p->nhat, out_a, 2 * (int64_t)p->M * Np, io_dtype, ptr_kind))) return rc;
  if ((rc = finish(p))) return rc;
  if ((rc = stage_out_real(p, p->in_c, out_b, 2 * (int64_t)p->M * Np, io_dtype, ptr_kind))) return rc;
  return finish(p);
}

int fit_physical_to_spectral(fit_plan* p, const void* u_phys, int ncomp, void* u_hat, int io_dtype, int ptr_kind) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!u_phys || !u_hat) return fail(FIT_E_ARGUMENT, "NULL array");
  if (io_dtype != FIT_F32 && io_dtype != FIT_F64) return fail(FIT_E_ARGUMENT, "io_dtype must be FIT_F32 or FIT_F64");
  if (ptr_kind != FIT_PTR_HOST && ptr_kind != FIT_PTR_DEVICE) return fail(FIT_E_ARGUMENT, "bad ptr_kind");
  if (ncomp < 1 || ncomp > std::max(p->M, 3))
    return fail(FIT_E_DIMENSION, "ncomp must be between 1 and max(M, 3) of the plan");
  if (p->dealias == FIT_DEALIAS_PADDED)
    return fail(FIT_E_UNSUPPORTED, "the physical-to-spectral transform runs on a plan without 3/2 padding");
  const size_t ib = io_dtype == FIT_F64 ? 8 : 4;
  const void* src = u_phys;
  void* tmp = nullptr;
  if (ptr_kind == FIT_PTR_HOST) {
    FIT_HIP(hipMalloc(&tmp, (size_t)ncomp * p->Np * ib));
    hipError_t e = hipMemcpyAsync(tmp, u_phys, (size_t)ncomp * p->Np * ib, hipMemcpyHostToDevice, p->stream);
    if (e != hipSuccess) { hipFree(tmp); return fail(FIT_E_DEVICE, hipGetErrorString(e)); }
    src = tmp;
  }
  // a device output of the compute precision is written directly; otherwise through nhat
  const bool direct = ptr_kind == FIT_PTR_DEVICE && io_dtype == p->dtype;
  void* dst = direct ? u_hat : p->nhat;
  rc = p->dtype == FIT_F64 ? phys_to_spec<double>(p, src, io_dtype, ncomp, dst)
                           : phys_to_spec<float>(p, src, io_dtype, ncomp, dst);
  if (!rc && !direct) rc = stage_out_real(p, p->nhat, u_hat, 2 * (int64_t)ncomp * p->Np, io_dtype, ptr_kind);
  if (!rc) rc = finish(p);
  if (tmp) hipFree(tmp);
  return rc;
}

int fit_helical_decompose(fit_plan* p, const void* u_hat, void* u_plus, void* u_minus, int io_dtype, int ptr_kind) {
  return fit_decompose(p, FIT_DECOMP_HELICAL, u_hat, u_plus, u_minus, io_dtype, ptr_kind);
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
  if (e->order_ev) hipEventDestroy(e->order_ev);
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
      if (option > FIT_ENS_OPT_PLAN) {
        // a plan option (FIT_OPT_*) on every group's plan
        for (EnsGroup& G : e->groups) {
          const int rc = fit_plan_set_option(G.p, option - FIT_ENS_OPT_PLAN, value);
          if (rc) return rc;
          G.ng = 0;
        }
        ens_drop_graph(e);
        return FIT_OK;
      }
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
  // groups on padded grids pass the snapshot pointers themselves as kernel arguments (the others
  // read them from the device table): a graph replays only for the pointers it was captured with
  std::vector<void*> ptrs;
  for (EnsGroup& G : e->groups)
    if (G.p->dealias == FIT_DEALIAS_PADDED) ptrs.insert(ptrs.end(), G.h_tab.begin(), G.h_tab.end());
  if (e->exec && e->graph_io == io_key && e->graph_ptrs == ptrs && !e->prof) {
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
      e->graph_ptrs = ptrs;
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

// per-kernel totals of an ensemble: its groups' plans merged by label
static std::vector<KernAgg> ens_kagg(const fit_ensemble* e) {
  std::vector<KernAgg> v;
  for (const EnsGroup& G : e->groups)
    for (const KernAgg& a : G.p->kagg) {
      KernAgg* k = nullptr;
      for (KernAgg& b : v)
        if (b.name == a.name) { k = &b; break; }
      if (!k) { v.push_back(KernAgg{a.name, a.cls}); k = &v.back(); }
      k->launches += a.launches;
      k->ms += a.ms;
      k->bytes += a.bytes;
    }
  return v;
}

int fit_ensemble_profile_kernel_count(const fit_ensemble* e) { return e ? (int)ens_kagg(e).size() : -1; }

int fit_ensemble_profile_kernel(const fit_ensemble* e, int i, char* name, int name_len, int* kernel_class,
                                int64_t* launches, double* total_ms, double* alg_bytes) {
  if (!e) return fail(FIT_E_ARGUMENT, "null ensemble");
  return kagg_out(ens_kagg(e), i, name, name_len, kernel_class, launches, total_ms, alg_bytes);
}

void* fit_ensemble_stream(const fit_ensemble* e) { return e ? (void*)e->stream : nullptr; }

int fit_ensemble_wait_stream(fit_ensemble* e, void* stream) {
  if (!e) return fail(FIT_E_ARGUMENT, "null ensemble");
  return order_after(e->device, &e->order_ev, e->stream, stream);
}

}  // extern "C"
