// fit_pipeline.h -- pass launchers and pipeline orchestration of the transfer path, as templates
// over the compute precision R.  Included only by fit_f32.hip and fit_f64.hip, which instantiate
// the entry points declared in fit_internal.h for float and double.
#pragma once
#include "fit_internal.h"

namespace fitx {

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
  prof_begin(p, FIT_K_FFT_COLUMN, (double)nx * nob * (rows_in + rows_out) * csz * nfields,
             (!p->plain && reg_supported(L)) ? "k_col_r" : "k_col", L,
             (std::string(SIGN > 0 ? ",inv" : ",fwd") + KTag<LD>::v + KTag<ST>::v).c_str());
  if (!p->plain && reg_supported(L) && p->col_pipe > 0) {
    // software-pipelined form: col_pipe tiles of the other index per workgroup
    auto run = [&](auto ctag) {
      using C = decltype(ctag);
      const size_t ccs = sizeof(cpx<C>);
      const int T = L / gE(L);
      int TX = std::min(p->col_tx > 0 ? p->col_tx : 16, 512 / T);
      const size_t lds_cap = p->col_tx > 0 ? (size_t)160 * 1024 : g_max_lds;
      while (TX > 1 && (size_t)L * TX * ccs > lds_cap) TX /= 2;
      const int gy = std::max(1, (nob + p->col_pipe - 1) / p->col_pipe);
      dim3 grid((nx + TX - 1) / TX, gy, nfields);
      const size_t lds = (size_t)L * TX * ccs;
      switch (L) {
#define FIT_CASE(LL)                                                                                          \
  case LL:                                                                                                    \
    hipLaunchKernelGGL((k_col_p<R, LL, SIGN, LD, ST, C>), grid, dim3(T * TX), lds, p->stream, ld, st,         \
                       twc<R, C>(p, axis), nx, TX, axis, ob, ib, wb, fband, fdiv, gother, nob);               \
    break;
        FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
        default: break;
      }
    };
    if (std::is_same<R, float>::value && (p->col_f64 & (SIGN > 0 ? 1 : (axis == 1 ? 2 : 4)))) run(0.0);
    else run((R)0);
  } else if (!p->plain && reg_supported(L)) {
    auto run = [&](auto ctag) {
      using C = decltype(ctag);
      const size_t ccs = sizeof(cpx<C>);
      const int T = L / colE<C>(L);
      // columns per workgroup: 16 (128-byte rows in FP32) within 64 KB of LDS, or the plan's cap
      // within the 160 KB a workgroup may hold
      int TX = std::min(p->col_tx > 0 ? p->col_tx : 16, col_threads<C>() / T);
      const size_t lds_cap = p->col_tx > 0 ? (size_t)160 * 1024 : g_max_lds;
      while (TX > 1 && (size_t)L * TX * ccs > lds_cap) TX /= 2;
      dim3 grid((nx + TX - 1) / TX, nob, nfields);
      const size_t lds = (size_t)L * TX * ccs;
      switch (L) {
#define FIT_CASE(LL)                                                                                          \
  case LL:                                                                                                    \
    hipLaunchKernelGGL((k_col_r<R, LL, SIGN, LD, ST, C>), grid, dim3(T * TX), lds, p->stream, ld, st,         \
                       twc<R, C>(p, axis), nx, TX, axis, ob, ib, wb, fband, fdiv, gother);                    \
    break;
        FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
        default: break;
      }
    };
    if (std::is_same<R, float>::value && (p->col_f64 & (SIGN > 0 ? 1 : (axis == 1 ? 2 : 4)))) run(0.0);
    else run((R)0);
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
             ((double)nin * csz + (double)L * sizeof(R) + extra_bytes_per_row) * nrows * nfields,
             (!p->plain && reg_supported(L)) ? "k_row_inv_r" : "k_row_inv", L, KTag<ST>::v);
  if (!p->plain && reg_supported(L)) {
    const int T = reg_T(L);
    const size_t LL = (size_t)lds_line(L, rp_E(L));
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
  prof_begin(p, FIT_K_FFT_ROW_FORWARD, ((double)nout * csz + (double)L * sizeof(R)) * nrows * nfields, "k_row_fwd",
             L, KTag<ST>::v);
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

inline const int* full_bi(fit_plan* p) {
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
  const bool c64 = std::is_same<R, float>::value && (p->col_f64 & 1);   // FP64 arithmetic (FIT_OPT_COL_F64)
  {
    const int L = p->G[ax];
    const Band ob = nd == 3 ? Band{bi[1], p->G[1]} : full_band(1);
    const Band ib = Band{bi[ax], L};
    const Band wb = full_band(L);
    const int nob = ob.count();
    prof_begin(p, FIT_K_FFT_COLUMN, (double)ngc * xa * nob * (2.0 * ib.count() + 2.0 * wb.count()) * csz,
               "k_grad1_r", L);
    auto run = [&](auto ctag) {
      using C = decltype(ctag);
      switch (L) {
#define FIT_CASE(LL)                                                                                          \
  case LL: {                                                                                                  \
    constexpr int TX = grad_tx<LL, C>();                                                                      \
    dim3 grid((xa + TX - 1) / TX, nob, ngc);                                                                  \
    hipLaunchKernelGGL((k_grad1_r<R, LL, C>), grid, dim3(two_T(LL) * TX), (size_t)L * TX * sizeof(cpx<C>),    \
                       p->stream, ls, p->M, giver0, st, twc<R, C>(p, ax), xa, TX, ax, ob, ib, wb, fband,      \
                       p->G[1]);                                                                              \
  } break;
        FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
        default: break;
      }
    };
    if (c64) run(0.0);
    else run((R)0);
    prof_end(p);
    FIT_HIP(hipGetLastError());
  }
  if (nd == 3) {
    const int L = p->G[1];
    const Band ob = full_band(p->G[2]);
    const Band ib = Band{bi[1], L};
    const Band wb = full_band(L);
    const bool split = p->grad_split < 0 ? L < 512 : p->grad_split != 0;
    prof_begin(p, FIT_K_FFT_COLUMN,
               (double)ngc * xa * ob.count() * ((split ? 1.0 : 2.0) * ib.count() + (split ? 2.0 : 3.0) * wb.count()) * csz,
               "k_grad2_r", L);
    auto run = [&](auto ctag) {
      using C = decltype(ctag);
      const int nob = ob.count();
      const int gy = p->col_pipe > 0 ? std::max(1, (nob + p->col_pipe - 1) / p->col_pipe) : nob;
      switch (L) {
#define FIT_CASE(LL)                                                                                          \
  case LL: {                                                                                                  \
    if (p->col_pipe > 0) {                                                                                    \
      const int TX = grad_tile(p, L, sizeof(cpx<C>));                                                         \
      dim3 grid((xa + TX - 1) / TX, gy, ngc);                                                                 \
      hipLaunchKernelGGL((k_grad2_p<R, LL, C>), grid, dim3(two_T(LL) * TX), (size_t)L * TX * sizeof(cpx<C>),  \
                         p->stream, st.p, st.sy, st.sz, st.sf, (const R*)p->kt[1], twc<R, C>(p, 1), xa, TX,   \
                         ob, ib, wb, fband, p->M, split ? 0 : 1, nob);                                        \
    } else {                                                                                                  \
      constexpr int TX = grad_tx<LL, C>();                                                                    \
      dim3 grid((xa + TX - 1) / TX, gy, ngc);                                                                 \
      hipLaunchKernelGGL((k_grad2_r<R, LL, C>), grid, dim3(two_T(LL) * TX), (size_t)L * TX * sizeof(cpx<C>),  \
                         p->stream, st.p, st.sy, st.sz, st.sf, (const R*)p->kt[1], twc<R, C>(p, 1), xa, TX,   \
                         ob, ib, wb, fband, p->M, split ? 0 : 1);                                             \
    }                                                                                                         \
  } break;
        FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
        default: break;
      }
    };
    if (c64) run(0.0);
    else run((R)0);
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
  // wave form when the velocity lines stay in registers (E <= 8); the workgroup form otherwise
  // (at 1024 points the wave form holds 16 points per lane in 164 VGPRs: 1.01 ms against
  // 0.34 ms for the workgroup form in the 2D ensemble)
  const bool wave_ok = reg_supported(L) && !p->xfused_block && wE(L) <= 8 &&
                       (size_t)4 * (64 / (L / wE(L))) * (size_t)lds_line(L, wE(L)) * csz <= g_max_lds;
  prof_begin(p, FIT_K_FUSED_X, bytes, wave_ok ? "k_xfused_w" : reg_supported(L) ? "k_xfused_r" : "k_xfused", L);
  if (wave_ok) {
    // wave-per-line kernel: 4 waves per workgroup, 64/T row pairs per wave
    const int T = L / wE(L);
    const int pairs_per_block = 4 * (64 / T);
    const size_t LL = (size_t)lds_line(L, wE(L));
    const size_t lds = (size_t)pairs_per_block * LL * csz;
    const int nrg = (npairs + pairs_per_block - 1) / pairs_per_block;
    const int ngiv = nfg / p->M;
    const int GC = ugs ? 1 : std::max(1, std::min(ngiv, p->xfused_gc > 0 ? p->xfused_gc : 4));
    const int nchunk = (ngiv + GC - 1) / GC;
    // only lengths with at most 8 points per lane are instantiated (the others never qualify)
    switch (L) {
#define FIT_CASE(LLEN)                                                                                            \
  case LLEN:                                                                                                      \
    if constexpr (wE(LLEN) <= 8) {                                                                                \
      hipLaunchKernelGGL((k_xfused_w<R, LLEN>), dim3((unsigned)((int64_t)nrg * nchunk)), dim3(256), lds,          \
                         p->stream, (const cpx<R>*)p->tw[0], (const cpx<R>*)p->work2, gsy, gsz, gsf, (const R*)u, \
                         usy, usz, p->Gp, (cpx<R>*)p->nxbuf, osy, osz, osf, p->nd, p->M, p->G[1], nrows,          \
                         (R)(1.0 / (double)p->Gp), xa - 1, fband, xb - 1, ngiv, GC, kxt, ugs);                     \
    } else {                                                                                                      \
      prof_end(p);                                                                                                \
      return fail(FIT_E_UNSUPPORTED, "wave-form x pass needs at most 8 points per lane");                         \
    }                                                                                                             \
    break;
      FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
      default: break;
    }
  } else if (reg_supported(L)) {
    // workgroup form: B row pairs per workgroup, single LDS exchange buffer
    const int T = xblk_T(L);
    const size_t LL = (size_t)lds_line(L, xE(L));
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
  prof_begin(p, FIT_K_FUSED_X, bytes,
             (std::is_same<R, float>::value && p->xstage && (L == 512 || L == 768)) ? "k_xfused_self_stg"
                                                                                     : "k_xfused_self",
             L);
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
    const size_t lds = (size_t)B * lds_line(LLEN, sE(LLEN)) * csz;                                                         \
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

template <typename R> int lastdens_launch(fit_plan* p, int gb, const int* bo, int mode, const DensArgs<R>& da,
                                          int nbins, double* out);
template <typename R> int carrier_half(fit_plan* p, int inv, const void* f, const void* const* srcs, int nb, void* eh);

// Padded plans: the Nyquist-plane extras of a half-grid N̂ (StNhatHalf), after its M components
// in the same buffer; nullptr without padding.
template <typename R>
cpx<R>* nhat_ext(fit_plan* p, void* nh, int64_t* esf) {
  *esf = 0;
  if (p->dealias != FIT_DEALIAS_PADDED) return nullptr;
  const int n1h = p->n[0] / 2 + 1, w1 = p->n[0] - n1h;
  const int64_t Hn = (int64_t)n1h * p->n[1] * p->n[2];
  *esf = (int64_t)w1 * (p->n[1] + p->n[2]);
  return (cpx<R>*)nh + (int64_t)p->M * Hn;
}

// Remaining forward passes of nxbuf (M fields, one field group per component) into N̂: the full
// original grid, or for the flux (nhat_half_req) the original half grid plus the padded Nyquist
// extras.
template <typename R>
int v2_forward_nhat(fit_plan* p, const int* bo) {
  const int xb = xext(p, bo[0]), pb = xpitch(xb);
  if (p->fused_ninv > 0) {
    // flux: the last forward pass computes the densities and shell sums (N̂ never stored)
    if (p->nd == 3) {
      const int64_t sy = pb, sz = (int64_t)pb * p->G[1], sf = sz * p->G[2];
      LdHalf<R> ld{(const cpx<R>*)p->nxbuf, sy, sz, sf};
      StHalf<R> st{(cpx<R>*)p->nxbuf, sy, sz, sf};
      int rc = launch_col<R, -1>(p, ld, st, 1, p->M, xb, full_band(p->G[2]), full_band(p->G[1]),
                                 Band{bo[1], p->G[1]});
      if (rc) return rc;
    }
    DensArgs<R> da{};
    da.g = grid_of<R>(p);
    da.car = (const cpx<R>*)p->fused_car;
    da.csf = p->Np;
    da.trunc = p->dealias == FIT_DEALIAS_TWOTHIRDS;
    da.ninv = p->fused_ninv;
    da.inv0 = p->fused_inv[0];
    da.inv1 = p->fused_inv[1];
    da.ks0 = (const R*)p->ks[0];
    da.ks1 = (const R*)p->ks[1];
    da.ks2 = (const R*)p->ks[2];
    da.shell = p->shell;
    const int nb = p->fused_ninv * p->nsh;
    int rc = grow(p, (void**)&p->dcolG, &p->dcolG_bytes, (size_t)nb * sizeof(double));
    if (rc) return rc;
    if ((rc = lastdens_launch<R>(p, 1, bo, kDensFull, da, nb, p->dcolG))) return rc;
    p->fused_done = true;
    return FIT_OK;
  }
  const int64_t sy = pb, sz = (int64_t)pb * p->G[1], sf = sz * p->G[2];
  LdHalf<R> ld{(const cpx<R>*)p->nxbuf, sy, sz, sf};
  StHalf<R> st{(cpx<R>*)p->nxbuf, sy, sz, sf};
  Grid g = grid_of<R>(p);
  const bool trunc = p->dealias == FIT_DEALIAS_TWOTHIRDS;
  const double scale = (double)p->Gp / ((double)p->Np * (double)p->Np);
  void* dst = p->nhat_dst ? p->nhat_dst : p->nhat;
  // the flux reads N̂ only through the pair density: then the original half grid suffices
  const bool half = p->nhat_half_req && !p->nhat_dst;
  const int64_t nsf = half ? (int64_t)(p->n[0] / 2 + 1) * p->n[1] * p->n[2] : p->Np;
  int rc;
  if (trunc) FIT_HIP(hipMemsetAsync(dst, 0, (size_t)p->M * nsf * sizeof(cpx<R>), p->stream));
  auto last = [&](auto sn) {
    if (p->nd == 3) {
      int rc2 = launch_col<R, -1>(p, ld, st, 1, p->M, xb, full_band(p->G[2]), full_band(p->G[1]),
                                  Band{bo[1], p->G[1]});
      if (rc2) return rc2;
      return launch_col<R, -1>(p, ld, sn, 2, p->M, xb, Band{bo[1], p->G[1]}, full_band(p->G[2]),
                               Band{bo[2], p->G[2]});
    }
    return launch_col<R, -1>(p, ld, sn, 1, p->M, xb, full_band(1), full_band(p->G[1]), Band{bo[1], p->G[1]});
  };
  if (half) {
    int64_t esf = 0;
    cpx<R>* ext = nhat_ext<R>(p, dst, &esf);
    if ((rc = last(StNhatHalf<R>{g, (cpx<R>*)dst, nsf, trunc, (R)scale, ext, esf}))) return rc;
    p->nhat_half_done = true;
    return FIT_OK;
  }
  return last(StNhat<R>{g, (cpx<R>*)dst, p->Np, trunc, (R)scale});
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
  const int dpriv = detbins_priv(nsh, kThreads, kDensBinBudget);
  const size_t lds = detbins_bytes(nsh, kThreads, dpriv);
  const size_t csz = sizeof(cpx<R>);
  prof_begin(p, FIT_K_DENSITY_SHELL,
             (double)p->Np * (2.0 * p->M * csz + (shells ? 4.0 : 0.0) + (t_out ? sizeof(R) : 0.0)), "k_density");
  hipLaunchKernelGGL(k_density<R>, dim3(nblk), dim3(kThreads), lds, p->stream, g, int4{inv, 0, 0, 0}, 1, p->M,
                     (const cpx<R>*)car, p->Np, (const cpx<R>*)nh, p->Np, (const R*)p->ks[0],
                     (const R*)p->ks[1], (const R*)p->ks[2], shells ? p->shell : nullptr, nsh,
                     p->partials, (R*)t_out, dpriv);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  if (shells) {
    prof_begin(p, FIT_K_REDUCE, (double)nblk * nsh * 8.0 + nsh * 8.0, "k_reduce_rows");
    launch_reduce(p, p->partials, nblk, nsh, 1, p->dcol);
    prof_end(p);
    FIT_HIP(hipGetLastError());
  }
  return FIT_OK;
}

// density + shell partial sums into buffers the caller provides (nothing grown or allocated):
// nblk partial rows of nsh doubles at partials, reduced in fixed order into out (nsh doubles)
template <typename R>
int density_shells_into(fit_plan* p, int inv, const void* car, const void* nh, double* partials, int nblk,
                        double* out) {
  Grid g = grid_of<R>(p);
  const int nsh = p->nsh;
  const int dpriv = detbins_priv(nsh, kThreads, kDensBinBudget);
  hipLaunchKernelGGL(k_density<R>, dim3(nblk), dim3(kThreads), detbins_bytes(nsh, kThreads, dpriv), p->stream, g,
                     int4{inv, 0, 0, 0}, 1, p->M, (const cpx<R>*)car, p->Np, (const cpx<R>*)nh, p->Np,
                     (const R*)p->ks[0], (const R*)p->ks[1], (const R*)p->ks[2], p->shell, nsh, partials,
                     (R*)nullptr, dpriv);
  FIT_HIP(hipGetLastError());
  launch_reduce(p, partials, nblk, nsh, 1, out);
  FIT_HIP(hipGetLastError());
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
  prof_begin(p, FIT_K_DENSITY_SHELL, (double)p->Np * (2.0 * p->M * csz + 4.0), "k_density");
  const int dpriv = detbins_priv(nb, kThreads, kDensBinBudget);
  hipLaunchKernelGGL(k_density<R>, dim3(nblk), dim3(kThreads), detbins_bytes(nb, kThreads, dpriv), p->stream, g, iv,
                     ninv, p->M, (const cpx<R>*)car, p->Np, (const cpx<R>*)nh, p->Np, (const R*)p->ks[0],
                     (const R*)p->ks[1], (const R*)p->ks[2], p->shell, nsh, p->partials2, (R*)nullptr, dpriv);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  prof_begin(p, FIT_K_REDUCE, (double)nblk * nb * 8.0 + nb * 8.0, "k_reduce_rows");
  launch_reduce(p, p->partials2, nblk, nb, 1, p->dcolG);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

// density_multi with N̂ on the original half grid (nh, M components of Hn bins): the pair density
// of every half-grid bin and its mirror into the shared shell (k_density_pair), into dcolG.
template <typename R>
int density_multi_pair(fit_plan* p, int ninv, const int* invs, const void* car, const void* nh) {
  Grid g = grid_of<R>(p);
  const int nsh = p->nsh, nb = ninv * nsh;
  const int nblk = p->nblk_density;
  const size_t csz = sizeof(cpx<R>);
  const int64_t Hn = (int64_t)(p->n[0] / 2 + 1) * p->n[1] * p->n[2];
  int rc;
  if ((rc = grow(p, (void**)&p->partials2, &p->partials2_bytes, (size_t)nblk * nb * sizeof(double)))) return rc;
  if ((rc = grow(p, (void**)&p->dcolG, &p->dcolG_bytes, (size_t)nb * sizeof(double)))) return rc;
  const int4 iv{invs[0], ninv > 1 ? invs[1] : 0, ninv > 2 ? invs[2] : 0, ninv > 3 ? invs[3] : 0};
  // carrier at k and -k (every full-grid bin about once), N̂ and the shell map on the half grid
  const int ncar = p->M;
  prof_begin(p, FIT_K_DENSITY_SHELL, (double)p->Np * ncar * csz + (double)Hn * (p->M * csz + 4.0), "k_density_pair");
  const int thr = kThreads;
  const int dpriv = detbins_priv(nb, thr, kDensBinBudget);
  int64_t esf = 0;
  const cpx<R>* ext = nhat_ext<R>(p, const_cast<void*>(nh), &esf);
  hipLaunchKernelGGL(k_density_pair<R>, dim3(nblk), dim3(thr), detbins_bytes(nb, thr, dpriv), p->stream, g,
                     iv, ninv, p->M, (const cpx<R>*)car, p->Np, (const cpx<R>*)nh, Hn, (const R*)p->ks[0],
                     (const R*)p->ks[1], (const R*)p->ks[2], p->shh, nsh, p->partials2, dpriv, ext, esf);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  prof_begin(p, FIT_K_REDUCE, (double)nblk * nb * 8.0 + nb * 8.0, "k_reduce_rows");
  launch_reduce(p, p->partials2, nblk, nb, 1, p->dcolG);
  prof_end(p);
  FIT_HIP(hipGetLastError());
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

// Shell sums of ninv (<= 4) invariants of the flux of f advected by a, into dcolG (rows of nsh
// doubles): N̂ is stored and one density pass reads it for all invariants, or (FIT_OPT_FUSED_FLUX)
// the last forward pass computes the densities itself and N̂ is never stored.  Measured on the
// 512³ flux with 3/2 padding: stored 3.5 ms for the last pass + density, fused 5.1 ms (the fused
// kernel holds N̂ of all components per point and gathers the carrier at both modes of a bin; one
// workgroup per CU at 178 VGPRs), so the stored form is the default.
template <typename R>
int flux_multi(fit_plan* p, int ninv, const int* invs, const void* f, const void* a) {
  const int lax = p->nd == 3 ? 2 : 1;
  const bool fusable = p->fused_flux == 1 && !p->plain && p->nd >= 2 && p->M <= 3 && ninv <= 2 && p->shell &&
                       reg_supported(p->G[lax]) && (size_t)ninv * p->nsh * 8 <= (size_t)96 * 1024;
  int rc;
  if (fusable) {
    p->fused_ninv = ninv;
    p->fused_inv[0] = invs[0];
    p->fused_inv[1] = ninv > 1 ? invs[1] : invs[0];
    p->fused_car = f;
    p->fused_done = false;
    rc = nonlinear_term_impl<R>(p, f, a);
    p->fused_ninv = 0;
    p->fused_car = nullptr;
    if (rc) return rc;
    if (p->fused_done) return FIT_OK;
    // the pipeline took an unfused route and stored N̂
    return density_multi<R>(p, ninv, invs, f, p->nhat);
  }
  // N̂ on the half grid and the pair density when both modes of a pair share a shell
  p->nhat_half_req = p->shh && p->shell_symmetric && p->M <= 3 && p->nd >= 2 && !p->plain;
  p->nhat_half_done = false;
  rc = nonlinear_term_impl<R>(p, f, a);
  p->nhat_half_req = false;
  if (rc) return rc;
  if (p->nhat_half_done) {
    p->nhat_half_done = false;
    return density_multi_pair<R>(p, ninv, invs, f, p->nhat);
  }
  return density_multi<R>(p, ninv, invs, f, p->nhat);
}

// Physical -> spectral front-end (PencilFFTsExt.jl:83-88): û_c = fft(u_c)/Np for ncomp real
// physical components (precision T, (ns..., ncomp) column-major) into the full complex spectral
// array out (compute precision, (ns..., ncomp)).  r2c row pass, column passes, then the full-grid
// storer (each original index from its own half-grid bin or the conjugate of its mirror).
template <typename R, typename T>
int phys_to_spec_t(fit_plan* p, const T* u, int ncomp, void* out) {
  Grid g = grid_of<R>(p);
  for (int c = 0; c < ncomp; ++c) {
    LdRealAs<R, T> lr{u + (int64_t)c * p->Np, (int64_t)p->n[0], (int64_t)p->n[0] * p->n[1], 0};
    StNhat<R> sn{g, (cpx<R>*)out + (int64_t)c * p->Np, 0, 0, (R)(1.0 / (double)p->Np)};
    int rc;
    if (p->nd == 1) {
      if ((rc = launch_row_fwd<R>(p, lr, sn, 1, p->h1))) return rc;
      continue;
    }
    if ((rc = launch_row_fwd<R>(p, lr, st_work<R>(p), 1, p->h1))) return rc;
    if (p->nd == 2) {
      if ((rc = launch_col_full<R, -1>(p, ld_work<R>(p), sn, 1, 1))) return rc;
      continue;
    }
    if ((rc = launch_col_full<R, -1>(p, ld_work<R>(p), st_work<R>(p), 1, 1))) return rc;
    if ((rc = launch_col_full<R, -1>(p, ld_work<R>(p), sn, 2, 1))) return rc;
  }
  return FIT_OK;
}

template <typename R>
int phys_to_spec(fit_plan* p, const void* u, int phys_dtype, int ncomp, void* out) {
  return phys_dtype == FIT_F64 ? phys_to_spec_t<R, double>(p, (const double*)u, ncomp, out)
                               : phys_to_spec_t<R, float>(p, (const float*)u, ncomp, out);
}

// Plain giver loop: one giver per pipeline pass, full transforms, N̂ through HBM.
template <typename R>
int s2s_plain(fit_plan* p, int inv, const void* f, std::vector<double>& Tm) {
  int rc;
  const int nsh = p->nsh;
  for (int m = p->giver_lo; m <= (p->giver_hi > 0 ? p->giver_hi : nsh); ++m) {
    if ((rc = nonlinear_from_u<R>(p, f, m))) return rc;
    if ((rc = density_launch<R>(p, inv, f, p->nhat, nullptr, true))) return rc;
    FIT_HIP(hipMemcpyAsync(Tm.data() + (size_t)(m - 1) * nsh, p->dcol, nsh * sizeof(double), hipMemcpyDeviceToHost,
                           p->stream));
    // dcol is reused by the next giver: order the copy before the next reduction
    FIT_HIP(hipStreamSynchronize(p->stream));
  }
  return FIT_OK;
}

// Fused last forward pass + density + shell partial sums (k_lastdens_r) over gb field groups of M
// components in nxbuf (pitch xpitch(xext(bo[0]))), already through every other forward pass;
// shell sums reduced in fixed order into out (gb rows of nbins doubles, device).
template <typename R>
int lastdens_launch(fit_plan* p, int gb, const int* bo, int mode, const DensArgs<R>& da, int nbins, double* out) {
  const size_t csz = sizeof(cpx<R>);
  const int lax = p->nd == 3 ? 2 : 1;
  const int L = p->G[lax];
  const int xb = std::max(1, xext(p, bo[0])), pb = xpitch(xb);
  const int64_t sy = pb, sz = (int64_t)pb * p->G[1], sf = sz * p->G[2];
  const Band ob = p->nd == 3 ? Band{bo[1], p->G[1]} : full_band(1);
  const Band wb = p->nd == 3 ? Band{bo[2], p->G[2]} : Band{bo[1], p->G[1]};
  const int T = L / ldE(L);
  const int nthr = std::is_same<R, double>::value ? lastdens_threads<double>() : lastdens_threads<float>();
  // columns per workgroup: 16 (128-byte rows in FP32) within the thread bound and 64 KB of LDS
  int TX = std::max(1, std::min(16, nthr / T));
  while (TX > 1 && (size_t)L * TX * csz > 65536) TX /= 2;
  // private bins per wave within the budget (and the workgroup's LDS), with half the columns per
  // workgroup if only that makes them fit
  const size_t bin_budget = (size_t)p->last_bins_kb * 1024;
  auto priv_fits = [&](int tx) {
    return detbins_priv(nbins, T * tx, bin_budget) &&
           (detbins_bytes(nbins, T * tx, 1) + 15) / 16 * 16 + (size_t)L * tx * csz <= g_max_lds;
  };
  int dpriv = priv_fits(TX) ? 1 : 0;
  if (!dpriv && TX > 1 && priv_fits(TX / 2)) {
    TX /= 2;
    dpriv = 1;
  }
  const int gx = (xb + TX - 1) / TX;
  const int nrow = ob.count();
  // rows per workgroup: enough workgroups to fill the chip, few partial rows to reduce
  const int gy = std::max(1, std::min(nrow, (2048 + gx * gb - 1) / (gx * gb)));
  const int nblk = gx * gy;
  int rc;
  if ((rc = grow(p, (void**)&p->partials2, &p->partials2_bytes, (size_t)gb * nblk * nbins * sizeof(double))))
    return rc;
  const size_t lds = (detbins_bytes(nbins, T * TX, dpriv) + 15) / 16 * 16 + (size_t)L * TX * csz;
  const double bytes = (double)gb * p->M * xb * nrow * L * csz +
                       (mode == kDensFull ? (double)p->M * p->Np * csz + 4.0 * p->Np
                                          : (double)gb * p->M * xb * nrow * wb.count() * csz * (da.srcs ? 2.0 : 1.0));
  prof_begin(p, FIT_K_DENSITY_SHELL, bytes, mode == kDensFull ? "k_lastdens_r_full" : "k_lastdens_r", L);
  const R scale = mode == kDensFull ? (R)((double)p->Gp / ((double)p->Np * (double)p->Np)) : (R)(1.0 / (double)p->Np);
  dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)gb), block((unsigned)(T * TX));
  switch (L) {
#define FIT_CASE(LL)                                                                                             \
  case LL:                                                                                                       \
    if (mode == kDensFull)                                                                                       \
      hipLaunchKernelGGL((k_lastdens_r<R, LL, kDensFull>), grid, block, lds, p->stream, (const cpx<R>*)p->nxbuf, sy, \
                         sz, sf, (const cpx<R>*)p->tw[lax], xb, TX, lax, ob, wb, p->M, scale, da, nbins, p->nsh,  \
                         p->partials2, dpriv);                                                                   \
    else                                                                                                         \
      hipLaunchKernelGGL((k_lastdens_r<R, LL, kDensPair>), grid, block, lds, p->stream, (const cpx<R>*)p->nxbuf, sy, \
                         sz, sf, (const cpx<R>*)p->tw[lax], xb, TX, lax, ob, wb, p->M, scale, da, nbins, p->nsh,  \
                         p->partials2, dpriv);                                                                   \
    break;
    FIT_REG_LENGTHS(FIT_CASE)
#undef FIT_CASE
    default:
      prof_end(p);
      return fail(FIT_E_UNSUPPORTED, "fused last pass needs a register FFT length");
  }
  prof_end(p);
  FIT_HIP(hipGetLastError());
  prof_begin(p, FIT_K_REDUCE, (double)gb * nblk * nbins * 8.0, "k_reduce_rows");
  launch_reduce(p, p->partials2, nblk, nbins, gb, out);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

// LDS-engine form of batched_lastdens for lengths without a register FFT: tail of a batched pass
// for gb field groups already through the fused x pass (nxbuf, pitch
// xpitch(xext(band_rx))): the remaining forward pass(es), the pair density against eh (+ g*ehgs for group
// g) and FP64 shell bins, reduced in fixed order into out (gb rows of nsh doubles, device).
template <typename R>
int batched_lastdens_lds(fit_plan* p, int gb, const void* eh, int64_t ehgs, double* out) {
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
             (double)gb * p->M * xb * gy * L * csz + (double)gb * p->M * xb * gy * wb.count() * csz, "k_lastdens", L);
  // per-thread accumulators: the smallest register array that covers TX*L points per workgroup
  const int qneed = (L * TX + kThreads - 1) / kThreads;
  const int dpriv = detbins_priv(nsh, kThreads, kFftBinBudget);
  const size_t lds = (size_t)2 * L * TX * csz + detbins_bytes(nsh, kThreads, dpriv);
#define FIT_LASTDENS(QQ)                                                                                         \
  hipLaunchKernelGGL((k_lastdens<R, QQ>), dim3(gx, gy, gb), dim3(kThreads), lds,                                  \
                     p->stream, p->fd[lax], (const cpx<R>*)p->tw[lax], (const cpx<R>*)p->nxbuf, sy, sz, sf, xb, TX, \
                     lax, ob, wb, p->M, (R)scale, (const cpx<R>*)eh, (int64_t)n1h, (int64_t)n1h * p->n[1], Hn,     \
                     p->shh, nsh, p->n[0], p->partials2, ehgs, dpriv)
  if (qneed <= 4) FIT_LASTDENS(4);
  else if (qneed <= 8) FIT_LASTDENS(8);
  else if (qneed <= 16) FIT_LASTDENS(16);
  else FIT_LASTDENS(32);
#undef FIT_LASTDENS
  prof_end(p);
  FIT_HIP(hipGetLastError());
  prof_begin(p, FIT_K_REDUCE, (double)gb * nblk * nsh * 8.0, "k_reduce_rows");
  launch_reduce(p, p->partials2, nblk, nsh, gb, out);
  prof_end(p);
  FIT_HIP(hipGetLastError());
  return FIT_OK;
}

// Tail of a batched pass for gb field groups already through the fused x pass (nxbuf, pitch
// xpitch(xext(band_rx))): the remaining forward pass(es), the pair density against eh (+ g*ehgs for
// group g; or, with srcs, the Hermitian part of snapshot g's own masked array, component stride
// ssf) and FP64 shell bins, reduced in fixed order into out (gb rows of nsh doubles, device).
template <typename R>
int batched_lastdens(fit_plan* p, int gb, const void* eh, int64_t ehgs, double* out,
                     const void* const* srcs = nullptr) {
  const int nsh = p->nsh;
  const int* bo = p->band_rx;
  const int n1h = p->n[0] / 2 + 1;
  const int xb = std::max(1, xext(p, bo[0])), pb = xpitch(xb);
  const int64_t sy = pb, sz = (int64_t)pb * p->G[1], sf = sz * p->G[2];
  int rc;
  if (!reg_supported(p->G[p->nd == 3 ? 2 : 1])) {
    if (srcs) {  // the LDS form reads a precomputed carrier
      const int64_t Hn = (int64_t)n1h * p->n[1] * p->n[2];
      if ((rc = grow(p, &p->eh, &p->eh_bytes, (size_t)gb * p->M * Hn * sizeof(cpx<R>)))) return rc;
      if ((rc = carrier_half<R>(p, FIT_KINETIC_ENERGY, nullptr, srcs, gb, p->eh))) return rc;
      eh = p->eh;
      ehgs = (int64_t)p->M * Hn;
    }
    return batched_lastdens_lds<R>(p, gb, eh, ehgs, out);
  }
  if (p->nd == 3) {
    LdHalf<R> ld{(const cpx<R>*)p->nxbuf, sy, sz, sf};
    StHalf<R> st{(cpx<R>*)p->nxbuf, sy, sz, sf};
    if ((rc = launch_col<R, -1>(p, ld, st, 1, gb * p->M, xb, full_band(p->G[2]), full_band(p->G[1]),
                                Band{bo[1], p->G[1]})))
      return rc;
  }
  DensArgs<R> da{};
  da.g = grid_of<R>(p);
  da.trunc = p->dealias == FIT_DEALIAS_TWOTHIRDS;
  da.shell = p->shh;
  da.eh = (const cpx<R>*)eh;
  da.ehsy = n1h;
  da.ehsz = (int64_t)n1h * p->n[1];
  da.ehsf = (int64_t)n1h * p->n[1] * p->n[2];
  da.ehgs = ehgs;
  da.srcs = (const cpx<R>* const*)srcs;
  da.ssf = p->Np;
  return lastdens_launch<R>(p, gb, bo, kDensPair, da, nsh, out);
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
  prof_begin(p, FIT_K_OTHER, (double)nb * Hn * p->M * csz * 3.0, "k_carrier_half");
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
  const int glo = p->giver_lo, ghi = p->giver_hi > 0 ? p->giver_hi : nsh;
  int bmax[3] = {-1, -1, -1};
  for (int s = glo - 1; s < ghi; ++s) {
    bmax[0] = std::max(bmax[0], p->shell_band[s].x);
    bmax[1] = std::max(bmax[1], p->shell_band[s].y);
    bmax[2] = std::max(bmax[2], p->shell_band[s].z);
  }
  const int xa_max = std::max(1, xext(p, bmax[0]));
  const double per_giver = batch_bytes_per_group<R>(p, xa_max);
  size_t freeb = 0, totalb = 0;
  FIT_HIP(hipMemGetInfo(&freeb, &totalb));
  const double budget = 0.6 * ((double)freeb + (double)p->work2_bytes + (double)p->nxbuf_bytes);
  int ng = (int)std::max(1.0, std::min((double)(ghi - glo + 1), budget / per_giver));
  if (p->giver_batch > 0) ng = std::min(ng, p->giver_batch);
  if ((rc = grow(p, (void**)&p->dcolG, &p->dcolG_bytes, (size_t)ng * nsh * sizeof(double)))) return rc;
  for (int m0 = glo; m0 <= ghi; m0 += ng) {
    const int gb = std::min(ng, ghi - m0 + 1);
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
  const size_t csz = sizeof(cpx<R>);
  const int nsh = p->nsh;
  for (const SubGrid& sg : p->subs) {
    fit_plan* c = sg.c;
    auto gather = [&](const void* src, int ncomp, void** dst, size_t* have) -> int {
      int r = grow(p, dst, have, (size_t)ncomp * c->Np * csz);
      if (r) return r;
      prof_begin(p, FIT_K_OTHER, (double)ncomp * c->Np * csz * 2.0, "k_subgrid_gather");
      hipLaunchKernelGGL(k_subgrid_gather<R>, dim3(2048), dim3(256), 0, p->stream, (const cpx<R>*)src, p->n[0],
                         p->n[1], p->n[2], p->Np, (cpx<R>*)*dst, c->n[0], c->n[1], c->n[2], c->Np, sg.q[0], sg.q[1],
                         sg.q[2], ncomp);
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
    for (size_t i = (size_t)(sg.m0 - 1) * nsh; i < (size_t)sg.m1 * nsh; ++i) Tm[i] *= sg.tscale;
  }
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
    prof_begin(p, FIT_K_OTHER, (double)p->Np * (2.0 * p->M * cs + sizeof(R)), "k_band_field");
    hipLaunchKernelGGL(k_band_field<R>, dim3(1024), dim3(256), 0, p->stream, p->Np, p->M, (const cpx<R>*)f, p->Np,
                       kc, cen, nb, inv2s2, m, (cpx<R>*)p->in_c);
    prof_end(p);
    FIT_HIP(hipGetLastError());
    rc = (p->plain || p->nd == 1) ? nonlinear_from_u<R>(p, p->in_c, 0) : v2_nonlinear<R>(p, p->in_c);
    if (rc) return rc;
    prof_begin(p, FIT_K_DENSITY_SHELL, (double)p->Np * (2.0 * p->M * cs + sizeof(R)), "k_band_density");
    const int dpriv = detbins_priv(nb, kThreads, kDensBinBudget);
    hipLaunchKernelGGL(k_band_density<R>, dim3(nblk), dim3(kThreads), detbins_bytes(nb, kThreads, dpriv), p->stream,
                       g, inv, p->M, (const cpx<R>*)f, p->Np, (const cpx<R>*)p->nhat, p->Np, (const R*)p->ks[0],
                       (const R*)p->ks[1], (const R*)p->ks[2], kc, cen, nb, inv2s2, p->partials2, dpriv);
    prof_end(p);
    FIT_HIP(hipGetLastError());
    prof_begin(p, FIT_K_REDUCE, (double)nblk * nb * 8.0, "k_reduce_rows");
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
  const bool pad = p->dealias == FIT_DEALIAS_PADDED;
  const int64_t Ap = pad ? p->Gp : Np;   // the triad table lives on the transform grid
  if (hipMalloc(&A, (size_t)p->nd * Ap * cs) != hipSuccess || hipMalloc(&E, (size_t)p->M * Np * cs) != hipSuccess ||
      hipMalloc(&netpart, (size_t)nchunk * Np * sizeof(double)) != hipSuccess ||
      hipMalloc(&netd, (size_t)Np * sizeof(double)) != hipSuccess) {
    cleanup();
    return fail(FIT_E_OUT_OF_MEMORY, "mode-to-mode scratch allocation failed");
  }
  if (ptr_kind == FIT_PTR_DEVICE) Sd = S;
  else if (hipMalloc(&Sd, (size_t)Np * Np * os) != hipSuccess) { cleanup(); return fail(FIT_E_OUT_OF_MEMORY, "S staging"); }
  const int trunc = p->dealias == FIT_DEALIAS_TWOTHIRDS;
  hipLaunchKernelGGL(k_m2m_prep<R>, dim3(512), dim3(256), 0, p->stream, g, trunc, inv, p->M, (const cpx<R>*)f, Np,
                     (const cpx<R>*)a, Np, (const R*)p->ks[0], (const R*)p->ks[1], (const R*)p->ks[2],
                     pad ? (cpx<R>*)nullptr : (cpx<R>*)A, (cpx<R>*)E);
  if (pad)
    hipLaunchKernelGGL(k_m2m_padA<R>, dim3(512), dim3(256), 0, p->stream, g, (const cpx<R>*)a, Np, (cpx<R>*)A);
  dim3 grid((unsigned)((Np + 255) / 256), nchunk);
  prof_begin(p, FIT_K_OTHER, (double)Np * Np * os, "k_m2m");
  auto launch = [&](auto out_tag, auto nd_tag) {
    using OUT = decltype(out_tag);
    constexpr int ND = decltype(nd_tag)::value;
    if (pad)
      hipLaunchKernelGGL((k_m2m<R, OUT, ND, true>), grid, dim3(256), 0, p->stream, g, 0, p->M, (const cpx<R>*)f, Np,
                         (const cpx<R>*)A, (const cpx<R>*)E, (const R*)p->ks[0], (const R*)p->ks[1],
                         (const R*)p->ks[2], pchunk, (OUT*)Sd, (double*)netpart);
    else
      hipLaunchKernelGGL((k_m2m<R, OUT, ND, false>), grid, dim3(256), 0, p->stream, g, trunc, p->M, (const cpx<R>*)f,
                         Np, (const cpx<R>*)A, (const cpx<R>*)E, (const R*)p->ks[0], (const R*)p->ks[1],
                         (const R*)p->ks[2], pchunk, (OUT*)Sd, (double*)netpart);
  };
  auto by_nd = [&](auto out_tag) {
    if (p->nd == 1) launch(out_tag, std::integral_constant<int, 1>());
    else if (p->nd == 2) launch(out_tag, std::integral_constant<int, 2>());
    else launch(out_tag, std::integral_constant<int, 3>());
  };
  if (io_dtype == FIT_F64) by_nd(0.0);
  else by_nd(0.0f);
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
    if ((rc = v2_inverse_grads<R>(p, nullptr, 0, gb, p->band_in, nullptr, tu))) return rc;
    if ((rc = v2_xfused<R>(p, gb, p->band_in, nullptr, p->band_rx, G.ub, ugs))) return rc;
    if ((rc = batched_lastdens<R>(p, gb, nullptr, 0, G.rows + (size_t)b0 * nsh, tu))) return rc;
  }
  if (e->with_scalar) {
    MGuard mg(p, 1);
    if ((rc = v2_inverse_grads<R>(p, nullptr, 0, gb, p->band_in, nullptr, tt))) return rc;
    if ((rc = v2_xfused<R>(p, gb, p->band_in, nullptr, p->band_rx, G.ub, ugs))) return rc;
    if ((rc = batched_lastdens<R>(p, gb, nullptr, 0, G.rows + (size_t)(B + b0) * nsh, tt))) return rc;
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
  per += (double)nd * p->Gp * sizeof(R);
  size_t freeb = 0, totalb = 0;
  FIT_HIP(hipMemGetInfo(&freeb, &totalb));
  int ng = (int)std::max(1.0, std::min((double)B, 0.5 * (double)freeb / per));
  if (e->batch_cap > 0) ng = std::min(ng, e->batch_cap);
  G.ng = ng;
  int rc;
  if ((rc = grow(p, &G.ub, &G.ub_bytes, (size_t)ng * nd * p->Gp * sizeof(R)))) return rc;
  return FIT_OK;
}

// A group on a 3/2-padded grid: each snapshot through the plan's own padded flux pipeline
// (velocities and products on the G-point grid, densities on the original modes), its rows copied
// from dcolG into the group's row table.  Kinetic energy: f = a = û; scalar: f = θ̂, a = û.
template <typename R>
int ens_padded(fit_ensemble* e, EnsGroup& G) {
  fit_plan* p = G.p;
  const int B = (int)G.snaps.size(), nsh = p->nsh;
  int rc;
  for (int b = 0; b < B; ++b) {
    const void* u = G.h_tab[(size_t)b];
    {
      const int inv = FIT_KINETIC_ENERGY;
      if ((rc = flux_multi<R>(p, 1, &inv, u, u))) return rc;
      FIT_HIP(hipMemcpyAsync(G.rows + (size_t)b * nsh, p->dcolG, (size_t)nsh * sizeof(double),
                             hipMemcpyDeviceToDevice, p->stream));
    }
    if (e->with_scalar) {
      MGuard mg(p, 1);
      const int inv = FIT_PASSIVE_SCALAR;
      if ((rc = flux_multi<R>(p, 1, &inv, G.h_tab[(size_t)B + b], u))) return rc;
      FIT_HIP(hipMemcpyAsync(G.rows + (size_t)(B + b) * nsh, p->dcolG, (size_t)nsh * sizeof(double),
                             hipMemcpyDeviceToDevice, p->stream));
    }
  }
  return FIT_OK;
}

template <typename R>
int ens_run(fit_ensemble* e) {
  for (EnsGroup& G : e->groups) {
    if (G.p->dealias == FIT_DEALIAS_PADDED) {
      int rc = ens_padded<R>(e, G);
      if (rc) return rc;
      continue;
    }
    int rc = ens_batch_size<R>(e, G);
    if (rc) return rc;
    const int B = (int)G.snaps.size();
    for (int b0 = 0; b0 < B; b0 += G.ng)
      if ((rc = ens_batch<R>(e, G, b0, std::min(G.ng, B - b0)))) return rc;
  }
  return FIT_OK;
}


}  // namespace fitx
