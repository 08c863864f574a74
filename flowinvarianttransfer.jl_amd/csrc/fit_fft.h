// fit_fft.h -- device-side building blocks: complex helpers, radix-2/3/4/5/8 butterflies and a
// mixed-radix Stockham FFT that runs B interleaved lines of length L out of LDS.
//
// LDS layout: point i of line l lives at buf[i*B + l] (lines interleaved), so a wavefront that
// walks l fastest touches consecutive LDS words and, in column passes, consecutive HBM words.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace fit {

template <typename R> struct CT;
template <> struct CT<float> { using T = float2; };
template <> struct CT<double> { using T = double2; };
template <typename R> using cpx = typename CT<R>::T;

template <typename R> __device__ __forceinline__ cpx<R> mk(R x, R y) { cpx<R> v; v.x = x; v.y = y; return v; }
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  V r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename V> __device__ __forceinline__ V cconj(V a) { V r; r.x = a.x; r.y = -a.y; return r; }
template <typename V, typename R> __device__ __forceinline__ V cscale(V a, R s) { V r; r.x = a.x * s; r.y = a.y * s; return r; }
// multiply by (SIGN * i):  SIGN=+1 -> i*a = (-a.y, a.x);  SIGN=-1 -> -i*a = (a.y, -a.x)
template <int SIGN, typename V> __device__ __forceinline__ V cmuli(V a) {
  V r;
  if (SIGN > 0) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
  return r;
}

constexpr int FIT_MAX_STAGES = 16;

// Radix plan of one transform length (host fills it; passed by value to kernels).
struct FftDesc {
  int L;
  int nst;
  int rad[FIT_MAX_STAGES];
};

// ---------------------------------------------------------------------------------------------
// Butterflies: y_q = Σ_k x_k exp(SIGN * 2πi q k / r)
// ---------------------------------------------------------------------------------------------

template <int SIGN, typename R> __device__ __forceinline__ void dft2(cpx<R>* v) {
  cpx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b); v[1] = csub(a, b);
}

template <int SIGN, typename R> __device__ __forceinline__ void dft3(cpx<R>* v) {
  const R h = (R)0.86602540378443864676;  // sqrt(3)/2
  cpx<R> t = cadd(v[1], v[2]);
  cpx<R> d = csub(v[1], v[2]);
  cpx<R> m = mk<R>(v[0].x - (R)0.5 * t.x, v[0].y - (R)0.5 * t.y);
  cpx<R> s = cscale(cmuli<SIGN>(d), h);
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, s);
  v[2] = csub(m, s);
}

template <int SIGN, typename R> __device__ __forceinline__ void dft4(cpx<R>* v) {
  cpx<R> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  cpx<R> c = cadd(v[1], v[3]), d = cmuli<SIGN>(csub(v[1], v[3]));
  v[0] = cadd(a, c); v[2] = csub(a, c);
  v[1] = cadd(b, d); v[3] = csub(b, d);
}

template <int SIGN, typename R> __device__ __forceinline__ void dft5(cpx<R>* v) {
  const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
  const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
  cpx<R> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  cpx<R> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  cpx<R> x0 = v[0];
  cpx<R> p1 = mk<R>(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  cpx<R> p2 = mk<R>(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  cpx<R> q1 = cmuli<SIGN>(mk<R>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  cpx<R> q2 = cmuli<SIGN>(mk<R>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = mk<R>(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
  v[1] = cadd(p1, q1); v[4] = csub(p1, q1);
  v[2] = cadd(p2, q2); v[3] = csub(p2, q2);
}

template <int SIGN, typename R> __device__ __forceinline__ void dft8(cpx<R>* v) {
  const R r2 = (R)0.70710678118654752440;
  cpx<R> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<R> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<SIGN, R>(e);
  dft4<SIGN, R>(o);
  // o[q] *= w8^q, w8 = exp(SIGN i pi/4)
  o[1] = mk<R>(r2 * (o[1].x - (R)SIGN * o[1].y), r2 * (o[1].y + (R)SIGN * o[1].x));
  o[2] = cmuli<SIGN>(o[2]);
  o[3] = mk<R>(r2 * (-o[3].x - (R)SIGN * o[3].y), r2 * (-o[3].y + (R)SIGN * o[3].x));
#pragma unroll
  for (int q = 0; q < 4; ++q) { v[q] = cadd(e[q], o[q]); v[q + 4] = csub(e[q], o[q]); }
}

template <int RADIX, int SIGN, typename R> __device__ __forceinline__ void dft(cpx<R>* v) {
  if constexpr (RADIX == 2) dft2<SIGN, R>(v);
  else if constexpr (RADIX == 3) dft3<SIGN, R>(v);
  else if constexpr (RADIX == 4) dft4<SIGN, R>(v);
  else if constexpr (RADIX == 5) dft5<SIGN, R>(v);
  else dft8<SIGN, R>(v);
}

// twiddle exp(SIGN 2πi idx / L) from the table tw[idx] = exp(-2πi idx / L)
template <int SIGN, typename R> __device__ __forceinline__ cpx<R> twiddle(const cpx<R>* __restrict__ tw, int idx) {
  cpx<R> w = tw[idx];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// One Stockham stage of radix RADIX over B interleaved lines: src -> dst.
template <int RADIX, int SIGN, typename R>
__device__ __forceinline__ void stockham_stage(const cpx<R>* src, cpx<R>* dst, int L, int B, int Ns,
                                               const cpx<R>* __restrict__ tw, int tid, int nth) {
  const int Lr = L / RADIX;
  const int tstep = L / (Ns * RADIX);
  const int nbf = Lr * B;
  for (int q = tid; q < nbf; q += nth) {
    const int l = q % B;
    const int j = q / B;
    const int m = j % Ns;
    cpx<R> v[RADIX];
#pragma unroll
    for (int k = 0; k < RADIX; ++k) v[k] = src[(j + k * Lr) * B + l];
    if (m > 0) {
#pragma unroll
      for (int k = 1; k < RADIX; ++k) v[k] = cmul(v[k], twiddle<SIGN, R>(tw, k * m * tstep));
    }
    dft<RADIX, SIGN, R>(v);
    const int o = (j - m) * RADIX + m;
#pragma unroll
    for (int k = 0; k < RADIX; ++k) dst[(o + k * Ns) * B + l] = v[k];
  }
}

// Full mixed-radix transform of B interleaved lines held in LDS buffer a (b = scratch of the
// same size).  Returns the buffer holding the natural-order result.  Caller must have synced
// after filling a; the routine ends with a barrier.
template <int SIGN, typename R>
__device__ cpx<R>* lds_fft(cpx<R>* a, cpx<R>* b, const FftDesc& d, int B, const cpx<R>* __restrict__ tw,
                           int tid, int nth) {
  int Ns = 1;
  const int L = d.L;
  for (int s = 0; s < d.nst; ++s) {
    const int r = d.rad[s];
    switch (r) {
      case 2: stockham_stage<2, SIGN, R>(a, b, L, B, Ns, tw, tid, nth); break;
      case 3: stockham_stage<3, SIGN, R>(a, b, L, B, Ns, tw, tid, nth); break;
      case 4: stockham_stage<4, SIGN, R>(a, b, L, B, Ns, tw, tid, nth); break;
      case 5: stockham_stage<5, SIGN, R>(a, b, L, B, Ns, tw, tid, nth); break;
      default: stockham_stage<8, SIGN, R>(a, b, L, B, Ns, tw, tid, nth); break;
    }
    __syncthreads();
    cpx<R>* t = a; a = b; b = t;
    Ns *= r;
  }
  return a;
}

// 2/3 rule on one axis (NonlinearTerm.jl:134-141): folded |k| >= n ÷ 3
__device__ __forceinline__ bool dealiased1(int i, int n) {
  const int k = i <= n / 2 ? i : n - i;
  return k >= n / 3;
}

// padded-grid index -> original-grid index (inverse of the fftshift embedding of
// FFTWExt.jl:134-149); false when the padded bin is outside the embedded block.
__host__ __device__ __forceinline__ bool pad_pre(int i, int n, int G, int& o) {
  if (G == n) { o = i; return true; }
  if (i < (n + 1) / 2) { o = i; return true; }
  if (i >= G - n / 2) { o = n - (G - i); return true; }
  return false;
}

}  // namespace fit

namespace fit {

// ---------------------------------------------------------------------------------------------
// Register-resident Stockham FFT for compile-time lengths L = 2^a (E = 8 points per thread) or
// 3·2^a (E = 24).  A line is held by T = L/E threads; thread t owns points t + T·e (e < E) before
// and after every stage, so only the Stockham output scatter goes through LDS (double-buffered:
// one barrier per stage).  Radix plan: 8s (when 8 divides E), then 4s (when 4 divides E), 2s,
// then 3s; every radix divides E.
// ---------------------------------------------------------------------------------------------
constexpr int rp_nst(int L, int E) {
  int n = 0, r = L;
  while (r % 8 == 0 && E % 8 == 0) { r /= 8; ++n; }
  while (r % 4 == 0 && E % 4 == 0) { r /= 4; ++n; }
  while (r % 2 == 0) { r /= 2; ++n; }
  while (r % 3 == 0) { r /= 3; ++n; }
  return n;
}
constexpr int rp_rad(int L, int E, int s) {
  int n = 0, r = L;
  while (r % 8 == 0 && E % 8 == 0) { if (n == s) return 8; r /= 8; ++n; }
  while (r % 4 == 0 && E % 4 == 0) { if (n == s) return 4; r /= 4; ++n; }
  while (r % 2 == 0) { if (n == s) return 2; r /= 2; ++n; }
  while (r % 3 == 0) { if (n == s) return 3; r /= 3; ++n; }
  return 1;
}
constexpr bool rp_pow2(int L) { return L > 0 && (L & (L - 1)) == 0; }
constexpr int rp_E(int L) { return rp_pow2(L) ? 8 : 24; }
// points per thread for kernels that keep two lines in registers: radix-4 stages for 3·2^a
// (16 at power-of-two lengths from 1024: a column then spans L/16 lanes, so a 512-thread
// workgroup holds four 2048-point or eight 1024-point columns and reads 32- or 64-byte row
// segments instead of 16 or 32)
constexpr int gE(int L) { return rp_pow2(L) ? (L >= 1024 ? 16 : 8) : 12; }
// points per thread of the fused last pass + density, which also holds E density accumulators:
// at 16 points per lane it spills under its 512-thread bound, so it keeps 8
constexpr int ldE(int L) { return rp_pow2(L) ? 8 : 12; }
// points per thread of the workgroup-form fused x pass (radix-4 stages for 3·2^a: at 768 they
// halve the registers of the radix-8 form and run 10% faster)
constexpr int xE(int L) { return rp_pow2(L) ? 8 : 12; }
// points per thread of the self-advection fused x pass (three velocity lines stay in registers)
constexpr int sE(int L) { return rp_pow2(L) ? 8 : 12; }

// LDS exchange layout.  LDS serves 8-byte accesses 16 lanes per cycle over 32 four-byte banks,
// so an exchange is conflict-free when every aligned group of 16 lanes touches 16 distinct
// 8-byte slots modulo 16.  Point pos of a line sits at pos + C * (pos >> 4): reads of 16
// consecutive points are conflict-free for any C.  The writes of a stage with stride NS and
// radix r are: NS = 1 (lanes r points apart) with C = 1; NS < 16 (groups of NS lanes, NS * r
// points apart) with C = 16 / r, which moves successive groups to distinct slots; NS >= 16
// (runs of 16 consecutive points) with any C, 0 chosen.  lds_line(L, E) holds the largest.
constexpr int lds_pad(int NS, int r) { return NS == 1 ? 1 : (NS < 16 ? 16 / r : 0); }
constexpr int lds_pad_max(int L, int E) {
  int m = 1, ns = 1;
  for (int s = 0; s + 1 < rp_nst(L, E); ++s) {
    const int r = rp_rad(L, E, s);
    m = lds_pad(ns, r) > m ? lds_pad(ns, r) : m;
    ns *= r;
  }
  return m;
}
constexpr int lds_line(int L, int E) { return L + (L / 16) * lds_pad_max(L, E); }

// Stage-major twiddle table: for every stage s >= 1 (stride NS, radix r), entry (k-1)*NS + m at
// offset tws_off(s) holds W_{NS r}^{k m} = tw[k m L / (NS r)], so the lanes of a stage read
// consecutive entries (no bank conflicts).  Sum over stages of (r-1)*NS = L - 1 entries.
constexpr int tws_off(int L, int E, int S) {
  int off = 0, ns = 1;
  for (int s = 0; s < S; ++s) {
    const int r = rp_rad(L, E, s);
    if (s >= 1) off += (r - 1) * ns;
    ns *= r;
  }
  return off;
}
template <int L, int E, int S, int NS, typename R>
__device__ __forceinline__ void tws_build_stage(cpx<R>* dst, const cpx<R>* __restrict__ tw, int tid, int nth) {
  if constexpr (S < rp_nst(L, E)) {
    constexpr int r = rp_rad(L, E, S);
    if constexpr (S >= 1) {
      constexpr int off = tws_off(L, E, S);
      for (int i = tid; i < (r - 1) * NS; i += nth) dst[off + i] = tw[(i / NS + 1) * (i % NS) * (L / (NS * r))];
    }
    tws_build_stage<L, E, S + 1, NS * r, R>(dst, tw, tid, nth);
  }
}
// fill dst[0 .. L-1) from the natural table tw[k] = exp(-2 pi i k / L) (caller synchronises)
template <int L, int E, typename R>
__device__ __forceinline__ void tws_build(cpx<R>* dst, const cpx<R>* __restrict__ tw, int tid, int nth) {
  tws_build_stage<L, E, 0, 1, R>(dst, tw, tid, nth);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): this wave's LDS accesses are complete
  __builtin_amdgcn_wave_barrier();      // no code motion across this point
}
// Exchange barrier: the workgroup's, or (WAVE) only this wave's when every line the wave
// touches is held by lanes of that wave alone (T divides 64 and lines are lane-contiguous).
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) wave_lds_sync();
  else __syncthreads();
}

// TWS: tw is the stage-major table of tws_build instead of the natural one.
template <int L, int E, int S, int NS, int SIGN, bool DB, typename R, class IDX, bool WAVE = false, bool TWS = false>
__device__ __forceinline__ void reg_stage(cpx<R> (&v)[E], cpx<R>* lds0, cpx<R>* lds1, const IDX& idx, int t,
                                          const cpx<R>* __restrict__ tw) {
  if constexpr (S < rp_nst(L, E)) {
    constexpr int r = rp_rad(L, E, S);
    constexpr int T = L / E;
    constexpr int NB = E / r;
    constexpr bool LAST = S + 1 == rp_nst(L, E);
    constexpr int C = lds_pad(NS, r);
    static_assert(E % r == 0, "radix must divide the points per thread");
    // twiddle W_{NS r}^{k m} from either table
    auto twk = [&](int k, unsigned m) {
      if constexpr (TWS) return twiddle<SIGN, R>(tw, (int)(tws_off(L, E, S) + (k - 1) * NS + m));
      else return twiddle<SIGN, R>(tw, (int)(k * m * (L / (NS * r))));
    };
    // m = j mod NS is the same for every butterfly of the thread when NS divides T: its
    // twiddles are loaded once per stage.  Index arithmetic is unsigned (t >= 0).
    constexpr bool MC = NS > 1 && T % NS == 0;
    cpx<R> wc[r];
    if constexpr (MC) {
      const unsigned mc = (unsigned)t % NS;
#pragma unroll
      for (int k = 1; k < r; ++k) wc[k] = twk(k, mc);
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const unsigned j = (unsigned)t + i * T;
      const unsigned m = j % NS;
      cpx<R> w[r];
#pragma unroll
      for (int k = 0; k < r; ++k) w[k] = v[i + k * NB];
      if constexpr (NS > 1) {
#pragma unroll
        for (int k = 1; k < r; ++k) w[k] = cmul(w[k], MC ? wc[k] : twk(k, m));
      }
      dft<r, SIGN, R>(w);
      if constexpr (LAST) {
        // last stage: NS = L/r, so output j + k NS = t + T (i + k NB) is this thread's own slot
#pragma unroll
        for (int k = 0; k < r; ++k) v[i + k * NB] = w[k];
      } else {
        cpx<R>* lds = (DB && (S & 1)) ? lds1 : lds0;
        if constexpr (!DB) { if (i == 0) lds_sync<WAVE>(); }  // single buffer: previous readers done
        const int o = (int)((j - m) * r + m);
        const int io = idx.template at<C>(o);
#pragma unroll
        for (int k = 0; k < r; ++k) lds[idx.template step<NS, r, C>(io, o, k)] = w[k];
      }
    }
    if constexpr (!LAST) {
      cpx<R>* lds = (DB && (S & 1)) ? lds1 : lds0;
      lds_sync<WAVE>();
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = lds[idx.template at<C>(t + T * e)];
      reg_stage<L, E, S + 1, NS * r, SIGN, DB, R, IDX, WAVE, TWS>(v, lds0, lds1, idx, t, tw);
    }
  }
}

// v[e] = x[t + T e] in, X[t + T e] out (natural order).  DB: lds0 / lds1 are two exchange
// buffers (one barrier per stage); otherwise lds0 is used alone (two barriers per stage).  The
// caller must place a barrier between two calls that share the buffers.
// For the long 3·2^a lengths the twiddle base is passed through an empty asm so that the compiler
// cannot hoist the twiddle loads of a transform out of the loops around it: kernels that run
// several transforms per line would otherwise keep every twiddle of every stage live in registers
// for their whole duration (256 VGPRs plus AGPR spill, one wave per SIMD, in the fused x pass at
// 768).  Shorter lines keep the hoisted twiddles (measured faster at 192).
template <int L>
__device__ __forceinline__ void tw_opaque(const void*& tw) {
  if constexpr (L >= 768 && !rp_pow2(L)) asm volatile("" : "+s"(tw));
}
template <int L, int SIGN, bool DB, typename R, class IDX>
__device__ __forceinline__ void reg_fft(cpx<R> (&v)[rp_E(L)], cpx<R>* lds0, cpx<R>* lds1, const IDX& idx, int t,
                                        const cpx<R>* tw) {
  const void* tv = tw;
  tw_opaque<L>(tv);
  reg_stage<L, rp_E(L), 0, 1, SIGN, DB, R, IDX>(v, lds0, lds1, idx, t, (const cpx<R>*)tv);
}
// The same with an explicit number of points per thread (T = L/E threads per line).
template <int L, int E, int SIGN, bool DB, typename R, class IDX>
__device__ __forceinline__ void reg_fft_e(cpx<R> (&v)[E], cpx<R>* lds0, cpx<R>* lds1, const IDX& idx, int t,
                                          const cpx<R>* tw) {
  const void* tv = tw;
  tw_opaque<L>(tv);
  reg_stage<L, E, 0, 1, SIGN, DB, R, IDX>(v, lds0, lds1, idx, t, (const cpx<R>*)tv);
}
// For kernels that run several transforms per lane: FENCE passes the lane index through an empty
// asm first.  Every LDS and twiddle offset of a stage derives from it, and without the fence the
// compiler computes them once and keeps all of them live across every transform of the kernel
// (about 100 VGPRs at E = 12: 171 instead of 90 in the two-line gradient pass at 768, two waves
// per SIMD instead of five).  Recomputing them costs a few integer operations per stage, which
// is a loss where the fence does not raise the occupancy.
// WAVE: the exchanges synchronise the wave only (see lds_sync).
template <int L, int E, int SIGN, bool FENCE, bool WAVE, typename R, class IDX>
__device__ __forceinline__ void reg_fft_m(cpx<R> (&v)[E], cpx<R>* lds, const IDX& idx, int t, const cpx<R>* tw) {
  if constexpr (FENCE) asm volatile("" : "+v"(t));
  const void* tv = tw;
  tw_opaque<L>(tv);
  reg_stage<L, E, 0, 1, SIGN, false, R, IDX, WAVE>(v, lds, lds, idx, t, (const cpx<R>*)tv);
}
// Row-pair lines of T = L/E consecutive lanes are wave-private when T divides the wave size.
template <int L, int E> constexpr bool wave_lines() { return (L / E) <= 64 && 64 % (L / E) == 0; }
// Fused x passes, by measurement on MI355X (FP32 flux of a self-advected field, both forms):
// the gradient-product pass runs fenced above 192 (3.6 vs 4.3 ms at 384, 27.0 vs 27.9 ms at
// 768, 0.50 vs 0.54 ms at 2048, equal elsewhere); the self-advection pass, which holds three
// velocity lines, only at 384, 512 and 3072 (at 768 the fence leaves it at two waves per SIMD
// and costs 1.7 ms of 27).
template <int L> constexpr bool xfence_r() { return L > 192; }
template <int L> constexpr bool xfence_s() { return L == 384 || L == 512 || L == 3072; }

// line-major LDS index with one pad slot every 16 points (breaks power-of-two strides)
// Line-major LDS index: point pos of the line at base + pos + C * (pos >> 4) (see lds_pad); the
// call operator is the C = 1 layout used outside the transforms.  step<NS, r, C>(io, o, k) =
// at<C>(o + k NS) given io = at<C>(o), for the r outputs of one Stockham butterfly: o = (j - m) r
// + m with m = j mod NS, so the outputs fill one aligned block of NS r slots.  Within a 16-point
// group (NS r dividing 16) the pad term is that of o; for NS a multiple of 16 it grows by
// C k NS / 16; otherwise it is recomputed.
struct IdxLine {
  int base;
  template <int C>
  __device__ __forceinline__ int at(int pos) const { return base + pos + C * (pos >> 4); }
  __device__ __forceinline__ int operator()(int pos) const { return at<1>(pos); }
  template <int NS, int r, int C>
  __device__ __forceinline__ int step(int io, int o, int k) const {
    if constexpr (NS % 16 == 0) return io + k * (NS + C * (NS / 16));
    else if constexpr (16 % (NS * r) == 0) return io + k * NS;
    else return at<C>(o + k * NS);
  }
};
// column tiles with a compile-time tile width: every LDS offset of an exchange is the lane's
// base plus an immediate
template <int TX>
struct IdxColC {
  int x;
  template <int C>
  __device__ __forceinline__ int at(int pos) const { return pos * TX + x; }
  __device__ __forceinline__ int operator()(int pos) const { return pos * TX + x; }
  template <int NS, int r, int C>
  __device__ __forceinline__ int step(int io, int o, int k) const { return io + k * NS * TX; }
};
// column tiles: point-major, TX lines interleaved (x fastest in both HBM and LDS)
struct IdxCol {
  int x, TX;
  template <int C>
  __device__ __forceinline__ int at(int pos) const { return pos * TX + x; }
  __device__ __forceinline__ int operator()(int pos) const { return pos * TX + x; }
  template <int NS, int r, int C>
  __device__ __forceinline__ int step(int io, int o, int k) const { return io + k * NS * TX; }
};


}  // namespace fit

namespace fit {

// ---------------------------------------------------------------------------------------------
// Wave-private register FFT: a line of length L is held by T = L/E <= 64 lanes of ONE wavefront,
// so the Stockham exchanges need no workgroup barrier -- only that the wave's own LDS writes
// have completed (s_waitcnt lgkmcnt(0)) before its lanes read them, plus a code-motion fence.
// ---------------------------------------------------------------------------------------------
constexpr int wE(int L) {
  return L == 128 ? 8 : L == 192 ? 24 : L == 256 ? 8 : L == 384 ? 24 : L == 512 ? 8 : L == 768 ? 24 : L == 1024 ? 16
       : L == 1536 ? 24 : L == 2048 ? 32 : L == 3072 ? 48 : 0;
}


template <int L, int E, int S, int NS, int SIGN, typename R, class IDX>
__device__ __forceinline__ void wave_stage(cpx<R> (&v)[E], cpx<R>* lds, const IDX& idx, int t,
                                           const cpx<R>* __restrict__ tw) {
  if constexpr (S < rp_nst(L, E)) {
    constexpr int r = rp_rad(L, E, S);
    constexpr int T = L / E;
    constexpr int NB = E / r;
    constexpr bool LAST = S + 1 == rp_nst(L, E);
    static_assert(E % r == 0, "radix must divide the points per lane");
    if constexpr (!LAST) wave_lds_sync();  // reads of the previous stage are complete
    constexpr bool MC = NS > 1 && T % NS == 0;   // same m (and twiddles) for every butterfly
    cpx<R> wc[r];
    if constexpr (MC) {
      const unsigned mc = (unsigned)t % NS;
#pragma unroll
      for (int k = 1; k < r; ++k) wc[k] = twiddle<SIGN, R>(tw, (int)(k * mc * (L / (NS * r))));
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const unsigned j = (unsigned)t + i * T;
      const unsigned m = j % NS;
      cpx<R> w[r];
#pragma unroll
      for (int k = 0; k < r; ++k) w[k] = v[i + k * NB];
      if constexpr (NS > 1) {
#pragma unroll
        for (int k = 1; k < r; ++k)
          w[k] = cmul(w[k], MC ? wc[k] : twiddle<SIGN, R>(tw, (int)(k * m * (L / (NS * r)))));
      }
      dft<r, SIGN, R>(w);
      if constexpr (LAST) {
#pragma unroll
        for (int k = 0; k < r; ++k) v[i + k * NB] = w[k];   // own slots: no exchange
      } else {
        const int o = (int)((j - m) * r + m);
        const int io = idx.template at<lds_pad(NS, r)>(o);
#pragma unroll
        for (int k = 0; k < r; ++k) lds[idx.template step<NS, r, lds_pad(NS, r)>(io, o, k)] = w[k];
      }
    }
    if constexpr (!LAST) {
      wave_lds_sync();
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = lds[idx.template at<lds_pad(NS, r)>(t + T * e)];
      wave_stage<L, E, S + 1, NS * r, SIGN, R, IDX>(v, lds, idx, t, tw);
    }
  }
}

// v[e] = x[t + T e] in, X[t + T e] out; lds: this line's private region (wave-owned).
template <int L, int SIGN, typename R, class IDX>
__device__ __forceinline__ void wave_fft(cpx<R> (&v)[wE(L)], cpx<R>* lds, const IDX& idx, int t,
                                         const cpx<R>* __restrict__ tw) {
  wave_stage<L, wE(L), 0, 1, SIGN, R, IDX>(v, lds, idx, t, tw);
  wave_lds_sync();
}

}  // namespace fit

namespace fit {

// ---------------------------------------------------------------------------------------------
// Wave-private FFT with the twiddles held in registers.  The twiddles a lane applies depend only
// on its index t, so a kernel that runs many FFTs per lane loads them once (tw_fill) and every
// stage reads them at compile-time offsets: twr[off(S) + i*(r-1) + k-1] = W_{NS r}^{k m},
// m = (t + i T) mod NS, in the forward sign (the inverse uses the conjugate).
// ---------------------------------------------------------------------------------------------
constexpr int tw_off(int L, int E, int S) {
  int off = 0;
  for (int s = 1; s < S; ++s) {
    const int r = rp_rad(L, E, s);
    off += (E / r) * (r - 1);
  }
  return off;
}
constexpr int tw_count(int L, int E) { return tw_off(L, E, rp_nst(L, E)) > 0 ? tw_off(L, E, rp_nst(L, E)) : 1; }

template <int L, int E, int S, int NS, typename R, int NTW>
__device__ __forceinline__ void tw_fill_stage(cpx<R> (&twr)[NTW], int t, const cpx<R>* __restrict__ tw) {
  if constexpr (S < rp_nst(L, E)) {
    constexpr int r = rp_rad(L, E, S);
    constexpr int T = L / E;
    constexpr int NB = E / r;
    if constexpr (S >= 1) {
      constexpr int off = tw_off(L, E, S);
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int m = (t + i * T) % NS;
#pragma unroll
        for (int k = 1; k < r; ++k) twr[off + i * (r - 1) + k - 1] = tw[k * m * (L / (NS * r))];
      }
    }
    tw_fill_stage<L, E, S + 1, NS * r, R, NTW>(twr, t, tw);
  }
}

template <int L, typename R>
__device__ __forceinline__ void tw_fill(cpx<R> (&twr)[tw_count(L, wE(L))], int t, const cpx<R>* __restrict__ tw) {
  tw_fill_stage<L, wE(L), 0, 1, R, tw_count(L, wE(L))>(twr, t, tw);
}

template <int L, int E, int S, int NS, int SIGN, typename R, class IDX, int NTW>
__device__ __forceinline__ void wave_stage_tw(cpx<R> (&v)[E], cpx<R>* lds, const IDX& idx, int t,
                                              const cpx<R> (&twr)[NTW]) {
  if constexpr (S < rp_nst(L, E)) {
    constexpr int r = rp_rad(L, E, S);
    constexpr int T = L / E;
    constexpr int NB = E / r;
    constexpr bool LAST = S + 1 == rp_nst(L, E);
    constexpr int off = tw_off(L, E, S);
    if constexpr (!LAST) wave_lds_sync();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const unsigned j = (unsigned)t + i * T;
      const unsigned m = j % NS;
      cpx<R> w[r];
#pragma unroll
      for (int k = 0; k < r; ++k) w[k] = v[i + k * NB];
      if constexpr (S >= 1) {
#pragma unroll
        for (int k = 1; k < r; ++k) {
          cpx<R> tk = twr[off + i * (r - 1) + k - 1];
          if (SIGN > 0) tk.y = -tk.y;
          w[k] = cmul(w[k], tk);
        }
      }
      dft<r, SIGN, R>(w);
      if constexpr (LAST) {
#pragma unroll
        for (int k = 0; k < r; ++k) v[i + k * NB] = w[k];
      } else {
        const int o = (int)((j - m) * r + m);
        const int io = idx.template at<lds_pad(NS, r)>(o);
#pragma unroll
        for (int k = 0; k < r; ++k) lds[idx.template step<NS, r, lds_pad(NS, r)>(io, o, k)] = w[k];
      }
    }
    if constexpr (!LAST) {
      wave_lds_sync();
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = lds[idx.template at<lds_pad(NS, r)>(t + T * e)];
      wave_stage_tw<L, E, S + 1, NS * r, SIGN, R, IDX, NTW>(v, lds, idx, t, twr);
    }
  }
}

template <int L, int SIGN, typename R, class IDX>
__device__ __forceinline__ void wave_fft_tw(cpx<R> (&v)[wE(L)], cpx<R>* lds, const IDX& idx, int t,
                                            const cpx<R> (&twr)[tw_count(L, wE(L))]) {
  wave_stage_tw<L, wE(L), 0, 1, SIGN, R, IDX, tw_count(L, wE(L))>(v, lds, idx, t, twr);
  wave_lds_sync();
}

}  // namespace fit
