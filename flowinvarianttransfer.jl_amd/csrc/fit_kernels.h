// fit_kernels.h -- HIP kernels of the pseudospectral transfer path (gfx950, wave64).
//
// Data in HBM (compute precision R = float or double):
//   API arrays    : full complex (n1, n2, n3, C), x fastest, components planar (Julia layout)
//   half spectra  : complex (h1, G2, G3), h1 = G1/2 + 1, on the transform grid G (= n, or the
//                   3/2-padded grid); inverse passes run z -> y -> x(c2r), forward x(r2c) -> y -> z
//   physical      : real (G1, G2, G3)
// Missing dimensions (nd < 3) have extent 1.
#pragma once
#include "fit_fft.h"

namespace fit {

// ---------------------------------------------------------------------------------------------
// Geometry shared by loaders / storers
// ---------------------------------------------------------------------------------------------
struct Grid {
  int nd;
  int n1, n2, n3;      // original grid
  int G1, G2, G3;      // transform grid
  int h1;              // G1/2 + 1
  int64_t Np;          // n1 n2 n3
  __device__ __forceinline__ bool dealiased(int ox, int oy, int oz) const {
    if (dealiased1(ox, n1)) return true;
    if (nd >= 2 && dealiased1(oy, n2)) return true;
    if (nd >= 3 && dealiased1(oz, n3)) return true;
    return false;
  }
  __device__ __forceinline__ bool pre(int x, int y, int z, int& ox, int& oy, int& oz) const {
    return pad_pre(x, n1, G1, ox) && pad_pre(y, n2, G2, oy) && pad_pre(z, n3, G3, oz);
  }
  __device__ __forceinline__ int64_t oidx(int ox, int oy, int oz) const {
    return (int64_t)ox + (int64_t)n1 * ((int64_t)oy + (int64_t)n2 * oz);
  }
};

// Workgroups go to the 8 XCDs round-robin by linear id and each XCD has its own L2.  This remaps
// the linear id so that each XCD runs one contiguous range of logical tiles (x tile fastest):
// neighbouring column tiles, which share 128-byte lines, then run on the same XCD close in time
// instead of each XCD fetching whole lines for half of them.
__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz) {
  constexpr unsigned X = 8;
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned n = gx * gy * gridDim.z;
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned q = n / X, r = n % X, xcd = b % X, loc = b / X;
  const unsigned l = xcd < r ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
  bx = (int)(l % gx);
  by = (int)((l / gx) % gy);
  bz = (int)(l / (gx * gy));
}

// One-dimensional form of xcd_tile: the logical workgroup index such that each XCD runs one
// contiguous range of logical indices, in order.
__device__ __forceinline__ int xcd_linear() {
  constexpr unsigned X = 8;
  const unsigned n = gridDim.x, b = blockIdx.x;
  const unsigned q = n / X, r = n % X, xcd = b % X, loc = b / X;
  return (int)(xcd < r ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc);
}

// Folded-index band {i : min(i, G - i) <= a} of one axis (all of it when 2a+1 >= G); used to
// prune FFT passes to the spectral support of their input and the modes their output feeds.
struct Band {
  int a, G;
  __host__ __device__ __forceinline__ bool full() const { return 2 * a + 1 >= G; }
  __host__ __device__ __forceinline__ int count() const { return full() ? G : (a < 0 ? 0 : 2 * a + 1); }
  __host__ __device__ __forceinline__ bool has(int i) const { return full() || i <= a || i >= G - a; }
  __host__ __device__ __forceinline__ int at(int t) const { return (full() || t <= a) ? t : G - (2 * a + 1 - t); }
  // an in-band index for every i (i itself when in band): a load from it stays in the lines the
  // band's own loads touch, so a line can be loaded unconditionally and the result selected
  __host__ __device__ __forceinline__ int clamp(int i) const { return has(i) ? i : (a < 0 ? 0 : a); }
};

// precision conversion of a complex value (load in storage precision, compute in C)
template <typename C, typename V> __device__ __forceinline__ cpx<C> cvt(V v) { return mk<C>((C)v.x, (C)v.y); }

// Deterministic FP64 LDS shell bins.  The lanes of one wave instruction that add to the same LDS
// address are applied in a fixed hardware order, and a wave's instructions in program order; only
// the interleaving of different waves varies from run to run.  So either every wave owns a private
// copy of the bins (priv, summed in wave order when the partial row is written), or the waves
// commit their contributions one after another between workgroup barriers.  With the fixed-order
// cross-workgroup reduction the shell sums are then bitwise reproducible.
struct DetBins {
  double* b;
  int nb, priv;
  __device__ __forceinline__ void zero() const {
    const int n = priv ? nb * (int)(blockDim.x >> 6) : nb;
    for (int s = threadIdx.x; s < n; s += blockDim.x) b[s] = 0.0;
  }
  // f(bins) adds the calling thread's contributions; every thread of the workgroup calls commit
  // the same number of times (it may hold workgroup barriers)
  template <class F> __device__ __forceinline__ void commit(F&& f) const {
    if (priv) {
      f(b + (threadIdx.x >> 6) * nb);
      return;
    }
    const int nw = (int)(blockDim.x >> 6), w = (int)(threadIdx.x >> 6);
    for (int i = 0; i < nw; ++i) {
      if (w == i) f(b);
      __syncthreads();
    }
  }
  // after a barrier: the workgroup's sum of bin s
  __device__ __forceinline__ double total(int s) const {
    if (!priv) return b[s];
    double v = 0.0;
    const int nw = (int)(blockDim.x >> 6);
    for (int w = 0; w < nw; ++w) v += b[w * nb + s];
    return v;
  }
};
// private copies when they fit the budget; the LDS bytes the bins take
__host__ __device__ inline int detbins_priv(int nb, int nthreads, size_t budget) {
  return (size_t)(nthreads / 64) * nb * sizeof(double) <= budget ? 1 : 0;
}
__host__ __device__ inline size_t detbins_bytes(int nb, int nthreads, int priv) {
  return (size_t)(priv ? nthreads / 64 : 1) * (nb > 0 ? nb : 1) * sizeof(double);
}

// ---------------------------------------------------------------------------------------------
// Loaders: value at logical (x, y, z) of field f
// ---------------------------------------------------------------------------------------------

template <typename R> struct LdHalf {
  const cpx<R>* p; int64_t sy, sz, sf;
  __device__ __forceinline__ cpx<R> operator()(int x, int y, int z, int f) const {
    return p[x + y * sy + z * sz + f * sf];
  }
};

template <typename R> struct LdReal {
  const R* p; int64_t sy, sz, sf;
  __device__ __forceinline__ R operator()(int x, int y, int z, int f) const {
    return p[x + y * sy + z * sz + f * sf];
  }
};

// real physical field stored in precision T, read in compute precision R
template <typename R, typename T> struct LdRealAs {
  const T* p; int64_t sy, sz, sf;
  __device__ __forceinline__ R operator()(int x, int y, int z, int f) const {
    return (R)p[x + y * sy + z * sz + f * sf];
  }
};

// Hermitian part, on the transform grid, of  mult(o) * src(o)  where o is the original-grid
// preimage of the bin (zero outside the embedded block) and
//   mult = keep(o) [2/3] * (i ks_j[o_j] if gradj >= 0) * 1[shell(o) == giver] (if giver > 0).
// This is exactly what real.(ifft(...)) of the reference sees (FFTWExt.jl:82-96, 171-177, 256).
template <typename R> struct LdSpectral {
  Grid g;
  const cpx<R>* src; int64_t sf;    // component f at src + f*sf (full original grid)
  int trunc;                        // Orszag 2/3 input truncation
  int gradj;                        // -1: no derivative; else derivative direction
  int gradj_per_field;              // 1: gradj = f % nd  (batched gradient fields)
  int src_per_field_div;            // component index = f / src_per_field_div
  const R* ks0; const R* ks1; const R* ks2;
  const int* shell; int giver;
  // ensembles: component `comp` of a batch lives in snapshot comp / cps at srcs[comp / cps] +
  // (comp % cps)*sf (srcs == nullptr: one array at src); batched field groups g read component
  // c + g*gstride (gstride = 0: every group reads the same field, as the giver shells do)
  const cpx<R>* const* srcs; int cps; int gstride;
  __device__ __forceinline__ const cpx<R>* base(int comp) const {
    return srcs ? srcs[comp / cps] + (int64_t)(comp % cps) * sf : src + (int64_t)comp * sf;
  }
  __device__ __forceinline__ cpx<R> G(int x, int y, int z, int comp, int gj) const {
    int ox, oy, oz;
    cpx<R> z0 = mk<R>((R)0, (R)0);
    if (!g.pre(x, y, z, ox, oy, oz)) return z0;
    if (trunc && g.dealiased(ox, oy, oz)) return z0;
    const int64_t o = g.oidx(ox, oy, oz);
    if (giver > 0 && shell[o] != giver) return z0;
    cpx<R> v = base(comp)[o];
    if (gj >= 0) {
      const R k = gj == 0 ? ks0[ox] : (gj == 1 ? ks1[oy] : ks2[oz]);
      v = mk<R>(-k * v.y, k * v.x);
    }
    return v;
  }
  // masked source value and the wavenumber of axis dk at its original index (0 outside)
  __device__ __forceinline__ cpx<R> Graw(int x, int y, int z, int comp, int dk, R& k) const {
    int ox, oy, oz;
    k = (R)0;
    if (!g.pre(x, y, z, ox, oy, oz)) return mk<R>((R)0, (R)0);
    if (trunc && g.dealiased(ox, oy, oz)) return mk<R>((R)0, (R)0);
    const int64_t o = g.oidx(ox, oy, oz);
    if (giver > 0 && shell[o] != giver) return mk<R>((R)0, (R)0);
    k = dk == 0 ? ks0[ox] : (dk == 1 ? ks1[oy] : ks2[oz]);
    return base(comp)[o];
  }
  __device__ __forceinline__ cpx<R> operator()(int x, int y, int z, int f) const {
    const int comp = f / src_per_field_div;
    const int gj = gradj_per_field ? (f % g.nd) : gradj;
    const cpx<R> a = G(x, y, z, comp, gj);
    const int mx = x == 0 ? 0 : g.G1 - x;
    const int my = y == 0 ? 0 : g.G2 - y;
    const int mz = z == 0 ? 0 : g.G3 - z;
    const cpx<R> b = G(mx, my, mz, comp, gj);
    return mk<R>((R)0.5 * (a.x + b.x), (R)0.5 * (a.y - b.y));
  }
};

// Batched gradient sources: field f = (giver g, component c, direction j), f = (g*M + c)*nd + j.
// Hermitian part of keep·(i ks_j)·1[shell == giver0 + g]·f̂_c on the transform grid.
template <typename R> struct LdSpectralB {
  LdSpectral<R> base;   // geometry, masks, ks, shell map
  int M, giver0;        // giver0 = 0: no shell filter
  __device__ __forceinline__ cpx<R> operator()(int x, int y, int z, int f) const {
    const int nd = base.g.nd;
    const int j = f % nd;
    const int c = (f / nd) % M;
    const int g = f / (nd * M);
    LdSpectral<R> s = base;
    s.giver = giver0 > 0 ? giver0 + g : 0;
    const int comp = c + g * s.gstride;
    const cpx<R> a = s.G(x, y, z, comp, j);
    const int mx = x == 0 ? 0 : s.g.G1 - x;
    const int my = y == 0 ? 0 : s.g.G2 - y;
    const int mz = z == 0 ? 0 : s.g.G3 - z;
    const cpx<R> b = s.G(mx, my, mz, comp, j);
    return mk<R>((R)0.5 * (a.x + b.x), (R)0.5 * (a.y - b.y));
  }
};

// ---------------------------------------------------------------------------------------------
// Storers
// ---------------------------------------------------------------------------------------------

template <typename R> struct StHalf {
  cpx<R>* p; int64_t sy, sz, sf;
  __device__ __forceinline__ void operator()(int x, int y, int z, int f, cpx<R> v) const {
    p[x + y * sy + z * sz + f * sf] = v;
  }
};

template <typename R> struct StReal {
  R* p; int64_t sy, sz, sf; R scale;
  __device__ __forceinline__ void operator()(int x, int y, int z, int f, R v) const {
    p[x + y * sy + z * sz + f * sf] = v * scale;
  }
};

// acc (+)= u_j * scale * v   -- the physical-space product Σ_j u_j ∂_j f_c (FFTWExt.jl:100-109)
template <typename R> struct StProduct {
  R* acc; const R* u; int64_t sy, sz; R scale; int first;
  __device__ __forceinline__ void operator()(int x, int y, int z, int f, R v) const {
    const int64_t i = x + y * sy + z * sz;
    const R g = u[i] * (v * scale);
    acc[i] = first ? g : acc[i] + g;
  }
};

// Final forward pass: transform-grid half-spectrum bin -> full original-grid N̂ (truncation of the
// padded spectrum, FFTWExt.jl:144-149 and :180; 2/3 output mask, :117), times `scale`.
// Every original index is produced exactly once: either from its own bin or, when its bin lies in
// the unstored half, as the conjugate of the mirrored stored bin.
template <typename R> struct StNhat {
  Grid g; cpx<R>* out; int64_t sf; int trunc; R scale;
  __device__ __forceinline__ void put(int ox, int oy, int oz, int f, cpx<R> v) const {
    if (trunc && g.dealiased(ox, oy, oz)) v = mk<R>((R)0, (R)0);
    out[g.oidx(ox, oy, oz) + f * sf] = cscale(v, scale);
  }
  __device__ __forceinline__ void operator()(int x, int y, int z, int f, cpx<R> v) const {
    int ox, oy, oz;
    if (g.pre(x, y, z, ox, oy, oz)) put(ox, oy, oz, f, v);
    const int mx = x == 0 ? 0 : g.G1 - x;
    if (mx > g.G1 / 2) {
      const int my = y == 0 ? 0 : g.G2 - y;
      const int mz = z == 0 ? 0 : g.G3 - z;
      if (g.pre(mx, my, mz, ox, oy, oz)) put(ox, oy, oz, f, cconj(v));
    }
  }
};

// N̂ on the original half grid x <= n1/2 (index x + (n1/2+1)(y + n2 z)): the transform-grid bins
// whose preimage lies there, plus, with padding, the Nyquist plane x = n1/2 reached through the
// mirror of the transform half grid.  With padding an even-length y (z) axis keeps its Nyquist
// index n/2 at padded frequency -n/2 for both k and -k, so there N̂(-k) is not conj N̂(k): the
// bins x > n1/2 of those planes are kept as well (ext, per component w1 = n1 - (n1/2+1) columns:
// y = n2/2 plane w1 × n3, then z = n3/2 plane w1 × n2; ext == nullptr without padding).
template <typename R> struct StNhatHalf {
  Grid g; cpx<R>* out; int64_t sf; int trunc; R scale; cpx<R>* ext; int64_t esf;
  __device__ __forceinline__ void put(int ox, int oy, int oz, int f, cpx<R> v) const {
    if (trunc && g.dealiased(ox, oy, oz)) v = mk<R>((R)0, (R)0);
    out[ox + (int64_t)(g.n1 / 2 + 1) * (oy + (int64_t)g.n2 * oz) + f * sf] = cscale(v, scale);
  }
  __device__ __forceinline__ void operator()(int x, int y, int z, int f, cpx<R> v) const {
    int ox, oy, oz;
    if (g.pre(x, y, z, ox, oy, oz) && 2 * ox <= g.n1) put(ox, oy, oz, f, v);
    const int mx = x == 0 ? 0 : g.G1 - x;
    if (mx > g.G1 / 2) {
      const int my = y == 0 ? 0 : g.G2 - y;
      const int mz = z == 0 ? 0 : g.G3 - z;
      if (g.pre(mx, my, mz, ox, oy, oz)) {
        if (2 * ox <= g.n1) {
          put(ox, oy, oz, f, cconj(v));
        } else if (ext) {
          const int n1h = g.n1 / 2 + 1, w1 = g.n1 - n1h;
          const cpx<R> w = cscale(cconj(v), scale);
          if (g.n2 % 2 == 0 && 2 * oy == g.n2) ext[(ox - n1h) + (int64_t)w1 * oz + f * esf] = w;
          if (g.n3 % 2 == 0 && 2 * oz == g.n3) ext[(int64_t)w1 * g.n3 + (ox - n1h) + (int64_t)w1 * oy + f * esf] = w;
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------------------------
// FFT pass kernels
// ---------------------------------------------------------------------------------------------

// Strided-axis pass: each workgroup transforms TX adjacent columns (consecutive x) of length L
// along y (axis 1, other index = z) or z (axis 2, other index = y).  Pruning: only columns in
// x < nx and other-index band `ob` run; input rows outside `ib` are zero (not loaded); output rows
// outside `wb` are not stored.  With fband != nullptr, field f uses the per-giver folded bounds
// fband[f / fdiv] = (ax, ay, az): whole tiles outside them exit before any barrier.
template <typename R, int SIGN, class LD, class ST>
__global__ void __launch_bounds__(256) k_col(LD ld, ST st, FftDesc d, const cpx<R>* __restrict__ tw,
                                             int nx, int TX, int axis, Band ob, Band ib, Band wb,
                                             const int4* __restrict__ fband, int fdiv, int Gother) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* a = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* b = a + (size_t)d.L * TX;
  const int L = d.L;
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int x0 = bx * TX;
  const int o = ob.at(by);
  const int f = bz;
  if (fband) {
    const int4 fb = fband[f / fdiv];
    if (x0 > fb.x) return;
    if (axis == 2) {
      if (!Band{fb.y, Gother}.has(o)) return;
      ib = Band{fb.z, L};
    } else {
      ib = Band{fb.y, L};
    }
  }
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int e = tid; e < L * TX; e += nth) {
    const int xl = e % TX, i = e / TX, x = x0 + xl;
    cpx<R> v = mk<R>((R)0, (R)0);
    if (x < nx && ib.has(i)) v = axis == 1 ? ld(x, i, o, f) : ld(x, o, i, f);
    a[i * TX + xl] = v;
  }
  __syncthreads();
  cpx<R>* res = lds_fft<SIGN, R>(a, b, d, TX, tw, tid, nth);
  for (int e = tid; e < L * TX; e += nth) {
    const int xl = e % TX, i = e / TX, x = x0 + xl;
    if (x < nx && wb.has(i)) {
      if (axis == 1) st(x, i, o, f, res[i * TX + xl]);
      else st(x, o, i, f, res[i * TX + xl]);
    }
  }
}

// Contiguous-axis inverse pass (c2r): two Hermitian half rows A, B of length h1 are packed into
// Z = A + iB (Hermitian extension to L points), one complex inverse FFT, a = Re z, b = Im z.
// The imaginary parts of the self-conjugate bins (k = 0, L/2) are dropped, as c2r / real() do.
template <typename R, class LD, class ST>
__global__ void __launch_bounds__(256) k_row_inv(LD ld, ST st, FftDesc d, const cpx<R>* __restrict__ tw,
                                                 int G2, int nrows, int P, int kxa) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* a = reinterpret_cast<cpx<R>*>(smem_raw);
  const int L = d.L;
  cpx<R>* b = a + (size_t)L * P;
  const int f = blockIdx.z;
  const int pair0 = blockIdx.x * P;
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int e = tid; e < L * P; e += nth) {
    const int k = e % L, p = e / L;
    const int r0 = 2 * (pair0 + p), r1 = r0 + 1;
    const int kk = k <= L / 2 ? k : L - k;
    cpx<R> A = mk<R>((R)0, (R)0), B = A;
    if (r0 < nrows && kk <= kxa) A = ld(kk, r0 % G2, r0 / G2, f);
    if (r1 < nrows && kk <= kxa) B = ld(kk, r1 % G2, r1 / G2, f);
    if (k == 0 || 2 * k == L) { A.y = (R)0; B.y = (R)0; }
    if (k > L / 2) { A.y = -A.y; B.y = -B.y; }
    a[k * P + p] = mk<R>(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  cpx<R>* res = lds_fft<1, R>(a, b, d, P, tw, tid, nth);
  for (int e = tid; e < L * P; e += nth) {
    const int x = e % L, p = e / L;
    const int r0 = 2 * (pair0 + p), r1 = r0 + 1;
    const cpx<R> z = res[x * P + p];
    if (r0 < nrows) st(x, r0 % G2, r0 / G2, f, z.x);
    if (r1 < nrows) st(x, r1 % G2, r1 / G2, f, z.y);
  }
}

// The same inverse pass with the register FFT: B row pairs per workgroup, T = L/E threads each
// (thread t holds points t + T e), one LDS exchange buffer per line.
template <typename R, int L, class LD, class ST>
__global__ void __launch_bounds__(256) k_row_inv_r(LD ld, ST st, const cpx<R>* __restrict__ tw, int G2,
                                                   int nrows, int B, int kxa) {
  constexpr int E = rp_E(L), T = L / E, LL = lds_line(L, E);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* lds0 = reinterpret_cast<cpx<R>*>(smem_raw);
  const int line = threadIdx.x / T, t = threadIdx.x % T;
  const int f = blockIdx.z;
  const int pair = blockIdx.x * B + line;
  const int r0 = 2 * pair, r1 = r0 + 1;
  const bool ok0 = r0 < nrows, ok1 = r1 < nrows;
  const int y0 = ok0 ? r0 % G2 : 0, z0 = ok0 ? r0 / G2 : 0, y1 = ok1 ? r1 % G2 : 0, z1 = ok1 ? r1 / G2 : 0;
  cpx<R> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + T * e;
    const int kk = k <= L / 2 ? k : L - k;
    cpx<R> A = mk<R>((R)0, (R)0), Bv = A;
    if (kk <= kxa) {
      if (ok0) A = ld(kk, y0, z0, f);
      if (ok1) Bv = ld(kk, y1, z1, f);
    }
    if (k == 0 || 2 * k == L) { A.y = (R)0; Bv.y = (R)0; }
    if (k > L / 2) { A.y = -A.y; Bv.y = -Bv.y; }
    v[e] = mk<R>(A.x - Bv.y, A.y + Bv.x);
  }
  reg_fft<L, 1, false, R>(v, lds0, lds0, IdxLine{line * LL}, t, tw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = t + T * e;
    if (ok0) st(x, y0, z0, f, v[e].x);
    if (ok1) st(x, y1, z1, f, v[e].y);
  }
}

// Contiguous-axis forward pass (r2c) on row pairs: z = a + ib, Z = FFT(z),
// A(k) = (Z(k) + conj Z(L-k))/2,  B(k) = (Z(k) - conj Z(L-k))/(2i),  k = 0..L/2.
template <typename R, class LD, class ST>
__global__ void __launch_bounds__(256) k_row_fwd(LD ld, ST st, FftDesc d, const cpx<R>* __restrict__ tw,
                                                 int G2, int nrows, int P, int h1) {
  // h1 here is the number of output bins stored (kx < h1): the output band of the pass
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* a = reinterpret_cast<cpx<R>*>(smem_raw);
  const int L = d.L;
  cpx<R>* b = a + (size_t)L * P;
  const int f = blockIdx.z;
  const int pair0 = blockIdx.x * P;
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int e = tid; e < L * P; e += nth) {
    const int x = e % L, p = e / L;
    const int r0 = 2 * (pair0 + p), r1 = r0 + 1;
    R va = (R)0, vb = (R)0;
    if (r0 < nrows) va = ld(x, r0 % G2, r0 / G2, f);
    if (r1 < nrows) vb = ld(x, r1 % G2, r1 / G2, f);
    a[x * P + p] = mk<R>(va, vb);
  }
  __syncthreads();
  cpx<R>* res = lds_fft<-1, R>(a, b, d, P, tw, tid, nth);
  for (int e = tid; e < h1 * P; e += nth) {
    const int k = e % h1, p = e / h1;
    const int r0 = 2 * (pair0 + p), r1 = r0 + 1;
    const cpx<R> Zk = res[k * P + p];
    const cpx<R> Zm = res[((L - k) % L) * P + p];
    const cpx<R> A = mk<R>((R)0.5 * (Zk.x + Zm.x), (R)0.5 * (Zk.y - Zm.y));
    const cpx<R> B = mk<R>((R)0.5 * (Zk.y + Zm.y), (R)-0.5 * (Zk.x - Zm.x));
    if (r0 < nrows) st(k, r0 % G2, r0 / G2, f, A);
    if (r1 < nrows) st(k, r1 % G2, r1 / G2, f, B);
  }
}

// ---------------------------------------------------------------------------------------------
// Fused contiguous-axis pass for one field group (giver g, component c) and a block of row pairs:
//   for j < nd:  ∂_j f_c rows = c2r of the half rows of gradient field (fg*nd + j)   [kx <= kxa_g]
//                N += u_j * ∂_j f_c                         (u_j physical, rows read from HBM)
//   then r2c of N and store the half rows kx <= kxb  (N never touches HBM)
// The two rows of a pair ride in the real / imaginary parts of one complex FFT both ways.
// ---------------------------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_xfused(FftDesc d, const cpx<R>* __restrict__ tw,
                                                const cpx<R>* __restrict__ grad, int64_t gsy, int64_t gsz,
                                                int64_t gsf, const R* __restrict__ u, int64_t usy, int64_t usz,
                                                int64_t usf, cpx<R>* __restrict__ out, int64_t osy, int64_t osz,
                                                int64_t osf, int nd, int M, int G2, int nrows, int P,
                                                R gscale, int kxa, const int4* __restrict__ fband, int kxb,
                                                int nfg, int64_t ugs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = d.L;
  cpx<R>* a = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* b = a + (size_t)L * P;
  cpx<R>* acc = b + (size_t)L * P;
  // field group fastest: consecutive workgroups share the same u rows (L2 reuse)
  const int fg = blockIdx.x % nfg;
  u += (int64_t)(fg / M) * ugs;  // ensembles: each batch index advects with its own velocity
  const int pair0 = (blockIdx.x / nfg) * P;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int ka = fband ? min(kxa, fband[fg / M].x) : kxa;
  for (int j = 0; j < nd; ++j) {
    const cpx<R>* gf = grad + (int64_t)(fg * nd + j) * gsf;
    for (int e = tid; e < L * P; e += nth) {
      const int k = e % L, p = e / L;
      const int r0 = 2 * (pair0 + p), r1 = r0 + 1;
      const int kk = k <= L / 2 ? k : L - k;
      cpx<R> A = mk<R>((R)0, (R)0), B = A;
      if (kk <= ka) {
        if (r0 < nrows) A = gf[kk + (r0 % G2) * gsy + (r0 / G2) * gsz];
        if (r1 < nrows) B = gf[kk + (r1 % G2) * gsy + (r1 / G2) * gsz];
      }
      if (k == 0 || 2 * k == L) { A.y = (R)0; B.y = (R)0; }
      if (k > L / 2) { A.y = -A.y; B.y = -B.y; }
      a[k * P + p] = mk<R>(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    cpx<R>* res = lds_fft<1, R>(a, b, d, P, tw, tid, nth);
    const R* uj = u + (int64_t)j * usf;
    for (int e = tid; e < L * P; e += nth) {
      const int x = e % L, p = e / L;
      const int r0 = 2 * (pair0 + p), r1 = r0 + 1;
      const cpx<R> z = res[x * P + p];
      const R u0 = r0 < nrows ? uj[x + (r0 % G2) * usy + (r0 / G2) * usz] : (R)0;
      const R u1 = r1 < nrows ? uj[x + (r1 % G2) * usy + (r1 / G2) * usz] : (R)0;
      const cpx<R> t = mk<R>(u0 * (z.x * gscale), u1 * (z.y * gscale));
      acc[x * P + p] = j == 0 ? t : cadd(acc[x * P + p], t);
    }
    __syncthreads();
  }
  cpx<R>* res = lds_fft<-1, R>(acc, a, d, P, tw, tid, nth);
  cpx<R>* of = out + (int64_t)fg * osf;
  const int nk = kxb + 1;
  for (int e = tid; e < nk * P; e += nth) {
    const int k = e % nk, p = e / nk;
    const int r0 = 2 * (pair0 + p), r1 = r0 + 1;
    const cpx<R> Zk = res[k * P + p];
    const cpx<R> Zm = res[((L - k) % L) * P + p];
    const cpx<R> A = mk<R>((R)0.5 * (Zk.x + Zm.x), (R)0.5 * (Zk.y - Zm.y));
    const cpx<R> B = mk<R>((R)0.5 * (Zk.y + Zm.y), (R)-0.5 * (Zk.x - Zm.x));
    if (r0 < nrows) of[k + (r0 % G2) * osy + (r0 / G2) * osz] = A;
    if (r1 < nrows) of[k + (r1 % G2) * osy + (r1 / G2) * osz] = B;
  }
}

// ---------------------------------------------------------------------------------------------
// Register-FFT variants (compile-time L) of the strided pass and of the fused x pass.
// ---------------------------------------------------------------------------------------------

// Column pass: thread (t, x) with x fastest; TX columns per workgroup; single LDS buffer.
// FP64 workgroups stay within 256 threads, so a lane may hold all 24 complex doubles of its line
// and the FFT temporaries (a 1024-thread bound caps a lane at 128 VGPRs and spills ~100).
// C: the precision the transform computes in (its LDS exchanges and twiddles included); fields are
// loaded and stored in R.  C = double on an FP32 field rounds each value once per pass instead of
// at every butterfly stage.
// FP64 arithmetic holds 12 points per lane at 3·2^a lengths (24 in FP32): 64 lanes per column
// keep four columns (256 threads) in the 48 KB of LDS that three workgroups per CU share.
template <typename R> constexpr int col_threads() { return sizeof(R) == 8 ? 256 : 1024; }
template <typename C> constexpr int colE(int L) { return (sizeof(C) == 8 || L >= 1024) ? gE(L) : rp_E(L); }
template <typename R, int L, int SIGN, class LD, class ST, typename C = R>
__global__ void __launch_bounds__(col_threads<C>()) k_col_r(LD ld, ST st, const cpx<C>* __restrict__ tw, int nx, int TX,
                                                int axis, Band ob, Band ib, Band wb,
                                                const int4* __restrict__ fband, int fdiv, int Gother) {
  constexpr int E = colE<C>(L), T = L / E;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<C>* lds = reinterpret_cast<cpx<C>*>(smem_raw);
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
  const int x0 = bx * TX;
  const int o = ob.at(by);
  const int f = bz;
  if (fband) {
    const int4 fb = fband[f / fdiv];
    if (x0 > fb.x) return;
    if (axis == 2) {
      if (!Band{fb.y, Gother}.has(o)) return;
      ib = Band{fb.z, L};
    } else {
      ib = Band{fb.y, L};
    }
  }
  const int x = x0 + xl;
  cpx<C> v[E];
  if constexpr (SIGN < 0) {
    // forward passes read dense bands: unconditional loads from a clamped column and in-band row,
    // results selected (a load under a divergent branch is waited for at the join, serialising
    // the line's loads: 2.8 -> 2.1 ms at 768 with FP64 arithmetic).  The inverse passes of the
    // giver loop read narrow bands, where skipping the loads wins.
    const int xs = x < nx ? x : nx - 1;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e, ic = ib.clamp(i);
      const cpx<R> a = axis == 1 ? ld(xs, ic, o, f) : ld(xs, o, ic, f);
      v[e] = (x < nx && ib.has(i)) ? cvt<C>(a) : mk<C>((C)0, (C)0);
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      v[e] = mk<C>((C)0, (C)0);
      if (x < nx && ib.has(i)) v[e] = cvt<C>(axis == 1 ? ld(x, i, o, f) : ld(x, o, i, f));
    }
  }
  reg_fft_e<L, E, SIGN, false, C>(v, lds, lds, IdxCol{xl, TX}, t, tw);
  if (x < nx) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      if (wb.has(i)) {
        if (axis == 1) st(x, i, o, f, cvt<R>(v[e]));
        else st(x, o, i, f, cvt<R>(v[e]));
      }
    }
  }
}

// Software-pipelined form of k_col_r: each workgroup walks the tiles ty = blockIdx.y,
// + gridDim.y, ... (< nob) of its columns and issues the loads of the next tile before it
// transforms the current one, so they are in flight during the transform's LDS exchanges (the
// exchange barriers wait on LDS traffic only).  The prefetched line is held in the storage
// precision R; 12 points per lane at 3·2^a lengths keep both lines in registers.
template <typename R, int L, int SIGN, class LD, class ST, typename C = R>
__global__ void __launch_bounds__(512) k_col_p(LD ld, ST st, const cpx<C>* __restrict__ tw, int nx, int TX,
                                               int axis, Band ob, Band ib, Band wb,
                                               const int4* __restrict__ fband, int fdiv, int Gother, int nob) {
  constexpr int E = gE(L), T = L / E;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<C>* lds = reinterpret_cast<cpx<C>*>(smem_raw);
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
  const int x0 = bx * TX;
  const int f = bz;
  bool ochk = false;         // axis 2 with per-field bounds: tiles outside the field's y band skip
  Band obf = ob;
  if (fband) {
    const int4 fb = fband[f / fdiv];
    if (x0 > fb.x) return;
    if (axis == 2) {
      ochk = true;
      obf = Band{fb.y, Gother};
      ib = Band{fb.z, L};
    } else {
      ib = Band{fb.y, L};
    }
  }
  const int x = x0 + xl;
  const int ys = (int)gridDim.y;
  auto next = [&](int ty) {
    while (ty < nob && ochk && !obf.has(ob.at(ty))) ty += ys;
    return ty;
  };
  cpx<R> vn[E];
  const int xs = x < nx ? x : nx - 1;
  auto load = [&](int ty) {
    const int o = ob.at(ty);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e, ic = ib.clamp(i);
      const cpx<R> a = axis == 1 ? ld(xs, ic, o, f) : ld(xs, o, ic, f);
      vn[e] = (x < nx && ib.has(i)) ? a : mk<R>((R)0, (R)0);
    }
  };
  int ty = next(by);
  if (ty >= nob) return;
  load(ty);
  while (true) {
    cpx<C> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cvt<C>(vn[e]);
    const int tn = next(ty + ys);
    if (tn < nob) load(tn);
    reg_fft_e<L, E, SIGN, false, C>(v, lds, lds, IdxCol{xl, TX}, t, tw);
    const int o = ob.at(ty);
    if (x < nx) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = t + T * e;
        if (wb.has(i)) {
          if (axis == 1) st(x, i, o, f, cvt<R>(v[e]));
          else st(x, o, i, f, cvt<R>(v[e]));
        }
      }
    }
    if (tn >= nob) break;
    ty = tn;
  }
}

// ---------------------------------------------------------------------------------------------
// Gradient inverse passes with shared transforms.  The Hermitian part of i k_d·f̂ equals
// i k_t(d)·h, h the Hermitian part of f̂, whenever the wavenumbers are odd under k -> -k on every
// bin the input reaches (no populated self-conjugate Nyquist bin); k_t is the transform-grid
// wavenumber (the original one inside the embedded block, minus the mirror's outside it).  As
// i k_t(d) is constant along every other axis it commutes with their transforms, so per field
// f_c the passes run
//   first pass (last axis a = nd-1):  slot 0 <- F_a^-1 h,   slot a <- F_a^-1 [exact i k_a f̂]
//   y pass (nd = 3):                  slot 0 <- F_y^-1 slot 0,  slot 1 <- F_y^-1 [i k_t(y) slot 0],
//                                     slot 2 <- F_y^-1 slot 2
// and the x pass multiplies slot 0 by i k_t(x): one source load and nd transforms per pass
// instead of nd loads of the source.  Fields are (g, c, j) = (g M + c) nd + j as before.
// Two lines in flight per thread: at most 512 threads per workgroup keeps them in registers.
// ---------------------------------------------------------------------------------------------
// Columns per workgroup of the two-line gradient passes, a compile-time function of the length and
// the arithmetic precision (16 within 512 threads and 64 KB of LDS), so that every LDS offset of
// their exchanges is an immediate.
template <int L, typename C> constexpr int grad_tx() {
  int tx = 512 / (L / gE(L));
  tx = tx > 16 ? 16 : tx;
  while (tx > 1 && (long)L * tx * 2 * (long)sizeof(C) > 65536) tx /= 2;
  return tx;
}

template <typename R, int L, typename C = R>
__global__ void __launch_bounds__(512) k_grad1_r(LdSpectral<R> ls, int M, int giver0, StHalf<R> st,
                                                  const cpx<C>* __restrict__ tw, int nx, int TXr, int axis,
                                                  Band ob, Band ib, Band wb, const int4* __restrict__ fband,
                                                  int Gother) {
  constexpr int E = gE(L), T = L / E, TX = grad_tx<L, C>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<C>* lds = reinterpret_cast<cpx<C>*>(smem_raw);
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
  const int x0 = bx * TX;
  const int o = ob.at(by);
  const int gc = bz;
  const int g = gc / M, c = gc % M;
  if (fband) {
    const int4 fb = fband[g];
    if (x0 > fb.x) return;
    if (axis == 2) {
      if (!Band{fb.y, Gother}.has(o)) return;
      ib = Band{fb.z, L};
    } else {
      ib = Band{fb.y, L};
    }
  }
  LdSpectral<R> s = ls;
  s.giver = giver0 > 0 ? giver0 + g : 0;
  const int nd = s.g.nd;
  const int x = x0 + xl;
  const int mx = x == 0 ? 0 : s.g.G1 - x;
  cpx<C> v0[E], v1[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = t + T * e;
    v0[e] = mk<C>((C)0, (C)0);
    v1[e] = v0[e];
    if (x < nx && ib.has(i)) {
      const int yy = axis == 1 ? i : o, zz = axis == 1 ? o : i;
      const int my = yy == 0 ? 0 : s.g.G2 - yy, mz = zz == 0 ? 0 : s.g.G3 - zz;
      R kar, kbr;
      const cpx<C> a = cvt<C>(s.Graw(x, yy, zz, c + g * s.gstride, axis, kar));
      const cpx<C> b = cvt<C>(s.Graw(mx, my, mz, c + g * s.gstride, axis, kbr));
      const C ka = (C)kar, kb = (C)kbr;
      v0[e] = mk<C>((C)0.5 * (a.x + b.x), (C)0.5 * (a.y - b.y));
      v1[e] = mk<C>((C)-0.5 * (ka * a.y + kb * b.y), (C)0.5 * (ka * a.x - kb * b.x));
    }
  }
  const int f0 = gc * nd, fa = gc * nd + nd - 1;
  reg_fft_m<L, E, 1, true, false, C>(v0, lds, IdxColC<TX>{xl}, t, tw);
  if (x < nx) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      if (wb.has(i)) {
        if (axis == 1) st(x, i, o, f0, cvt<R>(v0[e]));
        else st(x, o, i, f0, cvt<R>(v0[e]));
      }
    }
  }
  __syncthreads();
  reg_fft_m<L, E, 1, true, false, C>(v1, lds, IdxColC<TX>{xl}, t, tw);
  if (x < nx) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      if (wb.has(i)) {
        if (axis == 1) st(x, i, o, fa, cvt<R>(v1[e]));
        else st(x, o, i, fa, cvt<R>(v1[e]));
      }
    }
  }
}

// y pass of the shared-transform gradients (nd = 3), in place on the (g, c, j) fields.
template <typename R, int L, typename C = R>
__global__ void __launch_bounds__(512) k_grad2_r(cpx<R>* __restrict__ buf, int64_t sy, int64_t sz, int64_t sf,
                                                  const R* __restrict__ kyt, const cpx<C>* __restrict__ tw,
                                                  int nx, int TXr, Band ob, Band ib, Band wb,
                                                  const int4* __restrict__ fband, int M, int slot2) {
  constexpr int E = gE(L), T = L / E, TX = grad_tx<L, C>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<C>* lds = reinterpret_cast<cpx<C>*>(smem_raw);
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
  const int x0 = bx * TX;
  const int o = ob.at(by);
  const int gc = bz;
  if (fband) {
    const int4 fb = fband[gc / M];
    if (x0 > fb.x) return;
    ib = Band{fb.y, L};
  }
  const int x = x0 + xl;
  // the lane's row t of slot 0; element e is T*e rows further (a uniform offset)
  cpx<R>* f0 = buf + (int64_t)(gc * 3) * sf + x + o * sz + (int64_t)t * sy;
  cpx<R>* f1 = f0 + sf;
  cpx<R>* f2 = f1 + sf;
  const int64_t ry = (int64_t)T * sy;
  cpx<C> v0[E], v1[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = t + T * e;
    v0[e] = mk<C>((C)0, (C)0);
    if (x < nx && ib.has(i)) v0[e] = cvt<C>(f0[e * ry]);
    const C k = (C)kyt[i];
    v1[e] = mk<C>(-k * v0[e].y, k * v0[e].x);
  }
  reg_fft_m<L, E, 1, true, false, C>(v0, lds, IdxColC<TX>{xl}, t, tw);
  if (x < nx) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      if (wb.has(i)) f0[e * ry] = cvt<R>(v0[e]);
    }
  }
  __syncthreads();
  reg_fft_m<L, E, 1, true, false, C>(v1, lds, IdxColC<TX>{xl}, t, tw);
  if (x < nx) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      if (wb.has(i)) f1[e * ry] = cvt<R>(v1[e]);
    }
  }
  if (!slot2) return;  // slot 2 runs as a plain column pass
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = t + T * e;
    v0[e] = mk<C>((C)0, (C)0);
    if (x < nx && ib.has(i)) v0[e] = cvt<C>(f2[e * ry]);
  }
  __syncthreads();
  reg_fft_m<L, E, 1, true, false, C>(v0, lds, IdxColC<TX>{xl}, t, tw);
  if (x < nx) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      if (wb.has(i)) f2[e * ry] = cvt<R>(v0[e]);
    }
  }
}

// Software-pipelined form of k_grad2_r (see k_col_p): per tile, the slot-2 line and the next
// tile's slot-0 line are loaded before the transforms that precede their use.
template <typename R, int L, typename C = R>
__global__ void __launch_bounds__(512) k_grad2_p(cpx<R>* __restrict__ buf, int64_t sy, int64_t sz, int64_t sf,
                                                  const R* __restrict__ kyt, const cpx<C>* __restrict__ tw,
                                                  int nx, int TX, Band ob, Band ib, Band wb,
                                                  const int4* __restrict__ fband, int M, int slot2, int nob) {
  constexpr int E = gE(L), T = L / E;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<C>* lds = reinterpret_cast<cpx<C>*>(smem_raw);
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
  const int x0 = bx * TX;
  const int gc = bz;
  if (fband) {
    const int4 fb = fband[gc / M];
    if (x0 > fb.x) return;
    ib = Band{fb.y, L};
  }
  const int x = x0 + xl;
  const int ys = (int)gridDim.y;
  cpx<R>* fb0 = buf + (int64_t)(gc * 3) * sf + x;
  cpx<R> s0[E], s2[E];
  const int xs = x < nx ? x : nx - 1;
  auto load = [&](cpx<R> (&d)[E], int slot, int ty) {
    const cpx<R>* src = fb0 + (xs - x) + slot * sf + ob.at(ty) * sz;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      const cpx<R> a = src[(int64_t)ib.clamp(i) * sy];
      d[e] = (x < nx && ib.has(i)) ? a : mk<R>((R)0, (R)0);
    }
  };
  auto store = [&](const cpx<C> (&v)[E], int slot, int ty) {
    cpx<R>* dst = fb0 + slot * sf + ob.at(ty) * sz;
    if (x < nx) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = t + T * e;
        if (wb.has(i)) dst[i * sy] = cvt<R>(v[e]);
      }
    }
  };
  int ty = by;
  if (ty >= nob) return;
  load(s0, 0, ty);
  if (slot2) load(s2, 2, ty);
  while (true) {
    const int tn = ty + ys;
    cpx<C> v0[E], v1[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = t + T * e;
      v0[e] = cvt<C>(s0[e]);
      const C k = (C)kyt[i];
      v1[e] = mk<C>(-k * v0[e].y, k * v0[e].x);
    }
    if (tn < nob) load(s0, 0, tn);
    reg_fft_m<L, E, 1, true, false, C>(v0, lds, IdxCol{xl, TX}, t, tw);
    store(v0, 0, ty);
    reg_fft_m<L, E, 1, true, false, C>(v1, lds, IdxCol{xl, TX}, t, tw);
    store(v1, 1, ty);
    if (slot2) {
#pragma unroll
      for (int e = 0; e < E; ++e) v0[e] = cvt<C>(s2[e]);
      if (tn < nob) load(s2, 2, tn);
      reg_fft_m<L, E, 1, true, false, C>(v0, lds, IdxCol{xl, TX}, t, tw);
      store(v0, 2, ty);
    }
    if (tn >= nob) break;
    ty = tn;
  }
}

// Fused x pass (see k_xfused) with the register FFT: B row pairs per workgroup, T threads each,
// one LDS exchange buffer with workgroup barriers (fallback when the wave form's LDS does not fit).
// Each workgroup runs GCM consecutive field groups of its row block, so the u rows it reads stay
// in its own XCD's L2 instead of being fetched again by neighbouring workgroups on other XCDs.
template <typename R, int L>
__global__ void __launch_bounds__(256) k_xfused_r(const cpx<R>* __restrict__ tw, const cpx<R>* __restrict__ grad,
                                                  int64_t gsy, int64_t gsz, int64_t gsf, const R* __restrict__ u,
                                                  int64_t usy, int64_t usz, int64_t usf, cpx<R>* __restrict__ out,
                                                  int64_t osy, int64_t osz, int64_t osf, int nd, int M, int G2,
                                                  int nrows, int B, R gscale, int kxa,
                                                  const int4* __restrict__ fband, int kxb, int nfg,
                                                  const R* __restrict__ kxt, int64_t ugs, int GCM) {
  constexpr int E = xE(L), T = L / E, LL = lds_line(L, E);
  constexpr bool WV = wave_lines<L, E>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* lds0 = reinterpret_cast<cpx<R>*>(smem_raw);
  const int line = threadIdx.x / T, t = threadIdx.x % T;
  const int nchunk = (nfg + GCM - 1) / GCM;
  // the chunks of one row block read the same velocity rows: run them on one XCD (one L2)
  const int blk = xcd_linear();
  const int chunk = blk % nchunk;
  const int pair = (blk / nchunk) * B + line;
  const int r0 = 2 * pair, r1 = r0 + 1;
  const bool ok0 = r0 < nrows, ok1 = r1 < nrows;
  const int64_t ro0 = ok0 ? (int64_t)(r0 % G2) * gsy + (int64_t)(r0 / G2) * gsz : 0;
  const int64_t ro1 = ok1 ? (int64_t)(r1 % G2) * gsy + (int64_t)(r1 / G2) * gsz : 0;
  const int64_t uo0 = ok0 ? (int64_t)(r0 % G2) * usy + (int64_t)(r0 / G2) * usz : 0;
  const int64_t uo1 = ok1 ? (int64_t)(r1 % G2) * usy + (int64_t)(r1 / G2) * usz : 0;
  const int64_t oo0 = ok0 ? (int64_t)(r0 % G2) * osy + (int64_t)(r0 / G2) * osz : 0;
  const int64_t oo1 = ok1 ? (int64_t)(r1 % G2) * osy + (int64_t)(r1 / G2) * osz : 0;
  const IdxLine idx{line * LL};
  const int fg1 = min(nfg, (chunk + 1) * GCM);
  for (int fg = chunk * GCM; fg < fg1; ++fg) {
    const R* ug = u + (int64_t)(fg / M) * ugs;  // ensembles: each batch index has its own velocity
    const int ka = fband ? min(kxa, fband[fg / M].x) : kxa;
    cpx<R> acc[E];
    for (int j = 0; j < nd; ++j) {
      const cpx<R>* gf = grad + (int64_t)(fg * nd + j) * gsf;
      cpx<R> v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = t + T * e;
        const int kk = k <= L / 2 ? k : L - k;
        cpx<R> A = mk<R>((R)0, (R)0), Bv = A;
        if (kk <= ka) {
          if (ok0) A = gf[kk + ro0];
          if (ok1) Bv = gf[kk + ro1];
        }
        if (kxt && j == 0) {  // slot 0 holds h: the x derivative is i k_t(x) h
          const R kx = kxt[kk];
          A = mk<R>(-kx * A.y, kx * A.x);
          Bv = mk<R>(-kx * Bv.y, kx * Bv.x);
        }
        if (k == 0 || 2 * k == L) { A.y = (R)0; Bv.y = (R)0; }
        if (k > L / 2) { A.y = -A.y; Bv.y = -Bv.y; }
        v[e] = mk<R>(A.x - Bv.y, A.y + Bv.x);
      }
      const R* uj = ug + (int64_t)j * usf;
      if constexpr (WV) {
        // wave-synchronised exchanges: the velocity rows are requested before the transform,
        // which they then overlap (row offsets are 0 past the grid: in-bounds loads, selected
        // to zero).  192-point T(n,m) pass: 2.92 -> 2.78 ms; the workgroup-barrier form waits for
        // them at its first barrier and only pays the registers (2048 points: 1.41 -> 1.51 ms).
        R u0[E], u1[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int xx = t + T * e;
          u0[e] = uj[xx + uo0];
          u1[e] = uj[xx + uo1];
        }
        lds_sync<WV>();  // the previous transform / output loop is done with the LDS buffer
        reg_fft_m<L, E, 1, xfence_r<L>(), WV, R>(v, lds0, idx, t, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const cpx<R> tt = mk<R>((ok0 ? u0[e] : (R)0) * (v[e].x * gscale), (ok1 ? u1[e] : (R)0) * (v[e].y * gscale));
          acc[e] = j == 0 ? tt : cadd(acc[e], tt);
        }
      } else {
        lds_sync<WV>();  // the previous transform / output loop is done with the LDS buffer
        reg_fft_m<L, E, 1, xfence_r<L>(), WV, R>(v, lds0, idx, t, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int xx = t + T * e;
          const R u0 = ok0 ? uj[xx + uo0] : (R)0;
          const R u1 = ok1 ? uj[xx + uo1] : (R)0;
          const cpx<R> tt = mk<R>(u0 * (v[e].x * gscale), u1 * (v[e].y * gscale));
          acc[e] = j == 0 ? tt : cadd(acc[e], tt);
        }
      }
    }
    lds_sync<WV>();
    reg_fft_m<L, E, -1, xfence_r<L>(), WV, R>(acc, lds0, idx, t, tw);
    lds_sync<WV>();
#pragma unroll
    for (int e = 0; e < E; ++e) lds0[idx(t + T * e)] = acc[e];
    lds_sync<WV>();
    cpx<R>* of = out + (int64_t)fg * osf;
    for (int k = t; k <= kxb; k += T) {
      const cpx<R> Zk = lds0[idx(k)];
      const cpx<R> Zm = lds0[idx((L - k) % L)];
      const cpx<R> A = mk<R>((R)0.5 * (Zk.x + Zm.x), (R)0.5 * (Zk.y - Zm.y));
      const cpx<R> Bv = mk<R>((R)0.5 * (Zk.y + Zm.y), (R)-0.5 * (Zk.x - Zm.x));
      if (ok0) of[k + oo0] = A;
      if (ok1) of[k + oo1] = Bv;
    }
  }
}

// Fused x pass of a self-advected field (advecting = advected, shared-transform gradients):
// slot 0 of component j already holds h_j after the y/z inverse passes, so the advecting
// velocity u_j = c2r(h_j) is formed here, per row pair, instead of being stored to HBM and read
// back.  Phase 1: u_j for j < nd into registers (E points per thread, the two rows of the pair
// in the real / imaginary parts).  Phase 2: for each component c, N = Σ_j u_j · c2r(∂_j f_c) with
// ∂_x f_c = c2r(i k_t(x) h_c), then r2c of N into the half rows kx <= kxb.  u and N never touch
// HBM.  B row pairs per workgroup, T = L/E threads each, one LDS exchange buffer.
template <typename R, int L, int E>
__global__ void __launch_bounds__(256) k_xfused_self(const cpx<R>* __restrict__ tw, const cpx<R>* __restrict__ grad,
                                                     int64_t gsy, int64_t gsz, int64_t gsf, cpx<R>* __restrict__ out,
                                                     int64_t osy, int64_t osz, int64_t osf, int nd, int M, int G2,
                                                     int nrows, int B, R gscale, int kxa, int kxb,
                                                     const R* __restrict__ kxt) {
  constexpr int T = L / E, LL = lds_line(L, E);
  constexpr bool WV = wave_lines<L, E>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* lds0 = reinterpret_cast<cpx<R>*>(smem_raw);
  const int line = threadIdx.x / T, t = threadIdx.x % T;
  const int pair = blockIdx.x * B + line;
  const int r0 = 2 * pair, r1 = r0 + 1;
  const bool ok0 = r0 < nrows, ok1 = r1 < nrows;
  const int64_t ro0 = ok0 ? (int64_t)(r0 % G2) * gsy + (int64_t)(r0 / G2) * gsz : 0;
  const int64_t ro1 = ok1 ? (int64_t)(r1 % G2) * gsy + (int64_t)(r1 / G2) * gsz : 0;
  const int64_t oo0 = ok0 ? (int64_t)(r0 % G2) * osy + (int64_t)(r0 / G2) * osz : 0;
  const int64_t oo1 = ok1 ? (int64_t)(r1 % G2) * osy + (int64_t)(r1 / G2) * osz : 0;
  const IdxLine idx{line * LL};
  // packed half rows (kx <= kxa) of field f, optionally times i k_t(x), extended to L points
  auto load = [&](const cpx<R>* gf, bool dx, cpx<R> (&v)[E]) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + T * e;
      const int kk = k <= L / 2 ? k : L - k;
      cpx<R> A = mk<R>((R)0, (R)0), Bv = A;
      if (kk <= kxa) {
        if (ok0) A = gf[kk + ro0];
        if (ok1) Bv = gf[kk + ro1];
      }
      if (dx) {
        const R kx = kxt[kk];
        A = mk<R>(-kx * A.y, kx * A.x);
        Bv = mk<R>(-kx * Bv.y, kx * Bv.x);
      }
      if (k == 0 || 2 * k == L) { A.y = (R)0; Bv.y = (R)0; }
      if (k > L / 2) { A.y = -A.y; Bv.y = -Bv.y; }
      v[e] = mk<R>(A.x - Bv.y, A.y + Bv.x);
    }
  };
  cpx<R> ur[3][E];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j < nd) {
      cpx<R> v[E];
      load(grad + (int64_t)(j * nd) * gsf, false, v);
      lds_sync<WV>();  // the previous transform is done with the LDS buffer
      reg_fft_m<L, E, 1, xfence_s<L>(), WV, R>(v, lds0, idx, t, tw);
#pragma unroll
      for (int e = 0; e < E; ++e) ur[j][e] = mk<R>(v[e].x * gscale, v[e].y * gscale);
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) ur[j][e] = mk<R>((R)0, (R)0);
    }
  }
  for (int c = 0; c < M; ++c) {
    cpx<R> acc[E];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (j < nd) {
        cpx<R> v[E];
        load(grad + (int64_t)(c * nd + j) * gsf, j == 0, v);
        lds_sync<WV>();
        reg_fft_m<L, E, 1, xfence_s<L>(), WV, R>(v, lds0, idx, t, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const cpx<R> tt = mk<R>(ur[j][e].x * (v[e].x * gscale), ur[j][e].y * (v[e].y * gscale));
          acc[e] = j == 0 ? tt : cadd(acc[e], tt);
        }
      }
    }
    lds_sync<WV>();
    reg_fft_m<L, E, -1, xfence_s<L>(), WV, R>(acc, lds0, idx, t, tw);
    lds_sync<WV>();
#pragma unroll
    for (int e = 0; e < E; ++e) lds0[idx(t + T * e)] = acc[e];
    lds_sync<WV>();
    cpx<R>* of = out + (int64_t)c * osf;
    for (int k = t; k <= kxb; k += T) {
      const cpx<R> Zk = lds0[idx(k)];
      const cpx<R> Zm = lds0[idx((L - k) % L)];
      const cpx<R> A = mk<R>((R)0.5 * (Zk.x + Zm.x), (R)0.5 * (Zk.y - Zm.y));
      const cpx<R> Bv = mk<R>((R)0.5 * (Zk.y + Zm.y), (R)-0.5 * (Zk.x - Zm.x));
      if (ok0) of[k + oo0] = A;
      if (ok1) of[k + oo1] = Bv;
    }
  }
}

// LDS DMA: 16 bytes per lane from global memory into LDS at lds_base + 16·lane.  Issued from asm:
// the compiler cannot tell a staging buffer from the other LDS arrays, and for the builtin form
// it waits for the transfer (vmcnt(0)) before every later LDS access, which serialises the
// prefetch with the transforms it should overlap.  The caller waits (vmcnt) before reading.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void lds_dma16(const void* src, unsigned lds_base) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(lds_base)
               : "memory", "m0");
}
#pragma clang diagnostic pop
typedef __attribute__((address_space(3))) void lds_void;

// k_xfused_self for FP32 lines held by one full wave (T = 64: L = 512, 768): the next field's
// two half rows are staged into LDS by DMA while the current line is transformed, so a wave
// never waits on HBM between its twelve transforms.  Twiddles and k_t(x) are copied into LDS
// once per workgroup (a global twiddle load would wait behind the staged rows, as vmcnt retires
// in order).  4 row pairs (waves) per workgroup; exchanges synchronise the wave only.
template <int L, int E>
__global__ void __launch_bounds__(256) k_xfused_self_stg(const cpx<float>* __restrict__ tw,
                                                         const cpx<float>* __restrict__ grad, int64_t gsy,
                                                         int64_t gsz, int64_t gsf, cpx<float>* __restrict__ out,
                                                         int64_t osy, int64_t osz, int64_t osf, int nd, int M,
                                                         int G2, int nrows, float gscale, int kxa, int kxb,
                                                         const float* __restrict__ kxt) {
  using R = float;
  constexpr int T = L / E, LL = lds_line(L, E), B = 256 / T;
  constexpr int SROW = L / 2 + 2;  // staged half row: kx <= L/2 plus one pad point (16-byte units)
  static_assert(T == 64, "one line pair per wave");
  __shared__ __align__(16) cpx<R> xch[B * LL];
  __shared__ __align__(16) cpx<R> stg[B * 2 * SROW];
  __shared__ __align__(16) cpx<R> tws[L];
  __shared__ R kxs[L / 2 + 1];
  tws_build<L, E, R>(tws, tw, threadIdx.x, blockDim.x);   // stage-major: conflict-free twiddle reads
  for (int i = threadIdx.x; i <= L / 2; i += blockDim.x) kxs[i] = kxt[i];
  __syncthreads();
  const int line = threadIdx.x / T, t = threadIdx.x % T;
  const int pair = blockIdx.x * B + line;
  const int r0 = 2 * pair, r1 = r0 + 1;
  const bool ok0 = r0 < nrows, ok1 = r1 < nrows;
  const int64_t ro0 = ok0 ? (int64_t)(r0 % G2) * gsy + (int64_t)(r0 / G2) * gsz : 0;
  const int64_t ro1 = ok1 ? (int64_t)(r1 % G2) * gsy + (int64_t)(r1 / G2) * gsz : 0;
  const int64_t oo0 = ok0 ? (int64_t)(r0 % G2) * osy + (int64_t)(r0 / G2) * osz : 0;
  const int64_t oo1 = ok1 ? (int64_t)(r1 % G2) * osy + (int64_t)(r1 / G2) * osz : 0;
  const IdxLine idx{line * LL};
  cpx<R>* sl = stg + line * 2 * SROW;
  const int nun = (kxa + 2) / 2;  // 16-byte units (two points) covering kx <= kxa
  // the staged points kx > kxa read as zero: those past the last unit are cleared once here (no
  // transfer writes them), the one inside it (kx = kxa + 1 for even kxa) after every transfer
  for (int i = 2 * nun + t; i < SROW; i += T) {
    sl[i] = mk<R>((R)0, (R)0);
    sl[SROW + i] = mk<R>((R)0, (R)0);
  }
  // stage rows r0, r1 of field gf (kx <= kxa; rows start 128-byte aligned, the pitch covers
  // the odd point of the last unit).  A row past the grid stages row 0 (offset 0): finite values
  // whose transform is never stored, so the transforms need no row test.
  auto issue = [&](const cpx<R>* gf) {
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
      const cpx<R>* src = gf + (rr ? ro1 : ro0);
#pragma unroll
      for (int i = 0; i < (SROW / 2 + 63) / 64; ++i) {
        const int ui = i * 64 + t;
        if (ui < nun)
          lds_dma16(src + 2 * ui,
                    __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)(lds_void*)(sl + rr * SROW + i * 128)));
      }
    }
  };
  // the staged rows, extended to L points (times i k_t(x) for dx), packed row pair
  auto consume = [&](bool dx, cpx<R> (&v)[E]) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the staged rows have landed
    __builtin_amdgcn_wave_barrier();
    if (t == 0 && kxa + 1 <= L / 2) {
      sl[kxa + 1] = mk<R>((R)0, (R)0);
      sl[SROW + kxa + 1] = mk<R>((R)0, (R)0);
    }
    wave_lds_sync();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + T * e;
      const int kk = k <= L / 2 ? k : L - k;
      cpx<R> A = sl[kk], Bv = sl[SROW + kk];
      if (dx) {
        const R kx = kxs[kk];
        A = mk<R>(-kx * A.y, kx * A.x);
        Bv = mk<R>(-kx * Bv.y, kx * Bv.x);
      }
      if (k == 0 || 2 * k == L) { A.y = (R)0; Bv.y = (R)0; }
      if (k > L / 2) { A.y = -A.y; Bv.y = -Bv.y; }
      v[e] = mk<R>(A.x - Bv.y, A.y + Bv.x);
    }
    wave_lds_sync();  // staged reads done before the next DMA overwrites them
  };
  auto fft = [&](auto sg, cpx<R>(&v)[E]) {
    int tt = t;
    asm volatile("" : "+v"(tt));
    reg_stage<L, E, 0, 1, decltype(sg)::value, false, R, IdxLine, true, true>(v, xch, xch, idx, tt, tws);
  };
  cpx<R> ur[3][E];
  issue(grad);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j < nd) {
      cpx<R> v[E];
      consume(false, v);
      issue(grad + (int64_t)(j + 1 < nd ? (j + 1) * nd : 0) * gsf);
      fft(std::integral_constant<int, 1>{}, v);
#pragma unroll
      for (int e = 0; e < E; ++e) ur[j][e] = mk<R>(v[e].x * gscale, v[e].y * gscale);
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) ur[j][e] = mk<R>((R)0, (R)0);
    }
  }
  for (int c = 0; c < M; ++c) {
    cpx<R> acc[E];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (j < nd) {
        cpx<R> v[E];
        consume(j == 0, v);
        const int q = c * nd + j + 1;
        if (q < M * nd) issue(grad + (int64_t)q * gsf);
        fft(std::integral_constant<int, 1>{}, v);
        // the gradient's 1/Gp is applied once to the transformed sum (hs below)
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const cpx<R> tt = mk<R>(ur[j][e].x * v[e].x, ur[j][e].y * v[e].y);
          acc[e] = j == 0 ? tt : cadd(acc[e], tt);
        }
      }
    }
    wave_lds_sync();
    fft(std::integral_constant<int, -1>{}, acc);
    wave_lds_sync();
#pragma unroll
    for (int e = 0; e < E; ++e) xch[idx(t + T * e)] = acc[e];
    wave_lds_sync();
    cpx<R>* of = out + (int64_t)c * osf;
    const R hs = (R)0.5 * gscale;
    for (int k = t; k <= kxb; k += T) {
      const cpx<R> Zk = xch[idx(k)];
      const cpx<R> Zm = xch[idx((L - k) % L)];
      const cpx<R> A = mk<R>(hs * (Zk.x + Zm.x), hs * (Zk.y - Zm.y));
      const cpx<R> Bv = mk<R>(hs * (Zk.y + Zm.y), -hs * (Zk.x - Zm.x));
      if (ok0) of[k + oo0] = A;
      if (ok1) of[k + oo1] = Bv;
    }
  }
}

// Fused x pass, wave-per-line form: each wavefront owns 64/T row pairs; a workgroup (4 waves)
// loads its u rows once and then runs GC givers × M components through
//   N = Σ_j u_j · c2r(grad_{g,c,j}),  r2c(N) -> half rows kx <= kxb
// with no workgroup barrier.  Workgroups with the same rows and different giver chunks are
// adjacent in the launch order (giver chunk fastest) so the u rows are shared through L2.
template <typename R, int L>
__global__ void __launch_bounds__(256) k_xfused_w(const cpx<R>* __restrict__ tw, const cpx<R>* __restrict__ grad,
                                                  int64_t gsy, int64_t gsz, int64_t gsf, const R* __restrict__ u,
                                                  int64_t usy, int64_t usz, int64_t usf, cpx<R>* __restrict__ out,
                                                  int64_t osy, int64_t osz, int64_t osf, int nd, int M, int G2,
                                                  int nrows, R gscale, int kxa, const int4* __restrict__ fband,
                                                  int kxb, int ngiv, int GC, const R* __restrict__ kxt,
                                                  int64_t ugs) {
  constexpr int E = wE(L), T = L / E, LPW = 64 / T, LL = lds_line(L, E);
  constexpr bool UREG = E <= 8;                  // keep u in registers for short lines
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* ldsb = reinterpret_cast<cpx<R>*>(smem_raw);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane % T;
  const int line = wave * LPW + lane / T;
  const int nchunk = (ngiv + GC - 1) / GC;
  const int blk = xcd_linear();  // chunks of one row block share velocity rows: one XCD
  const int chunk = blk % nchunk;
  const int pair = (blk / nchunk) * (4 * LPW) + line;
  // ensembles (ugs != 0, launched with GC = 1): the chunk's batch index has its own velocity
  u += (int64_t)(chunk * GC) * ugs;
  const int r0 = 2 * pair, r1 = r0 + 1;
  const bool ok0 = r0 < nrows, ok1 = r1 < nrows;
  const int y0 = ok0 ? r0 % G2 : 0, z0 = ok0 ? r0 / G2 : 0, y1 = ok1 ? r1 % G2 : 0, z1 = ok1 ? r1 / G2 : 0;
  const int64_t ro0 = (int64_t)y0 * gsy + (int64_t)z0 * gsz, ro1 = (int64_t)y1 * gsy + (int64_t)z1 * gsz;
  const int64_t uo0 = (int64_t)y0 * usy + (int64_t)z0 * usz, uo1 = (int64_t)y1 * usy + (int64_t)z1 * usz;
  const int64_t oo0 = (int64_t)y0 * osy + (int64_t)z0 * osz, oo1 = (int64_t)y1 * osy + (int64_t)z1 * osz;
  cpx<R>* lds = ldsb + line * LL;
  const IdxLine idx{0};
  cpx<R> ur[UREG ? 3 : 1][UREG ? E : 1];
  if constexpr (UREG) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const R* uj = u + (int64_t)j * usf;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int xx = t + T * e;
        ur[j][e] = j < nd ? mk<R>(ok0 ? uj[xx + uo0] * gscale : (R)0, ok1 ? uj[xx + uo1] * gscale : (R)0)
                          : mk<R>((R)0, (R)0);
      }
    }
  }
  cpx<R> twr[tw_count(L, E)];
  tw_fill<L, R>(twr, t, tw);
  const int g0 = chunk * GC;
  const int g1 = min(ngiv, g0 + GC);
  for (int g = g0; g < g1; ++g) {
    const int ka = fband ? min(kxa, fband[g].x) : kxa;
    for (int c = 0; c < M; ++c) {
      const int fg = g * M + c;
      cpx<R> acc[E];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if (j >= nd) break;
        const cpx<R>* gf = grad + (int64_t)(fg * nd + j) * gsf;
        cpx<R> v[E];
        if (ka < T) {
          // band below T: only slot 0 (k = t) and slot E-1 (k = L - T + t, |k| = T - t) can hold
          // a nonzero bin; unused rows read row 0 (valid memory) and are zeroed by select
          const cpx<R> z0 = mk<R>((R)0, (R)0);
#pragma unroll
          for (int e = 1; e < E - 1; ++e) v[e] = z0;
          const int ka0 = t <= ka ? t : 0;
          const int ka1 = (t >= 1 && T - t <= ka) ? T - t : 0;
          cpx<R> A0 = gf[ka0 + (ok0 ? ro0 : 0)], B0 = gf[ka0 + (ok1 ? ro1 : 0)];
          cpx<R> A1 = gf[ka1 + (ok0 ? ro0 : 0)], B1 = gf[ka1 + (ok1 ? ro1 : 0)];
          const bool u0 = t <= ka, u1 = t >= 1 && T - t <= ka;
          if (!(u0 && ok0)) A0 = z0;
          if (!(u0 && ok1)) B0 = z0;
          if (!(u1 && ok0)) A1 = z0;
          if (!(u1 && ok1)) B1 = z0;
          if (kxt && j == 0) {  // slot 0 holds h: the x derivative is i k_t(x) h
            const R k0 = kxt[ka0], k1 = kxt[ka1];
            A0 = mk<R>(-k0 * A0.y, k0 * A0.x);
            B0 = mk<R>(-k0 * B0.y, k0 * B0.x);
            A1 = mk<R>(-k1 * A1.y, k1 * A1.x);
            B1 = mk<R>(-k1 * B1.y, k1 * B1.x);
          }
          if (t == 0) { A0.y = (R)0; B0.y = (R)0; }
          // slot E-1 lies above L/2: conjugate of the stored bin T - t
          v[0] = mk<R>(A0.x - B0.y, A0.y + B0.x);
          v[E - 1] = mk<R>(A1.x + B1.y, -A1.y + B1.x);
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int k = t + T * e;
            const int kk = k <= L / 2 ? k : L - k;
            cpx<R> A = mk<R>((R)0, (R)0), Bv = A;
            if (kk <= ka) {
              if (ok0) A = gf[kk + ro0];
              if (ok1) Bv = gf[kk + ro1];
            }
            if (kxt && j == 0) {
              const R kx = kxt[kk];
              A = mk<R>(-kx * A.y, kx * A.x);
              Bv = mk<R>(-kx * Bv.y, kx * Bv.x);
            }
            if (k == 0 || 2 * k == L) { A.y = (R)0; Bv.y = (R)0; }
            if (k > L / 2) { A.y = -A.y; Bv.y = -Bv.y; }
            v[e] = mk<R>(A.x - Bv.y, A.y + Bv.x);
          }
        }
        wave_fft_tw<L, 1, R>(v, lds, idx, t, twr);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          cpx<R> uu;
          if constexpr (UREG) {
            uu = ur[j][e];
          } else {
            const R* uj = u + (int64_t)j * usf;
            const int xx = t + T * e;
            uu = mk<R>(ok0 ? uj[xx + uo0] * gscale : (R)0, ok1 ? uj[xx + uo1] * gscale : (R)0);
          }
          const cpx<R> tt = mk<R>(uu.x * v[e].x, uu.y * v[e].y);
          acc[e] = j == 0 ? tt : cadd(acc[e], tt);
        }
      }
      wave_fft_tw<L, -1, R>(acc, lds, idx, t, twr);
#pragma unroll
      for (int e = 0; e < E; ++e) lds[idx(t + T * e)] = acc[e];
      wave_lds_sync();
      cpx<R>* of = out + (int64_t)fg * osf;
      for (int k = t; k <= kxb; k += T) {
        const cpx<R> Zk = lds[idx(k)];
        const cpx<R> Zm = lds[idx((L - k) % L)];
        const cpx<R> A = mk<R>((R)0.5 * (Zk.x + Zm.x), (R)0.5 * (Zk.y - Zm.y));
        const cpx<R> Bv = mk<R>((R)0.5 * (Zk.y + Zm.y), (R)-0.5 * (Zk.x - Zm.x));
        if (ok0) of[k + oo0] = A;
        if (ok1) of[k + oo1] = Bv;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Fused last forward pass + transfer density + shell sums (non-padded plans, where N̂ is
// Hermitian): for the M component lines of giver g at TX adjacent columns, N̂_c = scale·FFT(line)
// on the output band, t_pair = w(kx)·Re Σ_c conj(eh_c) N̂_c with eh the Hermitian part of the
// masked effective carrier (the pair k, -k at once: w = 2 off the kx = 0, n1/2 planes), and
// FP64 LDS shell bins -> one partial row per workgroup.  N̂ never touches HBM.
// ---------------------------------------------------------------------------------------------
template <typename R, int QMAX>
__global__ void __launch_bounds__(256) k_lastdens(FftDesc d, const cpx<R>* __restrict__ tw,
                                                  const cpx<R>* __restrict__ src, int64_t sy, int64_t sz,
                                                  int64_t sf, int nx, int TX, int axis, Band ob, Band wb, int M,
                                                  R scale, const cpx<R>* __restrict__ eh, int64_t ehsy,
                                                  int64_t ehsz, int64_t ehsf, const int* __restrict__ shh,
                                                  int nsh, int n1, double* __restrict__ partials,
                                                  int64_t ehgs, int dpriv) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = d.L;
  cpx<R>* a = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* b = a + (size_t)L * TX;
  double* bins = reinterpret_cast<double*>(b + (size_t)L * TX);
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int x0 = bx * TX;
  const int o = ob.at(by);
  const int g = bz;
  eh += (int64_t)g * ehgs;  // ensembles: each batch index carries its own carrier
  const int tid = threadIdx.x, nth = blockDim.x;
  const DetBins db{bins, nsh, dpriv};
  db.zero();
  R tq[QMAX];  // TX*L <= QMAX*256 per launch: one density accumulator per owned line point
#pragma unroll
  for (int q = 0; q < QMAX; ++q) tq[q] = (R)0;
  for (int c = 0; c < M; ++c) {
    const cpx<R>* fsrc = src + (int64_t)(g * M + c) * sf;
    for (int e = tid; e < L * TX; e += nth) {
      const int xl = e % TX, i = e / TX, x = x0 + xl;
      cpx<R> v = mk<R>((R)0, (R)0);
      if (x < nx) v = axis == 1 ? fsrc[x + i * sy + o * sz] : fsrc[x + o * sy + i * sz];
      a[i * TX + xl] = v;
    }
    __syncthreads();
    cpx<R>* res = lds_fft<-1, R>(a, b, d, TX, tw, tid, nth);
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int e = tid + q * nth;
      if (e < L * TX) {
        const int xl = e % TX, i = e / TX, x = x0 + xl;
        if (x < nx && wb.has(i)) {
          const cpx<R> nv = cscale(res[i * TX + xl], scale);
          const int64_t hi = axis == 1 ? (x + i * ehsy + o * ehsz) : (x + o * ehsy + i * ehsz);
          const cpx<R> ev = eh[hi + c * ehsf];
          tq[q] += ev.x * nv.x + ev.y * nv.y;
        }
      }
    }
    __syncthreads();  // a / b are refilled by the next component
  }
  db.commit([&](double* bb) {
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int e = tid + q * nth;
      if (e < L * TX) {
        const int xl = e % TX, i = e / TX, x = x0 + xl;
        if (x < nx && wb.has(i)) {
          const int64_t hi = axis == 1 ? (x + i * ehsy + o * ehsz) : (x + o * ehsy + i * ehsz);
          const int sh = shh[hi];
          if (sh > 0) {
            const double w = (x == 0 || 2 * x == n1) ? 1.0 : 2.0;
            atomicAdd(&bb[sh - 1], w * (double)tq[q]);
          }
        }
      }
    }
  });
  __syncthreads();
  const int64_t blk = (int64_t)g * gridDim.x * gridDim.y + (int64_t)by * gridDim.x + bx;
  for (int s = tid; s < nsh; s += nth) partials[blk * nsh + s] = db.total(s);
}

// ---------------------------------------------------------------------------------------------
// Fused last forward pass + transfer density + shell partial sums, register FFT along the last
// axis (z in 3D, y in 2D).  TX adjacent columns per workgroup (x fastest), T = L/E lanes per
// column; the workgroup walks the rows `row = blockIdx.y, + gridDim.y, ...` of the other-index
// band and bins every density into FP64 LDS bins: one partial row per workgroup.  N̂ never
// touches HBM.  Two density forms:
//   kDensFull : the flux of one field (any dealiasing).  A transform-grid half bin gives
//               N̂ = scale·v at its original index o and scale·conj(v) at the original index of its
//               mirrored bin (the StNhat rule, which keeps the reference's 3/2-padding Nyquist
//               placement), 0 where the 2/3 mask drops o; the density of up to 2 invariants
//               (Invariants.jl:40-104) at each such mode from the full carrier and the full
//               shell map.
//   kDensPair : batched giver / snapshot loops on plans without padding (N̂ Hermitian): the pair
//               density t(k) + t(-k) = w(kx)·2 Re Σ_c conj(eh_c) N̂_c, eh the Hermitian part of the
//               masked carrier (w = 1/2 on the kx = 0, n1/2 planes), from a precomputed half-grid
//               eh, or formed here from the snapshot arrays (ensembles: kinetic energy / scalar).
// ---------------------------------------------------------------------------------------------
constexpr int kDensFull = 0, kDensPair = 1;

template <typename R> struct DensArgs {
  Grid g;
  // kDensFull: carrier (M components at car + c*csf, full original grid), invariants, masks
  const cpx<R>* car; int64_t csf; int trunc; int ninv; int inv0, inv1;
  const R* ks0; const R* ks1; const R* ks2;
  const int* shell;                       // full grid (kDensFull) or half grid (kDensPair)
  // kDensPair: precomputed half-grid eh (batch stride ehgs), or srcs != nullptr: snapshot b's
  // array srcs[b] (component stride ssf) with the 2/3 mask when trunc
  const cpx<R>* eh; int64_t ehsy, ehsz, ehsf, ehgs;
  const cpx<R>* const* srcs; int64_t ssf;
};

// component c of the effective carrier e of invariant inv at mode o (t = Re Σ_c conj(e_c) N̂_c):
// ĉ_c for kinetic energy / scalar, (i k × û)_c for helicity, (i k × (i k × û))_c for 3D enstrophy,
// (i k_y ω̂, -i k_x ω̂) with ω̂ = i(k_x û_y - k_y û_x) for 2D enstrophy; 0 for components the
// invariant does not use
template <typename R>
__device__ __forceinline__ cpx<R> carrier_comp(int inv, int nd, int c, const cpx<R>* __restrict__ car, int64_t csf,
                                               int64_t o, R kx, R ky, R kz) {
  if (inv == 0 || inv == 3) return car[o + c * csf];
  if (inv == 2 && nd == 2) {
    if (c > 1) return mk<R>((R)0, (R)0);
    const cpx<R> ux = car[o], uy = car[o + csf];
    const cpx<R> w = cmuli<1>(csub(cscale(uy, kx), cscale(ux, ky)));
    return c == 0 ? cmuli<1>(cscale(w, ky)) : cmuli<-1>(cscale(w, kx));
  }
  if (c > 2) return mk<R>((R)0, (R)0);
  const cpx<R> ux = car[o], uy = car[o + csf], uz = car[o + 2 * csf];
  if (inv == 1) {
    if (c == 0) return cmuli<1>(csub(cscale(uz, ky), cscale(uy, kz)));
    if (c == 1) return cmuli<1>(csub(cscale(ux, kz), cscale(uz, kx)));
    return cmuli<1>(csub(cscale(uy, kx), cscale(ux, ky)));
  }
  const cpx<R> wx = cmuli<1>(csub(cscale(uz, ky), cscale(uy, kz)));
  const cpx<R> wy = cmuli<1>(csub(cscale(ux, kz), cscale(uz, kx)));
  const cpx<R> wz = cmuli<1>(csub(cscale(uy, kx), cscale(ux, ky)));
  if (c == 0) return cmuli<1>(csub(cscale(wz, ky), cscale(wy, kz)));
  if (c == 1) return cmuli<1>(csub(cscale(wx, kz), cscale(wz, kx)));
  return cmuli<1>(csub(cscale(wy, kx), cscale(wx, ky)));
}

template <typename R> constexpr int lastdens_threads() { return sizeof(R) == 8 ? 256 : 512; }

template <typename R, int L, int MODE>
__global__ void __launch_bounds__(lastdens_threads<R>()) k_lastdens_r(
    const cpx<R>* __restrict__ src, int64_t sy, int64_t sz, int64_t sf, const cpx<R>* __restrict__ tw, int nx,
    int TX, int lax, Band ob, Band wb, int M, R scale, DensArgs<R> da, int nbins, int nsh,
    double* __restrict__ partials, int dpriv) {
  constexpr int E = ldE(L), T = L / E;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* bins = reinterpret_cast<double*>(smem_raw);
  const DetBins db{bins, nbins, dpriv};
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(
      smem_raw + ((detbins_bytes(nbins, (int)blockDim.x, dpriv) + 15) / 16) * 16);
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
  const int x = bx * TX + xl;
  const int gi = bz;  // batch index (kDensPair); 0 for kDensFull
  db.zero();
  const Grid& g = da.g;
  const cpx<R>* ehb = da.eh + (int64_t)gi * da.ehgs;
  const cpx<R>* sb = da.srcs ? da.srcs[gi] : nullptr;
  const int n1h = g.n1 / 2 + 1;
  __syncthreads();
  const int nrow = ob.count();
  for (int row = by; row < nrow; row += gridDim.y) {
    const int o = ob.at(row);
    R ta[MODE == kDensFull ? E : 1][4];
    R tq[MODE == kDensPair ? E : 1];
    // line of component c at (x, o), transformed (scaled by the caller)
    auto column = [&](int c, cpx<R> (&v)[E]) {
      const cpx<R>* fs = src + (int64_t)(gi * M + c) * sf;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = t + T * e;
        v[e] = mk<R>((R)0, (R)0);
        if (x < nx) v[e] = lax == 1 ? fs[x + i * sy + o * sz] : fs[x + o * sy + i * sz];
      }
      __syncthreads();  // the previous transform is done with the exchange buffer
      reg_fft_m<L, E, -1, true, false, R>(v, lds, IdxCol{xl, TX}, t, tw);
    };
    if constexpr (MODE == kDensFull) {
      // densities of the direct (d) and mirrored (m) mode of each point, both invariants,
      // accumulated component by component (t is linear in N̂)
#pragma unroll
      for (int e = 0; e < E; ++e) { ta[e][0] = ta[e][1] = ta[e][2] = ta[e][3] = (R)0; }
      for (int c = 0; c < M; ++c) {
        cpx<R> v[E];
        column(c, v);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = t + T * e;
          if (x >= nx || !wb.has(i)) continue;
          const int yy = lax == 1 ? i : o, zz = lax == 1 ? o : i;
          const cpx<R> n = cscale(v[e], scale);
          int ox, oy, oz;
          if (g.pre(x, yy, zz, ox, oy, oz) && !(da.trunc && g.dealiased(ox, oy, oz))) {
            const int64_t om = g.oidx(ox, oy, oz);
            const R kx = da.ks0[ox], ky = g.nd >= 2 ? da.ks1[oy] : (R)0, kz = g.nd >= 3 ? da.ks2[oz] : (R)0;
            const cpx<R> e0 = carrier_comp<R>(da.inv0, g.nd, c, da.car, da.csf, om, kx, ky, kz);
            ta[e][0] += e0.x * n.x + e0.y * n.y;
            if (da.ninv > 1) {
              const cpx<R> e1 = carrier_comp<R>(da.inv1, g.nd, c, da.car, da.csf, om, kx, ky, kz);
              ta[e][1] += e1.x * n.x + e1.y * n.y;
            }
          }
          const int mx = x == 0 ? 0 : g.G1 - x;
          if (mx > g.G1 / 2) {
            const int my = yy == 0 ? 0 : g.G2 - yy, mz = zz == 0 ? 0 : g.G3 - zz;
            if (g.pre(mx, my, mz, ox, oy, oz) && !(da.trunc && g.dealiased(ox, oy, oz))) {
              const int64_t om = g.oidx(ox, oy, oz);
              const R kx = da.ks0[ox], ky = g.nd >= 2 ? da.ks1[oy] : (R)0, kz = g.nd >= 3 ? da.ks2[oz] : (R)0;
              const cpx<R> e0 = carrier_comp<R>(da.inv0, g.nd, c, da.car, da.csf, om, kx, ky, kz);
              ta[e][2] += e0.x * n.x - e0.y * n.y;   // Re{conj(e) conj(n)}
              if (da.ninv > 1) {
                const cpx<R> e1 = carrier_comp<R>(da.inv1, g.nd, c, da.car, da.csf, om, kx, ky, kz);
                ta[e][3] += e1.x * n.x - e1.y * n.y;
              }
            }
          }
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) tq[e] = (R)0;
      for (int c = 0; c < M; ++c) {
        cpx<R> v[E];
        column(c, v);
        if (sb) {
          // eh from the snapshot array: Hermitian part of the masked field at (x, yy, zz); both
          // loads of every point issued unconditionally from a clamped column and in-band row
          const int xs = x < nx ? x : nx - 1, mx = xs == 0 ? 0 : g.n1 - xs;
          cpx<R> a[E], b[E];
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int ic = wb.clamp(t + T * e);
            const int yy = lax == 1 ? ic : o, zz = lax == 1 ? o : ic;
            const int my = yy == 0 ? 0 : g.n2 - yy, mz = zz == 0 ? 0 : g.n3 - zz;
            a[e] = sb[g.oidx(xs, yy, zz) + c * da.ssf];
            b[e] = sb[g.oidx(mx, my, mz) + c * da.ssf];
          }
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int i = t + T * e;
            if (x < nx && wb.has(i)) {
              const int yy = lax == 1 ? i : o, zz = lax == 1 ? o : i;
              const bool keep = !(da.trunc && g.dealiased(x, yy, zz));
              const cpx<R> ev = keep ? mk<R>((R)0.5 * (a[e].x + b[e].x), (R)0.5 * (a[e].y - b[e].y))
                                     : mk<R>((R)0, (R)0);
              const cpx<R> w = cscale(v[e], scale);
              tq[e] += ev.x * w.x + ev.y * w.y;
            }
          }
        } else {
          // the precomputed carrier: every point's load issued unconditionally from a clamped
          // column and in-band row, then selected (a load under a divergent branch is waited for
          // at the join; 0.45 -> 0.36 ms for the 48-giver last pass at 192)
          const int xs = x < nx ? x : nx - 1;
          cpx<R> ev[E];
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int ic = wb.clamp(t + T * e);
            const int yy = lax == 1 ? ic : o, zz = lax == 1 ? o : ic;
            ev[e] = ehb[xs + yy * da.ehsy + zz * da.ehsz + c * da.ehsf];
          }
#pragma unroll
          for (int e = 0; e < E; ++e)
            if (x < nx && wb.has(t + T * e)) {
              const cpx<R> w = cscale(v[e], scale);
              tq[e] += ev[e].x * w.x + ev[e].y * w.y;
            }
        }
      }
    }
    // shell bins
    if constexpr (MODE == kDensPair) {
      if (!dpriv) {
        // shared bins: stage (shell, weighted density) of every point in the exchange buffer
        // (free after the last transform; a point takes at most one complex slot), then wave 0
        // adds them all in point order -- one barrier instead of one per wave
        int* ssh = reinterpret_cast<int*>(lds);
        R* sval = reinterpret_cast<R*>(ssh + L * TX);
        __syncthreads();  // every lane is done reading the exchange buffer
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = t + T * e;
          const int slot = i * TX + xl;
          int shv = 0;
          if (x < nx && wb.has(i)) {
            const int yy = lax == 1 ? i : o, zz = lax == 1 ? o : i;
            shv = da.shell[x + (int64_t)n1h * (yy + (int64_t)g.n2 * zz)];
          }
          ssh[slot] = shv;
          sval[slot] = (x == 0 || 2 * x == g.n1) ? tq[e] : (R)2 * tq[e];
        }
        __syncthreads();
        if (threadIdx.x < 64)
          for (int q = (int)threadIdx.x; q < L * TX; q += 64) {
            const int shv = ssh[q];
            if (shv > 0) atomicAdd(&bins[shv - 1], (double)sval[q]);
          }
        continue;   // the next row's first barrier orders its transform after these reads
      }
    }
    db.commit([&](double* bb) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = t + T * e;
        if (x >= nx || !wb.has(i)) continue;
        const int yy = lax == 1 ? i : o, zz = lax == 1 ? o : i;
        if constexpr (MODE == kDensPair) {
          const int shv = da.shell[x + (int64_t)n1h * (yy + (int64_t)g.n2 * zz)];
          if (shv > 0) {
            const double w = (x == 0 || 2 * x == g.n1) ? 1.0 : 2.0;
            atomicAdd(&bb[shv - 1], w * (double)tq[e]);
          }
        } else {
          int ox, oy, oz;
          if (g.pre(x, yy, zz, ox, oy, oz) && !(da.trunc && g.dealiased(ox, oy, oz))) {
            const int shv = da.shell[g.oidx(ox, oy, oz)];
            if (shv > 0) {
              atomicAdd(&bb[shv - 1], (double)ta[e][0]);
              if (da.ninv > 1) atomicAdd(&bb[nsh + shv - 1], (double)ta[e][1]);
            }
          }
          const int mx = x == 0 ? 0 : g.G1 - x;
          if (mx > g.G1 / 2) {
            const int my = yy == 0 ? 0 : g.G2 - yy, mz = zz == 0 ? 0 : g.G3 - zz;
            if (g.pre(mx, my, mz, ox, oy, oz) && !(da.trunc && g.dealiased(ox, oy, oz))) {
              const int shv = da.shell[g.oidx(ox, oy, oz)];
              if (shv > 0) {
                atomicAdd(&bb[shv - 1], (double)ta[e][2]);
                if (da.ninv > 1) atomicAdd(&bb[nsh + shv - 1], (double)ta[e][3]);
              }
            }
          }
        }
      }
    });
  }
  __syncthreads();
  const int64_t blk = ((int64_t)gi * gridDim.y + by) * gridDim.x + bx;
  for (int s = threadIdx.x; s < nbins; s += blockDim.x) partials[blk * nbins + s] = db.total(s);
}

// Hermitian part of the masked effective carrier on the original half grid (x <= n1/2), only on
// the output band: eh_c(k) = keep(k)·(e_c(k) + conj e_c(-k))/2, e from Invariants.jl:40-104 as
// t = Re Σ_c conj(e_c) N̂_c (KE/scalar: ĉ; helicity: i k×û; enstrophy: i k×(i k×û), 2D form).
// Ensembles: blockIdx.y = snapshot b reads srcs[b] (when srcs != nullptr) and writes eh + b*ehbs.
template <typename R>
__global__ void k_carrier_half(Grid g, int inv, int M, int trunc, const cpx<R>* __restrict__ f, int64_t fsf,
                               const R* __restrict__ ks0, const R* __restrict__ ks1, const R* __restrict__ ks2,
                               Band by, Band bz, int xb, cpx<R>* __restrict__ eh, int64_t ehsf,
                               const cpx<R>* const* __restrict__ srcs, int64_t ehbs) {
  if (srcs) f = srcs[blockIdx.y];
  eh += (int64_t)blockIdx.y * ehbs;
  const int n1h = g.n1 / 2 + 1;
  const int64_t Hn = (int64_t)n1h * g.n2 * g.n3;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t hidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; hidx < Hn; hidx += stride) {
    const int ix = (int)(hidx % n1h), iy = (int)((hidx / n1h) % g.n2), iz = (int)(hidx / ((int64_t)n1h * g.n2));
    if (ix > xb || !by.has(iy) || !bz.has(iz)) continue;
    const int mx = ix == 0 ? 0 : g.n1 - ix, my = iy == 0 ? 0 : g.n2 - iy, mz = iz == 0 ? 0 : g.n3 - iz;
    const bool keep = !(trunc && g.dealiased(ix, iy, iz));
    const int64_t op = g.oidx(ix, iy, iz), om = g.oidx(mx, my, mz);
    const cpx<R> z0 = mk<R>((R)0, (R)0);
    if (inv == 0 || inv == 3) {  // kinetic energy / scalar: every component is its own carrier
      for (int c = 0; c < M; ++c) {
        const cpx<R> a = f[op + c * fsf], b = f[om + c * fsf];
        eh[hidx + c * ehsf] = keep ? mk<R>((R)0.5 * (a.x + b.x), (R)0.5 * (a.y - b.y)) : z0;
      }
      continue;
    }
    // vorticity invariants: the carrier has 3 components (2 for 2D enstrophy); the density
    // ignores the others (Invariants.jl:61-104), so their carrier is zero
    const int ne = (inv == 1 || g.nd == 3) ? 3 : 2;
    cpx<R> ep[3], em[3];
    for (int side = 0; side < 2; ++side) {
      const int64_t o = side == 0 ? op : om;
      const int jx = side == 0 ? ix : mx, jy = side == 0 ? iy : my, jz = side == 0 ? iz : mz;
      cpx<R>* e = side == 0 ? ep : em;
      const R kx = ks0[jx], ky = g.nd >= 2 ? ks1[jy] : (R)0, kz = g.nd >= 3 ? ks2[jz] : (R)0;
      if (ne == 3) {
        const cpx<R> ux = f[o], uy = f[o + fsf], uz = f[o + 2 * fsf];
        cpx<R> wx = cmuli<1>(csub(cscale(uz, ky), cscale(uy, kz)));
        cpx<R> wy = cmuli<1>(csub(cscale(ux, kz), cscale(uz, kx)));
        cpx<R> wz = cmuli<1>(csub(cscale(uy, kx), cscale(ux, ky)));
        if (inv == 1) {
          e[0] = wx; e[1] = wy; e[2] = wz;
        } else {
          e[0] = cmuli<1>(csub(cscale(wz, ky), cscale(wy, kz)));
          e[1] = cmuli<1>(csub(cscale(wx, kz), cscale(wz, kx)));
          e[2] = cmuli<1>(csub(cscale(wy, kx), cscale(wx, ky)));
        }
      } else {
        const cpx<R> ux = f[o], uy = f[o + fsf];
        const cpx<R> w = cmuli<1>(csub(cscale(uy, kx), cscale(ux, ky)));
        e[0] = cmuli<1>(cscale(w, ky));
        e[1] = cmuli<-1>(cscale(w, kx));
        e[2] = z0;
      }
    }
    for (int c = 0; c < M; ++c) {
      cpx<R> v = z0;
      if (keep && c < ne) v = mk<R>((R)0.5 * (ep[c].x + em[c].x), (R)0.5 * (ep[c].y - em[c].y));
      eh[hidx + c * ehsf] = v;
    }
  }
}

// out[g*nsh + s] = Σ_b partials[(g*nblk + b)*nsh + s] in increasing b (deterministic FP64)
// Fixed-order sum of the nblk rows of partials (row-major nblk × nsh; group g = blockIdx.y at
// partials + g·nblk·nsh) into out[g·nsh + s].  A workgroup covers 64 bins × 16 row groups: group y
// adds rows y, y + 16, ... in order, then the 16 group sums are added in order.  The association
// is fixed by the indices alone, so the result does not depend on scheduling; 16 row groups keep
// 16 times more loads in flight than one thread per bin.
constexpr int kRedBins = 64, kRedRows = 16;
static __global__ void __launch_bounds__(kRedBins * kRedRows) k_reduce_rows(const double* __restrict__ partials, int nblk,
                                                                     int nsh, double* __restrict__ out) {
  __shared__ double part[kRedRows][kRedBins];
  const int tx = threadIdx.x % kRedBins, y = threadIdx.x / kRedBins;
  const int s = blockIdx.x * kRedBins + tx;
  const double* p = partials + (int64_t)blockIdx.y * nblk * nsh;
  double acc = 0.0;
  if (s < nsh)
    for (int b = y; b < nblk; b += kRedRows) acc += p[(int64_t)b * nsh + s];
  part[y][tx] = acc;
  __syncthreads();
  if (y == 0 && s < nsh) {
    double t = 0.0;
    for (int k = 0; k < kRedRows; ++k) t += part[k][tx];
    out[(int64_t)blockIdx.y * nsh + s] = t;
  }
}

// ---------------------------------------------------------------------------------------------
// Densities (Invariants.jl:40-104) written as t = Re Σ_c conj(e_c) N̂_c, on the full grid
// ---------------------------------------------------------------------------------------------

template <typename R>
__device__ __forceinline__ R density_at(int inv, int nd, int M, const cpx<R>* car, int64_t csf,
                                        const cpx<R>* nh, int64_t nsf, int64_t o, R kx, R ky, R kz) {
  if (inv == 0 || inv == 3) {  // kinetic energy / passive scalar: Σ_c Re{conj(c_c) N̂_c}
    R t = (R)0;
    for (int c = 0; c < M; ++c) {
      const cpx<R> u = car[o + c * csf], n = nh[o + c * nsf];
      t += u.x * n.x + u.y * n.y;
    }
    return t;
  }
  if (inv == 1) {  // helicity: ω̂ = i k × û
    const cpx<R> ux = car[o], uy = car[o + csf], uz = car[o + 2 * csf];
    const cpx<R> wx = csub(cscale(uz, ky), cscale(uy, kz));
    const cpx<R> wy = csub(cscale(ux, kz), cscale(uz, kx));
    const cpx<R> wz = csub(cscale(uy, kx), cscale(ux, ky));
    // ω = i w;  Re{conj(i w) N} = Re{-i conj(w) N} = Im{conj(w) N} = w.x N.y - w.y N.x
    const cpx<R> n0 = nh[o], n1 = nh[o + nsf], n2 = nh[o + 2 * nsf];
    return (wx.x * n0.y - wx.y * n0.x) + (wy.x * n1.y - wy.y * n1.x) + (wz.x * n2.y - wz.y * n2.x);
  }
  // enstrophy
  if (nd == 2) {
    const cpx<R> ux = car[o], uy = car[o + csf];
    const cpx<R> nx = nh[o], ny = nh[o + nsf];
    const cpx<R> w = csub(cscale(uy, kx), cscale(ux, ky));    // ω̂ = i w
    const cpx<R> v = csub(cscale(ny, kx), cscale(nx, ky));    // N̂_ω = i v
    return w.x * v.x + w.y * v.y;                              // Re{conj(iw) iv} = Re{conj(w) v}
  }
  const cpx<R> ux = car[o], uy = car[o + csf], uz = car[o + 2 * csf];
  const cpx<R> n0 = nh[o], n1 = nh[o + nsf], n2 = nh[o + 2 * nsf];
  const cpx<R> wx = csub(cscale(uz, ky), cscale(uy, kz));
  const cpx<R> wy = csub(cscale(ux, kz), cscale(uz, kx));
  const cpx<R> wz = csub(cscale(uy, kx), cscale(ux, ky));
  const cpx<R> vx = csub(cscale(n2, ky), cscale(n1, kz));
  const cpx<R> vy = csub(cscale(n0, kz), cscale(n2, kx));
  const cpx<R> vz = csub(cscale(n1, kx), cscale(n0, ky));
  return (wx.x * vx.x + wx.y * vx.y) + (wy.x * vy.x + wy.y * vy.y) + (wz.x * vz.x + wz.y * vz.y);
}

// density_at from values already in registers (M <= 3 components of carrier u and N̂ n)
template <typename R>
__device__ __forceinline__ R density_vals(int inv, int nd, int M, const cpx<R> (&u)[3], const cpx<R> (&n)[3], R kx,
                                          R ky, R kz) {
  if (inv == 0 || inv == 3) {
    R t = u[0].x * n[0].x + u[0].y * n[0].y;
    if (M > 1) t += u[1].x * n[1].x + u[1].y * n[1].y;
    if (M > 2) t += u[2].x * n[2].x + u[2].y * n[2].y;
    return t;
  }
  if (inv == 1) {
    const cpx<R> wx = csub(cscale(u[2], ky), cscale(u[1], kz));
    const cpx<R> wy = csub(cscale(u[0], kz), cscale(u[2], kx));
    const cpx<R> wz = csub(cscale(u[1], kx), cscale(u[0], ky));
    return (wx.x * n[0].y - wx.y * n[0].x) + (wy.x * n[1].y - wy.y * n[1].x) + (wz.x * n[2].y - wz.y * n[2].x);
  }
  if (nd == 2) {
    const cpx<R> w = csub(cscale(u[1], kx), cscale(u[0], ky));
    const cpx<R> v = csub(cscale(n[1], kx), cscale(n[0], ky));
    return w.x * v.x + w.y * v.y;
  }
  const cpx<R> wx = csub(cscale(u[2], ky), cscale(u[1], kz));
  const cpx<R> wy = csub(cscale(u[0], kz), cscale(u[2], kx));
  const cpx<R> wz = csub(cscale(u[1], kx), cscale(u[0], ky));
  const cpx<R> vx = csub(cscale(n[2], ky), cscale(n[1], kz));
  const cpx<R> vy = csub(cscale(n[0], kz), cscale(n[2], kx));
  const cpx<R> vz = csub(cscale(n[1], kx), cscale(n[0], ky));
  return (wx.x * vx.x + wx.y * vx.y) + (wy.x * vy.x + wy.y * vy.y) + (wz.x * vz.x + wz.y * vz.y);
}

// density over the full original grid; optional per-mode output t_out; optional shell partial
// sums (FP64 LDS bins, one row of nsh partials per workgroup, reduced later in fixed order).
template <typename R>
__global__ void __launch_bounds__(256) k_density(Grid g, int4 invs, int ninv, int M, const cpx<R>* __restrict__ car,
                                                 int64_t csf, const cpx<R>* __restrict__ nh, int64_t nsf,
                                                 const R* __restrict__ ks0, const R* __restrict__ ks1,
                                                 const R* __restrict__ ks2, const int* __restrict__ shell,
                                                 int nsh, double* __restrict__ partials, R* __restrict__ t_out,
                                                 int dpriv) {
  // ninv invariants (codes invs.x, .y, .z, .w) from one read of carrier and N̂: bins q*nsh + s;
  // t_out (per-mode output) holds the first invariant's density
  extern __shared__ double dbins[];
  const int nb = ninv * nsh;
  const DetBins db{dbins, nb, dpriv};
  if (shell) db.zero();
  __syncthreads();
  // x rows: the row's (y, z) and wavenumbers are found once per row instead of by 64-bit division
  // per mode.  With M <= 3, RP rows run per iteration and every load of all of them (carrier, N̂,
  // shell map) is issued before any arithmetic: RP times the bytes in flight per wave of this
  // latency-bound stream (the bins cap a CU at a few workgroups).
  constexpr int RP = 4;
  const int nrow = g.n2 * g.n3;
  const int rp = M <= 3 ? RP : 1;
  for (int row = blockIdx.x; row < nrow; row += rp * gridDim.x) {
    int rows[RP];
    bool live[RP];
    R ky[RP], kz[RP];
    int64_t base[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const int rr = row + r * (int)gridDim.x;
      live[r] = r < rp && rr < nrow;   // uniform
      rows[r] = live[r] ? rr : row;
      const int iy = rows[r] % g.n2, iz = rows[r] / g.n2;
      ky[r] = g.nd >= 2 ? ks1[iy] : (R)0;
      kz[r] = g.nd >= 3 ? ks2[iz] : (R)0;
      base[r] = (int64_t)rows[r] * g.n1;
    }
    for (int x0 = 0; x0 < g.n1; x0 += blockDim.x) {   // uniform trip count (commit may sync)
      const int ix = x0 + threadIdx.x;
      R tv[RP][4];
      int s[RP];
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        s[r] = 0;
        tv[r][0] = tv[r][1] = tv[r][2] = tv[r][3] = (R)0;
      }
      if (ix < g.n1) {
        const R kx = ks0[ix];
        if (rp == RP) {
          cpx<R> u[RP][3], n[RP][3];
#pragma unroll
          for (int r = 0; r < RP; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const int cc = c < M ? c : 0;
              u[r][c] = car[base[r] + ix + cc * csf];
              n[r][c] = nh[base[r] + ix + cc * nsf];
            }
#pragma unroll
          for (int r = 0; r < RP; ++r) s[r] = shell ? shell[base[r] + ix] : 0;
#pragma unroll
          for (int r = 0; r < RP; ++r)
            for (int q = 0; q < ninv; ++q) {
              const int inv = q == 0 ? invs.x : (q == 1 ? invs.y : (q == 2 ? invs.z : invs.w));
              tv[r][q] = density_vals<R>(inv, g.nd, M, u[r], n[r], kx, ky[r], kz[r]);
            }
        } else {
          const int64_t o = base[0] + ix;
          s[0] = shell ? shell[o] : 0;
          for (int q = 0; q < ninv; ++q) {
            const int inv = q == 0 ? invs.x : (q == 1 ? invs.y : (q == 2 ? invs.z : invs.w));
            tv[0][q] = density_at<R>(inv, g.nd, M, car, csf, nh, nsf, o, kx, ky[0], kz[0]);
          }
        }
        if (t_out)
#pragma unroll
          for (int r = 0; r < RP; ++r)
            if (live[r]) t_out[base[r] + ix] = tv[r][0];
      }
      if (shell)
        db.commit([&](double* bb) {
#pragma unroll
          for (int r = 0; r < RP; ++r)
            if (live[r] && s[r] > 0)
              for (int q = 0; q < ninv; ++q) atomicAdd(&bb[q * nsh + s[r] - 1], (double)tv[r][q]);
        });
    }
  }
  if (shell) {
    __syncthreads();
    for (int s = threadIdx.x; s < nb; s += blockDim.x) partials[(int64_t)blockIdx.x * nb + s] = db.total(s);
  }
}

// Component c of the effective carrier e (t = Re Σ_c conj(e_c) N̂_c) from the carrier's components
// u (M <= 3) at one bin with wavenumbers (kx, ky, kz): carrier_comp on register values.
template <typename R>
__device__ __forceinline__ cpx<R> carrier_vals(int inv, int nd, int c, const cpx<R> (&u)[3], R kx, R ky, R kz) {
  if (inv == 0 || inv == 3) return u[c];
  if (inv == 2 && nd == 2) {
    if (c > 1) return mk<R>((R)0, (R)0);
    const cpx<R> w = cmuli<1>(csub(cscale(u[1], kx), cscale(u[0], ky)));
    return c == 0 ? cmuli<1>(cscale(w, ky)) : cmuli<-1>(cscale(w, kx));
  }
  if (inv == 1) {
    if (c == 0) return cmuli<1>(csub(cscale(u[2], ky), cscale(u[1], kz)));
    if (c == 1) return cmuli<1>(csub(cscale(u[0], kz), cscale(u[2], kx)));
    return cmuli<1>(csub(cscale(u[1], kx), cscale(u[0], ky)));
  }
  const cpx<R> wx = cmuli<1>(csub(cscale(u[2], ky), cscale(u[1], kz)));
  const cpx<R> wy = cmuli<1>(csub(cscale(u[0], kz), cscale(u[2], kx)));
  const cpx<R> wz = cmuli<1>(csub(cscale(u[1], kx), cscale(u[0], ky)));
  if (c == 0) return cmuli<1>(csub(cscale(wz, ky), cscale(wy, kz)));
  if (c == 1) return cmuli<1>(csub(cscale(wx, kz), cscale(wz, kx)));
  return cmuli<1>(csub(cscale(wy, kx), cscale(wx, ky)));
}

// Pair density over the original half grid (N̂ stored there, M <= 3), into the shell of k (which
// −k shares: symmetric shell maps).  On the planes x = 0 and x = n1/2, where −k is a half-grid bin
// of its own, t(k) alone.  Elsewhere t(k) + t(−k) with e(−k) from the carrier at the mirror bin
// and its own wavenumbers and N̂(−k) = conj N̂(k), i.e. Re Σ_c [conj e_c(k) + e_c(−k)] N̂_c(k);
// on the Nyquist rows of a padded grid (ext != nullptr) N̂(−k) comes from ext (StNhatHalf).
// FP64 LDS bins as in k_density; two rows of loads in flight per lane (measured: 1.50 ms for the
// 512³ KE + helicity pair density, against 1.73 ms at four rows and 1.65 ms with one thread per
// half-row bin in 320-thread workgroups).
template <typename R>
__global__ void __launch_bounds__(256) k_density_pair(Grid g, int4 invs, int ninv, int M, const cpx<R>* __restrict__ car,
                                                      int64_t csf, const cpx<R>* __restrict__ nh, int64_t nsf,
                                                      const R* __restrict__ ks0, const R* __restrict__ ks1,
                                                      const R* __restrict__ ks2, const int* __restrict__ shh,
                                                      int nsh, double* __restrict__ partials, int dpriv,
                                                      const cpx<R>* __restrict__ ext, int64_t esf) {
  extern __shared__ double dbins[];
  const int nb = ninv * nsh;
  const DetBins db{dbins, nb, dpriv};
  db.zero();
  __syncthreads();
  constexpr int RP = 2;
  const int n1h = g.n1 / 2 + 1, w1 = g.n1 - n1h;
  const int nrow = g.n2 * g.n3;
  for (int row = blockIdx.x; row < nrow; row += RP * gridDim.x) {
    bool live[RP];
    R ky[RP], kz[RP], kym[RP], kzm[RP];
    int64_t base[RP], mbase[RP], hbase[RP], ebase[RP];
    bool eown[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const int rr = row + r * (int)gridDim.x;
      live[r] = rr < nrow;   // uniform
      const int rw = live[r] ? rr : row;
      const int iy = rw % g.n2, iz = rw / g.n2;
      const int my = iy == 0 ? 0 : g.n2 - iy, mz = iz == 0 ? 0 : g.n3 - iz;
      ky[r] = g.nd >= 2 ? ks1[iy] : (R)0;
      kz[r] = g.nd >= 3 ? ks2[iz] : (R)0;
      kym[r] = g.nd >= 2 ? ks1[my] : (R)0;
      kzm[r] = g.nd >= 3 ? ks2[mz] : (R)0;
      base[r] = (int64_t)rw * g.n1;
      mbase[r] = ((int64_t)my + (int64_t)g.n2 * mz) * g.n1;
      hbase[r] = (int64_t)rw * n1h;
      // Nyquist row of a padded grid: N̂(−k) at ext[ebase + n1 − x]
      eown[r] = false;
      ebase[r] = 0;
      if (ext && g.n2 % 2 == 0 && 2 * iy == g.n2) {
        eown[r] = true;
        ebase[r] = (int64_t)w1 * mz - n1h;
      } else if (ext && g.n3 % 2 == 0 && 2 * iz == g.n3) {
        eown[r] = true;
        ebase[r] = (int64_t)w1 * g.n3 + (int64_t)w1 * my - n1h;
      }
    }
    for (int x0 = 0; x0 < n1h; x0 += blockDim.x) {   // uniform trip count (commit may sync)
      const int ix = x0 + threadIdx.x;
      R tv[RP][4];
      int s[RP];
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        s[r] = 0;
        tv[r][0] = tv[r][1] = tv[r][2] = tv[r][3] = (R)0;
      }
      if (ix < n1h) {
        const int mx = ix == 0 ? 0 : g.n1 - ix;
        const R kx = ks0[ix], kxm = ks0[mx];
        const bool selfc = ix == 0 || 2 * ix == g.n1;
        cpx<R> up[RP][3], um[RP][3], n[RP][3];
#pragma unroll
        for (int r = 0; r < RP; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int cc = c < M ? c : 0;
            up[r][c] = car[base[r] + ix + cc * csf];
            um[r][c] = car[mbase[r] + mx + cc * csf];
            n[r][c] = nh[hbase[r] + ix + cc * nsf];
          }
#pragma unroll
        for (int r = 0; r < RP; ++r) s[r] = shh[hbase[r] + ix];
#pragma unroll
        for (int r = 0; r < RP; ++r) {
          const bool own = !selfc && eown[r];   // uniform in the row but for the two planes
          cpx<R> nm[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) nm[c] = own ? ext[ebase[r] + mx + (c < M ? c : 0) * esf] : mk<R>((R)0, (R)0);
          for (int q = 0; q < ninv; ++q) {
            const int inv = q == 0 ? invs.x : (q == 1 ? invs.y : (q == 2 ? invs.z : invs.w));
            R t = (R)0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              if (c >= M) break;
              const cpx<R> e = carrier_vals<R>(inv, g.nd, c, up[r], kx, ky[r], kz[r]);
              const cpx<R> em = carrier_vals<R>(inv, g.nd, c, um[r], kxm, kym[r], kzm[r]);
              const cpx<R> nc = n[r][c];
              if (selfc) t += e.x * nc.x + e.y * nc.y;                                   // t(k)
              else if (own) t += (e.x * nc.x + e.y * nc.y) + (em.x * nm[c].x + em.y * nm[c].y);
              else t += (e.x + em.x) * nc.x + (e.y - em.y) * nc.y;                   // t(k) + t(−k)
            }
            tv[r][q] = t;
          }
        }
      }
      db.commit([&](double* bb) {
#pragma unroll
        for (int r = 0; r < RP; ++r)
          if (live[r] && s[r] > 0)
            for (int q = 0; q < ninv; ++q) atomicAdd(&bb[q * nsh + s[r] - 1], (double)tv[r][q]);
      });
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nb; s += blockDim.x) partials[(int64_t)blockIdx.x * nb + s] = db.total(s);
}

// ---------------------------------------------------------------------------------------------
// Smooth band-to-band transfer (BandTransfer.jl:20-44, :66-113): graded band weights
// w_n(κ) = exp(-(ln(κ/c_n))²/(2σ²)) / Σ_n' exp(-(ln(κ/c_n'))²/(2σ²)), 0 for κ <= 0, in FP64.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double band_sum(double kap, const double* __restrict__ cen, int nb, double inv2s2) {
  double s = 0.0;
  for (int n = 0; n < nb; ++n) {
    const double r = log(kap / cen[n]);
    s += exp(-r * r * inv2s2);
  }
  return s;
}

// out_c = w_m(κ) f_c on the full original grid: the band-m giver field
template <typename R>
__global__ void k_band_field(int64_t Np, int M, const cpx<R>* __restrict__ f, int64_t fsf, const R* __restrict__ kc,
                             const double* __restrict__ cen, int nb, double inv2s2, int m, cpx<R>* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < Np; o += stride) {
    const double kap = (double)kc[o];
    double w = 0.0;
    if (kap > 0.0) {
      const double sum = band_sum(kap, cen, nb, inv2s2);
      const double r = log(kap / cen[m]);
      if (sum > 0.0) w = exp(-r * r * inv2s2) / sum;
    }
    for (int c = 0; c < M; ++c) out[o + c * Np] = cscale(f[o + c * fsf], (R)w);
  }
}

// band sums Σ_k w_n(k) t(k) of the density t = Re Σ_c conj(e_c) N̂_c over the full grid, for all
// nb bands: FP64 LDS bins, one row of nb partials per workgroup (reduced in fixed order later)
template <typename R>
__global__ void __launch_bounds__(256) k_band_density(Grid g, int inv, int M, const cpx<R>* __restrict__ car,
                                                      int64_t csf, const cpx<R>* __restrict__ nh, int64_t nsf,
                                                      const R* __restrict__ ks0, const R* __restrict__ ks1,
                                                      const R* __restrict__ ks2, const R* __restrict__ kc,
                                                      const double* __restrict__ cen, int nb, double inv2s2,
                                                      double* __restrict__ partials, int dpriv) {
  extern __shared__ double dbins[];
  const DetBins db{dbins, nb, dpriv};
  db.zero();
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t o0 = (int64_t)blockIdx.x * blockDim.x; o0 < g.Np; o0 += stride) {   // uniform trip count
    const int64_t o = o0 + threadIdx.x;
    double t = 0.0, kap = 0.0, sum = 0.0;
    if (o < g.Np) {
      kap = (double)kc[o];
      if (kap > 0.0) {
        const int ix = (int)(o % g.n1);
        const int iy = (int)((o / g.n1) % g.n2);
        const int iz = (int)(o / ((int64_t)g.n1 * g.n2));
        const R kx = ks0[ix];
        const R ky = g.nd >= 2 ? ks1[iy] : (R)0;
        const R kz = g.nd >= 3 ? ks2[iz] : (R)0;
        t = (double)density_at<R>(inv, g.nd, M, car, csf, nh, nsf, o, kx, ky, kz);
        if (t != 0.0) sum = band_sum(kap, cen, nb, inv2s2);
      }
    }
    db.commit([&](double* bb) {
      if (t != 0.0 && sum > 0.0)
        for (int n = 0; n < nb; ++n) {
          const double r = log(kap / cen[n]);
          atomicAdd(&bb[n], exp(-r * r * inv2s2) / sum * t);
        }
    });
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nb; s += blockDim.x) partials[(int64_t)blockIdx.x * nb + s] = db.total(s);
}

// out[s] = Σ_b partials[b][s], b in increasing order (deterministic FP64 reduction)

// ---------------------------------------------------------------------------------------------
// Precision conversion of API arrays
// ---------------------------------------------------------------------------------------------
template <typename A, typename B>
__global__ void k_convert(const A* __restrict__ in, B* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = (B)in[i];
}

// ---------------------------------------------------------------------------------------------
// Band-limited sub-grid: copy the folded box |k_d| <= q_d of C full-complex components from an
// API grid (n0, n1, n2) into a smaller transform grid (L0, L1, L2), zeros elsewhere.  Both grids
// use FFT index order (k >= 0 first, then the negative wavenumbers).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int sub_wavenumber(int i, int L) { return 2 * i <= L ? i : i - L; }

template <typename R>
__global__ void k_subgrid_gather(const cpx<R>* __restrict__ src, int n0, int n1, int n2, int64_t snp,
                                 cpx<R>* __restrict__ dst, int L0, int L1, int L2, int64_t dnp, int q0, int q1,
                                 int q2, int C) {
  const int64_t total = dnp * C;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int c = (int)(e / dnp);
    const int64_t o = e - (int64_t)c * dnp;
    const int i = (int)(o % L0);
    const int64_t r = o / L0;
    const int j = (int)(r % L1), l = (int)(r / L1);
    const int ki = sub_wavenumber(i, L0), kj = sub_wavenumber(j, L1), kl = sub_wavenumber(l, L2);
    cpx<R> v = mk<R>((R)0, (R)0);
    if (abs(ki) <= q0 && abs(kj) <= q1 && abs(kl) <= q2) {
      const int si = ki >= 0 ? ki : ki + n0, sj = kj >= 0 ? kj : kj + n1, sl = kl >= 0 ? kl : kl + n2;
      v = src[(int64_t)c * snp + si + (int64_t)n0 * (sj + (int64_t)n1 * sl)];
    }
    dst[e] = v;
  }
}

// ---------------------------------------------------------------------------------------------
// Mode-to-mode triads.  For giver p (single-mode field entering ifft through its Hermitian part):
//   N̂_p(k) = keep(k) keep(p) / (2 Np²) Σ_j p_j [ i f̂(p) A_j(k-p) - i conj(f̂(p)) A_j(k+p) ],
//   A_j = Hermitian part of keep·â_j,  S(k|p) = Re Σ_c conj(e_c(k)) N̂_{p,c}(k)
// which is what ScaleToScaleTransfer.jl:75-88 evaluates with one full FFT nonlinear term per p.
// ---------------------------------------------------------------------------------------------

// Hermitian part of keep·â_j over the full grid (nd arrays) and effective carrier e_c (M arrays)
template <typename R>
__global__ void k_m2m_prep(Grid g, int trunc, int inv, int M, const cpx<R>* __restrict__ f, int64_t fsf,
                           const cpx<R>* __restrict__ adv, int64_t asf, const R* ks0, const R* ks1,
                           const R* ks2, cpx<R>* __restrict__ Aout, cpx<R>* __restrict__ eout) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < g.Np; o += stride) {
    const int ix = (int)(o % g.n1), iy = (int)((o / g.n1) % g.n2), iz = (int)(o / ((int64_t)g.n1 * g.n2));
    const int mx = ix == 0 ? 0 : g.n1 - ix, my = iy == 0 ? 0 : g.n2 - iy, mz = iz == 0 ? 0 : g.n3 - iz;
    const int64_t om = g.oidx(mx, my, mz);
    const bool kp = !(trunc && g.dealiased(ix, iy, iz));
    const bool km = !(trunc && g.dealiased(mx, my, mz));
    for (int j = 0; Aout && j < g.nd; ++j) {
      const cpx<R> a = kp ? adv[o + j * asf] : mk<R>((R)0, (R)0);
      const cpx<R> b = km ? adv[om + j * asf] : mk<R>((R)0, (R)0);
      Aout[o + j * g.Np] = mk<R>((R)0.5 * (a.x + b.x), (R)0.5 * (a.y - b.y));
    }
    const R kx = ks0[ix], ky = g.nd >= 2 ? ks1[iy] : (R)0, kz = g.nd >= 3 ? ks2[iz] : (R)0;
    if (inv == 0 || inv == 3) {
      for (int c = 0; c < M; ++c) eout[o + c * g.Np] = f[o + c * fsf];
    } else if (inv == 1 || g.nd == 3) {
      const cpx<R> ux = f[o], uy = f[o + fsf], uz = f[o + 2 * fsf];
      cpx<R> wx = csub(cscale(uz, ky), cscale(uy, kz));
      cpx<R> wy = csub(cscale(ux, kz), cscale(uz, kx));
      cpx<R> wz = csub(cscale(uy, kx), cscale(ux, ky));
      wx = cmuli<1>(wx); wy = cmuli<1>(wy); wz = cmuli<1>(wz);                 // ω = i k × u
      if (inv == 1) {
        eout[o] = wx; eout[o + g.Np] = wy; eout[o + 2 * g.Np] = wz;
      } else {                                                                  // e = i k × ω
        const cpx<R> vx = csub(cscale(wz, ky), cscale(wy, kz));
        const cpx<R> vy = csub(cscale(wx, kz), cscale(wz, kx));
        const cpx<R> vz = csub(cscale(wy, kx), cscale(wx, ky));
        eout[o] = cmuli<1>(vx); eout[o + g.Np] = cmuli<1>(vy); eout[o + 2 * g.Np] = cmuli<1>(vz);
      }
    } else {                                                                    // 2D enstrophy
      const cpx<R> ux = f[o], uy = f[o + fsf];
      const cpx<R> w = cmuli<1>(csub(cscale(uy, kx), cscale(ux, ky)));        // ω̂ = i(kx uy - ky ux)
      eout[o] = cmuli<1>(cscale(w, ky));                                        // e_x =  i ky ω̂
      eout[o + g.Np] = cmuli<-1>(cscale(w, kx));                                // e_y = -i kx ω̂
    }
    // vorticity invariants use 3 carrier components (2 for 2D enstrophy): zero the rest
    if (inv == 1 || inv == 2) {
      const int ne = (inv == 1 || g.nd == 3) ? 3 : 2;
      for (int c = ne; c < M; ++c) eout[o + c * g.Np] = mk<R>((R)0, (R)0);
    }
  }
}

// PaddedThreeHalves: A_j on the padded grid G = Hermitian part of the fftshift-embedded â_j
// (FFTWExt.jl:134-141, 171): A(q) = (emb(q) + conj emb(-q)) / 2, emb(q) = â at the original
// index of q when q lies in the embedded block, else 0.  No 2/3 mask on the padded path.
template <typename R>
__global__ void k_m2m_padA(Grid g, const cpx<R>* __restrict__ adv, int64_t asf, cpx<R>* __restrict__ Aout) {
  const int64_t Gp = (int64_t)g.G1 * g.G2 * g.G3;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < Gp; q += stride) {
    const int qx = (int)(q % g.G1), qy = (int)((q / g.G1) % g.G2), qz = (int)(q / ((int64_t)g.G1 * g.G2));
    const int mx = qx == 0 ? 0 : g.G1 - qx, my = qy == 0 ? 0 : g.G2 - qy, mz = qz == 0 ? 0 : g.G3 - qz;
    int ox, oy, oz, px, py, pz;
    const bool in_q = g.pre(qx, qy, qz, ox, oy, oz);
    const bool in_m = g.pre(mx, my, mz, px, py, pz);
    for (int j = 0; j < g.nd; ++j) {
      const cpx<R> a = in_q ? adv[g.oidx(ox, oy, oz) + j * asf] : mk<R>((R)0, (R)0);
      const cpx<R> b = in_m ? adv[g.oidx(px, py, pz) + j * asf] : mk<R>((R)0, (R)0);
      Aout[q + j * Gp] = mk<R>((R)0.5 * (a.x + b.x), (R)0.5 * (a.y - b.y));
    }
  }
}

// Transform-grid index of original index i (the inverse of pad_pre): identity without padding,
// i - n + G for the upper half (the Nyquist bin of an even n lands at -n/2).
__device__ __forceinline__ int m2m_emb(int i, int n, int G) { return i < (n + 1) / 2 ? i : i - n + G; }

// One thread per receiver k, looping over a chunk of givers p.  S written column-major
// S[k + Np p] in OUT precision; FP64 partial net sums per chunk.  The giver's grid indices advance
// incrementally (no integer division in the loop) and the dimension count is a template argument
// so the wavenumber and carrier arrays stay in registers.
__device__ __forceinline__ bool m2m_keep1(int i, int n, int h, int t) { return (i <= h ? i : n - i) < t; }

// With PAD the triad sum runs on the padded grid: k and p are placed by the embedding, k ∓ p are
// taken mod G and A is the padded table (keep = 1, the scale stays 1/(2 Np²)).
template <typename R, typename OUT, int ND, bool PAD>
__global__ void __launch_bounds__(256) k_m2m(Grid g, int trunc, int M, const cpx<R>* __restrict__ f, int64_t fsf,
                                             const cpx<R>* __restrict__ A, const cpx<R>* __restrict__ e,
                                             const R* __restrict__ ks0, const R* __restrict__ ks1,
                                             const R* __restrict__ ks2, int64_t pchunk, OUT* __restrict__ S,
                                             double* __restrict__ netpart) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.Np) return;
  const int64_t p0 = (int64_t)blockIdx.y * pchunk;
  const int64_t p1 = p0 + pchunk < g.Np ? p0 + pchunk : g.Np;
  const int n1 = g.n1, n2 = ND >= 2 ? g.n2 : 1, n3 = ND >= 3 ? g.n3 : 1;
  const int h1 = n1 / 2, h2 = n2 / 2, h3 = n3 / 2, t1 = n1 / 3, t2 = n2 / 3, t3 = n3 / 3;
  const int np = (int)g.Np;
  const int G1 = PAD ? g.G1 : n1, G2 = PAD ? (ND >= 2 ? g.G2 : 1) : n2, G3 = PAD ? (ND >= 3 ? g.G3 : 1) : n3;
  const int ap = G1 * G2 * G3;    // component stride of A
  const int kx = (int)(k % n1), ky = (int)((k / n1) % n2), kz = (int)(k / ((int64_t)n1 * n2));
  const int ekx = PAD ? m2m_emb(kx, n1, G1) : kx, eky = PAD ? m2m_emb(ky, n2, G2) : ky,
            ekz = PAD ? m2m_emb(kz, n3, G3) : kz;
  const bool keepk = !trunc || (m2m_keep1(kx, n1, h1, t1) && (ND < 2 || m2m_keep1(ky, n2, h2, t2)) &&
                                (ND < 3 || m2m_keep1(kz, n3, h3, t3)));
  cpx<R> ek[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) ek[c] = c < M ? e[k + (int64_t)c * np] : mk<R>((R)0, (R)0);
  const R inv2 = (R)(0.5 / ((double)g.Np * (double)g.Np));
  int px = (int)(p0 % n1), py = (int)((p0 / n1) % n2), pz = (int)(p0 / ((int64_t)n1 * n2));
  OUT* __restrict__ sp = S + k + g.Np * p0;
  double net = 0.0;
  // keep(p) along y and z changes only when px wraps
  bool keep_yz = (ND < 2 || m2m_keep1(py, n2, h2, t2)) && (ND < 3 || m2m_keep1(pz, n3, h3, t3));
  for (int64_t p = p0; p < p1; ++p) {
    R s = (R)0;
    const bool keepp = !trunc || (keep_yz && m2m_keep1(px, n1, h1, t1));
    if (keepk && keepp) {
      const int epx = PAD ? m2m_emb(px, n1, G1) : px;
      const int ax = ekx - epx < 0 ? ekx - epx + G1 : ekx - epx;
      const int bx = ekx + epx >= G1 ? ekx + epx - G1 : ekx + epx;
      int om = ax, op = bx;
      if (ND >= 2) {
        const int epy = PAD ? m2m_emb(py, n2, G2) : py;
        const int ay = eky - epy < 0 ? eky - epy + G2 : eky - epy;
        const int by = eky + epy >= G2 ? eky + epy - G2 : eky + epy;
        om += G1 * ay; op += G1 * by;
      }
      if (ND >= 3) {
        const int epz = PAD ? m2m_emb(pz, n3, G3) : pz;
        const int az = ekz - epz < 0 ? ekz - epz + G3 : ekz - epz;
        const int bz = ekz + epz >= G3 ? ekz + epz - G3 : ekz + epz;
        om += G1 * G2 * az; op += G1 * G2 * bz;
      }
      R pj[3];
      pj[0] = ks0[px];
      if (ND >= 2) pj[1] = ks1[py];
      if (ND >= 3) pj[2] = ks2[pz];
      cpx<R> dm = mk<R>((R)0, (R)0), dp = dm;
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        dm = cadd(dm, cscale(A[om + j * ap], pj[j]));
        dp = cadd(dp, cscale(A[op + j * ap], pj[j]));
      }
      const int pi = (int)p;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c < M) {
          const cpx<R> fp = f[pi + c * fsf];
          // i dm fp - i dp conj(fp)
          const cpx<R> n = csub(cmuli<1>(cmul(dm, fp)), cmuli<1>(cmul(dp, cconj(fp))));
          s += ek[c].x * n.x + ek[c].y * n.y;
        }
      }
      for (int c = 4; c < M; ++c) {
        const cpx<R> fp = f[pi + c * fsf];
        const cpx<R> n = csub(cmuli<1>(cmul(dm, fp)), cmuli<1>(cmul(dp, cconj(fp))));
        const cpx<R> ec = e[k + (int64_t)c * np];
        s += ec.x * n.x + ec.y * n.y;
      }
      s *= inv2;
    }
    __builtin_nontemporal_store((OUT)s, sp);      // S is written once and not re-read here
    sp += g.Np;
    net += (double)s;
    if (++px == n1) {
      px = 0;
      if (ND >= 2 && ++py == n2) { py = 0; ++pz; }
      keep_yz = (ND < 2 || m2m_keep1(py, n2, h2, t2)) && (ND < 3 || m2m_keep1(pz, n3, h3, t3));
    }
  }
  netpart[(int64_t)blockIdx.y * g.Np + k] = net;
}

// ---------------------------------------------------------------------------------------------
// Helical split (Decomposition.jl:63-98)
// ---------------------------------------------------------------------------------------------
template <typename R>
__global__ void k_helical(Grid g, const cpx<R>* __restrict__ u, int64_t usf, const R* ks0, const R* ks1,
                          const R* ks2, cpx<R>* __restrict__ up, cpx<R>* __restrict__ um, int64_t osf) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const R is2 = (R)1 / sqrt((R)2);
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < g.Np; o += stride) {
    const int ix = (int)(o % g.n1), iy = (int)((o / g.n1) % g.n2), iz = (int)(o / ((int64_t)g.n1 * g.n2));
    const R kx = ks0[ix], ky = ks1[iy], kz = ks2[iz];
    const R kk = sqrt(kx * kx + ky * ky + kz * kz);
    cpx<R> zero = mk<R>((R)0, (R)0);
    if (kk == (R)0) {
      for (int c = 0; c < 3; ++c) { up[o + c * osf] = zero; um[o + c * osf] = zero; }
      continue;
    }
    const R hx = kx / kk, hy = ky / kk, hz = kz / kk;
    const bool usez = fabs(hz) < (R)0.9;
    const R rx = usez ? (R)0 : (R)1, ry = (R)0, rz = usez ? (R)1 : (R)0;
    R e1x = hy * rz - hz * ry, e1y = hz * rx - hx * rz, e1z = hx * ry - hy * rx;
    const R n1 = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
    e1x /= n1; e1y /= n1; e1z /= n1;
    const R e2x = hy * e1z - hz * e1y, e2y = hz * e1x - hx * e1z, e2z = hx * e1y - hy * e1x;
    const cpx<R> u1 = u[o], u2 = u[o + usf], u3 = u[o + 2 * usf];
    const cpx<R> ue1 = cadd(cadd(cscale(u1, e1x), cscale(u2, e1y)), cscale(u3, e1z));
    const cpx<R> ue2 = cadd(cadd(cscale(u1, e2x), cscale(u2, e2y)), cscale(u3, e2z));
    const cpx<R> upc = cscale(csub(ue1, cmuli<1>(ue2)), is2);
    const cpx<R> umc = cscale(cadd(ue1, cmuli<1>(ue2)), is2);
    const R e1[3] = {e1x, e1y, e1z}, e2[3] = {e2x, e2y, e2z};
    for (int c = 0; c < 3; ++c) {
      const cpx<R> hp = mk<R>(e1[c] * is2, e2[c] * is2);
      const cpx<R> hm = mk<R>(e1[c] * is2, -e2[c] * is2);
      up[o + c * osf] = cmul(upc, hp);
      um[o + c * osf] = cmul(umc, hm);
    }
  }
}


// ---------------------------------------------------------------------------------------------
// Toroidal / poloidal split (Decomposition.jl:107-141): e1 = k̂ × ẑ / |k̂ × ẑ| (horizontal), e2 =
// k̂ × e1; (x̂, ŷ) when k_⊥ = 0; tor = (û·e1) e1, pol = (û·e2) e2; zero at k = 0.
// ---------------------------------------------------------------------------------------------
template <typename R>
__global__ void k_torpol(Grid g, const cpx<R>* __restrict__ u, int64_t usf, const R* ks0, const R* ks1,
                         const R* ks2, cpx<R>* __restrict__ tor, cpx<R>* __restrict__ pol, int64_t osf) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < g.Np; o += stride) {
    const int ix = (int)(o % g.n1), iy = (int)((o / g.n1) % g.n2), iz = (int)(o / ((int64_t)g.n1 * g.n2));
    const R kx = ks0[ix], ky = ks1[iy], kz = ks2[iz];
    const R kk = sqrt(kx * kx + ky * ky + kz * kz);
    const cpx<R> zero = mk<R>((R)0, (R)0);
    if (kk == (R)0) {
      for (int c = 0; c < 3; ++c) { tor[o + c * osf] = zero; pol[o + c * osf] = zero; }
      continue;
    }
    const R hx = kx / kk, hy = ky / kk, hz = kz / kk;
    const R kperp = sqrt(kx * kx + ky * ky);
    R e1[3], e2[3];
    if (kperp > (R)0) {
      const R n1 = sqrt(hy * hy + hx * hx);
      e1[0] = hy / n1; e1[1] = -hx / n1; e1[2] = (R)0;
      e2[0] = hy * e1[2] - hz * e1[1]; e2[1] = hz * e1[0] - hx * e1[2]; e2[2] = hx * e1[1] - hy * e1[0];
    } else {
      e1[0] = (R)1; e1[1] = (R)0; e1[2] = (R)0;
      e2[0] = (R)0; e2[1] = (R)1; e2[2] = (R)0;
    }
    const cpx<R> u1 = u[o], u2 = u[o + usf], u3 = u[o + 2 * usf];
    const cpx<R> c1 = cadd(cadd(cscale(u1, e1[0]), cscale(u2, e1[1])), cscale(u3, e1[2]));
    const cpx<R> c2 = cadd(cadd(cscale(u1, e2[0]), cscale(u2, e2[1])), cscale(u3, e2[2]));
    for (int c = 0; c < 3; ++c) {
      tor[o + c * osf] = cscale(c1, e1[c]);
      pol[o + c * osf] = cscale(c2, e2[c]);
    }
  }
}

}  // namespace fit
