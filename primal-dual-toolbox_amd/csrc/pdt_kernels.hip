// pdt_kernels.hip -- CDNA4 (gfx950) kernels of the TV/TGV primal-dual solvers.
//
// Design (one launch per half-iteration, batched over slices):
//  * Denoising (TvOptimizer / TgvOptimizer): each half-iteration is one
//    kernel over row-block work items that fuses the forward/backward
//    difference stencils, the dual ascent with its pointwise ball projection
//    and the over-relaxed primal step (tgv_kernels.cuh:225-324,568-605,647-686).
//  * Cartesian MRI (TgvMriOptimizer and the with-operator solvers): three
//    kernels per iteration.
//      rows_fwd  : dual p (and q) update of image row y, then for every coil
//                  S_c * u_bar row, ifftshift, PE-FFT in LDS, fftshift, keep
//                  the sampled PE columns                  -> T1 [c][y][ks]
//      cols      : per coil and group of sampled columns: FE-FFT of the
//                  zero-padded (read-out oversampled) column, mask, data prox
//                  of r, mask, inverse FE-FFT, crop to the image window
//                                                           -> T2 [c][y][ks]
//      rows_adj  : inverse PE-FFT of every coil's row, conj(S_c) coil sum
//                  (deterministic, in registers), then the primal u (and v)
//                  step of row y.
//    This replaces prefft/applyMask/preifft/combineImg/padFOV/cropFOV, the
//    cuFFT calls and TGV_prox_r/TGV_dual_*/TGV_primal_* of the reference
//    (mricartesianoperator.cu:28-160,196-414; tgvmrioptimizer.cu:162-198).
//  * The FFT is a mixed-radix Stockham autosort in LDS with radix-2/3/4/5/8
//    register codelets and a generic odd-prime stage (e.g. 23 for 368/736).
//  * Gap terms are accumulated in double per work item on evaluation
//    iterations only (template GAP), and reduced per slice in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "pdt_desc.hpp"
#include "pdt_launch.hpp"

namespace pdt {

// ---------------------------------------------------------------------------
// Pixel arithmetic
// ---------------------------------------------------------------------------
template <class R>
struct alignas(2 * sizeof(R)) Cx {
  R re, im;
};

template <class R> __device__ __forceinline__ Cx<R> operator+(Cx<R> a, Cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R> __device__ __forceinline__ Cx<R> operator-(Cx<R> a, Cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R> __device__ __forceinline__ Cx<R> operator-(Cx<R> a) { return {-a.re, -a.im}; }
// Complex products with explicit fused multiply-adds: left to the compiler,
// the contraction of a.re b.im + a.im b.re fused one product in one kernel
// and the other in another (both are one rounding short of two), so the same
// product of the same values differed in the last bit between kernels.
template <class R> __device__ __forceinline__ Cx<R> operator*(Cx<R> a, Cx<R> b) {
  return {fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}
template <class R> __device__ __forceinline__ Cx<R> operator*(R s, Cx<R> a) { return {s * a.re, s * a.im}; }
template <class R> __device__ __forceinline__ Cx<R> operator*(Cx<R> a, R s) { return {s * a.re, s * a.im}; }
template <class R> __device__ __forceinline__ Cx<R> operator/(Cx<R> a, R s) { return {a.re / s, a.im / s}; }
// a * x + y with one rounding per component.  The denoising primal steps
// write div + lambda f with it, so that every kernel rounds it the same way
// (left to the compiler, a kernel that holds f across iterations computed
// lambda f once, rounded, and the one-iteration kernels fused it).
template <class R> __device__ __forceinline__ R fmadd(R a, R x, R y) { return fma(a, x, y); }
template <class R> __device__ __forceinline__ Cx<R> fmadd(R a, Cx<R> x, Cx<R> y) {
  return {fma(a, x.re, y.re), fma(a, x.im, y.im)};
}
// c ? a : b, component-wise (a select, not a branch)
template <class R> __device__ __forceinline__ Cx<R> csel(bool c, Cx<R> a, Cx<R> b) {
  return {c ? a.re : b.re, c ? a.im : b.im};
}
// a * conj(b)
template <class R> __device__ __forceinline__ Cx<R> mulconj(Cx<R> a, Cx<R> b) {
  return {fma(a.re, b.re, a.im * b.im), fma(a.im, b.re, -(a.re * b.im))};
}

template <class P> struct PT { using R = P; static constexpr bool cplx = false; };
template <class R> struct PT<Cx<R>> { using Rt = R; using R_ = R; static constexpr bool cplx = true; };
template <class P> struct RealOf { using T = P; };
template <class R> struct RealOf<Cx<R>> { using T = R; };

template <class P> __device__ __forceinline__ P pzero() { return P(0); }
template <> __device__ __forceinline__ Cx<float> pzero<Cx<float>>() { return {0.f, 0.f}; }
template <> __device__ __forceinline__ Cx<double> pzero<Cx<double>>() { return {0.0, 0.0}; }

__device__ __forceinline__ float abs2(float a) { return a * a; }
__device__ __forceinline__ double abs2(double a) { return a * a; }
template <class R> __device__ __forceinline__ R abs2(Cx<R> a) { return fma(a.re, a.re, a.im * a.im); }  // (as the products above)
__device__ __forceinline__ double dabs2(float a) { return double(a) * double(a); }
__device__ __forceinline__ double dabs2(double a) { return a * a; }
template <class R> __device__ __forceinline__ double dabs2(Cx<R> a) {
  return double(a.re) * double(a.re) + double(a.im) * double(a.im);
}
__device__ __forceinline__ double redot(float a, float b) { return double(a) * double(b); }
__device__ __forceinline__ double redot(double a, double b) { return a * b; }
template <class R> __device__ __forceinline__ double redot(Cx<R> a, Cx<R> b) {
  return double(a.re) * double(b.re) + double(a.im) * double(b.im);
}

// fp16 storage of the TGV duals (compute stays in the pixel type).
template <class P> struct Half;
template <> struct Half<float> { using T = _Float16; };
template <> struct Half<Cx<float>> { using T = Cx<_Float16>; };
template <> struct Half<double> { using T = double; };
template <> struct Half<Cx<double>> { using T = Cx<double>; };

// Device memory reached through the slice descriptors is addressed with
// global-address-space pointers (gptr): loads and stores through them compile
// to global_* instructions with a scalar base, where generic pointers give
// flat_* ones, which also count against the LDS wait counter (so every LDS
// wait of an FFT stage would wait for the outstanding register prefetches).
// (The host pass of the same source, which only checks device code, keeps
// generic pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using G = __attribute__((address_space(1))) T;
template <class T> struct NoAS { using type = T; };
template <class T> struct NoAS<G<T>> { using type = T; };
#else
template <class T> using G = T;
template <class T> struct NoAS { using type = T; };
#endif
template <class T> using Plain = std::remove_const_t<typename NoAS<T>::type>;
template <class T>
__device__ __forceinline__ G<T>* gptr(const void* p) {
  return (G<T>*)(const_cast<void*>(p));
}

// Element access of a plane stored as S (get/put take any pointer to S).
template <class P, class S> struct Ld {
  template <class A> static __device__ __forceinline__ P get(A* a, size_t i) { return a[i]; }
  template <class A> static __device__ __forceinline__ void put(A* a, size_t i, P v) { a[i] = v; }
};
// The value stored is the float result rounded to nearest half: the empty
// register constraint keeps the compiler from fusing the product that
// produced it with the conversion (a mixed-precision FMA rounds once, straight
// to half), so every kernel stores the same halves.
__device__ __forceinline__ _Float16 to_half(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(v));
#endif
  return _Float16(v);
}
template <> struct Ld<float, _Float16> {
  template <class A> static __device__ __forceinline__ float get(A* a, size_t i) { return float(a[i]); }
  template <class A> static __device__ __forceinline__ void put(A* a, size_t i, float v) { a[i] = to_half(v); }
};
template <> struct Ld<Cx<float>, Cx<_Float16>> {
  template <class A> static __device__ __forceinline__ Cx<float> get(A* a, size_t i) {
    Cx<_Float16> h = a[i];
    return {float(h.re), float(h.im)};
  }
  template <class A> static __device__ __forceinline__ void put(A* a, size_t i, Cx<float> v) {
    a[i] = Cx<_Float16>{to_half(v.re), to_half(v.im)};
  }
};

// ---------------------------------------------------------------------------
// Block reductions (wave64 shuffles, then the 4 waves in a fixed order)
// ---------------------------------------------------------------------------
template <int N, int NT = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[N], double* out) {
  __shared__ double red[(NT / 64) * kGapTerms];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < N; ++t) {
    double x = v[t];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) red[wave * kGapTerms + t] = x;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0;
    for (int w = 0; w < NT / 64; ++w) s += red[w * kGapTerms + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// block_sum for a block of blockDim.x threads (a multiple of 64, at most 1024).
template <int N>
__device__ __forceinline__ void block_sum_dyn(double (&v)[N], double* out) {
  __shared__ double red[16 * kGapTerms];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int t = 0; t < N; ++t) {
    double x = v[t];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) red[wave * kGapTerms + t] = x;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0;
    for (int w = 0; w < nw; ++w) s += red[w * kGapTerms + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

template <int NT = kBlock>
__device__ __forceinline__ double block_max(double v) {
  __shared__ double red[NT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < NT / 64; ++w) m = fmax(m, red[w]);
  return m;
}

// ---------------------------------------------------------------------------
// Mixed-radix Stockham FFT in LDS
//   stage radix r, p = product of earlier radices, nr = n / r:
//   for i in [0, nr): k = i mod p,
//     x_j = in[i + j nr] * w_n^(j k n/(p r)),  y = DFT_r(x),
//     out[(i - k) r + k + q p] = y_q.
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ Cx<R> twd(const Cx<R>* tw, int m, bool inv) {
  Cx<R> w = tw[m];
  if (inv) w.im = -w.im;
  return w;
}
// x * (s i), s = +1 / -1
template <class R> __device__ __forceinline__ Cx<R> muli(Cx<R> x, R s) { return {-s * x.im, s * x.re}; }

template <class R, int RR> struct Dft;
template <class R> struct Dft<R, 2> {
  static __device__ __forceinline__ void run(Cx<R>* x, R) {
    Cx<R> a = x[0], b = x[1];
    x[0] = a + b;
    x[1] = a - b;
  }
};
template <class R> struct Dft<R, 3> {
  static __device__ __forceinline__ void run(Cx<R>* x, R s) {
    const R h = R(0.86602540378443864676);
    Cx<R> t = x[1] + x[2];
    Cx<R> d = muli(x[1] - x[2], s) * h;
    Cx<R> m = x[0] - t * R(0.5);
    x[0] = x[0] + t;
    x[1] = m + d;
    x[2] = m - d;
  }
};
template <class R> struct Dft<R, 4> {
  static __device__ __forceinline__ void run(Cx<R>* x, R s) {
    Cx<R> t0 = x[0] + x[2], t1 = x[0] - x[2], t2 = x[1] + x[3], t3 = muli(x[1] - x[3], s);
    x[0] = t0 + t2;
    x[2] = t0 - t2;
    x[1] = t1 + t3;
    x[3] = t1 - t3;
  }
};
template <class R> struct Dft<R, 5> {
  static __device__ __forceinline__ void run(Cx<R>* x, R s) {
    const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
    const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
    Cx<R> t1 = x[1] + x[4], t2 = x[2] + x[3], d1 = x[1] - x[4], d2 = x[2] - x[3];
    Cx<R> a1 = x[0] + t1 * c1 + t2 * c2;
    Cx<R> a2 = x[0] + t1 * c2 + t2 * c1;
    Cx<R> b1 = muli(d1 * s1 + d2 * s2, s);
    Cx<R> b2 = muli(d1 * s2 - d2 * s1, s);
    x[0] = x[0] + t1 + t2;
    x[1] = a1 + b1;
    x[4] = a1 - b1;
    x[2] = a2 + b2;
    x[3] = a2 - b2;
  }
};
template <class R> struct Dft<R, 8> {
  static __device__ __forceinline__ void run(Cx<R>* x, R s) {
    Cx<R> e[4] = {x[0], x[2], x[4], x[6]};
    Cx<R> o[4] = {x[1], x[3], x[5], x[7]};
    Dft<R, 4>::run(e, s);
    Dft<R, 4>::run(o, s);
    const R c = R(0.70710678118654752440);
    Cx<R> w1 = {c, s * c}, w3 = {-c, s * c};
    Cx<R> o1 = o[1] * w1, o2 = muli(o[2], s), o3 = o[3] * w3;
    x[0] = e[0] + o[0];
    x[4] = e[0] - o[0];
    x[1] = e[1] + o1;
    x[5] = e[1] - o1;
    x[2] = e[2] + o2;
    x[6] = e[2] - o2;
    x[3] = e[3] + o3;
    x[7] = e[3] - o3;
  }
};

// Radices 6 and 10: even/odd split into two DFTs of half length m,
// X[k] = E[k] + w^k O[k], X[k + m] = E[k] - w^k O[k], w = exp(s i pi / m).
template <class R> struct Dft<R, 6> {
  static __device__ __forceinline__ void run(Cx<R>* x, R s) {
    Cx<R> e[3] = {x[0], x[2], x[4]};
    Cx<R> o[3] = {x[1], x[3], x[5]};
    Dft<R, 3>::run(e, s);
    Dft<R, 3>::run(o, s);
    const R h = R(0.86602540378443864676);  // sin(pi / 3)
    const Cx<R> t1 = o[1] * Cx<R>{R(0.5), s * h}, t2 = o[2] * Cx<R>{R(-0.5), s * h};
    x[0] = e[0] + o[0];
    x[3] = e[0] - o[0];
    x[1] = e[1] + t1;
    x[4] = e[1] - t1;
    x[2] = e[2] + t2;
    x[5] = e[2] - t2;
  }
};
template <class R> struct Dft<R, 10> {
  static __device__ __forceinline__ void run(Cx<R>* x, R s) {
    Cx<R> e[5] = {x[0], x[2], x[4], x[6], x[8]};
    Cx<R> o[5] = {x[1], x[3], x[5], x[7], x[9]};
    Dft<R, 5>::run(e, s);
    Dft<R, 5>::run(o, s);
    // cos and sin of k pi / 5, k = 1, 2
    const R c1 = R(0.80901699437494742410), c2 = R(0.30901699437494742410);
    const R s1 = R(0.58778525229247312917), s2 = R(0.95105651629515357212);
    Cx<R> t[5];
    t[0] = o[0];
    t[1] = o[1] * Cx<R>{c1, s * s1};
    t[2] = o[2] * Cx<R>{c2, s * s2};
    t[3] = o[3] * Cx<R>{-c2, s * s2};
    t[4] = o[4] * Cx<R>{-c1, s * s1};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      x[k] = e[k] + t[k];
      x[k + 5] = e[k] - t[k];
    }
  }
};

template <class R, int RR>
__device__ __forceinline__ void butterfly(const Cx<R>* in, Cx<R>* out, int nr, int p,
                                          const Cx<R>* tw, int ts, bool inv) {
  Cx<R> x[RR];
  x[0] = in[0];
#pragma unroll
  for (int j = 1; j < RR; ++j) x[j] = in[j * nr] * twd(tw, j * ts, inv);
  Dft<R, RR>::run(x, inv ? R(1) : R(-1));
#pragma unroll
  for (int q = 0; q < RR; ++q) out[q * p] = x[q];
}

// Generic radix (primes without a specialisation): direct DFT reading from
// LDS; w^(j q mod r) is stepped instead of reduced with a modulo.
template <class R>
__device__ void butterfly_gen(const Cx<R>* in, Cx<R>* out, int r, int nr, int p, const Cx<R>* tw,
                              int ts, bool inv) {
  for (int q = 0; q < r; ++q) {
    Cx<R> acc = in[0];
    int jq = 0;
    for (int j = 1; j < r; ++j) {
      jq += q;
      if (jq >= r) jq -= r;
      const Cx<R> v = in[j * nr] * twd(tw, j * ts, inv);
      acc = acc + v * twd(tw, jq * nr, inv);
    }
    out[q * p] = acc;
  }
}

// cos and sin of 2 pi j / n at compile time (Taylor series on |x| <= pi / 2
// after the symmetries of the circle), for the odd-prime butterflies.
constexpr double ct_taylor_sin(double x) {
  double t = x, r = x;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / double((2 * k) * (2 * k + 1));
    r += t;
  }
  return r;
}
constexpr double ct_taylor_cos(double x) {
  double t = 1, r = 1;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / double((2 * k - 1) * (2 * k));
    r += t;
  }
  return r;
}
constexpr double ct_sincos2pi(int j, int n, bool want_sin) {
  j %= n;
  double sg = 1;
  if (2 * j > n) {  // angle in (pi, 2 pi): sin changes sign, cos does not
    j = n - j;
    if (want_sin) sg = -1;
  }
  constexpr double kPi = 3.14159265358979323846264338327950288;
  double x = 2 * kPi * double(j) / double(n);  // [0, pi]
  if (2 * x > kPi) {                          // (pi / 2, pi]: reflect about pi / 2
    x = kPi - x;
    return want_sin ? sg * ct_taylor_sin(x) : -ct_taylor_cos(x);
  }
  return want_sin ? sg * ct_taylor_sin(x) : ct_taylor_cos(x);
}
template <int P>
struct OddRoots {
  double c[P] = {}, s[P] = {};
  constexpr OddRoots() {
    for (int j = 0; j < P; ++j) {
      c[j] = ct_sincos2pi(j, P, false);
      s[j] = ct_sincos2pi(j, P, true);
    }
  }
};
template <int P>
constexpr OddRoots<P> kOddRoots{};

// Odd radix P (7, 11, 13, 23, ...) in the symmetric form: with
// s_m = x_m + x_(P-m), d_m = x_m - x_(P-m), m = 1..h, h = (P-1)/2,
//   X_0 = x_0 + sum_m s_m,
//   X_k, X_(P-k) = x_0 + sum_m cos(2 pi m k / P) s_m  +/-  i sg sum_m sin(2 pi m k / P) d_m,
// sg = -1 forward, +1 inverse: h^2 complex-by-real products per output pair.
// Both loops unroll, so the coefficients are compile-time constants rather
// than loads from the twiddle table.  in[j * nr] are the inputs, out[q * p]
// the outputs; tw (length-n table) supplies the stage twiddles only.  With
// G > 1 the call computes only the output pairs k = g + 1 (mod G) (and X_0
// for g = 0): G callers share one butterfly.
template <class R, int P>
__device__ __forceinline__ void butterfly_odd(const Cx<R>* in, Cx<R>* out, int nr, int p, const Cx<R>* tw,
                                              int ts, bool inv, int G = 1, int g = 0) {
  constexpr int h = (P - 1) / 2;
  const Cx<R> x0 = in[0];
  Cx<R> s[h], d[h];
  Cx<R> sum = x0;
#pragma unroll
  for (int m = 1; m <= h; ++m) {
    const Cx<R> a = in[m * nr] * twd(tw, m * ts, inv);
    const Cx<R> b = in[(P - m) * nr] * twd(tw, (P - m) * ts, inv);
    s[m - 1] = a + b;
    d[m - 1] = a - b;
    sum = sum + s[m - 1];
  }
  if (g == 0) out[0] = sum;
  const R sg = inv ? R(1) : R(-1);
#pragma unroll
  for (int k = 1; k <= h; ++k) {
    if (((k - 1) & (G - 1)) != g) continue;  // G is 1, 2 or 4
    Cx<R> A = x0, B = pzero<Cx<R>>();
#pragma unroll
    for (int m = 1; m <= h; ++m) {
      const int jk = (m * k) % P;
      A = A + s[m - 1] * R(kOddRoots<P>.c[jk]);
      B = B + d[m - 1] * R(kOddRoots<P>.s[jk]);
    }
    const Cx<R> iB = muli(B, sg);
    out[k * p] = A + iB;
    out[(P - k) * p] = A - iB;
  }
}

// Exact a / b for 0 <= a < 2^24, 1 <= b < 2^12 from a float reciprocal.
__device__ __forceinline__ int qdiv(int a, int b, float rb) {
  int q = __float2int_rz(float(a) * rb);
  if (q * b > a) --q;
  else if ((q + 1) * b <= a) ++q;
  return q;
}
// a / n where n == NF when NF > 0 (a division by a constant), else qdiv.
template <int NF>
__device__ __forceinline__ int divn(int a, int n, float rn) {
  if constexpr (NF > 0) return a / NF;
  else return qdiv(a, n, rn);
}


// Transforms nseq sequences of length pl.n stored back to back in a; b is
// scratch of the same size; tw is the twiddle table (LDS or global).
// Returns the buffer that holds the result.  Ends with a barrier.  Inlined,
// so that a and b stay LDS pointers (a call would pass generic ones and turn
// every access into a flat one) and the caller's registers need no saving.
template <class R, int NT>
__device__ __forceinline__ Cx<R>* lds_fft(Cx<R>* a, Cx<R>* b, int nseq, const FftPlanD& pl, const Cx<R>* tw, bool inv) {
  const int n = pl.n;
  for (int s = 0; s < pl.ns; ++s) {
    const int r = pl.r[s], nr = pl.nr[s], p = pl.p[s], tws = pl.tws[s];
    const float rnr = pl.rnr[s], rp = pl.rp[s];
    const bool p2 = (p & (p - 1)) == 0;  // uniform: power-of-two radices come first
    const int total = nseq * nr;
    for (int t = threadIdx.x; t < total; t += NT) {
      const int seq = nseq == 1 ? 0 : qdiv(t, nr, rnr);
      const int i = t - seq * nr;
      const int k = p2 ? (i & (p - 1)) : i - qdiv(i, p, rp) * p;
      const Cx<R>* in = a + seq * n + i;
      Cx<R>* out = b + seq * n + (i - k) * r + k;
      const int ts = k * tws;
      switch (r) {
        case 2: butterfly<R, 2>(in, out, nr, p, tw, ts, inv); break;
        case 3: butterfly<R, 3>(in, out, nr, p, tw, ts, inv); break;
        case 4: butterfly<R, 4>(in, out, nr, p, tw, ts, inv); break;
        case 5: butterfly<R, 5>(in, out, nr, p, tw, ts, inv); break;
        case 8: butterfly<R, 8>(in, out, nr, p, tw, ts, inv); break;
        default: butterfly_gen<R>(in, out, r, nr, p, tw, ts, inv); break;
      }
    }
    __syncthreads();
    Cx<R>* t = a;
    a = b;
    b = t;
  }
  return a;
}

// ---------------------------------------------------------------------------
// Compile-time-planned Stockham FFT for the common transform lengths.  The
// stage formula of lds_fft, with radix, stride and twiddle steps
// constant-folded; lengths without a specialisation use lds_fft.
// ---------------------------------------------------------------------------
constexpr int ct_radix(int n, int s) {
  // 384 and 640: a last stage of radix 6 or 10 instead of a radix-2 stage
  // followed by a radix-3 or radix-5 one (the first stage stays radix 8,
  // as the padded LDS layout requires).
  if (n == 384) return s < 2 ? 8 : s == 2 ? 6 : 1;
  if (n == 640) return s < 2 ? 8 : s == 2 ? 10 : 1;
  int k = 0;
  while (n % 8 == 0 && n > 8) {
    if (k == s) return 8;
    n /= 8;
    ++k;
  }
  const int small[5] = {8, 4, 2, 3, 5};
  for (int j = 0; j < 5; ++j)
    while (n % small[j] == 0 && n >= small[j]) {
      if (k == s) return small[j];
      n /= small[j];
      ++k;
    }
  for (int p = 7; n > 1; p += 2)
    while (n % p == 0) {
      if (k == s) return p;
      n /= p;
      ++k;
    }
  return 1;
}
constexpr int ct_nstages(int n) {
  int s = 0;
  while (ct_radix(n, s) != 1) ++s;
  return s;
}
// The padded LDS layout (element e at e + e / 8) keeps every input and
// output of a butterfly at a constant padded stride when n, every n / r and
// every earlier-radix product p (or r, when p = 1) are multiples of 8.
constexpr bool ct_pad_ok(int n) {
  if (n % 8 != 0) return false;
  int p = 1;
  for (int s = 0; s < ct_nstages(n); ++s) {
    const int r = ct_radix(n, s);
    if ((n / r) % 8 != 0 || (p == 1 ? r != 8 : p % 8 != 0)) return false;
    p *= r;
  }
  return true;
}

// Unpadded FFT buffers (the wide adjoint row pass, whose two buffers fill
// the LDS share of a workgroup) permute their slots instead: bits 0-2 of an
// element index are XORed with bits 4-6 and bit 3 with bit 6, a permutation
// of every aligned block of 16 elements.  The Stockham writes at strides of 8
// and 64 elements then spread the lanes of a wave over distinct banks, while
// runs of 16 consecutive elements (the reads) stay on distinct banks.
__device__ __forceinline__ int swz(int e) { return e ^ ((e >> 4) & 7) ^ (((e >> 6) & 1) << 3); }

// LDS slot of element e of an FFT buffer.  Kernels with a compile-time length
// (NF > 0) pad one element after every 8: the Stockham stages write with
// strides of 8 elements (and multiples), which would otherwise put the 64
// lanes of a wave on a few LDS banks.  Without padding (LP = false) the
// slots are permuted by swz.
template <int NF, bool LP = true>
__device__ __forceinline__ int lp(int e) {
  if constexpr (NF > 0 && ct_pad_ok(NF)) return LP ? e + (e >> 3) : swz(e);
  else return e;
}

// TW: stage twiddles present (false for the first stage, where they are 1).
// nr, p: LDS strides of the inputs and outputs.
template <class R, int RR, bool TW>
__device__ __forceinline__ void butterfly_ct(const Cx<R>* in, Cx<R>* out, int ts, const Cx<R>* tw, bool inv,
                                             int nr, int p) {
  if constexpr (RR == 2 || RR == 3 || RR == 4 || RR == 5 || RR == 6 || RR == 8 || RR == 10) {
    Cx<R> x[RR];
    x[0] = in[0];
#pragma unroll
    for (int j = 1; j < RR; ++j) x[j] = TW ? in[j * nr] * twd(tw, j * ts, inv) : in[j * nr];
    Dft<R, RR>::run(x, inv ? R(1) : R(-1));
#pragma unroll
    for (int q = 0; q < RR; ++q) out[q * p] = x[q];
  } else {
    static_assert(RR % 2 == 1, "even radices are 2, 4, 6, 8 and 10");
    butterfly_odd<R, RR>(in, out, nr, p, tw, ts, inv);
  }
}

// butterfly_ct on swz-permuted buffers: inputs at logical elements
// ein + j nr of a, outputs at eout + q p of b.
template <class R, int RR, bool TW>
__device__ __forceinline__ void butterfly_ct_swz(const Cx<R>* a, Cx<R>* b, int ein, int eout, int ts,
                                                 const Cx<R>* tw, bool inv, int nr, int p) {
  static_assert(RR == 2 || RR == 3 || RR == 4 || RR == 5 || RR == 6 || RR == 8 || RR == 10, "radix");
  Cx<R> x[RR];
  x[0] = a[swz(ein)];
#pragma unroll
  for (int j = 1; j < RR; ++j) x[j] = TW ? a[swz(ein + j * nr)] * twd(tw, j * ts, inv) : a[swz(ein + j * nr)];
  Dft<R, RR>::run(x, inv ? R(1) : R(-1));
#pragma unroll
  for (int q = 0; q < RR; ++q) b[swz(eout + q * p)] = x[q];
}

template <class R, int N, int S, int P, int NT, bool INV, bool LP>
__device__ __forceinline__ Cx<R>* fft_ct_stage(Cx<R>* a, Cx<R>* b, int nseq, const Cx<R>* tw) {
  if constexpr (S == ct_nstages(N)) {
    return a;
  } else {
    constexpr int r = ct_radix(N, S), nr = N / r, tws = N / (P * r);
    // padded layout (ct_pad_ok): every input and output of a butterfly
    // sits at a constant padded stride from its first one
    constexpr bool PAD = LP && ct_pad_ok(N);
    constexpr int NP = PAD ? N + N / 8 : N, nrp = PAD ? nr + nr / 8 : nr, pp = PAD && P > 1 ? P + P / 8 : P;
    const int total = nseq * nr;
    if constexpr (r > 8 && r % 2 == 1) {
      // Odd prime radix: a long chain of products per butterfly and few
      // butterflies per block, so the output pairs of each butterfly are
      // split over G waves (G = 4, 2 or 1, as many as keep every butterfly
      // in one pass).  Lane l of a wave takes butterfly 64 c + l and the
      // wave-uniform part g, so every part keeps constant coefficients.
      constexpr int NW = NT / 64;
      const int chunks = (total + 63) / 64;
      const int G = chunks * 4 <= NW ? 4 : chunks * 2 <= NW ? 2 : 1;
      const int lane = threadIdx.x & 63;
      const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // a scalar, so the parts branch per wave
      for (int q = wave; q < chunks * G; q += NW) {
        const int c = q / G, g = q - c * G;
        const int t = c * 64 + lane;
        if (t < total) {
          const int seq = t / nr;
          const int i = t - seq * nr;
          const int k = i % P;
          const int o = (i - k) * r + k;
          const int ip = PAD ? i + (i >> 3) : i, op = PAD ? o + (o >> 3) : o;
          butterfly_odd<R, r>(a + seq * NP + ip, b + seq * NP + op, nrp, pp, tw, k * tws, INV, G, g);
        }
      }
    } else {
      for (int t = threadIdx.x; t < total; t += NT) {
        const int seq = t / nr;
        const int i = t - seq * nr;
        const int k = i % P;
        const int o = (i - k) * r + k;
        if constexpr (!LP && ct_pad_ok(N)) {
          butterfly_ct_swz<R, r, (P > 1)>(a, b, seq * N + i, seq * N + o, k * tws, tw, INV, nr, P);
        } else {
          const int ip = PAD ? i + (i >> 3) : i, op = PAD ? o + (o >> 3) : o;
          butterfly_ct<R, r, (P > 1)>(a + seq * NP + ip, b + seq * NP + op, k * tws, tw, INV, nrp, pp);
        }
      }
    }
    __syncthreads();
    return fft_ct_stage<R, N, S + 1, P * r, NT, INV, LP>(b, a, nseq, tw);
  }
}

// FFT of length NF with a compile-time plan, or of length pl.n with the
// runtime plan when NF == 0.  Each FFT kernel is instantiated per length, so
// its register budget is that of its own plan only (the runtime plan with
// its radix switch needs several times the registers of a fixed one).
template <class R, int NT, bool INV, int NF, bool LP = true>
__device__ __forceinline__ Cx<R>* fft_n(Cx<R>* a, Cx<R>* b, int nseq, const FftPlanD& pl, const Cx<R>* tw) {
  if constexpr (NF > 0) return fft_ct_stage<R, NF, 0, 1, NT, INV, LP>(a, b, nseq, tw);
  else return lds_fft<R, NT>(a, b, nseq, pl, tw, INV);
}

// One sequence of the compile-time plan transformed by one wavefront alone,
// in place on its padded slots (lengths with ct_pad_ok, at most 64 M
// butterflies per stage): each stage reads every input of the wave's
// butterflies into registers before it writes an output, and a wave's LDS
// accesses take effect in issue order, so no workgroup barrier separates the
// stages.  The butterflies are fft_ct_stage's (butterfly_ct's operations),
// so the result is that of fft_n on padded buffers.  seq: the sequence's
// first slot; call with S = 0, P = 1.
template <class R, int N, int S, int P, bool INV>
__device__ __forceinline__ void fft_wave(Cx<R>* seq, const Cx<R>* tw) {
  if constexpr (S < ct_nstages(N)) {
    constexpr int r = ct_radix(N, S), nr = N / r, tws = N / (P * r);
    static_assert(ct_pad_ok(N) && r <= 10, "wave transforms: padded lengths, register codelets");
    constexpr int nrp = nr + nr / 8, pp = P > 1 ? P + P / 8 : P;
    constexpr int M = (nr + 63) / 64;  // butterflies per lane
    const int lane = threadIdx.x & 63;
    Cx<R> x[M][r];
    int ob[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = lane + 64 * m;
      ob[m] = -1;
      if (i < nr) {
        const int k = i % P;
        const int o = (i - k) * r + k;
        const Cx<R>* in = seq + i + (i >> 3);
        x[m][0] = in[0];
#pragma unroll
        for (int j = 1; j < r; ++j) x[m][j] = P > 1 ? in[j * nrp] * twd(tw, j * (k * tws), INV) : in[j * nrp];
        Dft<R, r>::run(x[m], INV ? R(1) : R(-1));
        ob[m] = o + (o >> 3);
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int m = 0; m < M; ++m)
      if (ob[m] >= 0) {
#pragma unroll
        for (int q = 0; q < r; ++q) seq[ob[m] + q * pp] = x[m][q];
      }
    __builtin_amdgcn_wave_barrier();
    fft_wave<R, N, S + 1, P * r, INV>(seq, tw);
  }
}

// numpy shift maps: ifftshift moves i to (i + ceil(n/2)) % n (the prefft /
// preifft index map of mricartesianoperator.cu:44-49); fftshift output k
// reads input (k - n/2) mod n.
// Both take 0 <= i < n and avoid the integer modulo.
__device__ __forceinline__ int ishift_dst(int i, int n) {
  const int j = i + (n + 1) / 2;
  return j >= n ? j - n : j;
}
__device__ __forceinline__ int fshift_src(int k, int n) {
  const int j = k + n - n / 2;
  return j >= n ? j - n : j;
}
// ifftshift gather: position j of ifftshift(z) holds z[(j + n/2) mod n].
__device__ __forceinline__ int ishift_src(int j, int n) {
  const int i = j + n / 2;
  return i >= n ? i - n : i;
}

// ---------------------------------------------------------------------------
// Stencil helpers on plane pointers (tgv_kernels.cuh:25-135)
// ---------------------------------------------------------------------------
// Both neighbours are loaded unconditionally (the first row / column reads
// its own element twice) and the boundary cases selected afterwards, so the
// loads issue together instead of behind branches.
template <class P, class S>
__device__ __forceinline__ P dmx(const S* a, int x, size_t i, int W) {
  using L = Ld<P, Plain<S>>;
  const P c = L::get(a, i), m = L::get(a, x > 0 ? i - 1 : i);
  const P d = (x < W - 1) ? c - m : -m;
  return x > 0 ? d : c;
}
template <class P, class S>
__device__ __forceinline__ P dmy(const S* a, int y, size_t i, int W, int H) {
  using L = Ld<P, Plain<S>>;
  const P c = L::get(a, i), m = L::get(a, y > 0 ? i - W : i);
  const P d = (y < H - 1) ? c - m : -m;
  return y > 0 ? d : c;
}

// 1 / x: the hardware reciprocal (within 1 ulp) in single precision, a
// correctly rounded division in double precision.
template <class R> __device__ __forceinline__ R rcp(R x) { return R(1) / x; }
template <> __device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

template <class R>
__device__ __forceinline__ R sched_alpha(const SliceDev& d, double a_start, double a_end, int k) {
  if (!d.schedule) return R(a_start);
  // tgvmrioptimizer.cu:150-155, evaluated in the solver's precision
  const R dk = R(k), dmax = R(d.kmax);
  return exp(dk / dmax * log(R(a_end)) + (dmax - dk) / dmax * log(R(a_start)));
}

// The arithmetic of the dual step at one pixel, on values: c = ubar, uxp /
// uyp = ubar at the forward neighbours (c itself at the last column / row);
// TGV: vx0, vy0 = vbar, vxx = vbar_x and vyx = vbar_y at the x neighbour,
// vxy, vyy at the y neighbour.  Updates p (and q) in place, with optional
// gap terms t[0] += |grad ub - vb|, t[1] += |E vb|_F, t[2] += Re<grad ub -
// vb, p> + Re<E vb, q>_F (TGV_dual_p_kernel / TGV_dual_q_kernel /
// TV_dual_p_kernel).  ia0, ia1 = 1 / alpha0, 1 / alpha1 (TGV only).
template <class P, bool TGV, bool GAP>
__device__ __forceinline__ void dual_math(P c, P uxp, P uyp, P vx0, P vy0, P vxx, P vyx, P vxy, P vyy, P& px, P& py,
                                          P& qxx, P& qyy, P& qxy, typename RealOf<P>::T sigma,
                                          typename RealOf<P>::T ia0, typename RealOf<P>::T ia1, double* t) {
  using R = typename RealOf<P>::T;
  P tx = uxp - c, ty = uyp - c;
  if (TGV) {
    tx = tx - vx0;
    ty = ty - vy0;
    const P exx = vxx - vx0;
    const P eyy = vyy - vy0;
    const P exy = R(0.5) * (vxy - vx0 + vyx - vy0);
    const P bxx = qxx + sigma * exx;
    const P byy = qyy + sigma * eyy;
    const P bxy = qxy + sigma * exy;
    const R qn = sqrt(abs2(bxx) + abs2(byy) + R(2) * abs2(bxy));
    const R sq = rcp(fmax(R(1), qn * ia0));
    qxx = bxx * sq;
    qyy = byy * sq;
    qxy = bxy * sq;
    if (GAP) {
      t[1] += sqrt(dabs2(exx) + dabs2(eyy) + 2.0 * dabs2(exy));
      t[2] += redot(exx, qxx) + redot(eyy, qyy) + 2.0 * redot(exy, qxy);
    }
  }
  const P ax = px + sigma * tx;
  const P ay = py + sigma * ty;
  const R pn = sqrt(abs2(ax) + abs2(ay));
  const R sp = rcp(fmax(R(1), TGV ? pn * ia1 : pn));
  px = ax * sp;
  py = ay * sp;
  if (GAP) {
    t[0] += sqrt(dabs2(tx) + dabs2(ty));
    t[2] += redot(tx, px) + redot(ty, py);
  }
}

// Dual ascent + projection of p (and q) at one pixel (dual_math on the
// pixel's values in memory).
template <class P, class S, bool TGV, bool GAP>
__device__ __forceinline__ void dual_pixel(const SliceDev& d, int x, int y, typename RealOf<P>::T sigma,
                                           typename RealOf<P>::T ia0, typename RealOf<P>::T ia1,
                                           double* t) {
  using L = Ld<P, S>;
  const int W = d.W, H = d.H;
  const size_t n = size_t(W) * H;
  const size_t i = size_t(y) * W + x;
  const size_t ixp = i + (x < W - 1), iyp = i + ((y < H - 1) ? W : 0);
  const auto ub = gptr<const P>(d.ub);
  const auto px = gptr<S>(d.p);
  const auto py = px + n;
  P vx0 = pzero<P>(), vy0 = vx0, vxx = vx0, vyx = vx0, vxy = vx0, vyy = vx0;
  P qxx = vx0, qyy = vx0, qxy = vx0;
  const auto qxxp = gptr<S>(d.q);
  const auto qyyp = qxxp + n;
  const auto qxyp = qyyp + n;
  if (TGV) {
    const auto vbx = gptr<const P>(d.vb);
    const auto vby = vbx + n;
    vx0 = vbx[i];
    vy0 = vby[i];
    vxx = vbx[ixp];
    vyx = vby[ixp];
    vxy = vbx[iyp];
    vyy = vby[iyp];
    qxx = L::get(qxxp, i);
    qyy = L::get(qyyp, i);
    qxy = L::get(qxyp, i);
  }
  P pxv = L::get(px, i), pyv = L::get(py, i);
  dual_math<P, TGV, GAP>(ub[i], ub[ixp], ub[iyp], vx0, vy0, vxx, vyx, vxy, vyy, pxv, pyv, qxx, qyy, qxy, sigma, ia0,
                         ia1, t);
  if (TGV) {
    L::put(qxxp, i, qxx);
    L::put(qyyp, i, qyy);
    L::put(qxyp, i, qxy);
  }
  L::put(px, i, pxv);
  L::put(py, i, pyv);
}

// v-step of TGV_primal_v_kernel (tgv_kernels.cuh:647-686) at one pixel;
// returns |p + div_sym q|^2 for the gap.
template <class P, class S>
__device__ __forceinline__ double vstep_pixel(const SliceDev& d, int x, int y, typename RealOf<P>::T tau,
                                              typename RealOf<P>::T theta) {
  using R = typename RealOf<P>::T;
  using L = Ld<P, S>;
  const int W = d.W, H = d.H;
  const size_t n = size_t(W) * H;
  const size_t i = size_t(y) * W + x;
  const auto px = gptr<const S>(d.p);
  const auto py = px + n;
  const auto qxx = gptr<const S>(d.q);
  const auto qyy = qxx + n;
  const auto qxy = qyy + n;
  const auto vx = gptr<P>(d.v);
  const auto vy = vx + n;
  const auto vbx = gptr<P>(d.vb);
  const auto vby = vbx + n;
  const P dsx = dmx<P>(qxx, x, i, W) + dmy<P>(qxy, y, i, W, H);
  const P dsy = dmy<P>(qyy, y, i, W, H) + dmx<P>(qxy, x, i, W);
  const P pxi = L::get(px, i), pyi = L::get(py, i);
  const P ox = vx[i], oy = vy[i];
  const P nx = ox + tau * (dsx + pxi);
  const P ny = oy + tau * (dsy + pyi);
  vx[i] = nx;
  vy[i] = ny;
  vbx[i] = (R(1) + theta) * nx - theta * ox;
  vby[i] = (R(1) + theta) * ny - theta * oy;
  return dabs2(pxi + dsx) + dabs2(pyi + dsy);
}

__device__ __forceinline__ bool slice_active(const SliceDev& d) { return gptr<const int>(d.ctl)[CTL_ACTIVE] != 0; }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Give XCD x a contiguous range of logical work items, so that items
// touching neighbouring memory (adjacent column groups, adjacent rows) share
// one L2.  Bijective on [0, n).
__device__ __forceinline__ int xcd_item(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
  return x * q + min(x, r) + k;
}

// ---------------------------------------------------------------------------
// Denoising half-iterations.  Item = (slice, first row); rpi rows per item.
// ---------------------------------------------------------------------------

// TV_dual_p_kernel (tgv_kernels.cuh:294-324).  Gap: P(ubar) terms.
template <class P, bool GAP, int NT>
__global__ __launch_bounds__(NT) void k_tv_dual(const SliceDev* sl, const int2* items, int rpi,
                                                    double* part) {
  const int bi = xcd_item(blockIdx.x, gridDim.x);  // neighbouring rows on one XCD
  const int2 it = items ? items[bi] : make_int2(0, bi * rpi);
  const SliceDev& d = sl[it.x];
  if (!slice_active(d)) return;
  using R = typename RealOf<P>::T;
  const int W = d.W, H = d.H;
  const int ya = it.y, yb = min(H, ya + rpi);
  const R sigma = R(d.sigma);
  const auto ub = gptr<const P>(d.ub);
  const auto f = gptr<const P>(d.f);
  double t[kGapTerms] = {0};
  const int cnt = (yb - ya) * W;
  for (int idx = threadIdx.x; idx < cnt; idx += NT) {
    const int yy = idx / W;
    const int x = idx - yy * W, y = ya + yy;
    dual_pixel<P, P, false, GAP>(d, x, y, sigma, R(0), R(0), t);
    if (GAP) {
      const size_t i = size_t(y) * W + x;
      t[3] += dabs2(ub[i] - f[i]);
    }
  }
  if (GAP) block_sum<4, NT>(reinterpret_cast<double(&)[4]>(t), part + size_t(bi) * kGapTerms);
}

// TV_primal_u_kernel 2D (tgv_kernels.cuh:568-605).  Gap: ROF dual terms
// Re<div p, f>, ||div p||^2.
template <class P, bool GAP, int NT>
__global__ __launch_bounds__(NT) void k_tv_primal(const SliceDev* sl, const int2* items, int rpi,
                                                      double* part) {
  const int bi = xcd_item(blockIdx.x, gridDim.x);  // neighbouring rows on one XCD
  const int2 it = items ? items[bi] : make_int2(0, bi * rpi);
  const SliceDev& d = sl[it.x];
  if (!slice_active(d)) return;
  using R = typename RealOf<P>::T;
  const int W = d.W, H = d.H;
  const size_t n = size_t(W) * H;
  const int ya = it.y, yb = min(H, ya + rpi);
  const R tau = R(d.tau), lambda = R(d.lambda), theta = R(d.theta);
  const auto u = gptr<P>(d.u);
  const auto ub = gptr<P>(d.ub);
  const auto f = gptr<const P>(d.f);
  const auto px = gptr<const P>(d.p);
  const auto py = px + n;
  double t[2] = {0, 0};
  const int cnt = (yb - ya) * W;
  for (int idx = threadIdx.x; idx < cnt; idx += NT) {
    const int yy = idx / W;
    const int x = idx - yy * W, y = ya + yy;
    const size_t i = size_t(y) * W + x;
    const P uo = u[i];
    const P div = dmx<P>(px, x, i, W) + dmy<P>(py, y, i, W, H);
    P un = uo + tau * fmadd(lambda, P(f[i]), div);
    un = un * rcp(R(1) + tau * lambda);
    u[i] = un;
    ub[i] = (R(1) + theta) * un - theta * uo;
    if (GAP) {
      t[0] += redot(div, f[i]);
      t[1] += dabs2(div);
    }
  }
  if (GAP) block_sum<2, NT>(t, part + size_t(bi) * kGapTerms);
  if (ya == 0 && threadIdx.x == 0) gptr<int>(d.ctl)[CTL_ITER] += 1;
}

// Temporally blocked ROF iterations: tb_k times TV_dual_p_kernel then
// TV_primal_u_kernel (tgv_kernels.cuh:294-324,568-605) per launch, on a tile
// held by the workgroup.  Block (bx, by) owns the interior [x0, x1) x [y0, y1)
// and loads u, ubar, p (2 planes) and f on the interior grown by k pixels on
// every side (clipped to the image).  Each iteration shrinks the region
// where the values are exact by one pixel on each side that is not an image
// border: the dual step at (x, y) needs ubar at x + 1 and y + 1 (forward
// differences, none at the last column/row), the primal step p at x - 1 and
// y - 1 (backward differences, none at the first).  After k iterations the
// interior is exact and is written to the other copy of the state (src ->
// dst), so that no block reads a halo that another block of the same launch
// writes.  Each thread owns up to tv_block_own pixels of the tile for the
// whole launch and keeps their u, f and p in registers (only the pixel
// itself reads u and f, and its own p in the dual step); ubar and p, which
// neighbours read, are the three LDS planes.  The per-pixel arithmetic is
// that of k_tv_dual / k_tv_primal, in the same order, so the iterates equal
// theirs bit for bit.
constexpr int tv_block_own(size_t) { return 4; }  // (block_threads: 4 pixels per thread up to 4096)

template <class P, int NT>
__global__ __launch_bounds__(NT) void k_tv_block(const SliceDev* sl, int tx, int ty, int k, int dir) {
  extern __shared__ char smem_raw[];
  constexpr int J = tv_block_own(sizeof(P));
  const SliceDev& d = sl[0];
  if (!slice_active(d)) return;
  using R = typename RealOf<P>::T;
  const int W = d.W, H = d.H;
  const int nbx = (W + tx - 1) / tx;
  const int b = xcd_item(blockIdx.x, gridDim.x);  // neighbouring tiles on one XCD
  const int by = b / nbx, bx = b - by * nbx;
  const int x0 = bx * tx, x1 = min(W, x0 + tx), y0 = by * ty, y1 = min(H, y0 + ty);
  const int xa = max(0, x0 - k), xb = min(W, x1 + k), ya = max(0, y0 - k), yb = min(H, y1 + k);
  const int LW = xb - xa, LN = LW * (yb - ya);  // LN <= J * NT (block_fits)
  P* sB = reinterpret_cast<P*>(smem_raw);
  P* sX = sB + LN;
  P* sY = sX + LN;
  const size_t n = size_t(W) * H;
  // the owned pixels e = threadIdx.x + j NT of the tile (x = -1: none)
  P ru[J], rf[J], rpx[J], rpy[J];
  int rx[J], ry[J];
  {
    const auto gu = gptr<const P>(dir ? d.u2 : d.u);
    const auto gb = gptr<const P>(dir ? d.ub2 : d.ub);
    const auto gp = gptr<const P>(dir ? d.p2 : d.p);
    const auto gf = gptr<const P>(d.f);
    const float rLW = 1.0f / float(LW);
    P vb[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int e = threadIdx.x + j * NT;
      rx[j] = -1;
      ry[j] = -1;
      if (e < LN) {
        const int ly = qdiv(e, LW, rLW), lx = e - ly * LW;
        rx[j] = xa + lx;
        ry[j] = ya + ly;
        const size_t gi = size_t(ry[j]) * W + rx[j];
        ru[j] = gu[gi];
        vb[j] = gb[gi];
        rpx[j] = gp[gi];
        rpy[j] = gp[gi + n];
        rf[j] = gf[gi];
      }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int e = threadIdx.x + j * NT;
      if (e < LN) {
        sB[e] = vb[j];
        sX[e] = rpx[j];
        sY[e] = rpy[j];
      }
    }
  }
  __syncthreads();
  const R sigma = R(d.sigma), tau = R(d.tau), lambda = R(d.lambda), theta = R(d.theta);
  int uxa = xa, uxb = xb, uya = ya, uyb = yb;  // where u, ubar are exact
  for (int it = 0; it < k; ++it) {
    // dual step where ubar and its forward neighbours are exact
    const int pxa = uxa, pxb = uxb == W ? W : uxb - 1, pya = uya, pyb = uyb == H ? H : uyb - 1;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = rx[j], y = ry[j];
      if (x >= pxa && x < pxb && y >= pya && y < pyb) {
        const int l = threadIdx.x + j * NT;
        const P c = sB[l], ex = sB[l + (x < W - 1)], ey = sB[l + ((y < H - 1) ? LW : 0)];
        const P tx_ = ex - c, ty_ = ey - c;
        const P ax = rpx[j] + sigma * tx_;
        const P ay = rpy[j] + sigma * ty_;
        const R pn = sqrt(abs2(ax) + abs2(ay));
        const R sp = rcp(fmax(R(1), pn));
        rpx[j] = ax * sp;
        rpy[j] = ay * sp;
        sX[l] = rpx[j];
        sY[l] = rpy[j];
      }
    }
    __syncthreads();
    // primal step where p and its backward neighbours are exact
    uxa = pxa == 0 ? 0 : pxa + 1;
    uxb = pxb;
    uya = pya == 0 ? 0 : pya + 1;
    uyb = pyb;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = rx[j], y = ry[j];
      if (x >= uxa && x < uxb && y >= uya && y < uyb) {
        const int l = threadIdx.x + j * NT;
        const P cx = rpx[j], mx = sX[x > 0 ? l - 1 : l];
        const P dx = (x < W - 1) ? cx - mx : -mx;
        const P cy = rpy[j], my = sY[y > 0 ? l - LW : l];
        const P dy = (y < H - 1) ? cy - my : -my;
        const P div = (x > 0 ? dx : cx) + (y > 0 ? dy : cy);
        const P uo = ru[j];
        P un = uo + tau * fmadd(lambda, rf[j], div);
        un = un * rcp(R(1) + tau * lambda);
        ru[j] = un;
        sB[l] = (R(1) + theta) * un - theta * uo;
      }
    }
    __syncthreads();
  }
  const auto ou = gptr<P>(dir ? d.u : d.u2);
  const auto ob = gptr<P>(dir ? d.ub : d.ub2);
  const auto op = gptr<P>(dir ? d.p : d.p2);
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int x = rx[j], y = ry[j];
    if (x >= x0 && x < x1 && y >= y0 && y < y1) {
      const int l = threadIdx.x + j * NT;
      const size_t gi = size_t(y) * W + x;
      ou[gi] = ru[j];
      ob[gi] = sB[l];
      op[gi] = rpx[j];
      op[gi + n] = rpy[j];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) gptr<int>(d.ctl)[CTL_ITER] += k;
}

// The value a dual field holds after a store to its storage type S and a
// load back (S = P: itself; fp16 storage: rounded to half).
template <class P, class S>
__device__ __forceinline__ P storage_round(P v) {
  S s[1];
  Ld<P, S>::put(s, 0, v);
  return Ld<P, S>::get(s, 0);
}

// Temporally blocked TGV denoising: tb_k times the primal step
// (TV_primal_u_kernel + TGV_primal_v_kernel) then the dual step
// (TGV_dual_p_kernel + TGV_dual_q_kernel) per launch, in the reference's
// order (tgvoptimizer.cu:119-135; tgv_kernels.cuh:225-292,568-605,647-686),
// on a tile.  Block (bx, by) owns the interior [x0, x1) x [y0, y1) and loads
// u, v (2 planes), p (2), q (3) and f on the interior grown by k pixels on
// every side (clipped to the image).  The primal step at (x, y) needs p, q at
// x - 1 and y - 1 (backward differences, none at the first column / row), the
// dual step ubar, vbar at x + 1 and y + 1 (forward differences, none at the
// last): the region where the values are exact shrinks by one pixel per step
// on each side that is not an image border, and after k iterations the
// interior is exact.  It is written with ubar and vbar to the other copy of
// the state (src -> dst).  Each thread owns up to tgv_block_own pixels of the
// tile for the whole launch and keeps their u, v and f, which only the
// pixel itself reads, in registers; ubar, vbar, p and q, which neighbours
// read, live in LDS (two pixels per thread for 16-byte values, in
// workgroups of at most 512 threads, whose registers would not hold more).
// The per-pixel arithmetic
// is that of k_tgv_primal /
// k_tgv_dual in the same order, and the duals are rounded to their storage
// type after every dual step (as a store and a load would round them), so
// the iterates equal theirs bit for bit.
constexpr int tgv_block_own(size_t value_bytes) { return value_bytes > 8 ? 2 : 3; }

template <class P, class S, int NT>
__global__ __launch_bounds__(NT) void k_tgv_block(const SliceDev* sl, int tx, int ty, int k, int dir) {
  extern __shared__ char smem_raw[];
  constexpr int J = tgv_block_own(sizeof(P));
  using R = typename RealOf<P>::T;
  using L = Ld<P, S>;
  const SliceDev& d = sl[0];
  if (!slice_active(d)) return;
  const int W = d.W, H = d.H;
  const int nbx = (W + tx - 1) / tx;
  const int b = xcd_item(blockIdx.x, gridDim.x);  // neighbouring tiles on one XCD
  const int by = b / nbx, bx = b - by * nbx;
  const int x0 = bx * tx, x1 = min(W, x0 + tx), y0 = by * ty, y1 = min(H, y0 + ty);
  const int xa = max(0, x0 - k), xb = min(W, x1 + k), ya = max(0, y0 - k), yb = min(H, y1 + k);
  const int LW = xb - xa, LN = LW * (yb - ya);  // LN <= J * NT (block_fits)
  P* sUB = reinterpret_cast<P*>(smem_raw);
  P* sVBX = sUB + LN;
  P* sVBY = sVBX + LN;
  P* sPX = sVBY + LN;
  P* sPY = sPX + LN;
  P* sQXX = sPY + LN;
  P* sQYY = sQXX + LN;
  P* sQXY = sQYY + LN;
  const size_t n = size_t(W) * H;
  // the owned pixels e = threadIdx.x + j NT of the tile (x = -1: none)
  P ru[J], rvx[J], rvy[J], rf[J];
  int rx[J], ry[J];
  {
    const auto gu = gptr<const P>(dir ? d.u2 : d.u);
    const auto gv = gptr<const P>(dir ? d.v2 : d.v);
    const auto gp = gptr<const S>(dir ? d.p2 : d.p);
    const auto gq = gptr<const S>(dir ? d.q2 : d.q);
    const auto gf = gptr<const P>(d.f);
    const float rLW = 1.0f / float(LW);
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int e = threadIdx.x + j * NT;
      rx[j] = -1;
      ry[j] = -1;
      if (e < LN) {
        const int ly = qdiv(e, LW, rLW), lx = e - ly * LW;
        rx[j] = xa + lx;
        ry[j] = ya + ly;
        const size_t gi = size_t(ry[j]) * W + rx[j];
        ru[j] = gu[gi];
        rvx[j] = gv[gi];
        rvy[j] = gv[gi + n];
        rf[j] = gf[gi];
        sPX[e] = L::get(gp, gi);
        sPY[e] = L::get(gp, gi + n);
        sQXX[e] = L::get(gq, gi);
        sQYY[e] = L::get(gq, gi + n);
        sQXY[e] = L::get(gq, gi + 2 * n);
      }
    }
  }
  __syncthreads();
  const R sigma = R(d.sigma), tau = R(d.tau), lambda = R(d.lambda), theta = R(d.theta);
  const R ia0 = rcp(R(d.alpha0)), ia1 = rcp(R(d.alpha1));
  // backward differences dmx / dmy on the tile
  auto bdx = [&](const P* a, int l, int x) {
    const P c = a[l], m = a[x > 0 ? l - 1 : l];
    const P dd = (x < W - 1) ? c - m : -m;
    return x > 0 ? dd : c;
  };
  auto bdy = [&](const P* a, int l, int y) {
    const P c = a[l], m = a[y > 0 ? l - LW : l];
    const P dd = (y < H - 1) ? c - m : -m;
    return y > 0 ? dd : c;
  };
  int pxa = xa, pxb = xb, pya = ya, pyb = yb;  // where p, q are exact
  for (int it = 0; it < k; ++it) {
    // primal step where p, q and their backward neighbours are exact
    const int uxa = pxa == 0 ? 0 : pxa + 1, uxb = pxb, uya = pya == 0 ? 0 : pya + 1, uyb = pyb;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = rx[j], y = ry[j];
      if (x >= uxa && x < uxb && y >= uya && y < uyb) {
        const int l = threadIdx.x + j * NT;
        const P uo = ru[j];
        const P div = bdx(sPX, l, x) + bdy(sPY, l, y);
        P un = uo + tau * fmadd(lambda, rf[j], div);
        un = un * rcp(R(1) + tau * lambda);
        ru[j] = un;
        sUB[l] = (R(1) + theta) * un - theta * uo;
        const P dsx = bdx(sQXX, l, x) + bdy(sQXY, l, y);
        const P dsy = bdy(sQYY, l, y) + bdx(sQXY, l, x);
        const P ox = rvx[j], oy = rvy[j];
        const P nx = ox + tau * (dsx + sPX[l]);
        const P ny = oy + tau * (dsy + sPY[l]);
        rvx[j] = nx;
        rvy[j] = ny;
        sVBX[l] = (R(1) + theta) * nx - theta * ox;
        sVBY[l] = (R(1) + theta) * ny - theta * oy;
      }
    }
    __syncthreads();
    // dual step where ubar, vbar and their forward neighbours are exact
    pxa = uxa;
    pxb = uxb == W ? W : uxb - 1;
    pya = uya;
    pyb = uyb == H ? H : uyb - 1;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = rx[j], y = ry[j];
      if (x >= pxa && x < pxb && y >= pya && y < pyb) {
        const int l = threadIdx.x + j * NT;
        const int lx = l + (x < W - 1), ly = l + ((y < H - 1) ? LW : 0);
        P px = sPX[l], py = sPY[l], qxx = sQXX[l], qyy = sQYY[l], qxy = sQXY[l];
        dual_math<P, true, false>(sUB[l], sUB[lx], sUB[ly], sVBX[l], sVBY[l], sVBX[lx], sVBY[lx], sVBX[ly], sVBY[ly],
                                  px, py, qxx, qyy, qxy, sigma, ia0, ia1, nullptr);
        sPX[l] = storage_round<P, S>(px);
        sPY[l] = storage_round<P, S>(py);
        sQXX[l] = storage_round<P, S>(qxx);
        sQYY[l] = storage_round<P, S>(qyy);
        sQXY[l] = storage_round<P, S>(qxy);
      }
    }
    __syncthreads();
  }
  const auto ou = gptr<P>(dir ? d.u : d.u2);
  const auto oub = gptr<P>(dir ? d.ub : d.ub2);
  const auto ov = gptr<P>(dir ? d.v : d.v2);
  const auto ovb = gptr<P>(dir ? d.vb : d.vb2);
  const auto op = gptr<S>(dir ? d.p : d.p2);
  const auto oq = gptr<S>(dir ? d.q : d.q2);
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int x = rx[j], y = ry[j];
    if (x >= x0 && x < x1 && y >= y0 && y < y1) {
      const int l = threadIdx.x + j * NT;
      const size_t gi = size_t(y) * W + x;
      ou[gi] = ru[j];
      oub[gi] = sUB[l];
      ov[gi] = rvx[j];
      ov[gi + n] = rvy[j];
      ovb[gi] = sVBX[l];
      ovb[gi + n] = sVBY[l];
      L::put(op, gi, sPX[l]);
      L::put(op, gi + n, sPY[l]);
      L::put(oq, gi, sQXX[l]);
      L::put(oq, gi + n, sQYY[l]);
      L::put(oq, gi + 2 * n, sQXY[l]);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) gptr<int>(d.ctl)[CTL_ITER] += k;
}

// TGV denoising primal half (tgvoptimizer.cu:119-127): TV_primal_u_kernel
// then TGV_primal_v_kernel at every pixel.
template <class P, class S, int NT>
__global__ __launch_bounds__(NT) void k_tgv_primal(const SliceDev* sl, const int2* items, int rpi) {
  const int bi = xcd_item(blockIdx.x, gridDim.x);  // neighbouring rows on one XCD
  const int2 it = items ? items[bi] : make_int2(0, bi * rpi);
  const SliceDev& d = sl[it.x];
  if (!slice_active(d)) return;
  using R = typename RealOf<P>::T;
  const int W = d.W, H = d.H;
  const size_t n = size_t(W) * H;
  const int ya = it.y, yb = min(H, ya + rpi);
  const R tau = R(d.tau), lambda = R(d.lambda), theta = R(d.theta);
  const auto u = gptr<P>(d.u);
  const auto ub = gptr<P>(d.ub);
  const auto f = gptr<const P>(d.f);
  const auto px = gptr<const S>(d.p);
  const auto py = px + n;
  const int cnt = (yb - ya) * W;
  for (int idx = threadIdx.x; idx < cnt; idx += NT) {
    const int yy = idx / W;
    const int x = idx - yy * W, y = ya + yy;
    const size_t i = size_t(y) * W + x;
    const P uo = u[i];
    const P div = dmx<P>(px, x, i, W) + dmy<P>(py, y, i, W, H);
    P un = uo + tau * fmadd(lambda, P(f[i]), div);
    un = un * rcp(R(1) + tau * lambda);
    u[i] = un;
    ub[i] = (R(1) + theta) * un - theta * uo;
    vstep_pixel<P, S>(d, x, y, tau, theta);
  }
}

// TGV denoising dual half (tgvoptimizer.cu:130-136).  Gap: P(ubar, vbar).
template <class P, class S, bool GAP, int NT>
__global__ __launch_bounds__(NT) void k_tgv_dual(const SliceDev* sl, const int2* items, int rpi,
                                                     double* part) {
  const int bi = xcd_item(blockIdx.x, gridDim.x);  // neighbouring rows on one XCD
  const int2 it = items ? items[bi] : make_int2(0, bi * rpi);
  const SliceDev& d = sl[it.x];
  if (!slice_active(d)) return;
  using R = typename RealOf<P>::T;
  const int W = d.W, H = d.H;
  const int ya = it.y, yb = min(H, ya + rpi);
  const R sigma = R(d.sigma), a0 = R(d.alpha0), a1 = R(d.alpha1);
  const auto ub = gptr<const P>(d.ub);
  const auto f = gptr<const P>(d.f);
  double t[kGapTerms] = {0};
  const int cnt = (yb - ya) * W;
  for (int idx = threadIdx.x; idx < cnt; idx += NT) {
    const int yy = idx / W;
    const int x = idx - yy * W, y = ya + yy;
    dual_pixel<P, S, true, GAP>(d, x, y, sigma, rcp(a0), rcp(a1), t);
    if (GAP) {
      const size_t i = size_t(y) * W + x;
      t[3] += dabs2(ub[i] - f[i]);
    }
  }
  if (GAP) block_sum<4, NT>(reinterpret_cast<double(&)[4]>(t), part + size_t(bi) * kGapTerms);
  if (ya == 0 && threadIdx.x == 0) gptr<int>(d.ctl)[CTL_ITER] += 1;
}

// Two pixels of storage type S side by side (one 8- or 16-byte access).
template <class S> struct alignas(2 * sizeof(S)) Pair {
  S a, b;
};

// k_tgv_dual with two horizontally adjacent pixels per thread (W even): the
// duals, ubar and vbar move as 2-pixel vectors (8 bytes per p, q plane access
// with fp16 storage, 16 with float2), the arithmetic is dual_math per pixel,
// so the iterates equal k_tgv_dual's bit for bit.
template <class P, class S, bool GAP, int NT>
__global__ __launch_bounds__(NT) void k_tgv_dual2(const SliceDev* sl, const int2* items, int rpi, double* part) {
  const int bi = xcd_item(blockIdx.x, gridDim.x);  // neighbouring rows on one XCD
  const int2 it = items ? items[bi] : make_int2(0, bi * rpi);
  const SliceDev& d = sl[it.x];
  if (!slice_active(d)) return;
  using R = typename RealOf<P>::T;
  using L = Ld<P, S>;
  const int W = d.W, H = d.H;
  const size_t n = size_t(W) * H;
  const int ya = it.y, yb = min(H, ya + rpi);
  const R sigma = R(d.sigma), ia0 = rcp(R(d.alpha0)), ia1 = rcp(R(d.alpha1));
  const auto ub = gptr<const P>(d.ub);
  const auto ubp = gptr<const Pair<P>>(d.ub);
  const auto vbx = gptr<const P>(d.vb);
  const auto vbxp = gptr<const Pair<P>>(d.vb);
  const auto f = gptr<const P>(d.f);
  const auto pp = gptr<Pair<S>>(d.p);
  const auto qp = gptr<Pair<S>>(d.q);
  const size_t n2 = n / 2;  // pairs per plane
  double t[kGapTerms] = {0};
  const int W2 = W / 2, cnt = (yb - ya) * W2;
  for (int idx = threadIdx.x; idx < cnt; idx += NT) {
    const int yy = idx / W2;
    const int x = 2 * (idx - yy * W2), y = ya + yy;
    const size_t i = size_t(y) * W + x, i2 = i / 2;
    const bool last = x + 2 >= W, bottom = y >= H - 1;
    const size_t dn2 = bottom ? i2 : i2 + W2;  // pair below (the pair itself on the last row)
    const Pair<P> u0 = ubp[i2], ud = ubp[dn2];
    const P u2 = last ? u0.b : ub[i + 2];
    const Pair<P> vx = vbxp[i2], vy = vbxp[n2 + i2], vxd = vbxp[dn2], vyd = vbxp[n2 + dn2];
    const P vx2 = last ? vx.b : vbx[i + 2], vy2 = last ? vy.b : vbx[n + i + 2];
    Pair<S> ps[2] = {pp[i2], pp[n2 + i2]};
    Pair<S> qs[3] = {qp[i2], qp[n2 + i2], qp[2 * n2 + i2]};
    P pxa = L::get(&ps[0].a, 0), pya = L::get(&ps[1].a, 0);
    P qxxa = L::get(&qs[0].a, 0), qyya = L::get(&qs[1].a, 0), qxya = L::get(&qs[2].a, 0);
    P pxb = L::get(&ps[0].b, 0), pyb = L::get(&ps[1].b, 0);
    P qxxb = L::get(&qs[0].b, 0), qyyb = L::get(&qs[1].b, 0), qxyb = L::get(&qs[2].b, 0);
    // pixel x (never the last column), then x + 1
    dual_math<P, true, GAP>(u0.a, u0.b, bottom ? u0.a : ud.a, vx.a, vy.a, vx.b, vy.b, bottom ? vx.a : vxd.a,
                            bottom ? vy.a : vyd.a, pxa, pya, qxxa, qyya, qxya, sigma, ia0, ia1, t);
    dual_math<P, true, GAP>(u0.b, last ? u0.b : u2, bottom ? u0.b : ud.b, vx.b, vy.b, last ? vx.b : vx2,
                            last ? vy.b : vy2, bottom ? vx.b : vxd.b, bottom ? vy.b : vyd.b, pxb, pyb, qxxb, qyyb,
                            qxyb, sigma, ia0, ia1, t);
    L::put(&ps[0].a, 0, pxa);
    L::put(&ps[0].b, 0, pxb);
    L::put(&ps[1].a, 0, pya);
    L::put(&ps[1].b, 0, pyb);
    L::put(&qs[0].a, 0, qxxa);
    L::put(&qs[0].b, 0, qxxb);
    L::put(&qs[1].a, 0, qyya);
    L::put(&qs[1].b, 0, qyyb);
    L::put(&qs[2].a, 0, qxya);
    L::put(&qs[2].b, 0, qxyb);
    pp[i2] = ps[0];
    pp[n2 + i2] = ps[1];
    qp[i2] = qs[0];
    qp[n2 + i2] = qs[1];
    qp[2 * n2 + i2] = qs[2];
    if (GAP) {
      t[3] += dabs2(u0.a - f[i]);
      t[3] += dabs2(u0.b - f[i + 1]);
    }
  }
  if (GAP) block_sum<4, NT>(reinterpret_cast<double(&)[4]>(t), part + size_t(bi) * kGapTerms);
  if (ya == 0 && threadIdx.x == 0) gptr<int>(d.ctl)[CTL_ITER] += 1;
}

// TGV dual restoration, step 1: d = div_sym q into tmp, per-item max |d|.
template <class P, class S>
__global__ __launch_bounds__(kBlock) void k_tgv_restore1(const SliceDev* sl, const int2* items, int rpi,
                                                         double* part) {
  const int bi = xcd_item(blockIdx.x, gridDim.x);  // neighbouring rows on one XCD
  const int2 it = items ? items[bi] : make_int2(0, bi * rpi);
  const SliceDev& d = sl[it.x];
  if (!slice_active(d)) return;
  const int W = d.W, H = d.H;
  const size_t n = size_t(W) * H;
  const int ya = it.y, yb = min(H, ya + rpi);
  const auto qxx = gptr<const S>(d.q);
  const auto qyy = qxx + n;
  const auto qxy = qyy + n;
  const auto dx = gptr<P>(d.tmp);
  const auto dy = dx + n;
  double m = 0;
  const int cnt = (yb - ya) * W;
  for (int idx = threadIdx.x; idx < cnt; idx += kBlock) {
    const int yy = idx / W;
    const int x = idx - yy * W, y = ya + yy;
    const size_t i = size_t(y) * W + x;
    const P a = dmx<P>(qxx, x, i, W) + dmy<P>(qxy, y, i, W, H);
    const P b = dmy<P>(qyy, y, i, W, H) + dmx<P>(qxy, x, i, W);
    dx[i] = a;
    dy[i] = b;
    m = fmax(m, sqrt(dabs2(a) + dabs2(b)));
  }
  m = block_max(m);
  if (threadIdx.x == 0) part[size_t(bi) * kGapTerms + 0] = m;
}

// Step 2: g = div d, Re<g, f>, ||g||^2.
template <class P>
__global__ __launch_bounds__(kBlock) void k_tgv_restore2(const SliceDev* sl, const int2* items, int rpi,
                                                         double* part) {
  const int bi = xcd_item(blockIdx.x, gridDim.x);  // neighbouring rows on one XCD
  const int2 it = items ? items[bi] : make_int2(0, bi * rpi);
  const SliceDev& d = sl[it.x];
  if (!slice_active(d)) return;
  const int W = d.W, H = d.H;
  const size_t n = size_t(W) * H;
  const int ya = it.y, yb = min(H, ya + rpi);
  const auto dx = gptr<const P>(d.tmp);
  const auto dy = dx + n;
  const auto f = gptr<const P>(d.f);
  double t[2] = {0, 0};
  const int cnt = (yb - ya) * W;
  for (int idx = threadIdx.x; idx < cnt; idx += kBlock) {
    const int yy = idx / W;
    const int x = idx - yy * W, y = ya + yy;
    const size_t i = size_t(y) * W + x;
    const P g = dmx<P>(dx, x, i, W) + dmy<P>(dy, y, i, W, H);
    t[0] += redot(g, f[i]);
    t[1] += dabs2(g);
  }
  block_sum<2>(t, part + size_t(bi) * kGapTerms);
}

// ---------------------------------------------------------------------------
// Cartesian MRI passes.  MODE: 0 operator only, 1 TV-with-op, 2 TGV
// (TGV-MRI and TGV-with-op).  Row items are (slice, image row).
// ---------------------------------------------------------------------------
enum { M_OP = 0, M_TV = 1, M_TGV = 2 };


// LDS layout of the row FFT passes: twiddles (W), two FFT buffers (cb * W
// each), listed columns (nc ints).
template <class C>
struct RowLds {
  C* tw;
  C* a;
  C* b;
  int* cols;
  __device__ RowLds(char* raw, int W, int cb, bool padded = true) {
    const int n = padded ? lds_pad(cb * W) : cb * W;
    tw = reinterpret_cast<C*>(raw);
    a = tw + W;
    b = a + n;
    cols = reinterpret_cast<int*>(b + n);
  }
};

// A short global array (twiddles, the column list) staged through registers:
// with a compile-time length the loads are issued at the start of a kernel
// together with its other loads and written to LDS after the first wait,
// instead of a load-wait-write loop that costs a memory latency of its own.
template <class T, int NF, int NT>
struct Staged {
  static constexpr int K = NF > 0 ? (NF + NT - 1) / NT : 1;
  T v[K];
  // branch-free (lanes past n load element n - 1, never stored), so that the
  // compiler's waits count these loads exactly
  template <class A>
  __device__ __forceinline__ void load(A* src, int n) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = src[min(int(threadIdx.x) + k * NT, n - 1)];
  }
  __device__ __forceinline__ void store(T* dst, int n) const {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < n) dst[e] = v[k];
    }
  }
};
template <class T, int NF, int NT, class A>
__device__ __forceinline__ void copy_to_lds(T* dst, A* src, int n) {
  for (int e = threadIdx.x; e < n; e += NT) dst[e] = src[e];
}

// Row pass, forward.  Item = (slice, y * ngr + g): coil group g of image row
// y.  The groups of row y also perform the dual step of row y (TV or TGV),
// each on its own range of pixels.  Then for
// each coil of the group: S_c * src row -> ifftshift -> PE-FFT -> fftshift
// -> sampled columns -> T1 [c][y][ks].
// With ndual > 0 (a single slice, dense items) the dual step is left to
// ndual further workgroups after the nitems row items, each on a contiguous
// range of pixels (gap terms at part[nitems + b]), so that the FFT chains of
// the row items do not wait for it.
// The gap-evaluation variant holds extra FP64 accumulators; it is asked to
// fit five waves per SIMD, as many as the plain variant, so that the same
// tile shapes keep every block resident.
// Dual step (TV or TGV, MODE) of a single dense slice done by workgroups of
// their own, launched after the item workgroups of a pass: workgroup b of nd
// takes a contiguous pixel range, and the ranges go to XCDs in consecutive
// runs, so that the next row's pixels (forward differences) are mostly in
// the same L2.  Gap terms (GAP) to part[b].
template <class R, int MODE, bool GAP, int NT>
__device__ __forceinline__ void dual_workgroup(const SliceDev& d, int nd, double* part) {
  using C = Cx<R>;
  const int b = xcd_item(int(blockIdx.x) - int(gridDim.x) + nd, nd), W = d.W, WH = W * d.H;
  const int chunk = (WH + nd - 1) / nd;
  const int pa = min(WH, b * chunk), pe = min(WH, pa + chunk);
  const int k = gptr<int>(d.ctl)[CTL_ITER];
  const R sigma = R(d.sigma);
  const R ia0 = rcp(sched_alpha<R>(d, d.alpha0, d.alpha0_end, k));
  const R ia1 = rcp(sched_alpha<R>(d, d.alpha1, d.alpha1_end, k));
  const float rW = 1.0f / float(W);
  double t[kGapTerms] = {0};
  for (int i = pa + threadIdx.x; i < pe; i += NT) {
    const int y = qdiv(i, W, rW), x = i - y * W;
    dual_pixel<C, C, MODE == M_TGV, GAP>(d, x, y, sigma, ia0, ia1, t);
  }
  if (GAP) block_sum<3, NT>(reinterpret_cast<double(&)[3]>(t), part + size_t(b) * kGapTerms);
}

template <class R, int MODE, bool GAP, int NT, int NF>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(GAP ? 5 : 1))) void k_rows_fwd(const SliceDev* sl, const int2* items, int nitems, int cb,
                                                 double* part, int ndual) {
  extern __shared__ char smem_raw[];
  using C = Cx<R>;
  if (MODE != M_OP && int(blockIdx.x) >= nitems) {
    const SliceDev& d = sl[0];
    if (!slice_active(d)) return;
    dual_workgroup<R, MODE, GAP, NT>(d, ndual, part + size_t(nitems) * kGapTerms);
    return;
  }
  const int li = xcd_item(blockIdx.x, nitems);
  const int2 it = items ? items[li] : make_int2(0, li);  // null: the dense single-slice list
  const SliceDev& d = sl[it.x];
  const int W = NF > 0 ? NF : d.W, H = d.H, nc = d.ncols;
  const int y = it.y / d.ngr, g = it.y - y * d.ngr;
  const int c0 = g * cb, ncb = min(cb, d.C - c0);
  RowLds<C> L(smem_raw, W, cb);
  const auto gtw = gptr<const C>(d.planW.tw);
  const auto gcols = gptr<const int>(d.cols);
  Staged<C, NF, NT> s_tw;
  Staged<int, NF, NT> s_cols;
  if (NF > 0) {
    s_tw.load(gtw, W);
    s_cols.load(gcols, nc);
  } else {
    copy_to_lds<C, NF, NT>(L.tw, gtw, W);
    copy_to_lds<int, NF, NT>(L.cols, gcols, nc);
  }
  const auto src = gptr<const C>(MODE == M_OP ? d.op_in : d.ub) + size_t(y) * W;
  const auto S = gptr<const C>(d.S);
  const int G = W * d.Hk;
  const auto Sg = S ? S + size_t(c0) * G + size_t(y + d.y0) * W : S;  // coil c0, row y + y0
  const float rW = 1.0f / float(W), rnc = 1.0f / float(max(nc, 1));
  // With a compile-time length and at most kPrefetch elements per thread the
  // row and sensitivity values are loaded into registers first and combined
  // after the dual step, so that both sets of loads are in flight together.
  C ps[kPrefetch], pS[kPrefetch];
  const bool pf = NF > 0 && ncb * W <= kPrefetch * NT;
  if (pf) {
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncb * W) {
        const int cl = divn<NF>(e, W, rW), x = e - cl * W;
        ps[k] = src[x];
        if (S) pS[k] = Sg[cl * G + x];
      }
    }
  } else {
    for (int e = threadIdx.x; e < ncb * W; e += NT) {
      const int cl = divn<NF>(e, W, rW), x = e - cl * W;
      C z = src[x];
      if (S) z = z * Sg[cl * G + x];
      L.a[lp<NF>(cl * W + ishift_dst(x, W))] = z;
    }
  }
  // A finished slice's blocks leave before their first store; the check comes
  // after the loads above are issued, so its control-word load overlaps them.
  if (!slice_active(d)) return;
  double t[kGapTerms] = {0};
  if (MODE != M_OP && ndual == 0) {
    // The dual step of row y is shared out over the row's coil groups in
    // pixel ranges of a multiple of 32, so no group carries all of it.
    const int chunk = ((W + d.ngr - 1) / d.ngr + 31) & ~31;
    const int xa = min(W, g * chunk), xe = min(W, xa + chunk);
    const int k = gptr<int>(d.ctl)[CTL_ITER];
    const R sigma = R(d.sigma);
    const R a0 = sched_alpha<R>(d, d.alpha0, d.alpha0_end, k);
    const R a1 = sched_alpha<R>(d, d.alpha1, d.alpha1_end, k);
    const R ia0 = rcp(a0), ia1 = rcp(a1);
    for (int x = xa + threadIdx.x; x < xe; x += NT) dual_pixel<C, C, MODE == M_TGV, GAP>(d, x, y, sigma, ia0, ia1, t);
  }
  if (NF > 0) {
    s_tw.store(L.tw, W);
    s_cols.store(L.cols, nc);
  }
  if (pf) {
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncb * W) {
        const int cl = divn<NF>(e, W, rW), x = e - cl * W;
        L.a[lp<NF>(cl * W + ishift_dst(x, W))] = S ? ps[k] * pS[k] : ps[k];
      }
    }
  }
  __syncthreads();
  C* res = fft_n<R, NT, false, NF>(L.a, L.b, ncb, d.planW, L.tw);
  // T1 is row-major per coil, [c][y][ks], like T2: a row writes nc
  // contiguous entries per coil (the column pass gathers with a stride)
  const auto T1 = gptr<C>(d.T1) + (size_t(c0) * H + y) * nc;  // coil c0, row y
  for (int e = threadIdx.x; e < ncb * nc; e += NT) {
    const int cl = qdiv(e, nc, rnc), ks = e - cl * nc;
    T1[size_t(cl) * H * nc + ks] = res[lp<NF>(cl * W + fshift_src(L.cols[ks], W))];
  }
  if (GAP) block_sum<3, NT>(reinterpret_cast<double(&)[3]>(t), part + size_t(li) * kGapTerms);
}

// Column pass.  CMODE: 0 solver (forward, r prox, adjoint), 1 operator
// forward (writes op_out {W, Hk, C}), 2 operator adjoint (reads op_in).
// PF (solver, compile-time length, at most kColPrefetchCols columns per
// item): the gather and the prox operands go through registers.  It is a
// template argument so that the other gather, a loop, is not compiled into
// the same kernel: the compiler's waits after the join of the two would count
// the prox loads as outstanding before the FFT.
enum { CM_SOLVE = 0, CM_FWD = 1, CM_ADJ = 2 };

// DUAL (M_TV or M_TGV; a single dense slice): ndual workgroups after the
// column items do the dual step of the iteration (dual_workgroup, gap terms
// to part_dual), which reads and writes none of what the column items do;
// with the adjoint row pass also writing the next T1 (k_rows_adj NEXT) an
// iteration is these two launches.
template <class R, int CMODE, bool GAP, int NT, int NF, bool PF = false, int DUAL = 0>
__global__ __launch_bounds__(NT) void k_cols(const SliceDev* sl, const int2* items, int nitems, int cg,
                                             double* part, int ndual, double* part_dual) {
  static_assert(!PF || (NF > 0 && CMODE == CM_SOLVE), "register prefetch needs a compile-time length");
  static_assert(DUAL == 0 || CMODE == CM_SOLVE, "dual workgroups: solver iterations");
  extern __shared__ char smem_raw[];
  if (DUAL != 0 && int(blockIdx.x) >= nitems) {
    const SliceDev& d = sl[0];
    if (!slice_active(d)) return;
    dual_workgroup<R, DUAL, GAP, NT>(d, ndual, part_dual);
    return;
  }
  const int li = xcd_item(blockIdx.x, nitems);
  const int2 it = items ? items[li] : make_int2(0, li);  // null: the dense single-slice list
  const SliceDev& d = sl[it.x];
  using C = Cx<R>;
  const int W = d.W, H = d.H, Hk = NF > 0 ? NF : d.Hk, y0 = d.y0, nc = d.ncols;
  const int c = it.y / d.ngroups, g = it.y - c * d.ngroups;
  const int ks0 = g * cg, ncg = min(cg, nc - ks0);
  const R scale = R(1) / sqrt(R(W) * R(Hk));
  const auto Mc = gptr<const R>(d.Mc);
  C* tw = reinterpret_cast<C*>(smem_raw);
  C* bufA = tw + Hk;
  C* bufB = bufA + lds_pad(cg * Hk);
  const auto gtw = gptr<const C>(d.planH.tw);
  Staged<C, NF, NT> s_tw;
  if (NF > 0) s_tw.load(gtw, Hk);
  else copy_to_lds<C, NF, NT>(tw, gtw, Hk);
  const float rHk = 1.0f / float(Hk), rncg = 1.0f / float(ncg);
  C* res;
  double t[2] = {0, 0};
  // TGV_prox_r_kernel (tgv_kernels.cuh:212-223) operands; columns ks0.. of
  // coil c are contiguous in the compact layout.  With a compile-time length
  // they are loaded into registers, issued after the T1 gather: vector memory
  // loads return in issue order, so the gather's wait before the LDS stores
  // leaves them in flight, and they arrive during the forward FFT.
  const R sigma = R(d.sigma), lambda = R(d.lambda), inv_sl = rcp(R(1) + sigma * lambda);
  const auto fk = gptr<const C>(d.fk) + (size_t(c) * nc + ks0) * Hk;
  const auto rk = gptr<C>(d.rk) + (size_t(c) * nc + ks0) * Hk;
  const auto mk = Mc + size_t(ks0) * Hk;
  constexpr int KP = NF > 0 ? (kColPrefetchCols * NF + NT - 1) / NT : 1;
  C pf_f[KP], pf_r[KP];
  R pf_m[KP];
  constexpr bool pf = PF;
  // the gather of the zero-padded, ifftshifted columns of T1 (coil c, columns
  // ks0.., a stride of nc in y): into registers with a compile-time length
  const auto T1 = gptr<const C>(d.T1) + size_t(c) * H * nc + ks0;
  C pf_t[KP];
  if (pf) {
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncg * Hk) {
        const int j = qdiv(e, ncg, rncg), cl = e - j * ncg;
        const int yy = ishift_src(j, Hk) - y0;
        pf_t[k] = (yy >= 0 && yy < H) ? C(T1[yy * nc + cl]) : pzero<C>();
      }
    }
    // branch-free (lanes past the columns load the last element, unused), so
    // that the compiler's wait for the gather counts these loads as issued
    const int last = ncg * Hk - 1;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int e = min(int(threadIdx.x) + k * NT, last);
      pf_f[k] = fk[e];
      pf_r[k] = rk[e];
      pf_m[k] = mk[e];
    }
  }
  if (!slice_active(d)) return;  // after the loads are issued, before any store
  if (CMODE != CM_ADJ) {
    // registers -> LDS with the prefetch, else the gather straight into LDS
    if (pf) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int e = threadIdx.x + k * NT;
        if (e < ncg * Hk) {
          const int j = qdiv(e, ncg, rncg), cl = e - j * ncg;
          bufA[lp<NF>(cl * Hk + j)] = pf_t[k];
        }
      }
    } else {
      for (int e = threadIdx.x; e < ncg * Hk; e += NT) {
        const int j = qdiv(e, ncg, rncg), cl = e - j * ncg;
        const int fe = ishift_src(j, Hk);
        const int yy = fe - y0;
        bufA[lp<NF>(cl * Hk + j)] = (yy >= 0 && yy < H) ? C(T1[yy * nc + cl]) : pzero<C>();
      }
    }
    if (NF > 0) s_tw.store(tw, Hk);
    __syncthreads();
    res = fft_n<R, NT, false, NF>(bufA, bufB, ncg, d.planH, tw);
    C* other = (res == bufA) ? bufB : bufA;
    if (CMODE == CM_FWD) {
      const auto out = gptr<C>(d.op_out);
      for (int e = threadIdx.x; e < ncg * Hk; e += NT) {
        const int cl = divn<NF>(e, Hk, rHk), kfe = e - cl * Hk;
        const R m = Mc[size_t(ks0 + cl) * Hk + kfe];
        out[(size_t(c) * Hk + kfe) * W + gptr<const int>(d.cols)[ks0 + cl]] = (m * scale) * res[lp<NF>(cl * Hk + fshift_src(kfe, Hk))];
      }
      return;
    }
    // prox of r on A ubar - f
    auto prox = [&](int e, R m, C f, C r) {
      const int cl = divn<NF>(e, Hk, rHk), kfe = e - cl * Hk;
      const C a = (m * scale) * res[lp<NF>(cl * Hk + fshift_src(kfe, Hk))];
      const C tmp = a - f;
      const C rn = (r + sigma * tmp) * inv_sl;
      rk[e] = rn;
      if (GAP) {
        t[0] += dabs2(tmp);
        t[1] += dabs2(tmp - lambda * rn);
      }
      other[lp<NF>(cl * Hk + ishift_dst(kfe, Hk))] = m * rn;
    };
    if (pf) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int e = threadIdx.x + k * NT;
        if (e < ncg * Hk) prox(e, pf_m[k], pf_f[k], pf_r[k]);
      }
    } else {
      for (int e = threadIdx.x; e < ncg * Hk; e += NT) prox(e, mk[e], fk[e], rk[e]);
    }
    __syncthreads();
    res = fft_n<R, NT, true, NF>(other, res, ncg, d.planH, tw);
  } else {
    if (NF > 0) s_tw.store(tw, Hk);
    const auto in = gptr<const C>(d.op_in);
    for (int e = threadIdx.x; e < ncg * Hk; e += NT) {
      const int cl = divn<NF>(e, Hk, rHk), kfe = e - cl * Hk;
      const R m = Mc[size_t(ks0 + cl) * Hk + kfe];
      bufA[lp<NF>(cl * Hk + ishift_dst(kfe, Hk))] = m * in[(size_t(c) * Hk + kfe) * W + gptr<const int>(d.cols)[ks0 + cl]];
    }
    __syncthreads();
    res = fft_n<R, NT, true, NF>(bufA, bufB, ncg, d.planH, tw);
  }
  const auto T2 = gptr<C>(d.T2) + size_t(c) * H * nc + ks0;
  for (int e = threadIdx.x; e < ncg * H; e += NT) {
    const int yy = qdiv(e, ncg, rncg), cl = e - yy * ncg;
    T2[yy * nc + cl] = scale * res[lp<NF>(cl * Hk + fshift_src(yy + y0, Hk))];
  }
  if (GAP) block_sum<2, NT>(t, part + size_t(li) * kGapTerms);
}

// Primal step of pixel (x, y) given w = (A^* r)(x, y):
// TGV_primal_u_noprox_kernel (tgv_kernels.cuh:432-467) and, for TGV,
// TGV_primal_v_kernel (:647-686).  Adds |w - div p|^2 and |p + div_sym q|^2
// to t for the gap.
template <class R, int MODE, bool GAP>
__device__ __forceinline__ void mri_primal_pixel(const SliceDev& d, int x, int y, Cx<R> w, R tau, R theta,
                                                 double* t) {
  using C = Cx<R>;
  const int W = d.W, H = d.H;
  const size_t n = size_t(W) * H, i = size_t(y) * W + x;
  const auto u = gptr<C>(d.u);
  const auto ub = gptr<C>(d.ub);
  const auto px = gptr<const C>(d.p);
  const auto py = px + n;
  const C uo = u[i];
  const C div = dmx<C>(px, x, i, W) + dmy<C>(py, y, i, W, H);
  const C un = uo + tau * (div - w);
  u[i] = un;
  ub[i] = (R(1) + theta) * un - theta * uo;
  if (GAP) t[0] += dabs2(w - div);
  if (MODE == M_TGV) {
    const double rv = vstep_pixel<C, C>(d, x, y, tau, theta);
    if (GAP) t[1] += rv;
  }
}

// Register slots per thread of the adjoint row pass: the 1024-thread variant
// must stay within 64 VGPRs (two blocks per CU), so it holds fewer.
template <int NT> constexpr int rows_adj_ps() { return NT >= 1024 ? 6 : kPrefetch; }
template <int NT> constexpr int rows_adj_pt() { return NT >= 1024 ? 2 : 4; }
// the straight-line path loads its sensitivity slots unconditionally, so it
// holds exactly as many as it may use (5 coils per thread group: 15 coils
// on rows of 320, 20 on rows of 256)
constexpr int kRowsAdjFusedSlots = 5;

// The adjoint row pass of one row y holding all coils (FUSED in k_rows_adj):
// T2 row -> ifftshift -> inverse PE-FFT -> fftshift -> * conj(S_c), summed
// over the coils (thread group gi takes coils [gi ch, (gi + 1) ch) of pixel
// gx, the group sums added in group order), then the u step of row y
// (TGV_primal_u_noprox_kernel, tgv_kernels.cuh:432-467).  Load order: the
// twiddles, the T2 gather, u and p of the primal step (their arithmetic
// waits for them), then the sensitivities, which the FFT does not need.
template <class R, int MODE, bool GAP, int NT, int NF, bool NEXT>
__device__ __forceinline__ void rows_adj_fused(const SliceDev& d, int y, double* part_rows, char* smem_raw) {
  using C = Cx<R>;
  constexpr int W = NF, PS = kRowsAdjFusedSlots, PT = rows_adj_pt<NT>();
  // Lengths with the padded layout: one padded slot range per coil, whose
  // T2 scatter, transforms (fft_wave) and T1 entries one wavefront does
  // alone (coils wave, wave + NW, ...; lb holds the coil-group sums), so the
  // block waits at four barriers.  Other lengths: two unpadded buffers and
  // workgroup-wide stages (lp without padding is the identity for them).
  constexpr bool IP = ct_pad_ok(NF);
  constexpr int NW = NT / 64, NPW = W + W / 8;  // waves; padded slots per coil
  constexpr int SPW = (kRowsAdjFusedSlots * (NT / NF) + NW - 1) / NW;  // coils per wave
  const int H = d.H, nc = d.ncols, ncb = d.C;
  const int ngs = min(NT / W, ncb), ch = (ncb + ngs - 1) / ngs;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  C* const tw = reinterpret_cast<C*>(smem_raw);
  C* const la = tw + W;
  C* const lb = la + (IP ? ncb * NPW : ncb * W);
  Staged<C, NF, NT> s_tw;
  s_tw.load(gptr<const C>(d.planW.tw), W);
  if constexpr (!IP)
    for (int e = threadIdx.x; e < ncb * W; e += NT) la[e] = pzero<C>();
  // T2 row y of every coil and the PE column of each sampled entry (padded
  // layout: entries ks = lane + 64 k of the wave's coils, branch-free with
  // clamped indices; else entry e = tid + k NT of the row)
  const float rnc = 1.0f / float(max(nc, 1));
  const auto T2 = gptr<const C>(d.T2) + size_t(y) * nc;
  const auto gcols = gptr<const int>(d.cols);
  const int cstride = H * nc;
  C pT[IP ? SPW : 1][PT];
  int pK[PT];
  if constexpr (IP) {
#pragma unroll
    for (int k = 0; k < PT; ++k) pK[k] = gcols[min(lane + 64 * k, nc - 1)];
#pragma unroll
    for (int sq = 0; sq < SPW; ++sq)
#pragma unroll
      for (int k = 0; k < PT; ++k) pT[sq][k] = T2[min(wave + NW * sq, ncb - 1) * cstride + min(lane + 64 * k, nc - 1)];
  } else {
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncb * nc) {
        const int cl = qdiv(e, nc, rnc), ks = e - cl * nc;
        pT[0][k] = T2[cl * cstride + ks];
        pK[k] = gcols[ks];
      }
    }
  }
  // u(x, y) and the neighbours of div p (p is final for this iteration),
  // loaded branch-free (threads past the row load pixel 0, unused)
  const R tau = R(d.tau), theta = R(d.theta);
  // (NEXT: every thread group takes the u step of its own pixel gx, each
  // forming the same values, so that its forward products need no exchange
  // of the new ubar through LDS)
  const int gi = threadIdx.x / W, gx = threadIdx.x - gi * W;
  const int x = threadIdx.x;
  const bool upix = NEXT ? gi < ngs : x < W;
  const int xr = NEXT ? (gi < ngs ? gx : 0) : (x < W ? x : 0);
  const size_t n = size_t(W) * H, i = size_t(y) * W + xr;
  const auto pp = gptr<const C>(d.p);
  const C e_uo = gptr<const C>(d.u)[i];
  const C cx = pp[i], mx = pp[xr > 0 ? i - 1 : i];
  const C cy = pp[n + i], my = pp[n + (y > 0 ? i - W : i)];
  // sensitivities, branch-free (spare slots load a valid element, unused)
  const int G = W * d.Hk;
  const auto Sg = gptr<const C>(d.S) + size_t(y + d.y0) * W + gx;
  C pS[PS];
#pragma unroll
  for (int k = 0; k < PS; ++k) pS[k] = Sg[size_t(min(gi * ch + k, ncb - 1)) * G];
  if (!slice_active(d)) return;  // after the loads are issued, before any store
  // descriptor fields of the stores, read before the asm below (after it
  // the compiler no longer takes the descriptor for unmodified, and would
  // read them with vector loads at the end of the block)
  const auto gu = gptr<C>(d.u);
  const auto gub = gptr<C>(d.ub);
  const auto gctl = gptr<int>(d.ctl);
  const int row_beg = d.row_beg;
  // div p = dmx + dmy (the same operations as dmx / dmy), formed here: the
  // empty asm keeps the compiler from sinking it past the FFT, where the
  // four neighbours would hold registers instead of the one sum
  const C dx = csel(xr > 0, csel(xr < W - 1, cx - mx, -mx), cx);
  const C dy = csel(y > 0, csel(y < H - 1, cy - my, -my), cy);
  C e_div = dx + dy;
  asm volatile("" : "+v"(e_div.re), "+v"(e_div.im));
  C* res;
  C* other;
  if constexpr (IP) {
    // each wave: its coils' slots zeroed, their T2 entries at the
    // ifftshifted PE slots, then (twiddles in) their inverse transforms
#pragma unroll
    for (int sq = 0; sq < SPW; ++sq) {
      const int cl = wave + NW * sq;
      if (cl < ncb) {
        C* const sp = la + cl * NPW;
        for (int e = lane; e < NPW; e += 64) sp[e] = pzero<C>();
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < PT; ++k) {
          const int j = ishift_dst(pK[k], W);
          if (lane + 64 * k < nc) sp[j + (j >> 3)] = pT[sq][k];
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
    s_tw.store(tw, W);
    __syncthreads();  // twiddles
#pragma unroll
    for (int sq = 0; sq < SPW; ++sq) {
      const int cl = wave + NW * sq;
      if (cl < ncb) fft_wave<R, NF, 0, 1, true>(la + cl * NPW, tw);
    }
    res = la;
    other = lb;
  } else {
    s_tw.store(tw, W);
    __syncthreads();  // zero fill, twiddles
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncb * nc) {
        const int cl = qdiv(e, nc, rnc);
        la[lp<NF, false>(cl * W + ishift_dst(pK[k], W))] = pT[0][k];
      }
    }
    __syncthreads();
    res = fft_n<R, NT, true, NF, false>(la, lb, ncb, d.planW, tw);
    other = (res == la) ? lb : la;
  }
  __syncthreads();  // every coil's inverse transform (fft_n: its last stage's)
  if (gi < ngs) {
    C acc = pzero<C>();
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      const int cl = gi * ch + k;
      if (k < ch && cl < ncb) acc = acc + mulconj(res[lp<NF, IP>(cl * W + fshift_src(gx, W))], pS[k]);
    }
    other[gi * W + gx] = acc;
  }
  __syncthreads();
  double t[2] = {0, 0};
  if (upix) {
    C w = other[xr];
    for (int q = 1; q < ngs; ++q) w = w + other[q * W + xr];
    const C un = e_uo + tau * (e_div - w);
    const C ubn = (R(1) + theta) * un - theta * e_uo;
    if (!NEXT || gi == 0) {  // (NEXT: one thread group stores and counts the pixel)
      gu[i] = un;
      gub[i] = ubn;
      if (GAP) t[0] += dabs2(w - e_div);
    }
    if constexpr (NEXT) {
      // the next forward row pass's input (what k_rows_fwd's row items form
      // from ubar, the same operations): S_c ubar, ifftshifted, into the
      // inverse transform's buffer (consumed above the barrier)
#pragma unroll
      for (int k = 0; k < PS; ++k) {
        const int cl = gi * ch + k;
        if (k < ch && cl < ncb) res[lp<NF, IP>(cl * W + ishift_dst(gx, W))] = ubn * pS[k];
      }
    }
  }
  if (GAP) block_sum<2, NT>(t, part_rows + size_t(row_beg + y) * kGapTerms);
  if (y == 0 && threadIdx.x == 0) {
    gctl[CTL_ITER] += 1;
    gctl[CTL_RUPD] += 1;
  }
  if constexpr (NEXT) {
    // The rest of the next iteration's forward row pass on the new ubar row:
    // PE-FFT -> fftshift -> the sampled columns -> T1 [c][y][ks] (the
    // entries of the T2 gather, in the same layout).
    __syncthreads();  // products in res; the coil sums in other are read
    const auto T1 = gptr<C>(d.T1) + size_t(y) * nc;
    if constexpr (IP) {
#pragma unroll
      for (int sq = 0; sq < SPW; ++sq) {
        const int cl = wave + NW * sq;
        if (cl < ncb) {
          C* const sp = res + cl * NPW;
          fft_wave<R, NF, 0, 1, false>(sp, tw);
#pragma unroll
          for (int k = 0; k < PT; ++k) {
            const int ks = lane + 64 * k, j = fshift_src(pK[k], W);
            if (ks < nc) T1[size_t(cl) * H * nc + ks] = sp[j + (j >> 3)];
          }
        }
      }
    } else {
      C* fres = fft_n<R, NT, false, NF, false>(res, other, ncb, d.planW, tw);
#pragma unroll
      for (int k = 0; k < PT; ++k) {
        const int e = threadIdx.x + k * NT;
        if (e < ncb * nc) {
          const int cl = qdiv(e, nc, rnc);
          T1[size_t(cl) * H * nc + (e - cl * nc)] = fres[lp<NF, false>(cl * W + fshift_src(pK[k], W))];
        }
      }
    }
  }
}

// Row pass, adjoint.  Item = (slice, y * ngr_a + g): for each coil of group g,
// T2 row -> ifftshift -> inverse PE-FFT -> fftshift -> * conj(S_c), summed
// over the group's coils in order.  With several coil groups (ngr_a > 1) the
// sum goes to wpart[g][y][x] and k_mri_primal completes the step; with one
// group the block holds all of (A^* r)(., y) and performs the primal step of
// row y itself (MODE M_OP: writes A^* r to op_out).  The 1024-thread
// variant on swz-permuted buffers is held to 64 VGPRs (its slot arithmetic
// would take more), so that two workgroups share a CU.
// FUSED (host-checked, see rows_adj_fused_ok): a compile-time row length, one
// coil group per row holding every coil (the u step follows in the block),
// 2 W <= NT, the coil products in registers (every thread group's coils fit
// its slots), the T2 row in registers, sensitivities present, and no v step
// in the row blocks (TGV: separate workgroups).  That path is straight-line,
// so the compiler's waits can leave the sensitivity loads, issued after the
// T2 gather and the primal-step loads, in flight during the inverse FFT (the
// other paths' runtime branches make it wait for every load before it).
// NEXT (with FUSED): the row blocks also compute the next iteration's T1
// from their new ubar row, which k_rows_fwd's row items would otherwise do
// (rows_adj_fused).  Its gap-evaluation variant (one iteration in
// gap_every) is not held to 64 VGPRs: within them it spilled.
template <class R, int MODE, bool GAP, int NT, int NF, bool FUSED = false, bool NEXT = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT >= 1024 && NF > 0 && ct_pad_ok(NF) && !(GAP && NEXT) ? 8 : 1))) void k_rows_adj(const SliceDev* sl, const int2* items, int nitems, int cb,
                                                 double* part_rows, int nvstep) {
  extern __shared__ char smem_raw[];
  using C = Cx<R>;
  static_assert(!FUSED || (NF > 0 && MODE != M_OP && NT == kRowAdjThreadsWide), "fused adjoint row path");
  static_assert(!NEXT || FUSED, "next forward rows: the fused path");
  if (int(blockIdx.x) >= nitems) {
    // nvstep workgroups after the row items (a single slice): the v step of
    // TGV_primal_v_kernel on contiguous pixel ranges, where the row items
    // would do it before their FFT (gap terms at part_rows[row_beg + H + b]);
    // elsewhere they only write zero gap terms.
    const SliceDev& d = sl[0];
    if (!slice_active(d)) return;
    // consecutive pixel ranges on one XCD (backward differences of q)
    const int b = xcd_item(int(blockIdx.x) - nitems, nvstep), W = d.W, WH = W * d.H;
    double t[2] = {0, 0};
    if (MODE == M_TGV && NF > 0 && d.ngr_a == 1 && 2 * W <= NT) {
      const int chunk = (WH + nvstep - 1) / nvstep;
      const int pa = min(WH, b * chunk), pe = min(WH, pa + chunk);
      const R tau = R(d.tau), theta = R(d.theta);
      const float rW = 1.0f / float(W);
      for (int i = pa + threadIdx.x; i < pe; i += NT) {
        const int y = qdiv(i, W, rW), x = i - y * W;
        const double rv = vstep_pixel<C, C>(d, x, y, tau, theta);
        if (GAP) t[1] += rv;
      }
    }
    if (GAP) block_sum<2, NT>(t, part_rows + size_t(d.row_beg + d.H + b) * kGapTerms);
    return;
  }
  const int li = xcd_item(blockIdx.x, nitems);
  const int2 it = items ? items[li] : make_int2(0, li);  // null: the dense single-slice list
  const SliceDev& d = sl[it.x];
  if constexpr (FUSED) {
    rows_adj_fused<R, MODE, GAP, NT, NF, NEXT>(d, it.y, part_rows, smem_raw);
    return;
  }
  const int W = NF > 0 ? NF : d.W, H = d.H, nc = d.ncols;
  const int y = it.y / d.ngr_a, g = it.y - y * d.ngr_a;
  const int c0 = g * cb, ncb = min(cb, d.C - c0);
  // the wide variant keeps two blocks per CU only without the LDS padding
  constexpr bool LP = NT != kRowAdjThreadsWide;
  RowLds<C> L(smem_raw, W, cb, LP);
  const auto gtw = gptr<const C>(d.planW.tw);
  const auto gcols = gptr<const int>(d.cols);
  Staged<C, NF, NT> s_tw;
  if (NF > 0) s_tw.load(gtw, W);
  else copy_to_lds<C, NF, NT>(L.tw, gtw, W);
  for (int e = threadIdx.x; e < (LP ? lds_pad(ncb * W) : ncb * W); e += NT) L.a[e] = pzero<C>();
  const float rW = 1.0f / float(W), rnc = 1.0f / float(max(nc, 1));
  // sensitivities of the group's coils on row y + y0: with a compile-time
  // length loaded into registers before the T2 gather (one latency for both)
  const auto S = gptr<const C>(d.S);
  const int G = W * d.Hk;
  const auto Sg = S ? S + size_t(c0) * G + size_t(y + d.y0) * W : S;
  constexpr int PS = rows_adj_ps<NT>(), PT = rows_adj_pt<NT>();
  C pS[PS];
  // Coil sum in registers (one block holds the whole row, u step fused):
  // thread t = gi W + x takes pixel x of the coils [gi ch, (gi + 1) ch) of
  // the row, multiplies and sums them after the FFT, and the ng partial sums
  // of a pixel are added in group order.  Otherwise every thread takes
  // elements e = cl W + x in turn and one thread per pixel sums all coils.
  const int ngs = min(NT / W, ncb), ch = (ncb + ngs - 1) / max(ngs, 1);
  const bool gsum = NF > 0 && Sg && MODE != M_OP && d.ngr_a == 1 && 2 * W <= NT && ch <= PS;
  const int gi = threadIdx.x / W, gx = threadIdx.x - gi * W;
  if (gsum) {
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      const int cl = gi * ch + k;
      if (gi < ngs && k < ch && cl < ncb) pS[k] = Sg[cl * G + gx];
    }
  }
  const bool pf = !gsum && NF > 0 && Sg && ncb * W <= PS * NT;
  if (pf) {
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncb * W) {
        const int cl = divn<NF>(e, W, rW), x = e - cl * W;
        pS[k] = Sg[cl * G + x];
      }
    }
  }
  // the T2 gather (values and their PE columns) into registers with a
  // compile-time length, issued with the sensitivity loads
  const auto T2 = gptr<const C>(d.T2) + (size_t(c0) * H + y) * nc;  // row y of coil c0
  const int cstride = H * nc;
  C pT[PT];
  int pK[PT];
  const bool pft = NF > 0 && ncb * nc <= PT * NT;
  if (pft) {
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncb * nc) {
        const int cl = qdiv(e, nc, rnc), ks = e - cl * nc;
        pT[k] = T2[cl * cstride + ks];
        pK[k] = gcols[ks];
      }
    }
  } else {
    copy_to_lds<int, NF, NT>(L.cols, gcols, nc);
  }
  if (!slice_active(d)) return;  // after the loads above are issued, before any store
  // With all coils of the row in one block of at least 2 W threads, the
  // parts of the primal step that do not need A^* r run before the FFT, so
  // their loads overlap those of S and T2: threads x < W load u(x, y) and
  // form div p (p is final for this iteration), threads W + x perform the
  // whole v step of pixel x (TGV_primal_v_kernel, tgv_kernels.cuh:647-686,
  // needs only p and q).
  const R tau = R(d.tau), theta = R(d.theta);
  double t[2] = {0, 0};
  const bool early = NF > 0 && MODE != M_OP && d.ngr_a == 1 && 2 * W <= NT;
  C e_uo = pzero<C>(), e_div = pzero<C>();
  if (early) {
    const int x = threadIdx.x;
    if (x < W) {
      const size_t n = size_t(W) * H, i = size_t(y) * W + x;
      const auto px = gptr<const C>(d.p);
      e_uo = gptr<const C>(d.u)[i];
      e_div = dmx<C>(px, x, i, W) + dmy<C>(px + n, y, i, W, H);
    } else if (MODE == M_TGV && nvstep == 0 && x < 2 * W) {
      const double rv = vstep_pixel<C, C>(d, x - W, y, tau, theta);
      if (GAP) t[1] += rv;
    }
  }
  if (NF > 0) s_tw.store(L.tw, W);
  __syncthreads();  // zero fill, twiddles, column list
  if (pft) {
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncb * nc) {
        const int cl = qdiv(e, nc, rnc);
        L.a[lp<NF, LP>(cl * W + ishift_dst(pK[k], W))] = pT[k];
      }
    }
  } else {
    for (int e = threadIdx.x; e < ncb * nc; e += NT) {
      const int cl = qdiv(e, nc, rnc), ks = e - cl * nc;
      L.a[lp<NF, LP>(cl * W + ishift_dst(L.cols[ks], W))] = T2[cl * cstride + ks];
    }
  }
  __syncthreads();
  C* res = fft_n<R, NT, true, NF, LP>(L.a, L.b, ncb, d.planW, L.tw);
  C* other = (res == L.a) ? L.b : L.a;
  if (gsum) {
    if (gi < ngs) {
      C acc = pzero<C>();
#pragma unroll
      for (int k = 0; k < PS; ++k) {
        const int cl = gi * ch + k;
        if (k < ch && cl < ncb) acc = acc + mulconj(res[lp<NF, LP>(cl * W + fshift_src(gx, W))], pS[k]);
      }
      other[gi * W + gx] = acc;
    }
  } else if (pf) {
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < ncb * W) {
        const int cl = divn<NF>(e, W, rW), x = e - cl * W;
        other[lp<NF, LP>(cl * W + x)] = mulconj(res[lp<NF, LP>(cl * W + fshift_src(x, W))], pS[k]);
      }
    }
  } else {
    for (int e = threadIdx.x; e < ncb * W; e += NT) {
      const int cl = divn<NF>(e, W, rW), x = e - cl * W;
      const C v = res[lp<NF, LP>(cl * W + fshift_src(x, W))];
      other[lp<NF, LP>(cl * W + x)] = Sg ? mulconj(v, Sg[cl * G + x]) : v;
    }
  }
  __syncthreads();
  if (d.ngr_a > 1) {
    const auto wp = gptr<C>(d.wpart) + (size_t(g) * H + y) * W;
    for (int x = threadIdx.x; x < W; x += NT) {
      C acc = other[lp<NF, LP>(x)];
      for (int cl = 1; cl < ncb; ++cl) acc = acc + other[lp<NF, LP>(cl * W + x)];
      wp[x] = acc;
    }
    return;
  }
  if (early) {
    const int x = threadIdx.x;
    if (x < W) {
      C w;
      if (gsum) {
        w = other[x];
        for (int q = 1; q < ngs; ++q) w = w + other[q * W + x];
      } else {
        w = other[lp<NF, LP>(x)];
        for (int cl = 1; cl < ncb; ++cl) w = w + other[lp<NF, LP>(cl * W + x)];
      }
      const size_t i = size_t(y) * W + x;
      const C un = e_uo + tau * (e_div - w);
      gptr<C>(d.u)[i] = un;
      gptr<C>(d.ub)[i] = (R(1) + theta) * un - theta * e_uo;
      if (GAP) t[0] += dabs2(w - e_div);
    }
  } else {
    for (int x = threadIdx.x; x < W; x += NT) {
      C w = other[lp<NF, LP>(x)];
      for (int cl = 1; cl < ncb; ++cl) w = w + other[lp<NF, LP>(cl * W + x)];
      if (MODE == M_OP) gptr<C>(d.op_out)[size_t(y) * W + x] = w;
      else mri_primal_pixel<R, MODE, GAP>(d, x, y, w, tau, theta, t);
    }
  }
  if (MODE == M_OP) return;
  if (GAP) block_sum<2, NT>(t, part_rows + size_t(d.row_beg + y) * kGapTerms);
  if (y == 0 && threadIdx.x == 0) {
    gptr<int>(d.ctl)[CTL_ITER] += 1;
    gptr<int>(d.ctl)[CTL_RUPD] += 1;
  }
}

// Primal half of the MRI solvers for slices with several coil groups:
// w = sum_g wpart[g] (A^* r), then the primal step of each pixel.
// Item = (slice, first row), rpi rows; the block is sized to cover them in
// one pass (all loads of a thread issued together).  MODE M_OP writes w to
// op_out instead (operator adjoint).  Slices with one coil group are done in
// k_rows_adj.
template <class R, int MODE, bool GAP>
__global__ __launch_bounds__(1024) void k_mri_primal(const SliceDev* sl, const int2* items, int nitems, int rpi,
                                                     double* part) {
  const int li = xcd_item(blockIdx.x, nitems);
  const int2 it = items ? items[li] : make_int2(0, li * rpi);  // null: the dense single-slice list
  const SliceDev& d = sl[it.x];
  if (!slice_active(d) || d.ngr_a == 1) return;
  using C = Cx<R>;
  const int W = d.W, H = d.H, ngr = d.ngr_a;
  const size_t n = size_t(W) * H;
  const int ya = it.y, yb = min(H, ya + rpi);
  const auto wp = gptr<const C>(d.wpart);
  const R tau = R(d.tau), theta = R(d.theta);
  double t[2] = {0, 0};
  const int cnt = (yb - ya) * W;
  for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
    const int yy = idx / W;
    const int x = idx - yy * W, y = ya + yy;
    const size_t i = size_t(y) * W + x;
    C w = wp[i];
    for (int g = 1; g < ngr; ++g) w = w + wp[size_t(g) * n + i];
    if (MODE == M_OP) gptr<C>(d.op_out)[i] = w;
    else mri_primal_pixel<R, MODE, GAP>(d, x, y, w, tau, theta, t);
  }
  if (MODE == M_OP) return;
  if (GAP) block_sum_dyn<2>(t, part + size_t(li) * kGapTerms);
  if (ya == 0 && threadIdx.x == 0) {
    gptr<int>(d.ctl)[CTL_ITER] += 1;
    gptr<int>(d.ctl)[CTL_RUPD] += 1;
  }
}

// ---------------------------------------------------------------------------
// Gap reduction: one block per slice, fixed summation order.
// Denoising: part[0] dual, part[1] primal, part[2] restore1, part[3] restore2
// over the row items.  MRI: part[0] rows_fwd, part[2] rows_adj over the row
// items, part[1] cols over the column items.
// ---------------------------------------------------------------------------
constexpr int kGapThreads = kBlock;  // one workgroup per slice (1024 threads measured slower)

template <class R>
__global__ __launch_bounds__(kGapThreads) void k_gap_reduce(const SliceDev* sl, int kind, const double* p0,
                                                       const double* p1, const double* p2,
                                                       const double* p3) {
  const SliceDev& d = sl[blockIdx.x];
  if (!slice_active(d)) return;
  const bool mri = kind == K_TGV_MRI || kind == K_TV_OP || kind == K_TGV_OP;
  double s[kGapTerms] = {0};
  double mx = 0;
  if (mri) {
    // four items per step, their loads issued before the additions (which
    // keep the order of a plain strided loop)
    for (int i0 = d.rg_beg + threadIdx.x; i0 < d.rg_end; i0 += 4 * kGapThreads) {
      double a[4][3];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u * kGapThreads;
        const bool in = i < d.rg_end;
        const double* q = p0 + size_t(in ? i : i0) * kGapTerms;
        a[u][0] = in ? q[0] : 0.0;
        a[u][1] = in ? q[1] : 0.0;
        a[u][2] = in ? q[2] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s[0] += a[u][0];
        s[1] += a[u][1];
        s[2] += a[u][2];
      }
    }
  }
  for (int i = d.row_beg + threadIdx.x; i < d.row_end; i += kGapThreads) {
    const double* a = p0 + size_t(i) * kGapTerms;
    if (kind == K_TV) {
      s[0] += a[0];
      s[1] += a[3];
      s[2] += p1[size_t(i) * kGapTerms + 0];
      s[3] += p1[size_t(i) * kGapTerms + 1];
    } else if (kind == K_TGV) {
      s[0] += a[0];
      s[1] += a[1];
      s[2] += a[3];
      mx = fmax(mx, p2[size_t(i) * kGapTerms + 0]);
      s[4] += p3[size_t(i) * kGapTerms + 0];
      s[5] += p3[size_t(i) * kGapTerms + 1];
    } else {
      s[5] += p2[size_t(i) * kGapTerms + 0];
      s[6] += p2[size_t(i) * kGapTerms + 1];
    }
  }
  if (mri) {
    for (int i0 = d.col_beg + threadIdx.x; i0 < d.col_end; i0 += 4 * kGapThreads) {
      double a[4][2];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u * kGapThreads;
        const bool in = i < d.col_end;
        const double* q = p1 + size_t(in ? i : i0) * kGapTerms;
        a[u][0] = in ? q[0] : 0.0;
        a[u][1] = in ? q[1] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s[3] += a[u][0];
        s[4] += a[u][1];
      }
    }
  }
  __shared__ double tot[kGapTerms];
  block_sum<kGapTerms, kGapThreads>(s, tot);
  if (kind == K_TGV) mx = block_max<kGapThreads>(mx);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double lam = d.lambda;
  double P = 0, D = 0, c0 = 0, c1 = 0, c2 = 0;
  if (kind == K_TV) {
    // ROF: P(ubar) = sum|grad ubar| + lambda/2 ||ubar - f||^2,
    //      D(p) = -Re<div p, f> - ||div p||^2 / (2 lambda).
    P = tot[0] + 0.5 * lam * tot[1];
    D = -tot[2] - tot[3] / (2.0 * lam);
  } else if (kind == K_TGV) {
    P = d.alpha1 * tot[0] + d.alpha0 * tot[1] + 0.5 * lam * tot[2];
    const double a = tot[4], b = tot[5];
    const double smax = mx > 0 ? fmin(1.0, d.alpha1 / mx) : 1.0;
    const double so = b > 0 ? fmin(fmax(lam * a / b, 0.0), smax) : smax;
    D = so * a - so * so * b / (2.0 * lam);
    c0 = mx;
  } else {
    // Partial gap on the ball of radius (rho_u, rho_v) around (ubar, vbar):
    // sum of Fenchel-Young gaps plus rho * stationarity residuals.
    const int k = gptr<int>(d.ctl)[CTL_ITER] - 1;
    const bool tgv = kind != K_TV_OP;
    const double a0 = tgv ? double(sched_alpha<R>(d, d.alpha0, d.alpha0_end, k)) : 0.0;
    const double a1 = tgv ? double(sched_alpha<R>(d, d.alpha1, d.alpha1_end, k)) : 1.0;
    const double bk = pow(1.0 + d.sigma * lam, -double(gptr<int>(d.ctl)[CTL_RUPD]));
    const double preg = a1 * tot[0] + a0 * tot[1];
    P = (tot[3] + d.F0) / (2.0 * lam) + preg;
    const double fy = (tot[4] + d.F0 * bk * bk) / (2.0 * lam) + preg - tot[2];
    c1 = sqrt(tot[5]);
    c2 = sqrt(tot[6]);
    const double G = fy + d.radius_u * c1 + (tgv ? d.radius_v * c2 : 0.0);
    D = P - G;
    c0 = fy;
  }
  const double rel = (P - D) / fmax(fabs(P), 1e-300);
  d.stats[ST_P] = P;
  d.stats[ST_D] = D;
  d.stats[ST_GAP] = P - D;
  d.stats[ST_REL] = rel;
  d.stats[ST_C0] = c0;
  d.stats[ST_C1] = c1;
  d.stats[ST_C2] = c2;
  gptr<int>(d.ctl)[CTL_NGAP] += 1;
  if (d.gap_tol > 0 && rel <= d.gap_tol) {
    gptr<int>(d.ctl)[CTL_ACTIVE] = 0;
    gptr<int>(d.ctl)[CTL_DONE_AT] = gptr<int>(d.ctl)[CTL_ITER];
  }
}

// ---------------------------------------------------------------------------
// Setup kernels
// ---------------------------------------------------------------------------
// Gather f {W, Hk, C} into [c][ks][kfe], zero r; one block per (slice, c, kfe
// row) item given as (slice, c * Hk + kfe).  Writes |f|^2 over unlisted
// columns of that row to part.
template <class R>
__global__ __launch_bounds__(kBlock) void k_compact(const SliceDev* sl, const int2* items, const void* f_full,
                                                    double* part) {
  const int2 it = items[blockIdx.x];
  const SliceDev& d = sl[it.x];
  using C = Cx<R>;
  const int W = d.W, Hk = d.Hk, nc = d.ncols;
  const int c = it.y / Hk, kfe = it.y - c * Hk;
  const auto f = gptr<const C>(f_full) + (size_t(c) * Hk + kfe) * W;
  const auto fk = gptr<C>(d.fk);
  const auto rk = gptr<C>(d.rk);
  double t[1] = {0};
  for (int kx = threadIdx.x; kx < W; kx += kBlock) {
    const int ks = gptr<const int>(d.colmap)[kx];
    const C v = f[kx];
    if (ks >= 0) {
      const size_t idx = (size_t(c) * nc + ks) * Hk + kfe;
      fk[idx] = v;
      rk[idx] = pzero<C>();
    } else {
      t[0] += dabs2(v);
    }
  }
  block_sum<1>(t, part + size_t(blockIdx.x) * kGapTerms);
}

// Writes a slice descriptor passed by value to device memory and opens its
// control words (active, iteration 0): the set-up of an operator application
// that copies nothing from the host.
__global__ __launch_bounds__(64) void k_store_desc(SliceDev d, SliceDev* dst, int* ctl) {
  if (threadIdx.x == 0) *dst = d;
  if (threadIdx.x < CTL_WORDS) ctl[threadIdx.x] = threadIdx.x == CTL_ACTIVE ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_zero(uint32_t* p, size_t n) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
#define PDT_LAUNCH(kern, grid, lds, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, st, __VA_ARGS__)

// LDS of a CU on gfx950, its allocation granule, and an upper bound on the
// static LDS of any kernel here (block reductions).
constexpr size_t kLdsPerCu = 160 * 1024, kLdsGranule = 512, kStaticLdsMax = 2048;

size_t lds_occupancy(size_t need, int wpc, size_t limit) {
  if (wpc <= 0) return need;
  // wpc workgroups of `share` bytes fit, wpc + 1 do not (wpc <= 6, so the
  // static part's slack cannot let a further one in)
  const size_t share = (kLdsPerCu / size_t(wpc)) / kLdsGranule * kLdsGranule;
  if (share <= kStaticLdsMax) return need;
  const size_t want = share - kStaticLdsMax;
  return want > limit ? need : (want > need ? want : need);
}

size_t lds_rows_bytes(int dtype, int cb, int W, int nc, bool padded) {
  const size_t e = (dtype == DT_C128 || dtype == DT_F64) ? 16 : 8;
  return (size_t(W) + 2 * size_t(padded ? lds_pad(cb * W) : cb * W)) * e + size_t(nc) * sizeof(int);
}
size_t lds_cols_bytes(int dtype, int cg, int Hk) {
  const size_t e = (dtype == DT_C128 || dtype == DT_F64) ? 16 : 8;
  return (size_t(Hk) + 2 * size_t(lds_pad(cg * Hk))) * e;
}

int num_stages(int kind) { return (kind == K_TV || kind == K_TGV) ? 2 : 4; }

const char* stage_name(int kind, int stage) {
  switch (kind) {
    case K_TV: return stage == 0 ? "k_tv_dual" : "k_tv_primal";
    case K_TGV: return stage == 0 ? "k_tgv_primal" : "k_tgv_dual";
    default: return stage == 0 ? "k_rows_fwd" : stage == 1 ? "k_cols" : stage == 2 ? "k_rows_adj" : "k_mri_primal";
  }
}

// Denoising stages: TV = {dual, primal}; TGV = {primal, dual (+ restoration
// on gap iterations)}.
template <class P, class S>
static void denoise_stage(const Launch& L, int kind, bool gap, int stage) {
  const int n = L.nrows;
  if (n == 0) return;
  const int2* rows = L.dense ? nullptr : L.rows;
  // occupancy cap: LDS only reserved (the kernels use none dynamically),
  // within the 64 KiB a launch may ask for without a raised attribute
  const size_t lds = lds_occupancy(0, L.wpc[WPC_DEN], 64 * 1024);
  // the half-iteration kernels with L.den_threads threads per workgroup
  auto launch = [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    if (kind == K_TV) {
      if (stage == 0) {
        if (gap) hipLaunchKernelGGL((k_tv_dual<P, true, NT>), dim3(n), dim3(NT), lds, L.stream, L.sl, rows, L.rpi, L.part[0]);
        else hipLaunchKernelGGL((k_tv_dual<P, false, NT>), dim3(n), dim3(NT), lds, L.stream, L.sl, rows, L.rpi, L.part[0]);
      } else {
        if (gap) hipLaunchKernelGGL((k_tv_primal<P, true, NT>), dim3(n), dim3(NT), lds, L.stream, L.sl, rows, L.rpi, L.part[1]);
        else hipLaunchKernelGGL((k_tv_primal<P, false, NT>), dim3(n), dim3(NT), lds, L.stream, L.sl, rows, L.rpi, L.part[1]);
      }
    } else if (stage == 0) {
      hipLaunchKernelGGL((k_tgv_primal<P, S, NT>), dim3(n), dim3(NT), lds, L.stream, L.sl, rows, L.rpi);
    } else if (L.den_pair) {
      // two pixels per thread: half the threads of the one-pixel launch (at least 256)
      constexpr int NP = NT > kBlock ? NT / 2 : kBlock;
      if (gap) hipLaunchKernelGGL((k_tgv_dual2<P, S, true, NP>), dim3(n), dim3(NP), lds, L.stream, L.sl, rows, L.rpi, L.part[0]);
      else hipLaunchKernelGGL((k_tgv_dual2<P, S, false, NP>), dim3(n), dim3(NP), lds, L.stream, L.sl, rows, L.rpi, L.part[0]);
    } else if (gap) {
      hipLaunchKernelGGL((k_tgv_dual<P, S, true, NT>), dim3(n), dim3(NT), lds, L.stream, L.sl, rows, L.rpi, L.part[0]);
    } else {
      hipLaunchKernelGGL((k_tgv_dual<P, S, false, NT>), dim3(n), dim3(NT), lds, L.stream, L.sl, rows, L.rpi, L.part[0]);
    }
  };
  if (L.den_threads >= 1024) launch(std::integral_constant<int, 1024>{});
  else if (L.den_threads >= 512) launch(std::integral_constant<int, 512>{});
  else launch(std::integral_constant<int, kBlock>{});
  if (kind == K_TGV && stage == 1 && gap) {
    PDT_LAUNCH((k_tgv_restore1<P, S>), n, 0, L.stream, L.sl, rows, L.rpi, L.part[2]);
    PDT_LAUNCH((k_tgv_restore2<P>), n, 0, L.stream, L.sl, rows, L.rpi, L.part[3]);
  }
}

#define PDT_LAUNCH_NT(kern, grid, nt, lds, st, ...) \
  hipLaunchKernelGGL(kern, dim3(grid), dim3(nt), lds, st, __VA_ARGS__)

// FFT lengths with kernels built on a compile-time plan (single precision);
// any other length, and every double-precision launch, uses the runtime plan.
constexpr int kFftLengths[] = {256, 320, 368, 384, 512, 640, 736, 768};

// Calls fn(std::integral_constant<int, NF>) with NF = n when n has a
// compile-time plan for R, else NF = 0.
template <class R, class Fn>
static void with_fft_len(int n, Fn&& fn) {
  if constexpr (std::is_same<R, float>::value) {
    switch (n) {
      case 256: fn(std::integral_constant<int, 256>{}); return;
      case 320: fn(std::integral_constant<int, 320>{}); return;
      case 368: fn(std::integral_constant<int, 368>{}); return;
      case 384: fn(std::integral_constant<int, 384>{}); return;
      case 512: fn(std::integral_constant<int, 512>{}); return;
      case 640: fn(std::integral_constant<int, 640>{}); return;
      case 736: fn(std::integral_constant<int, 736>{}); return;
      case 768: fn(std::integral_constant<int, 768>{}); return;
      default: break;
    }
  }
  fn(std::integral_constant<int, 0>{});
}

size_t block_lds(int kind, int dtype, int tx, int ty, int k) {
  const size_t e = dtype == DT_F32 ? 4 : (dtype == DT_C128 ? 16 : 8);
  const size_t planes = kind == K_TGV ? 8 : 3;  // TGV: ubar, vbar, p, q; ROF: ubar, p
  return planes * size_t(tx + 2 * k) * size_t(ty + 2 * k) * e;
}

// Workgroup size of a blocked launch: a few owned pixels per thread of the
// loaded tile (ROF: up to 4 below 4096 pixels, TGV: one to tgv_block_own),
// 256 to 1024 threads.
static int block_threads(int kind, int dtype, int tx, int ty, int k) {
  const int px = (tx + 2 * k) * (ty + 2 * k);
  if (kind == K_TGV) return px >= 1024 && dtype != DT_C128 ? 1024 : px >= 384 ? 512 : 256;
  return px > 2048 ? 1024 : px > 1024 ? 512 : 256;
}

bool block_fits(int kind, int dtype, int tx, int ty, int k, size_t lds_limit) {
  if (tx <= 0 || ty <= 0 || k <= 0) return false;
  const int px = (tx + 2 * k) * (ty + 2 * k);
  // every tile pixel owned by a thread (TGV: tgv_block_own per thread at most)
  const size_t e = dtype == DT_F32 ? 4 : (dtype == DT_C128 ? 16 : 8);
  if (kind == K_TGV && px > tgv_block_own(e) * block_threads(kind, dtype, tx, ty, k)) return false;
  // ROF: likewise, tv_block_own pixels per thread at most
  if (kind != K_TGV && px > tv_block_own(e) * block_threads(kind, dtype, tx, ty, k)) return false;
  return block_lds(kind, dtype, tx, ty, k) <= lds_limit;
}

void run_block(const Launch& L, int kind, int dtype, bool fp16, int dir) {
  const int tx = L.tb_x, ty = L.tb_y, k = L.tb_k;
  const size_t lds = block_lds(kind, dtype, tx, ty, k);
  const int grid = ((L.tb_w + tx - 1) / tx) * ((L.tb_h + ty - 1) / ty);
  auto launch = [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
#define PDT_BLOCK(P, S)                                                                                             \
  if (kind == K_TGV) hipLaunchKernelGGL((k_tgv_block<P, S, NT>), dim3(grid), dim3(NT), lds, L.stream, L.sl, tx, ty, k, dir); \
  else hipLaunchKernelGGL((k_tv_block<P, NT>), dim3(grid), dim3(NT), lds, L.stream, L.sl, tx, ty, k, dir)
    switch (dtype) {
      case DT_F32:
        if (fp16 && kind == K_TGV) hipLaunchKernelGGL((k_tgv_block<float, _Float16, NT>), dim3(grid), dim3(NT), lds, L.stream, L.sl, tx, ty, k, dir);
        else PDT_BLOCK(float, float);
        break;
      case DT_C64:
        if (fp16 && kind == K_TGV) hipLaunchKernelGGL((k_tgv_block<Cx<float>, Cx<_Float16>, NT>), dim3(grid), dim3(NT), lds, L.stream, L.sl, tx, ty, k, dir);
        else PDT_BLOCK(Cx<float>, Cx<float>);
        break;
      case DT_F64: PDT_BLOCK(double, double); break;
      case DT_C128: PDT_BLOCK(Cx<double>, Cx<double>); break;
    }
#undef PDT_BLOCK
  };
  const int nt = block_threads(kind, dtype, tx, ty, k);
  if (nt == 1024) launch(std::integral_constant<int, 1024>{});
  else if (nt == 512) launch(std::integral_constant<int, 512>{});
  else launch(std::integral_constant<int, 256>{});
}

bool fft_len_compiled(int dtype, int n) {
  if (dtype != DT_C64 && dtype != DT_F32) return false;
  for (int m : kFftLengths)
    if (m == n) return true;
  return false;
}

bool rows_adj_fused_ok(int dtype, int W, int C, int nc) {
  constexpr int NW = kRowAdjThreadsWide;
  if (!fft_len_compiled(dtype, W) || 2 * W > NW || C < 1) return false;
  const int ngs = std::min(NW / W, C), ch = (C + ngs - 1) / ngs;
  // (padded lengths: a wave holds the listed columns of its coils, 64 per slot)
  const bool cols_fit = ct_pad_ok(W) ? nc <= 64 * rows_adj_pt<NW>() : C * nc <= rows_adj_pt<NW>() * NW;
  return ch <= kRowsAdjFusedSlots && cols_fit;
}

// Adjoint row pass with L.adj_threads threads per block (kRowThreads, or
// kRowAdjThreadsWide for items holding many coils).
template <class R, int MODE, int NF>
static void launch_rows_adj(const Launch& L, bool gap) {
  constexpr int NR = kRowThreads, NW = kRowAdjThreadsWide;
  const size_t lds = lds_occupancy(L.lds_rows_a, L.wpc[WPC_ADJ], init_kernel_attrs());
  if (L.adj_threads == NW) {
    const int nv = L.dense ? L.nvstep : 0;
    if constexpr (NF > 0 && MODE != M_OP) {
      if (L.adj_fused) {
        if (gap) PDT_LAUNCH_NT((k_rows_adj<R, MODE, true, NW, NF, true>), L.nrg_a + nv, NW, lds, L.stream, L.sl, L.dense ? nullptr : L.rowgroups_a, L.nrg_a, L.cb_a, L.gap_rows, nv);
        else PDT_LAUNCH_NT((k_rows_adj<R, MODE, false, NW, NF, true>), L.nrg_a + nv, NW, lds, L.stream, L.sl, L.dense ? nullptr : L.rowgroups_a, L.nrg_a, L.cb_a, L.gap_rows, nv);
        return;
      }
    }
    if (gap) PDT_LAUNCH_NT((k_rows_adj<R, MODE, true, NW, NF>), L.nrg_a + nv, NW, lds, L.stream, L.sl, L.dense ? nullptr : L.rowgroups_a, L.nrg_a, L.cb_a, L.gap_rows, nv);
    else PDT_LAUNCH_NT((k_rows_adj<R, MODE, false, NW, NF>), L.nrg_a + nv, NW, lds, L.stream, L.sl, L.dense ? nullptr : L.rowgroups_a, L.nrg_a, L.cb_a, L.gap_rows, nv);
  } else {
    if (gap) PDT_LAUNCH_NT((k_rows_adj<R, MODE, true, NR, NF>), L.nrg_a, NR, lds, L.stream, L.sl, L.dense ? nullptr : L.rowgroups_a, L.nrg_a, L.cb_a, L.gap_rows, 0);
    else PDT_LAUNCH_NT((k_rows_adj<R, MODE, false, NR, NF>), L.nrg_a, NR, lds, L.stream, L.sl, L.dense ? nullptr : L.rowgroups_a, L.nrg_a, L.cb_a, L.gap_rows, 0);
  }
}

// MRI stages: 0 rows_fwd, 1 cols, 2 rows_adj, 3 primal.  The FFT kernels are
// picked by L.fft_w (row length) and L.fft_h (column length).
template <class R, int MODE>
static void mri_stage_m(const Launch& L, bool gap, int stage) {
  constexpr int NR = kRowThreads, NC = kColThreads;
  if (stage == 0) {
    const size_t lds = lds_occupancy(L.lds_rows, L.wpc[WPC_FWD], init_kernel_attrs());
    with_fft_len<R>(L.fft_w, [&](auto nf) {
      constexpr int NF = decltype(nf)::value;
      const int nd = L.dense ? L.ndual : 0;
      if (gap) PDT_LAUNCH_NT((k_rows_fwd<R, MODE, true, NR, NF>), L.nrg + nd, NR, lds, L.stream, L.sl, L.dense ? nullptr : L.rowgroups, L.nrg, L.cb, L.part[0], nd);
      else PDT_LAUNCH_NT((k_rows_fwd<R, MODE, false, NR, NF>), L.nrg + nd, NR, lds, L.stream, L.sl, L.dense ? nullptr : L.rowgroups, L.nrg, L.cb, L.part[0], nd);
    });
  } else if (stage == 1) {
    if (L.ncolitems == 0) return;
    const size_t lds = lds_occupancy(L.lds_cols, L.wpc[WPC_COLS], init_kernel_attrs());
    with_fft_len<R>(L.fft_h, [&](auto nf) {
      constexpr int NF = decltype(nf)::value;
      const int2* items = L.dense ? nullptr : L.cols;
      if constexpr (NF > 0) {
        if (L.cg <= kColPrefetchCols) {
          if (gap) PDT_LAUNCH_NT((k_cols<R, CM_SOLVE, true, NC, NF, true>), L.ncolitems, NC, lds, L.stream, L.sl, items, L.ncolitems, L.cg, L.part[1], 0, nullptr);
          else PDT_LAUNCH_NT((k_cols<R, CM_SOLVE, false, NC, NF, true>), L.ncolitems, NC, lds, L.stream, L.sl, items, L.ncolitems, L.cg, L.part[1], 0, nullptr);
          return;
        }
      }
      if (gap) PDT_LAUNCH_NT((k_cols<R, CM_SOLVE, true, NC, NF>), L.ncolitems, NC, lds, L.stream, L.sl, items, L.ncolitems, L.cg, L.part[1], 0, nullptr);
      else PDT_LAUNCH_NT((k_cols<R, CM_SOLVE, false, NC, NF>), L.ncolitems, NC, lds, L.stream, L.sl, items, L.ncolitems, L.cg, L.part[1], 0, nullptr);
    });
  } else if (stage == 2) {
    with_fft_len<R>(L.fft_w, [&](auto nf) {
      constexpr int NF = decltype(nf)::value;
      launch_rows_adj<R, MODE, NF>(L, gap);
    });
  } else {
    if (!L.split_primal) return;
    if (gap) PDT_LAUNCH_NT((k_mri_primal<R, MODE, true>), L.nrows, L.primal_threads, 0, L.stream, L.sl, L.rows, L.nrows, L.rpi, L.part[2]);
    else PDT_LAUNCH_NT((k_mri_primal<R, MODE, false>), L.nrows, L.primal_threads, 0, L.stream, L.sl, L.rows, L.nrows, L.rpi, L.part[2]);
  }
}

template <class R>
static void mri_stage(const Launch& L, int kind, bool gap, int stage) {
  if (kind == K_TV_OP) mri_stage_m<R, M_TV>(L, gap, stage);
  else mri_stage_m<R, M_TGV>(L, gap, stage);
}

// The two launches of a paired iteration (L.pair): 0 the column pass with
// the dual step's workgroups, 1 the fused adjoint row pass that also writes
// the next T1.  With the gap evaluation the partial sums go where the three
// passes put them: the dual step's after the forward row items' (which a
// paired iteration does not run; the engine zeroes theirs once), the column
// items' and the row items' as in those passes.
template <class R, int MODE, bool GAP>
static void mri_pair_m(const Launch& L, int which) {
  constexpr int NC = kColThreads, NW = kRowAdjThreadsWide;
  if (which == 0) {
    const size_t lds = lds_occupancy(L.lds_cols, L.wpc[WPC_COLS], init_kernel_attrs());
    double* const pd = L.part[0] + size_t(L.nrg) * kGapTerms;
    with_fft_len<R>(L.fft_h, [&](auto nf) {
      constexpr int NF = decltype(nf)::value;
      if constexpr (NF > 0) {
        if (L.cg <= kColPrefetchCols) {
          PDT_LAUNCH_NT((k_cols<R, CM_SOLVE, GAP, NC, NF, true, MODE>), L.ncolitems + L.ndual, NC, lds, L.stream, L.sl, nullptr, L.ncolitems, L.cg, L.part[1], L.ndual, pd);
          return;
        }
      }
      PDT_LAUNCH_NT((k_cols<R, CM_SOLVE, GAP, NC, NF, false, MODE>), L.ncolitems + L.ndual, NC, lds, L.stream, L.sl, nullptr, L.ncolitems, L.cg, L.part[1], L.ndual, pd);
    });
  } else {
    const size_t lds = lds_occupancy(L.lds_rows_a, L.wpc[WPC_ADJ], init_kernel_attrs());
    with_fft_len<R>(L.fft_w, [&](auto nf) {
      constexpr int NF = decltype(nf)::value;
      if constexpr (NF > 0)
        PDT_LAUNCH_NT((k_rows_adj<R, MODE, GAP, NW, NF, true, true>), L.nrg_a + L.nvstep, NW, lds, L.stream, L.sl, nullptr, L.nrg_a, L.cb_a, L.gap_rows, L.nvstep);
    });
  }
}

void run_pair_stage(const Launch& L, int kind, int dtype, int which, bool gap) {
  if (dtype != DT_C64) return;  // pairs need the compile-time row plans (single precision)
  if (kind == K_TV_OP) {
    if (gap) mri_pair_m<float, M_TV, true>(L, which);
    else mri_pair_m<float, M_TV, false>(L, which);
  } else {
    if (gap) mri_pair_m<float, M_TGV, true>(L, which);
    else mri_pair_m<float, M_TGV, false>(L, which);
  }
}

void run_stage(const Launch& L, int kind, int dtype, bool fp16, bool gap, int stage) {
  if (kind == K_TV || kind == K_TGV) {
    switch (dtype) {
      case DT_F32:
        if (fp16) denoise_stage<float, _Float16>(L, kind, gap, stage);
        else denoise_stage<float, float>(L, kind, gap, stage);
        break;
      case DT_C64:
        if (fp16) denoise_stage<Cx<float>, Cx<_Float16>>(L, kind, gap, stage);
        else denoise_stage<Cx<float>, Cx<float>>(L, kind, gap, stage);
        break;
      case DT_F64: denoise_stage<double, double>(L, kind, gap, stage); break;
      case DT_C128: denoise_stage<Cx<double>, Cx<double>>(L, kind, gap, stage); break;
    }
  } else {
    if (dtype == DT_C64) mri_stage<float>(L, kind, gap, stage);
    else mri_stage<double>(L, kind, gap, stage);
  }
}

void run_iteration(const Launch& L, int kind, int dtype, bool fp16, bool gap) {
  for (int s = 0; s < num_stages(kind); ++s) run_stage(L, kind, dtype, fp16, gap, s);
}

void run_gap_reduce(const Launch& L, int kind, int dtype) {
  if (dtype == DT_F32 || dtype == DT_C64)
    hipLaunchKernelGGL((k_gap_reduce<float>), dim3(L.nsl), dim3(kGapThreads), 0, L.stream, L.sl, kind, L.part[0],
                       L.part[1], L.part[2], L.part[3]);
  else
    hipLaunchKernelGGL((k_gap_reduce<double>), dim3(L.nsl), dim3(kGapThreads), 0, L.stream, L.sl, kind, L.part[0],
                       L.part[1], L.part[2], L.part[3]);
}

template <class R>
static void op_forward(const Launch& L) {
  constexpr int NR = kRowThreads, NC = kColThreads;
  with_fft_len<R>(L.fft_w, [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    PDT_LAUNCH_NT((k_rows_fwd<R, M_OP, false, NR, NF>), L.nrg, NR, L.lds_rows, L.stream, L.sl, L.rowgroups, L.nrg, L.cb, L.part[0], 0);
  });
  with_fft_len<R>(L.fft_h, [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    PDT_LAUNCH_NT((k_cols<R, CM_FWD, false, NC, NF>), L.ncolitems, NC, L.lds_cols, L.stream, L.sl, L.cols, L.ncolitems, L.cg, L.part[1], 0, nullptr);
  });
}

template <class R>
static void op_adjoint(const Launch& L) {
  constexpr int NC = kColThreads;
  with_fft_len<R>(L.fft_h, [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    PDT_LAUNCH_NT((k_cols<R, CM_ADJ, false, NC, NF>), L.ncolitems, NC, L.lds_cols, L.stream, L.sl, L.cols, L.ncolitems, L.cg, L.part[1], 0, nullptr);
  });
  with_fft_len<R>(L.fft_w, [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    launch_rows_adj<R, M_OP, NF>(L, false);
  });
  if (L.split_primal)
    PDT_LAUNCH_NT((k_mri_primal<R, M_OP, false>), L.nrows, L.primal_threads, 0, L.stream, L.sl, L.rows, L.nrows, L.rpi, L.part[2]);
}

// Raises the dynamic LDS limit of every FFT kernel instantiation to b bytes.
template <class R>
static bool set_lds_attrs(int b) {
  constexpr int NR = kRowThreads, NC = kColThreads;
  bool ok = true;
  auto set = [&](const void* f) {
    ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, b) == hipSuccess;
  };
  auto one = [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    set((const void*)&k_rows_fwd<R, M_OP, false, NR, NF>);
    set((const void*)&k_rows_fwd<R, M_TV, false, NR, NF>);
    set((const void*)&k_rows_fwd<R, M_TV, true, NR, NF>);
    set((const void*)&k_rows_fwd<R, M_TGV, false, NR, NF>);
    set((const void*)&k_rows_fwd<R, M_TGV, true, NR, NF>);
    set((const void*)&k_rows_adj<R, M_OP, false, NR, NF>);
    set((const void*)&k_rows_adj<R, M_TV, false, NR, NF>);
    set((const void*)&k_rows_adj<R, M_TV, true, NR, NF>);
    set((const void*)&k_rows_adj<R, M_TGV, false, NR, NF>);
    set((const void*)&k_rows_adj<R, M_TGV, true, NR, NF>);
    set((const void*)&k_rows_adj<R, M_OP, false, kRowAdjThreadsWide, NF>);
    set((const void*)&k_rows_adj<R, M_TV, false, kRowAdjThreadsWide, NF>);
    set((const void*)&k_rows_adj<R, M_TV, true, kRowAdjThreadsWide, NF>);
    set((const void*)&k_rows_adj<R, M_TGV, false, kRowAdjThreadsWide, NF>);
    set((const void*)&k_rows_adj<R, M_TGV, true, kRowAdjThreadsWide, NF>);
    if constexpr (NF > 0) {
      set((const void*)&k_rows_adj<R, M_TV, false, kRowAdjThreadsWide, NF, true>);
      set((const void*)&k_rows_adj<R, M_TV, true, kRowAdjThreadsWide, NF, true>);
      set((const void*)&k_rows_adj<R, M_TGV, false, kRowAdjThreadsWide, NF, true>);
      set((const void*)&k_rows_adj<R, M_TGV, true, kRowAdjThreadsWide, NF, true>);
      if constexpr (std::is_same<R, float>::value) {
        set((const void*)&k_rows_adj<R, M_TV, false, kRowAdjThreadsWide, NF, true, true>);
        set((const void*)&k_rows_adj<R, M_TGV, false, kRowAdjThreadsWide, NF, true, true>);
        set((const void*)&k_rows_adj<R, M_TV, true, kRowAdjThreadsWide, NF, true, true>);
        set((const void*)&k_rows_adj<R, M_TGV, true, kRowAdjThreadsWide, NF, true, true>);
      }
    }
    set((const void*)&k_cols<R, CM_SOLVE, false, NC, NF>);
    set((const void*)&k_cols<R, CM_SOLVE, true, NC, NF>);
    if constexpr (std::is_same<R, float>::value) {
      set((const void*)&k_cols<R, CM_SOLVE, false, NC, NF, false, M_TV>);
      set((const void*)&k_cols<R, CM_SOLVE, false, NC, NF, false, M_TGV>);
      set((const void*)&k_cols<R, CM_SOLVE, true, NC, NF, false, M_TV>);
      set((const void*)&k_cols<R, CM_SOLVE, true, NC, NF, false, M_TGV>);
    }
    if constexpr (NF > 0) {
      set((const void*)&k_cols<R, CM_SOLVE, false, NC, NF, true>);
      set((const void*)&k_cols<R, CM_SOLVE, true, NC, NF, true>);
      if constexpr (std::is_same<R, float>::value) {
        set((const void*)&k_cols<R, CM_SOLVE, false, NC, NF, true, M_TV>);
        set((const void*)&k_cols<R, CM_SOLVE, false, NC, NF, true, M_TGV>);
        set((const void*)&k_cols<R, CM_SOLVE, true, NC, NF, true, M_TV>);
        set((const void*)&k_cols<R, CM_SOLVE, true, NC, NF, true, M_TGV>);
      }
    }
    set((const void*)&k_cols<R, CM_FWD, false, NC, NF>);
    set((const void*)&k_cols<R, CM_ADJ, false, NC, NF>);
  };
  set((const void*)&k_tv_block<R, 256>);
  set((const void*)&k_tv_block<Cx<R>, 256>);
  set((const void*)&k_tv_block<R, 512>);
  set((const void*)&k_tv_block<Cx<R>, 512>);
  set((const void*)&k_tv_block<R, 1024>);
  set((const void*)&k_tv_block<Cx<R>, 1024>);
  set((const void*)&k_tgv_block<R, R, 256>);
  set((const void*)&k_tgv_block<Cx<R>, Cx<R>, 256>);
  set((const void*)&k_tgv_block<R, R, 512>);
  set((const void*)&k_tgv_block<Cx<R>, Cx<R>, 512>);
  set((const void*)&k_tgv_block<R, R, 1024>);
  set((const void*)&k_tgv_block<Cx<R>, Cx<R>, 1024>);
  if constexpr (std::is_same<R, float>::value) {
    set((const void*)&k_tgv_block<float, _Float16, 256>);
    set((const void*)&k_tgv_block<Cx<float>, Cx<_Float16>, 256>);
    set((const void*)&k_tgv_block<float, _Float16, 512>);
    set((const void*)&k_tgv_block<Cx<float>, Cx<_Float16>, 512>);
    set((const void*)&k_tgv_block<float, _Float16, 1024>);
    set((const void*)&k_tgv_block<Cx<float>, Cx<_Float16>, 1024>);
  }
  with_fft_len<R>(0, one);
  for (int n : kFftLengths)
    if (std::is_same<R, float>::value) with_fft_len<R>(n, one);
  return ok;
}

size_t init_kernel_attrs() {
  static size_t budget = 0;
  if (budget) return budget;
  // the block reductions use up to 2 KiB of static LDS
  const size_t want = kLdsMax - 4096;
  if (set_lds_attrs<float>(int(want)) && set_lds_attrs<double>(int(want))) {
    budget = want;
  } else {
    budget = 64 * 1024 - 4096;
  }
  (void)hipGetLastError();  // a refused attribute must not surface at the next launch check
  return budget;
}

void run_op_forward(const Launch& L, int dtype) {
  if (dtype == DT_C64) op_forward<float>(L);
  else op_forward<double>(L);
}

void run_op_adjoint(const Launch& L, int dtype) {
  if (dtype == DT_C64) op_adjoint<float>(L);
  else op_adjoint<double>(L);
}

int run_compact_data(const Launch& L, int dtype, const void* f_full, double* f0_part) {
  // L.rows here holds (slice, c * Hk + kfe) items
  if (dtype == DT_C64)
    PDT_LAUNCH((k_compact<float>), L.nrows, 0, L.stream, L.sl, L.rows, f_full, f0_part);
  else
    PDT_LAUNCH((k_compact<double>), L.nrows, 0, L.stream, L.sl, L.rows, f_full, f0_part);
  return L.nrows;
}

// Single-block reductions for set-up quantities, in a fixed order (each
// thread sums a strided subsequence, then a tree over the block): out[0] =
// sum_i a[i * stride] for i < n, and out[0] = sum_i |z_i|^2 over n pixels.
__device__ __forceinline__ void tree_sum_1024(double v, double* out) {
  __shared__ double red[1024];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}
// (Eight strided elements per step, loaded before they are added in order.)
__global__ __launch_bounds__(1024) void k_sum_strided(const double* a, int n, int stride, double* out) {
  double v = 0;
  for (int i0 = threadIdx.x; i0 < n; i0 += 8 * 1024) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * 1024;
      x[u] = i < n ? a[size_t(i) * stride] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) v += x[u];
  }
  tree_sum_1024(v, out);
}
template <class R>
__global__ __launch_bounds__(1024) void k_sumsq(const Cx<R>* z, int64_t n, double* out) {
  double v = 0;
  for (int64_t i0 = threadIdx.x; i0 < n; i0 += 8 * 1024) {
    Cx<R> x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t i = i0 + u * 1024;
      x[u] = i < n ? z[i] : pzero<Cx<R>>();
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) v += dabs2(x[u]);
  }
  tree_sum_1024(v, out);
}

void run_sum_strided(hipStream_t st, const double* a, int n, int stride, double* out) {
  PDT_LAUNCH_NT(k_sum_strided, 1, 1024, 0, st, a, n, stride, out);
}
void run_sumsq(hipStream_t st, int dtype, const void* z, int64_t n, double* out) {
  if (dtype == DT_C128) PDT_LAUNCH_NT((k_sumsq<double>), 1, 1024, 0, st, static_cast<const Cx<double>*>(z), n, out);
  else PDT_LAUNCH_NT((k_sumsq<float>), 1, 1024, 0, st, static_cast<const Cx<float>*>(z), n, out);
}

void run_store_desc(hipStream_t st, const SliceDev& d, SliceDev* dst, int* ctl) {
  hipLaunchKernelGGL(k_store_desc, dim3(1), dim3(64), 0, st, d, dst, ctl);
}

// Scratch of the allocation-free operator application: one array per GPU,
// part of the code object (no allocation), held until the process ends.
// Host code never passes the array's name to a kernel or a copy (on the host
// it is a stand-in, not a device address); it uses the per-device address
// that hipGetSymbolAddress returns, taken once per device with that device
// current.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

void* noalloc_scratch(int device) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) return nullptr;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    addr[device] = p;
  }
  return addr[device];
}

void run_zero(hipStream_t st, void* ptr, size_t bytes) {
  const size_t n = bytes / 4;
  if (n == 0) return;
  size_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  PDT_LAUNCH(k_zero, blocks, 0, st, static_cast<uint32_t*>(ptr), n);
}

}  // namespace pdt
