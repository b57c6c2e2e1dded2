// pdt_oracle.hpp -- CPU restatement of the reference TV/TGV primal-dual solvers.
//
// TEST INFRASTRUCTURE ONLY.  This header is the oracle: the parity tests, the
// smoke check and the CPU-baseline leg of the benchmark call it as the checker.
// The product (libpdt_hip.so) never links or calls it.
//
// The code restates, per pixel and in the same order, what the reference
// computes.  Every function names the reference file:line it follows
// (paths relative to the primal-dual-toolbox source tree).  It is templated on
// the real type R (double = the FP64 oracle, float = the FP32 OpenMP CPU
// baseline) and on the pixel type P (R for real images, std::complex<R> for
// complex ones).
//
// Array layout follows the reference's iu arrays: size {W, H[, C]}, x (dim 0)
// fastest, so element (x, y, c) lives at x + W * (y + H * c).  A numpy array of
// shape (C, H, W) in C order has the same memory image.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace pdto {

template <class R> using cx = std::complex<R>;

template <class P> struct pix;
template <> struct pix<float> { using R = float; static constexpr bool cplx = false; };
template <> struct pix<double> { using R = double; static constexpr bool cplx = false; };
template <> struct pix<cx<float>> { using R = float; static constexpr bool cplx = true; };
template <> struct pix<cx<double>> { using R = double; static constexpr bool cplx = true; };

// |a|^2 with the complex modulus coupling real and imaginary parts
// (iu::type_trait<P>::abs squared, tgv_kernels.cuh:137-154).
template <class R> inline R abs2(R a) { return a * a; }
template <class R> inline R abs2(cx<R> a) { return a.real() * a.real() + a.imag() * a.imag(); }
// Re<a, b> = Re(conj(a) * b), accumulated in double.
template <class R> inline double redot(R a, R b) { return double(a) * double(b); }
template <class R> inline double redot(cx<R> a, cx<R> b) {
  return double(a.real()) * double(b.real()) + double(a.imag()) * double(b.imag());
}
template <class R> inline double abs2d(R a) { return double(a) * double(a); }
template <class R> inline double abs2d(cx<R> a) {
  return double(a.real()) * double(a.real()) + double(a.imag()) * double(a.imag());
}

struct Grid {
  int W, H;
  size_t n() const { return size_t(W) * size_t(H); }
  size_t at(int x, int y) const { return size_t(x) + size_t(W) * size_t(y); }
};

// ---------------------------------------------------------------------------
// Finite differences (tgv_kernels.cuh)
// ---------------------------------------------------------------------------

// dp: forward differences, zero on the last column/row (tgv_kernels.cuh:40-50).
template <class P>
inline void dp(const P* u, Grid g, int x, int y, P& gx, P& gy) {
  const int xp = x + (x < g.W - 1);
  const int yp = y + (y < g.H - 1);
  gx = u[g.at(xp, y)] - u[g.at(x, y)];
  gy = u[g.at(x, yp)] - u[g.at(x, y)];
}

// One backward-difference term of dm (tgv_kernels.cuh:72-81): the first
// index keeps p(0), the last one is -p(n-2), interior p(i) - p(i-1).
template <class P>
inline P dmx(const P* a, Grid g, int x, int y) {
  return (x > 0) ? ((x < g.W - 1) ? a[g.at(x, y)] - a[g.at(x - 1, y)] : -a[g.at(x - 1, y)])
                 : a[g.at(x, y)];
}
template <class P>
inline P dmy(const P* a, Grid g, int x, int y) {
  return (y > 0) ? ((y < g.H - 1) ? a[g.at(x, y)] - a[g.at(x, y - 1)] : -a[g.at(x, y - 1)])
                 : a[g.at(x, y)];
}

// dm: divergence (tgv_kernels.cuh:68-82).
template <class P>
inline P dm(const P* px, const P* py, Grid g, int x, int y) {
  P d = dmx(px, g, x, y);
  d += dmy(py, g, x, y);
  return d;
}

// dp_sym: symmetrised gradient (tgv_kernels.cuh:25-38).
template <class P>
inline void dp_sym(const P* vx, const P* vy, Grid g, int x, int y, P& xx, P& yy, P& xy) {
  const int xp = x + (x < g.W - 1);
  const int yp = y + (y < g.H - 1);
  const size_t i = g.at(x, y);
  xx = vx[g.at(xp, y)] - vx[i];
  xy = typename pix<P>::R(0.5) * (vx[g.at(x, yp)] - vx[i] + vy[g.at(xp, y)] - vy[i]);
  yy = vy[g.at(x, yp)] - vy[i];
}

// dm_sym: symmetrised divergence (tgv_kernels.cuh:105-135).
template <class P>
inline void dm_sym(const P* qxx, const P* qyy, const P* qxy, Grid g, int x, int y, P& dx, P& dy) {
  dx = dmx(qxx, g, x, y) + dmy(qxy, g, x, y);
  dy = dmy(qyy, g, x, y) + dmx(qxy, g, x, y);
}

// norm (tgv_kernels.cuh:147-154) and norm_sym with weight 2 on xy (:137-145).
template <class P>
inline typename pix<P>::R norm2(P a, P b) {
  return std::sqrt(abs2(a) + abs2(b));
}
template <class P>
inline typename pix<P>::R norm_sym(P xx, P yy, P xy) {
  using R = typename pix<P>::R;
  return std::sqrt(abs2(xx) + abs2(yy) + R(2) * abs2(xy));
}

// ---------------------------------------------------------------------------
// Pointwise kernels, one function per reference __global__ (2D variants)
// ---------------------------------------------------------------------------

// TV_dual_p_kernel 2D (tgv_kernels.cuh:294-324): p <- (p + s grad ubar)/max(1,|.|).
template <class P>
void tv_dual_p(P* px, P* py, const P* ub, Grid g, typename pix<P>::R sigma) {
  using R = typename pix<P>::R;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      P gx, gy;
      dp(ub, g, x, y, gx, gy);
      const size_t i = g.at(x, y);
      P ax = px[i] + sigma * gx, ay = py[i] + sigma * gy;
      R s = R(1) / std::max(R(1), norm2(ax, ay));
      px[i] = ax * s;
      py[i] = ay * s;
    }
}

// TGV_dual_p_kernel (tgv_kernels.cuh:260-292).
template <class P>
void tgv_dual_p(P* px, P* py, const P* ub, const P* vbx, const P* vby, Grid g,
                typename pix<P>::R sigma, typename pix<P>::R alpha1) {
  using R = typename pix<P>::R;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      P gx, gy;
      dp(ub, g, x, y, gx, gy);
      const size_t i = g.at(x, y);
      P ax = px[i] + sigma * (gx - vbx[i]);
      P ay = py[i] + sigma * (gy - vby[i]);
      R s = R(1) / std::max(R(1), norm2(ax, ay) / alpha1);
      px[i] = ax * s;
      py[i] = ay * s;
    }
}

// TGV_dual_q_kernel (tgv_kernels.cuh:225-258).
template <class P>
void tgv_dual_q(P* qxx, P* qyy, P* qxy, const P* vbx, const P* vby, Grid g,
                typename pix<P>::R sigma, typename pix<P>::R alpha0) {
  using R = typename pix<P>::R;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      P exx, eyy, exy;
      dp_sym(vbx, vby, g, x, y, exx, eyy, exy);
      const size_t i = g.at(x, y);
      P axx = qxx[i] + sigma * exx, ayy = qyy[i] + sigma * eyy, axy = qxy[i] + sigma * exy;
      R s = R(1) / std::max(R(1), norm_sym(axx, ayy, axy) / alpha0);
      qxx[i] = axx * s;
      qyy[i] = ayy * s;
      qxy[i] = axy * s;
    }
}

// TV_primal_u_kernel 2D (tgv_kernels.cuh:568-605).
template <class P>
void tv_primal_u(P* u, P* ub, const P* f, const P* px, const P* py, Grid g,
                 typename pix<P>::R tau, typename pix<P>::R lambda, typename pix<P>::R theta) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      const size_t i = g.at(x, y);
      P uo = u[i];
      P div = dm(px, py, g, x, y);
      P un = uo + tau * (div + lambda * f[i]);
      un /= (typename pix<P>::R(1) + tau * lambda);
      u[i] = un;
      ub[i] = (typename pix<P>::R(1) + theta) * un - theta * uo;
    }
}

// TGV_primal_u_noprox_kernel 2D (tgv_kernels.cuh:432-467).  On entry ub holds
// the precomputed adjoint A^* r; on exit it holds the over-relaxed u.
template <class P>
void tgv_primal_u_noprox(P* u, P* ub, const P* px, const P* py, Grid g,
                         typename pix<P>::R tau, typename pix<P>::R theta) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      const size_t i = g.at(x, y);
      P uo = u[i];
      P w = ub[i];
      P div = dm(px, py, g, x, y);
      P un = uo + tau * (div - w);
      u[i] = un;
      ub[i] = (typename pix<P>::R(1) + theta) * un - theta * uo;
    }
}

// TGV_primal_v_kernel (tgv_kernels.cuh:647-686).
template <class P>
void tgv_primal_v(P* vx, P* vy, P* vbx, P* vby, const P* px, const P* py, const P* qxx,
                  const P* qyy, const P* qxy, Grid g, typename pix<P>::R tau,
                  typename pix<P>::R theta) {
  using R = typename pix<P>::R;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      const size_t i = g.at(x, y);
      P ox = vx[i], oy = vy[i];
      P dx, dy;
      dm_sym(qxx, qyy, qxy, g, x, y, dx, dy);
      P nx = ox + tau * (dx + px[i]);
      P ny = oy + tau * (dy + py[i]);
      vx[i] = nx;
      vy[i] = ny;
      vbx[i] = (R(1) + theta) * nx - theta * ox;
      vby[i] = (R(1) + theta) * ny - theta * oy;
    }
}

// TGV_prox_r_kernel (tgv_kernels.cuh:212-223): r <- (r + s tmp)/(1 + s lambda).
template <class P>
void prox_r(P* r, const P* tmp, size_t n, typename pix<P>::R sigma, typename pix<P>::R lambda) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i)
    r[i] = (r[i] + sigma * tmp[i]) / (typename pix<P>::R(1) + sigma * lambda);
}

// ---------------------------------------------------------------------------
// FFT: mixed-radix Stockham, any length, unnormalised.
// sign = -1 forward (exp(-2 pi i jk/n)), +1 inverse.  Stands in for the
// cuFFT calls that the reference makes through ImageUtilities
// (mricartesianoperator.cu:219,253); the transform is fixed to numpy's
// convention, which the tests pin with numpy golden vectors.
// ---------------------------------------------------------------------------

template <class R>
struct FftPlan {
  int n = 0;
  std::vector<int> radix;
  std::vector<cx<R>> tw;  // exp(-2 pi i m / n), m in [0, n), evaluated in long double
  // per stage s: stage twiddles w^(j k n / (p r)) for k < p, 1 <= j < r, at
  // stw[s][k * (r - 1) + j - 1] (forward sign); and the radix-r roots
  // exp(-2 pi i m / r) at root[s][m]
  std::vector<std::vector<cx<R>>> stw, root;
};

inline std::vector<int> factorize(int n) {
  std::vector<int> f;
  while (n % 4 == 0 && n > 4) { f.push_back(4); n /= 4; }
  for (int p : {2, 3, 4, 5}) while (n % p == 0 && n >= p) { f.push_back(p); n /= p; }
  for (int p = 7; n > 1; p += 2) while (n % p == 0) { f.push_back(p); n /= p; }
  return f;
}

template <class R>
const FftPlan<R>& fft_plan(int n) {
  // one plan per length, built once; the thread-local map makes the lookup
  // lock-free after the first call of a thread
  thread_local std::map<int, const FftPlan<R>*> local;
  auto hit = local.find(n);
  if (hit != local.end()) return *hit->second;
  static std::mutex mu;
  static std::map<int, std::unique_ptr<FftPlan<R>>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(n);
  if (it == cache.end()) {
    auto pl = std::make_unique<FftPlan<R>>();
    pl->n = n;
    pl->radix = factorize(n);
    pl->tw.resize(n);
    const long double twopi = 6.283185307179586476925286766559L;
    for (int m = 0; m < n; ++m) {
      long double a = -twopi * (long double)m / (long double)n;
      pl->tw[m] = cx<R>(R(std::cos(a)), R(std::sin(a)));
    }
    int p = 1;
    for (int r : pl->radix) {
      const int nr = n / r, ts = n / (p * r);
      std::vector<cx<R>> st(size_t(p) * (r - 1)), rt(r);
      for (int k = 0; k < p; ++k)
        for (int j = 1; j < r; ++j) st[size_t(k) * (r - 1) + j - 1] = pl->tw[size_t((long long)j * k * ts % n)];
      for (int m = 0; m < r; ++m) rt[m] = pl->tw[size_t(m) * nr];
      pl->stw.push_back(std::move(st));
      pl->root.push_back(std::move(rt));
      p *= r;
    }
    it = cache.emplace(n, std::move(pl)).first;
  }
  local[n] = it->second.get();
  return *it->second;
}

// Radix-r DFT of x (in place), sign -1 forward / +1 inverse; root[m] =
// exp(-2 pi i m / r).  Radices 2-5 with their closed forms, others direct.
template <class R>
inline void small_dft(cx<R>* x, int r, int sign, const cx<R>* root) {
  const R s = R(sign);
  auto rot = [&](cx<R> v) { return cx<R>(-s * v.imag(), s * v.real()); };  // v * (s i)
  switch (r) {
    case 2: {
      const cx<R> a = x[0], b = x[1];
      x[0] = a + b;
      x[1] = a - b;
      return;
    }
    case 3: {
      const R h = R(0.86602540378443864676372317075294L);  // sin(2 pi / 3)
      const cx<R> t = x[1] + x[2], d = rot(x[1] - x[2]) * h, m = x[0] - t * R(0.5);
      x[0] = x[0] + t;
      x[1] = m + d;
      x[2] = m - d;
      return;
    }
    case 4: {
      const cx<R> t0 = x[0] + x[2], t1 = x[0] - x[2], t2 = x[1] + x[3], t3 = rot(x[1] - x[3]);
      x[0] = t0 + t2;
      x[2] = t0 - t2;
      x[1] = t1 + t3;
      x[3] = t1 - t3;
      return;
    }
    case 5: {
      const R c1 = R(0.30901699437494742410229341718281906L), c2 = R(-0.80901699437494742410229341718281906L);
      const R s1 = R(0.95105651629515357211643933337938214L), s2 = R(0.58778525229247312916870595463907277L);
      const cx<R> t1 = x[1] + x[4], t2 = x[2] + x[3], d1 = x[1] - x[4], d2 = x[2] - x[3];
      const cx<R> a1 = x[0] + t1 * c1 + t2 * c2, a2 = x[0] + t1 * c2 + t2 * c1;
      const cx<R> b1 = rot(d1 * s1 + d2 * s2), b2 = rot(d1 * s2 - d2 * s1);
      x[0] = x[0] + t1 + t2;
      x[1] = a1 + b1;
      x[4] = a1 - b1;
      x[2] = a2 + b2;
      x[3] = a2 - b2;
      return;
    }
    default: {
      cx<R> y[64];
      for (int q = 0; q < r; ++q) {
        cx<R> acc = x[0];
        int m = 0;
        for (int j = 1; j < r; ++j) {
          m += q;
          if (m >= r) m -= r;
          const cx<R> w = sign > 0 ? std::conj(root[m]) : root[m];
          acc += x[j] * w;
        }
        y[q] = acc;
      }
      for (int q = 0; q < r; ++q) x[q] = y[q];
    }
  }
}

// In place on a (contiguous, length n); work must hold n elements.
// Stockham autosort: stage radix r, p = product of the earlier radices,
// nr = n / r; for i < nr, k = i mod p: x_j = in[i + j nr] w^(j k n/(p r)),
// y = DFT_r(x), out[(i - k) r + k + q p] = y_q.
template <class R>
void fft1(cx<R>* a, int n, int sign, cx<R>* work) {
  if (n <= 1) return;
  const FftPlan<R>& pl = fft_plan<R>(n);
  cx<R>* in = a;
  cx<R>* out = work;
  int p = 1;
  cx<R> x[64];
  for (size_t s = 0; s < pl.radix.size(); ++s) {
    const int r = pl.radix[s], nr = n / r;
    const cx<R>* st = pl.stw[s].data();
    const cx<R>* rt = pl.root[s].data();
    for (int i0 = 0; i0 < nr; i0 += p)  // i = i0 + k, k < p (p divides nr)
      for (int k = 0; k < p; ++k) {
        const int i = i0 + k;
        x[0] = in[i];
        const cx<R>* w = st + size_t(k) * (r - 1);
        for (int j = 1; j < r; ++j) x[j] = in[i + j * nr] * (sign > 0 ? std::conj(w[j - 1]) : w[j - 1]);
        small_dft(x, r, sign, rt);
        cx<R>* o = out + size_t(i0) * r + k;
        for (int q = 0; q < r; ++q) o[q * p] = x[q];
      }
    std::swap(in, out);
    p *= r;
  }
  if (in != a) std::copy(in, in + n, a);
}

// numpy-convention shifts.  ifftshift: element i moves to (i + ceil(n/2)) % n,
// which is the prefft/preifft index map of mricartesianoperator.cu:44-49,88-93.
// fftshift: element i moves to (i + n/2) % n.
inline int ifftshift_dst(int i, int n) { return (i + (n + 1) / 2) % n; }
inline int fftshift_dst(int i, int n) { return (i + n / 2) % n; }

// Centered orthonormal 2D transform of one W x H plane:
// out = fftshift(FFT2(ifftshift(in))) / sqrt(W H) (sign -1), inverse for +1.
template <class R>
void fft2c_plane(const cx<R>* in, cx<R>* out, int W, int H, int sign) {
  std::vector<cx<R>> t(size_t(W) * H), buf(std::max(W, H)), wk(std::max(W, H));
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      t[size_t(ifftshift_dst(x, W)) + size_t(W) * ifftshift_dst(y, H)] = in[size_t(x) + size_t(W) * y];
  for (int y = 0; y < H; ++y) fft1(&t[size_t(W) * y], W, sign, wk.data());
  for (int x = 0; x < W; ++x) {
    for (int y = 0; y < H; ++y) buf[y] = t[size_t(x) + size_t(W) * y];
    fft1(buf.data(), H, sign, wk.data());
    for (int y = 0; y < H; ++y) t[size_t(x) + size_t(W) * y] = buf[y];
  }
  const R sc = R(1.0 / std::sqrt((double)W * (double)H));
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      out[size_t(fftshift_dst(x, W)) + size_t(W) * fftshift_dst(y, H)] = t[size_t(x) + size_t(W) * y] * sc;
}

// ---------------------------------------------------------------------------
// Cartesian MRI operators (mricartesianoperator.cu)
// ---------------------------------------------------------------------------

enum OpKind { OP_CARTESIAN = 0, OP_CARTESIAN_ROOS = 1, OP_SAMPLING = 2 };

template <class R>
struct CartOp {
  int kind = OP_CARTESIAN;
  int W = 0, Hk = 0, C = 1;     // k-space grid {W (PE), Hk (FE), C}
  const cx<R>* S = nullptr;     // {W, Hk, C} coil sensitivities (null for sampling)
  const R* M = nullptr;         // {W, Hk} real mask
  int H() const { return kind == OP_CARTESIAN_ROOS ? Hk / 2 : Hk; }
  // FE_idx of mricartesianoperator.cu:341,381: unsigned(size * 0.25f + 1).
  int y0() const { return kind == OP_CARTESIAN_ROOS ? int((unsigned)(float(Hk) * 0.25f + 1)) : 0; }
  size_t img_n() const { return size_t(W) * H(); }
  size_t ks_n() const { return size_t(W) * Hk * C; }
};

// Forward: executeForward (mricartesianoperator.cu:197-227, ROOS :321-365,
// sampling mrisamplingoperator.cu:59-70).
template <class R>
void cart_forward(const CartOp<R>& op, const cx<R>* x, cx<R>* y) {
  const int W = op.W, Hk = op.Hk, H = op.H(), y0 = op.y0();
  const size_t G = size_t(W) * Hk;
#pragma omp parallel for schedule(dynamic)
  for (int c = 0; c < op.C; ++c) {
    std::vector<cx<R>> z(G, cx<R>(0));
    for (int yy = 0; yy < H; ++yy)
      for (int xx = 0; xx < W; ++xx) {
        size_t gi = size_t(xx) + size_t(W) * (yy + y0);
        cx<R> s = op.S ? op.S[gi + G * c] : cx<R>(1);
        z[gi] = x[size_t(xx) + size_t(W) * yy] * s;
      }
    fft2c_plane(z.data(), y + G * c, W, Hk, -1);
    for (size_t i = 0; i < G; ++i) y[G * c + i] *= op.M[i];
  }
}

// Adjoint: executeAdjoint (mricartesianoperator.cu:230-265, ROOS :368-414,
// sampling mrisamplingoperator.cu:73-88).  The coil combination is
// sum_c img_c * conj(S_c) (combineImg_kernel, :100-126).
template <class R>
void cart_adjoint(const CartOp<R>& op, const cx<R>* y, cx<R>* x) {
  const int W = op.W, Hk = op.Hk, H = op.H(), y0 = op.y0();
  const size_t G = size_t(W) * Hk;
  std::vector<cx<R>> acc(size_t(op.C) * G);
#pragma omp parallel for schedule(dynamic)
  for (int c = 0; c < op.C; ++c) {
    std::vector<cx<R>> z(G);
    for (size_t i = 0; i < G; ++i) z[i] = y[G * c + i] * op.M[i];
    fft2c_plane(z.data(), &acc[G * c], W, Hk, +1);
  }
#pragma omp parallel for schedule(static)
  for (int yy = 0; yy < H; ++yy)
    for (int xx = 0; xx < W; ++xx) {
      size_t gi = size_t(xx) + size_t(W) * (yy + y0);
      cx<R> s = 0;
      for (int c = 0; c < op.C; ++c) {
        cx<R> sc = op.S ? op.S[gi + G * c] : cx<R>(1);
        s += acc[G * c + gi] * std::conj(sc);
      }
      x[size_t(xx) + size_t(W) * yy] = s;
    }
}

// ---------------------------------------------------------------------------
// Solvers
// ---------------------------------------------------------------------------

enum SolverKind { S_TV = 0, S_TGV = 1, S_TGV_MRI = 2, S_TV_OP = 3, S_TGV_OP = 4 };

struct Params {
  double lambda = 1, alpha0 = 1, alpha1 = 1, reduction = 1;
  unsigned max_iter = 1, check = 100;
  double gap_tol = 0;      // 0: run exactly max_iter + 1 iterations
  unsigned gap_every = 0;  // evaluate the gap every gap_every iterations (0: never)
  double radius_u = 0, radius_v = 0;  // bounded-domain radii (0: defaults)
};

struct Stats {
  unsigned iters = 0;
  double primal = 0, dual = 0, gap = 0, rel_gap = 0;
  double seconds = 0;
  unsigned n_gap = 0;
};

}  // namespace pdto
