// pdt_oracle.cpp -- C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Built twice into one library: R = double is the FP64 oracle the parity tests
// check the GPU against; R = float (OpenMP) is the reported CPU baseline.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
#include <chrono>
#include <cstring>
#include <omp.h>

#include "pdt_oracle.hpp"

namespace pdto {

template <class P>
struct State {
  using R = typename pix<P>::R;
  Grid g{0, 0};
  std::vector<P> u, ub, vx, vy, vbx, vby, px, py, qxx, qyy, qxy;
  void init(Grid gg, const P* u0, const P* v0, bool tgv) {
    g = gg;
    const size_t n = g.n();
    u.assign(u0, u0 + n);
    ub = u;
    px.assign(n, P(0));
    py.assign(n, P(0));
    if (tgv) {
      vx.assign(n, P(0));
      vy.assign(n, P(0));
      if (v0) {
        std::copy(v0, v0 + n, vx.begin());
        std::copy(v0 + n, v0 + 2 * n, vy.begin());
      }
      vbx = vx;
      vby = vy;
      qxx.assign(n, P(0));
      qyy.assign(n, P(0));
      qxy.assign(n, P(0));
    }
  }
};

// ---- primal-dual gap helpers (definitions shared with the GPU library) ----

// sum_px norm(grad u - v) (v may be null) and sum_px norm_sym(E v).
template <class P>
void reg_terms(const State<P>& s, bool tgv, double& tv1, double& tv0) {
  const Grid g = s.g;
  double a = 0, b = 0;
#pragma omp parallel for reduction(+ : a, b) schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      P gx, gy;
      dp(s.ub.data(), g, x, y, gx, gy);
      const size_t i = g.at(x, y);
      if (tgv) {
        gx -= s.vbx[i];
        gy -= s.vby[i];
        P exx, eyy, exy;
        dp_sym(s.vbx.data(), s.vby.data(), g, x, y, exx, eyy, exy);
        b += std::sqrt(abs2d(exx) + abs2d(eyy) + 2.0 * abs2d(exy));
      }
      a += std::sqrt(abs2d(gx) + abs2d(gy));
    }
  tv1 = a;
  tv0 = b;
}

// ||ub - f||^2
template <class P>
double sqdist(const std::vector<P>& a, const P* f) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long long i = 0; i < (long long)a.size(); ++i) s += abs2d(a[i] - f[i]);
  return s;
}

// ROF dual value D(p) = -Re<div p, f> - ||div p||^2 / (2 lambda).
template <class P>
double rof_dual(const State<P>& s, const P* f, double lambda) {
  const Grid g = s.g;
  double a = 0, b = 0;
#pragma omp parallel for reduction(+ : a, b) schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      P d = dm(s.px.data(), s.py.data(), g, x, y);
      a += redot(d, f[g.at(x, y)]);
      b += abs2d(d);
    }
  return -a - b / (2.0 * lambda);
}

// TGV denoising dual value with the feasible restoration p~ = -s div_sym q:
// with d = div_sym q, m = max |d|, g = div d, a = Re<g, f>, b = ||g||^2,
// D(s) = s a - s^2 b / (2 lambda), s in [0, min(1, alpha1 / m)], maximised.
template <class P>
double tgv_dual_restored(const State<P>& s, const P* f, double lambda, double alpha1) {
  const Grid g = s.g;
  const size_t n = g.n();
  std::vector<P> dx(n), dy(n);
  double m = 0;
#pragma omp parallel for reduction(max : m) schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      const size_t i = g.at(x, y);
      dm_sym(s.qxx.data(), s.qyy.data(), s.qxy.data(), g, x, y, dx[i], dy[i]);
      m = std::max(m, std::sqrt(abs2d(dx[i]) + abs2d(dy[i])));
    }
  double a = 0, b = 0;
#pragma omp parallel for reduction(+ : a, b) schedule(static)
  for (int y = 0; y < g.H; ++y)
    for (int x = 0; x < g.W; ++x) {
      P gg = dm(dx.data(), dy.data(), g, x, y);
      a += redot(gg, f[g.at(x, y)]);
      b += abs2d(gg);
    }
  double smax = m > 0 ? std::min(1.0, alpha1 / m) : 1.0;
  double sopt = b > 0 ? std::min(std::max(lambda * a / b, 0.0), smax) : smax;
  return sopt * a - sopt * sopt * b / (2.0 * lambda);
}

template <class R>
double l2norm(const cx<R>* a, size_t n) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long long i = 0; i < (long long)n; ++i) s += abs2d(a[i]);
  return std::sqrt(s);
}

// Radius of the partial-gap ball, as a fraction of ||A^* f||.
constexpr double kRadiusFrac = 0.01;

inline bool gap_due(const Params& p, unsigned k) {
  return p.gap_every > 0 && ((k + 1) % p.gap_every) == 0;
}

inline double rel_gap(double P, double D) {
  return (P - D) / std::max(std::fabs(P), 1e-300);
}

// Gap history rows: P, D, rel gap, then up to three solver-specific
// components (MRI: smooth dual part, ||A^*r - div p||, ||p + div_sym q||).
constexpr int kHistCols = 6;
struct GapLog {
  double* hist;
  unsigned cap;
  void put(Stats& st, double P, double D, double c0 = 0, double c1 = 0, double c2 = 0) {
    st.primal = P;
    st.dual = D;
    st.gap = P - D;
    st.rel_gap = rel_gap(P, D);
    if (hist && st.n_gap < cap) {
      double* h = hist + kHistCols * st.n_gap;
      h[0] = P; h[1] = D; h[2] = st.rel_gap; h[3] = c0; h[4] = c1; h[5] = c2;
    }
    st.n_gap++;
  }
};

// A.1: TvOptimizer::solve (tvoptimizer.cu:77-125).
template <class P>
void solve_tv(Grid g, const P* f, const P* u0, const Params& prm, P* u_out, GapLog gl, Stats& st) {
  using R = typename pix<P>::R;
  State<P> s;
  s.init(g, u0 ? u0 : f, nullptr, false);
  const R tau = R(1.0 / std::sqrt(8.0)), sigma = R(1.0 / std::sqrt(8.0)), theta = R(1);
  const R lambda = R(prm.lambda);
  auto t0 = std::chrono::steady_clock::now();
  unsigned k = 0;
  for (k = 0; k <= prm.max_iter; ++k) {
    const bool chk = gap_due(prm, k);
    double Pv = 0;
    if (chk) {
      double a, b;
      reg_terms(s, false, a, b);
      Pv = a + 0.5 * prm.lambda * sqdist(s.ub, f);
    }
    tv_dual_p(s.px.data(), s.py.data(), s.ub.data(), g, sigma);
    tv_primal_u(s.u.data(), s.ub.data(), f, s.px.data(), s.py.data(), g, tau, lambda, theta);
    if (chk) {
      double Dv = rof_dual(s, f, prm.lambda);
      gl.put(st, Pv, Dv);
      if (prm.gap_tol > 0 && st.rel_gap <= prm.gap_tol) { ++k; break; }
    }
  }
  st.iters = std::min(k, prm.max_iter + 1);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::copy(s.u.begin(), s.u.end(), u_out);
}

// A.2: TgvOptimizer::solve (tgvoptimizer.cu:81-141), primal step first.
template <class P>
void solve_tgv(Grid g, const P* f, const P* u0, const P* v0, const Params& prm, P* u_out,
               P* v_out, GapLog gl, Stats& st) {
  using R = typename pix<P>::R;
  State<P> s;
  s.init(g, u0 ? u0 : f, v0, true);
  const R tau = R(1.0 / std::sqrt(12.0)), sigma = R(1.0 / std::sqrt(12.0)), theta = R(1);
  const R lambda = R(prm.lambda), a0 = R(prm.alpha0), a1 = R(prm.alpha1);
  auto t0 = std::chrono::steady_clock::now();
  unsigned k = 0;
  for (k = 0; k <= prm.max_iter; ++k) {
    const bool chk = gap_due(prm, k);
    tv_primal_u(s.u.data(), s.ub.data(), f, s.px.data(), s.py.data(), g, tau, lambda, theta);
    tgv_primal_v(s.vx.data(), s.vy.data(), s.vbx.data(), s.vby.data(), s.px.data(), s.py.data(),
                 s.qxx.data(), s.qyy.data(), s.qxy.data(), g, tau, theta);
    double Pv = 0;
    if (chk) {
      double t1, t0v;
      reg_terms(s, true, t1, t0v);
      Pv = prm.alpha1 * t1 + prm.alpha0 * t0v + 0.5 * prm.lambda * sqdist(s.ub, f);
    }
    tgv_dual_p(s.px.data(), s.py.data(), s.ub.data(), s.vbx.data(), s.vby.data(), g, sigma, a1);
    tgv_dual_q(s.qxx.data(), s.qyy.data(), s.qxy.data(), s.vbx.data(), s.vby.data(), g, sigma, a0);
    if (chk) {
      double Dv = tgv_dual_restored(s, f, prm.lambda, prm.alpha1);
      gl.put(st, Pv, Dv);
      if (prm.gap_tol > 0 && st.rel_gap <= prm.gap_tol) { ++k; break; }
    }
  }
  st.iters = std::min(k, prm.max_iter + 1);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::copy(s.u.begin(), s.u.end(), u_out);
  if (v_out) {
    std::copy(s.vx.begin(), s.vx.end(), v_out);
    std::copy(s.vy.begin(), s.vy.end(), v_out + g.n());
  }
}

// A.3 / A.4 / A.5: TgvMriOptimizer::solve (tgvmrioptimizer.cu:94-203),
// TgvOptimizerWithOp::solve (tgvoptimizerwithop.cu:90-179) and
// TvOptimizerWithOp::solve (tvoptimizerwithop.cu:87-164).
template <class R>
void solve_op(int kind, const CartOp<R>& op, const cx<R>* f, const cx<R>* u0, const cx<R>* v0,
              const Params& prm, cx<R>* u_out, cx<R>* v_out, GapLog gl, Stats& st) {
  using P = cx<R>;
  const Grid g{op.W, op.H()};
  const size_t nk = op.ks_n();
  const bool tgv = kind != S_TV_OP;
  std::vector<P> atf(g.n());
  cart_adjoint(op, f, atf.data());  // u0 = A^* f (tgvmrioptimizer.cu:112-118)
  State<P> s;
  s.init(g, u0 ? u0 : atf.data(), v0, tgv);
  std::vector<P> r(nk, P(0)), tmp(nk), w(g.n());
  R tau, sigma, lambda;
  if (kind == S_TGV_MRI) {
    tau = R(1.0 / 16.0); sigma = R(1.0 / 8.0); lambda = R(1);
  } else {
    tau = R(1.0 / 8.0); sigma = R(1.0 / 8.0); lambda = R(prm.lambda);
  }
  const R theta = R(1);
  // Partial-gap ball radii around the current primal iterate (rho_u, rho_v).
  const double Ru = prm.radius_u > 0 ? prm.radius_u : kRadiusFrac * l2norm(atf.data(), g.n());
  const double Rv = prm.radius_v > 0 ? prm.radius_v : std::sqrt(8.0) * Ru;
  // alpha continuation (tgvmrioptimizer.cu:127-130,142-155)
  const R a00 = R(prm.alpha0), a10 = R(prm.alpha1);
  const R a01 = R(prm.alpha0 * prm.reduction), a11 = R(prm.alpha1 * prm.reduction);
  const R dmax = R(prm.max_iter);
  auto t0 = std::chrono::steady_clock::now();
  unsigned k = 0;
  for (k = 0; k <= prm.max_iter; ++k) {
    R alpha0 = a00, alpha1 = a10;
    if (kind == S_TGV_MRI) {
      const R dk = R(k);
      alpha0 = std::exp(dk / dmax * std::log(a01) + (dmax - dk) / dmax * std::log(a00));
      alpha1 = std::exp(dk / dmax * std::log(a11) + (dmax - dk) / dmax * std::log(a10));
    }
    const bool chk = gap_due(prm, k);
    cart_forward(op, s.ub.data(), tmp.data());
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nk; ++i) tmp[i] = tmp[i] - f[i];
    double Pv = 0, Preg = 0;
    if (chk) {
      double t1, t0v;
      reg_terms(s, tgv, t1, t0v);
      double dsq = 0;
      for (size_t i = 0; i < nk; ++i) dsq += abs2d(tmp[i]);
      Preg = tgv ? double(alpha1) * t1 + double(alpha0) * t0v : t1;
      Pv = dsq / (2.0 * double(lambda)) + Preg;
    }
    prox_r(r.data(), tmp.data(), nk, sigma, lambda);
    if (tgv) {
      tgv_dual_p(s.px.data(), s.py.data(), s.ub.data(), s.vbx.data(), s.vby.data(), g, sigma, alpha1);
      tgv_dual_q(s.qxx.data(), s.qyy.data(), s.qxy.data(), s.vbx.data(), s.vby.data(), g, sigma, alpha0);
    } else {
      tv_dual_p(s.px.data(), s.py.data(), s.ub.data(), g, sigma);
    }
    cart_adjoint(op, r.data(), w.data());
    if (chk) {
      // Fenchel-Young gaps at (ubar, vbar; r, p, q) and the stationarity
      // residuals A^*r - div p, p + div_sym q.
      double fy_d = 0, lin = 0, res_u = 0, res_v = 0;
      for (size_t i = 0; i < nk; ++i) {
        P a = tmp[i] - lambda * r[i];  // (A ubar - f) - lambda r
        fy_d += abs2d(a);
      }
      fy_d /= 2.0 * double(lambda);
      for (int y = 0; y < g.H; ++y)
        for (int x = 0; x < g.W; ++x) {
          const size_t i = g.at(x, y);
          P gx, gy;
          dp(s.ub.data(), g, x, y, gx, gy);
          if (tgv) { gx -= s.vbx[i]; gy -= s.vby[i]; }
          lin += redot(gx, s.px[i]) + redot(gy, s.py[i]);
          if (tgv) {
            P exx, eyy, exy;
            dp_sym(s.vbx.data(), s.vby.data(), g, x, y, exx, eyy, exy);
            lin += redot(exx, s.qxx[i]) + redot(eyy, s.qyy[i]) + 2.0 * redot(exy, s.qxy[i]);
          }
          P d = dm(s.px.data(), s.py.data(), g, x, y);
          res_u += abs2d(w[i] - d);
          if (tgv) {
            P dx, dy;
            dm_sym(s.qxx.data(), s.qyy.data(), s.qxy.data(), g, x, y, dx, dy);
            res_v += abs2d(s.px[i] + dx) + abs2d(s.py[i] + dy);
          }
        }
      const double fy_reg = Preg - lin;
      const double G = fy_d + fy_reg + Ru * std::sqrt(res_u) + (tgv ? Rv * std::sqrt(res_v) : 0.0);
      gl.put(st, Pv, Pv - G, fy_d + fy_reg, std::sqrt(res_u), std::sqrt(res_v));
    }
    s.ub = w;
    tgv_primal_u_noprox(s.u.data(), s.ub.data(), s.px.data(), s.py.data(), g, tau, theta);
    if (tgv)
      tgv_primal_v(s.vx.data(), s.vy.data(), s.vbx.data(), s.vby.data(), s.px.data(), s.py.data(),
                   s.qxx.data(), s.qyy.data(), s.qxy.data(), g, tau, theta);
    if (chk && prm.gap_tol > 0 && st.rel_gap <= prm.gap_tol) { ++k; break; }
  }
  st.iters = std::min(k, prm.max_iter + 1);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::copy(s.u.begin(), s.u.end(), u_out);
  if (v_out && tgv) {
    std::copy(s.vx.begin(), s.vx.end(), v_out);
    std::copy(s.vy.begin(), s.vy.end(), v_out + g.n());
  }
}

}  // namespace pdto

// ---------------------------------------------------------------------------
// C interface (ctypes).  prec = 64 (double / complex128) or 32 (float /
// complex64).  Complex arrays are interleaved (re, im).
// ---------------------------------------------------------------------------

struct pdto_opdesc {
  int kind, W, Hk, C;
  const void* S;  // complex {W, Hk, C}, may be null for kind 2 (sampling)
  const void* M;  // real {W, Hk}
};

struct pdto_params {
  double lambda, alpha0, alpha1, reduction;
  unsigned max_iter, check;
  double gap_tol;
  unsigned gap_every;
  double radius_u, radius_v;
};

struct pdto_stats {
  unsigned iters, n_gap;
  double primal, dual, gap, rel_gap, seconds;
};


template <class R>
static pdto::CartOp<R> mkop(const pdto_opdesc* d) {
  pdto::CartOp<R> op;
  op.kind = d->kind;
  op.W = d->W;
  op.Hk = d->Hk;
  op.C = d->kind == pdto::OP_SAMPLING ? 1 : d->C;
  op.S = d->kind == pdto::OP_SAMPLING ? nullptr : static_cast<const pdto::cx<R>*>(d->S);
  op.M = static_cast<const R*>(d->M);
  return op;
}

static pdto::Params mkprm(const pdto_params* p) {
  pdto::Params q;
  q.lambda = p->lambda;
  q.alpha0 = p->alpha0;
  q.alpha1 = p->alpha1;
  q.reduction = p->reduction;
  q.max_iter = p->max_iter;
  q.check = p->check;
  q.gap_tol = p->gap_tol;
  q.gap_every = p->gap_every;
  q.radius_u = p->radius_u;
  q.radius_v = p->radius_v;
  return q;
}

static void put_stats(const pdto::Stats& s, pdto_stats* o) {
  if (!o) return;
  o->iters = s.iters;
  o->n_gap = s.n_gap;
  o->primal = s.primal;
  o->dual = s.dual;
  o->gap = s.gap;
  o->rel_gap = s.rel_gap;
  o->seconds = s.seconds;
}

template <class R>
static int solve_t(int kind, int cplx, const pdto_opdesc* opd, int W, int H, const void* f,
                   const void* u0, const void* v0, const pdto_params* prm, void* u_out,
                   void* v_out, double* hist, unsigned cap, pdto_stats* st) {
  using namespace pdto;
  Params p = mkprm(prm);
  Stats s;
  GapLog gl{hist, cap};
  Grid g{W, H};
  if (kind == S_TV || kind == S_TGV) {
    if (cplx) {
      using P = cx<R>;
      if (kind == S_TV)
        solve_tv<P>(g, (const P*)f, (const P*)u0, p, (P*)u_out, gl, s);
      else
        solve_tgv<P>(g, (const P*)f, (const P*)u0, (const P*)v0, p, (P*)u_out, (P*)v_out, gl, s);
    } else {
      using P = R;
      if (kind == S_TV)
        solve_tv<P>(g, (const P*)f, (const P*)u0, p, (P*)u_out, gl, s);
      else
        solve_tgv<P>(g, (const P*)f, (const P*)u0, (const P*)v0, p, (P*)u_out, (P*)v_out, gl, s);
    }
  } else {
    if (!opd || !cplx) return 1;
    CartOp<R> op = mkop<R>(opd);
    solve_op<R>(kind, op, (const cx<R>*)f, (const cx<R>*)u0, (const cx<R>*)v0, p, (cx<R>*)u_out,
                (cx<R>*)v_out, gl, s);
  }
  put_stats(s, st);
  return 0;
}

extern "C" {

void pdto_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int pdto_max_threads(void) { return omp_get_max_threads(); }

int pdto_solve(int prec, int kind, int cplx, const pdto_opdesc* op, int W, int H, const void* f,
               const void* u0, const void* v0, const pdto_params* prm, void* u_out, void* v_out,
               double* gap_hist, unsigned gap_cap, pdto_stats* st) {
  if (prec == 64)
    return solve_t<double>(kind, cplx, op, W, H, f, u0, v0, prm, u_out, v_out, gap_hist, gap_cap, st);
  if (prec == 32)
    return solve_t<float>(kind, cplx, op, W, H, f, u0, v0, prm, u_out, v_out, gap_hist, gap_cap, st);
  return 1;
}

int pdto_op_forward(int prec, const pdto_opdesc* d, const void* x, void* y) {
  if (prec == 64) pdto::cart_forward(mkop<double>(d), (const pdto::cx<double>*)x, (pdto::cx<double>*)y);
  else if (prec == 32) pdto::cart_forward(mkop<float>(d), (const pdto::cx<float>*)x, (pdto::cx<float>*)y);
  else return 1;
  return 0;
}

int pdto_op_adjoint(int prec, const pdto_opdesc* d, const void* y, void* x) {
  if (prec == 64) pdto::cart_adjoint(mkop<double>(d), (const pdto::cx<double>*)y, (pdto::cx<double>*)x);
  else if (prec == 32) pdto::cart_adjoint(mkop<float>(d), (const pdto::cx<float>*)y, (pdto::cx<float>*)x);
  else return 1;
  return 0;
}

// In-place unnormalised 1D transform (sign -1 forward, +1 inverse), FP64.
int pdto_fft1(int n, int sign, void* data) {
  std::vector<pdto::cx<double>> wk(n);
  pdto::fft1((pdto::cx<double>*)data, n, sign, wk.data());
  return 0;
}

// Centered orthonormal 2D transform of one {W, H} plane, FP64.
int pdto_fft2c(int W, int H, int sign, const void* in, void* out) {
  pdto::fft2c_plane((const pdto::cx<double>*)in, (pdto::cx<double>*)out, W, H, sign);
  return 0;
}

// Stencil operators on FP64 arrays (cplx selects complex128 pixels).
// which: 0 grad u -> (gx, gy); 1 div (px, py) -> d; 2 symgrad (vx, vy) ->
// (xx, yy, xy); 3 symdiv (qxx, qyy, qxy) -> (dx, dy).
int pdto_stencil(int which, int cplx, int W, int H, const void* in, void* out) {
  using namespace pdto;
  Grid g{W, H};
  const size_t n = g.n();
  auto run = [&](auto tag) {
    using P = decltype(tag);
    const P* a = (const P*)in;
    P* o = (P*)out;
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const size_t i = g.at(x, y);
        switch (which) {
          case 0: dp(a, g, x, y, o[i], o[i + n]); break;
          case 1: o[i] = dm(a, a + n, g, x, y); break;
          case 2: dp_sym(a, a + n, g, x, y, o[i], o[i + n], o[i + 2 * n]); break;
          case 3: dm_sym(a, a + n, a + 2 * n, g, x, y, o[i], o[i + n]); break;
        }
      }
  };
  if (cplx) run(cx<double>(0)); else run(double(0));
  return 0;
}

}  // extern "C"
