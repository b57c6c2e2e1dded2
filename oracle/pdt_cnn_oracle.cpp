// pdt_cnn_oracle.cpp -- FP64 restatement of the warm-start network
// (TEST INFRASTRUCTURE ONLY: loaded by tests/ to check the HIP kernels).
//
// Network: nl conv3x3 layers (zero padding, stride 1), channel counts
// ch[0..nl], ReLU after every layer but the last.  Tensors NHWC (channel
// fastest).  Parameters per layer: W[tap][co][ci] (tap = 3 ky + kx), b[co].
// Loss 1/(2m) sum (y - t)^2 over the m outputs.
//
// With bf16 = 1 the values the GPU rounds to bfloat16 (round to nearest
// even) are rounded at the same points: the input, the weights, every hidden
// activation, the output gradient and every propagated gradient; sums stay in
// FP64.  With bf16 = 0 everything is FP64 (finite-difference checks).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double bf(double v) {
  float f = float(v);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return double(f);
}

struct Net {
  int nl;
  std::vector<int> ch;
  std::vector<size_t> off;
  const double* p;
  bool b16;
  int n, h, w;
  size_t P() const { return size_t(n) * h * w; }
  double W(int l, int tap, int co, int ci) const {
    const double v = p[off[l] + (size_t(tap) * ch[l + 1] + co) * ch[l] + ci];
    return b16 ? bf(v) : v;
  }
  double B(int l, int co) const { return p[off[l] + size_t(9) * ch[l] * ch[l + 1] + co]; }
  // neighbour pixel of p at tap, or -1 outside the image
  int64_t nbr(size_t pix, int tap) const {
    const size_t img = pix / (size_t(h) * w), r = pix - img * size_t(h) * w;
    const int y = int(r / w) + tap / 3 - 1, x = int(r % w) + tap % 3 - 1;
    if (y < 0 || y >= h || x < 0 || x >= w) return -1;
    return int64_t(img * size_t(h) * w + size_t(y) * w + x);
  }
  // z = conv_l(a)
  void conv(int l, const std::vector<double>& a, std::vector<double>& z) const {
    const int ci_n = ch[l], co_n = ch[l + 1];
    z.assign(P() * co_n, 0.0);
    for (size_t q = 0; q < P(); ++q)
      for (int co = 0; co < co_n; ++co) {
        double s = B(l, co);
        for (int tap = 0; tap < 9; ++tap) {
          const int64_t nb = nbr(q, tap);
          if (nb < 0) continue;
          for (int ci = 0; ci < ci_n; ++ci) s += W(l, tap, co, ci) * a[size_t(nb) * ci_n + ci];
        }
        z[q * co_n + co] = s;
      }
  }
  // acts[0] = x, acts[l + 1] = layer l output (post-ReLU, rounded) ; y last
  void forward(const double* x, std::vector<std::vector<double>>& acts) const {
    acts.assign(nl + 1, {});
    acts[0].assign(x, x + P() * ch[0]);
    if (b16)
      for (double& v : acts[0]) v = bf(v);
    for (int l = 0; l < nl; ++l) {
      conv(l, acts[l], acts[l + 1]);
      if (l + 1 < nl)
        for (double& v : acts[l + 1]) {
          v = v > 0 ? v : 0.0;
          if (b16) v = bf(v);
        }
    }
  }
};

Net make(int nl, const int* ch, const double* params, int n, int h, int w, int b16) {
  Net net;
  net.nl = nl;
  net.ch.assign(ch, ch + nl + 1);
  size_t o = 0;
  for (int l = 0; l < nl; ++l) {
    net.off.push_back(o);
    o += size_t(9) * ch[l] * ch[l + 1] + ch[l + 1];
  }
  net.p = params;
  net.b16 = b16 != 0;
  net.n = n;
  net.h = h;
  net.w = w;
  return net;
}

}  // namespace

extern "C" {

int pdto_cnn_num_params(int nl, const int* ch) {
  size_t o = 0;
  for (int l = 0; l < nl; ++l) o += size_t(9) * ch[l] * ch[l + 1] + ch[l + 1];
  return int(o);
}

int pdto_cnn_forward(int nl, const int* ch, const double* params, int n, int h, int w, int b16, const double* x,
                     double* y) {
  const Net net = make(nl, ch, params, n, h, w, b16);
  std::vector<std::vector<double>> acts;
  net.forward(x, acts);
  std::memcpy(y, acts[nl].data(), acts[nl].size() * sizeof(double));
  return 0;
}

// Backpropagation: dz_L = (y - t) / m; for layer l: dW = sum_p dz_l[p] a_l[p + off]^T,
// db = sum_p dz_l[p], and dz_{l-1} = (conv_l^T dz_l) * [a_l > 0].
int pdto_cnn_loss_grad(int nl, const int* ch, const double* params, int n, int h, int w, int b16, const double* x,
                       const double* t, double* loss, double* grad) {
  const Net net = make(nl, ch, params, n, h, w, b16);
  std::vector<std::vector<double>> acts;
  net.forward(x, acts);
  const size_t P = net.P(), M = P * ch[nl];
  std::vector<double> dz(M);
  double ss = 0;
  for (size_t i = 0; i < M; ++i) {
    const double d = acts[nl][i] - t[i];
    ss += d * d;
    dz[i] = d / double(M);
    if (b16) dz[i] = bf(dz[i]);
  }
  *loss = ss / (2.0 * double(M));
  for (int l = nl - 1; l >= 0; --l) {
    const int ci_n = ch[l], co_n = ch[l + 1];
    const std::vector<double>& a = acts[l];
    double* gW = grad + net.off[l];
    double* gb = gW + size_t(9) * ci_n * co_n;
    for (size_t i = 0; i < size_t(9) * ci_n * co_n; ++i) gW[i] = 0;
    for (int co = 0; co < co_n; ++co) gb[co] = 0;
    std::vector<double> din(l > 0 ? P * ci_n : 0, 0.0);
    for (size_t q = 0; q < P; ++q)
      for (int co = 0; co < co_n; ++co) {
        const double g = dz[q * co_n + co];
        gb[co] += g;
        for (int tap = 0; tap < 9; ++tap) {
          const int64_t nb = net.nbr(q, tap);
          if (nb < 0) continue;
          for (int ci = 0; ci < ci_n; ++ci) {
            gW[(size_t(tap) * co_n + co) * ci_n + ci] += g * a[size_t(nb) * ci_n + ci];
            if (l > 0) din[size_t(nb) * ci_n + ci] += net.W(l, tap, co, ci) * g;
          }
        }
      }
    if (l > 0) {
      for (size_t i = 0; i < din.size(); ++i) {
        din[i] = a[i] > 0 ? din[i] : 0.0;
        if (b16) din[i] = bf(din[i]);
      }
      dz.swap(din);
    }
  }
  return 0;
}

// Adam (Kingma & Ba 2015, Algorithm 1), step counted from 1.
void pdto_adam(int64_t np, double* w, const double* g, double* m, double* v, double lr, double b1, double b2,
               double eps, int step) {
  const double c1 = 1.0 - std::pow(b1, step), c2 = 1.0 - std::pow(b2, step);
  for (int64_t i = 0; i < np; ++i) {
    m[i] = b1 * m[i] + (1.0 - b1) * g[i];
    v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
    w[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
  }
}

}  // extern "C"
