// Uses the header-only C++ mirror the way the reference's C++ callers use
// src/optimizer and src/operator: a TV denoising solve and a Cartesian
// operator adjointness check.  Exit status 0 on success; prints the relative
// difference of the TV result to a plain CPU restatement of the same 50
// iterations (tvoptimizer.cu:107-122) computed here in double.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "primaldualtoolbox/pdt.hpp"

using Arr2f = pdt::DeviceArray<float, 2>;
using Arr2c = pdt::DeviceArray<std::complex<float>, 2>;
using Arr3c = pdt::DeviceArray<std::complex<float>, 3>;

static std::vector<double> tv_cpu(const std::vector<float>& f, int W, int H, double lambda, int iters) {
  const double tau = 1.0 / std::sqrt(8.0), sigma = tau;
  std::vector<double> u(f.begin(), f.end()), ub = u, px(W * H, 0.0), py(W * H, 0.0);
  for (int k = 0; k < iters; ++k) {
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        int i = x + W * y;
        double gx = ub[(x < W - 1 ? i + 1 : i)] - ub[i], gy = ub[(y < H - 1 ? i + W : i)] - ub[i];
        double ax = px[i] + sigma * gx, ay = py[i] + sigma * gy;
        double s = 1.0 / std::max(1.0, std::sqrt(ax * ax + ay * ay));
        px[i] = ax * s;
        py[i] = ay * s;
      }
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        int i = x + W * y;
        double d = (x > 0 ? (x < W - 1 ? px[i] - px[i - 1] : -px[i - 1]) : px[i]) +
                   (y > 0 ? (y < H - 1 ? py[i] - py[i - W] : -py[i - W]) : py[i]);
        double uo = u[i], un = (uo + tau * (d + lambda * f[i])) / (1 + tau * lambda);
        u[i] = un;
        ub[i] = 2 * un - uo;
      }
  }
  return u;
}

int main() {
  try {
    const int W = 24, H = 20;
    std::mt19937 rng(7);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::vector<float> f(W * H);
    for (auto& v : f) v = nd(rng);
    auto fd = std::make_shared<Arr2f>(pdt::Size<2>{W, H});
    fd->upload(f);
    pdt::TvOptimizer<Arr2f, Arr2f> opt;
    opt.setNoisyData(fd);
    opt.getParameters().lambda = 5;
    opt.getParameters().max_iter = 49;
    opt.solve();
    std::vector<float> u = opt.getResult()->download();
    std::vector<double> ref = tv_cpu(f, W, H, 5.0, 50);
    double num = 0, den = 0;
    for (size_t i = 0; i < u.size(); ++i) {
      num += (u[i] - ref[i]) * (u[i] - ref[i]);
      den += ref[i] * ref[i];
    }
    const double err = std::sqrt(num / den);
    std::printf("TV 24x20, 50 iterations: rel L2 vs CPU = %.3e\n", err);

    pdt::MriCartesianOperator<Arr2c, Arr3c> op;
    std::uniform_real_distribution<float> ud(0.f, 1.f);
    std::vector<std::complex<float>> S(10 * 10 * 5);
    for (auto& s : S) s = {ud(rng), ud(rng)};
    std::vector<float> M(10 * 10);
    for (auto& m : M) m = ud(rng) < 0.5f ? 0.f : 1.f;
    op.addConstant(S, pdt::Size<3>{10, 10, 5});
    op.addConstant(M, pdt::Size<2>{10, 10});
    const double adj = op.adjointnessCheck();
    std::printf("%s adjointness rel. error = %.3e\n", op.print().c_str(), adj);

    // warm start with an all-zero network: u0 = A^*f exactly, v0 = 0
    auto mop = std::make_shared<pdt::MriCartesianOperator<Arr2c, Arr3c>>();
    mop->addConstant(S, pdt::Size<3>{10, 10, 5});
    mop->addConstant(M, pdt::Size<2>{10, 10});
    std::vector<std::complex<float>> fk(10 * 10 * 5);
    for (size_t i = 0; i < fk.size(); ++i) fk[i] = {M[i % 100] * nd(rng), M[i % 100] * nd(rng)};
    auto fdev = std::make_shared<Arr3c>(pdt::Size<3>{10, 10, 5});
    fdev->upload(fk);
    Arr2c atf(pdt::Size<2>{10, 10});
    mop->adjoint(*fdev, atf);
    pdt::TgvMriOptimizer<Arr2c, Arr3c> mri;
    mri.setOperator(mop);
    mri.setNoisyData(fdev);
    pdt::WarmStartNetwork net;
    net.setParams(std::vector<float>(net.numParams(), 0.f));
    net.warmStart(mri);
    const std::vector<std::complex<float>> u0 = mri.getResult()->download(), a = atf.download();
    double wnum = 0, wden = 0;
    for (size_t i = 0; i < a.size(); ++i) {
      wnum += std::norm(u0[i] - a[i]);
      wden += std::norm(a[i]);
    }
    const double werr = std::sqrt(wnum / wden);
    std::printf("warm start (zero network, %zu parameters): rel. difference to A^*f = %.3e\n", net.numParams(), werr);
    return (err < 1e-5 && adj < 1e-5 && werr < 1e-6) ? 0 : 1;
  } catch (const pdt::Exception& e) {
    std::printf("pdt::Exception (%d): %s\n", e.code(), e.what());
    return 2;
  }
}
