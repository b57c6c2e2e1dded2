// Runs the MATLAB entry functions of primaldualtoolbox/matlab.hpp the way a
// mexFunction would: column-major double arrays in, column-major arrays out.
//
// Usage: matlab_entry <in.bin> <out.bin>
//   in:  int64 nfe_k, npe, nc, remove_roos, max_iter; double alpha0, alpha1,
//        reduction; then complex sensitivities (nfe_k, npe, nc), complex image
//        (nfe, npe), complex rawdata (nfe_k, npe, nc) and real mask
//        (nfe_k, npe), each column-major, complex values interleaved.
//   out: forward(image), adjoint(rawdata), TGV reconstruction of rawdata.
#include <cstdio>
#include <vector>

#include "primaldualtoolbox/matlab.hpp"

using cd = std::complex<double>;

template <class T>
static bool rd(std::FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  std::FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t h[5];
  double a[3];
  if (!rd(f, h, 5) || !rd(f, a, 3)) return 2;
  const int64_t nfe_k = h[0], npe = h[1], nc = h[2];
  const bool roos = h[3] != 0;
  const int64_t nfe = roos ? nfe_k / 2 : nfe_k;
  std::vector<cd> sens(size_t(nfe_k * npe * nc)), img(size_t(nfe * npe)), raw(sens.size());
  std::vector<double> mask(size_t(nfe_k * npe));
  if (!rd(f, sens.data(), sens.size()) || !rd(f, img.data(), img.size()) || !rd(f, raw.data(), raw.size()) ||
      !rd(f, mask.data(), mask.size()))
    return 2;
  std::fclose(f);
  try {
    pdt::matlab::TgvParams prm;
    prm.alpha0 = a[0];
    prm.alpha1 = a[1];
    prm.reduction = a[2];
    prm.max_iter = unsigned(h[4]);
    const auto y = pdt::matlab::cartesianForward(roos, sens.data(), img.data(), mask.data(), nfe_k, npe, nc);
    const auto x = pdt::matlab::cartesianAdjoint(roos, sens.data(), raw.data(), mask.data(), nfe_k, npe, nc);
    const auto u = pdt::matlab::tgvMriReconstruct(roos, sens.data(), raw.data(), mask.data(), nfe_k, npe, nc, prm);
    std::FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(y.data(), sizeof(cd), y.size(), o);
    std::fwrite(x.data(), sizeof(cd), x.size(), o);
    std::fwrite(u.data(), sizeof(cd), u.size(), o);
    std::fclose(o);
    std::printf("matlab entries: forward %zu, adjoint %zu, tgv %zu values\n", y.size(), x.size(), u.size());
    return 0;
  } catch (const pdt::Exception& e) {
    std::printf("pdt::Exception (%d): %s\n", e.code(), e.what());
    return 1;
  }
}
