// matlab.hpp -- the reference's MATLAB entry points (matlab/CMakeLists.txt:14-30)
// as host functions a mexFunction can call directly.
//
// Arrays are MATLAB's column-major host arrays: sensitivities and rawdata of
// size (nfe, npe, nc), images and masks (nfe, npe), FE fastest, complex
// values interleaved (the R2018a complex API, mxGetComplexDoubles).  They are
// reordered here to the library's layout (PE fastest, iu size {npe, nfe, nc}),
// the layout in which the reference's Python module and data files hold the
// same arrays (bin/mri_cartesian_test.py:29-36 vs bin/mri_cartesian_test.m:25-45).
// As in the mex sources, the operators compute in single precision
// (mexMriCartesianFwdOp.cpp:43) and the reconstructions in double precision
// (mexMriCartesianReconstructTgv.cpp:44).  Errors are pdt::Exception.
#pragma once

#include <complex>
#include <cstdint>
#include <memory>
#include <vector>

#include "primaldualtoolbox/pdt.hpp"

namespace pdt {
namespace matlab {

using cd = std::complex<double>;

// The four fields the reconstruction mex files read from tgv_params
// (mexMriCartesianReconstructTgv.cpp:79-89).
struct TgvParams {
  double alpha0 = 1, alpha1 = 1, reduction = 1;
  unsigned int max_iter = 1;
};

namespace detail {
// MATLAB (fe, pe, c) at fe + nfe * (pe + npe * c)  <->  library (pe, fe, c)
// at pe + npe * (fe + nfe * c).
template <class To, class From>
std::vector<To> from_colmajor(const From* a, int64_t nfe, int64_t npe, int64_t nc) {
  std::vector<To> out(size_t(nfe * npe * nc));
  for (int64_t c = 0; c < nc; ++c)
    for (int64_t pe = 0; pe < npe; ++pe)
      for (int64_t fe = 0; fe < nfe; ++fe) out[size_t(pe + npe * (fe + nfe * c))] = To(a[fe + nfe * (pe + npe * c)]);
  return out;
}
template <class From>
std::vector<cd> to_colmajor(const std::vector<From>& a, int64_t nfe, int64_t npe, int64_t nc) {
  std::vector<cd> out(a.size());
  for (int64_t c = 0; c < nc; ++c)
    for (int64_t pe = 0; pe < npe; ++pe)
      for (int64_t fe = 0; fe < nfe; ++fe) out[size_t(fe + nfe * (pe + npe * c))] = cd(a[size_t(pe + npe * (fe + nfe * c))]);
  return out;
}

template <class R>
using Img = DeviceArray<std::complex<R>, 2>;
template <class R>
using Ksp = DeviceArray<std::complex<R>, 3>;

// Operator with constants [0] = sensitivities (nfe_k, npe, nc), [1] = mask
// (nfe_k, npe) in the reference's constant order (mricartesianoperator.h:25-30).
template <class R>
std::shared_ptr<OperatorBase<Img<R>, Ksp<R>>> make_op(bool remove_roos, const cd* sens, const double* mask,
                                                     int64_t nfe_k, int64_t npe, int64_t nc) {
  std::shared_ptr<OperatorBase<Img<R>, Ksp<R>>> op;
  if (remove_roos) op = std::make_shared<MriCartesianRemoveROOSOperator<Img<R>, Ksp<R>>>();
  else op = std::make_shared<MriCartesianOperator<Img<R>, Ksp<R>>>();
  op->addConstant(from_colmajor<std::complex<R>>(sens, nfe_k, npe, nc), Size<3>{npe, nfe_k, nc});
  op->addConstant(from_colmajor<R>(mask, nfe_k, npe, 1), Size<2>{npe, nfe_k});
  return op;
}

inline int64_t image_fe(bool remove_roos, int64_t nfe_k) { return remove_roos ? nfe_k / 2 : nfe_k; }
}  // namespace detail

// gpuMriCartesianFwd / gpuMriCartesianRemoveROOSFwd (sensitivities, image,
// mask): image (nfe_k or nfe_k / 2, npe) -> rawdata (nfe_k, npe, nc).
inline std::vector<cd> cartesianForward(bool remove_roos, const cd* sens, const cd* image, const double* mask,
                                        int64_t nfe_k, int64_t npe, int64_t nc) {
  auto op = detail::make_op<float>(remove_roos, sens, mask, nfe_k, npe, nc);
  const int64_t nfe = detail::image_fe(remove_roos, nfe_k);
  detail::Img<float> x(Size<2>{npe, nfe});
  x.upload(detail::from_colmajor<std::complex<float>>(image, nfe, npe, 1));
  detail::Ksp<float> y(op->getOutputSize(x));
  op->forward(x, y);
  return detail::to_colmajor(y.download(), nfe_k, npe, nc);
}

// gpuMriCartesianAdj / gpuMriCartesianRemoveROOSAdj (sensitivities, rawdata,
// mask): rawdata (nfe_k, npe, nc) -> image (nfe_k or nfe_k / 2, npe).
inline std::vector<cd> cartesianAdjoint(bool remove_roos, const cd* sens, const cd* rawdata, const double* mask,
                                        int64_t nfe_k, int64_t npe, int64_t nc) {
  auto op = detail::make_op<float>(remove_roos, sens, mask, nfe_k, npe, nc);
  detail::Ksp<float> y(Size<3>{npe, nfe_k, nc});
  y.upload(detail::from_colmajor<std::complex<float>>(rawdata, nfe_k, npe, nc));
  detail::Img<float> x(op->getInputSize(y));
  op->adjoint(y, x);
  return detail::to_colmajor(x.download(), detail::image_fe(remove_roos, nfe_k), npe, 1);
}

// gpuTgvMriOptimizer_cartesian / _cartesianRemoveROOS (sensitivities,
// rawdata, mask, tgv_params): TGV-MRI reconstruction in double precision.
inline std::vector<cd> tgvMriReconstruct(bool remove_roos, const cd* sens, const cd* rawdata, const double* mask,
                                         int64_t nfe_k, int64_t npe, int64_t nc, const TgvParams& prm) {
  auto op = detail::make_op<double>(remove_roos, sens, mask, nfe_k, npe, nc);
  auto f = std::make_shared<detail::Ksp<double>>(Size<3>{npe, nfe_k, nc});
  f->upload(detail::from_colmajor<cd>(rawdata, nfe_k, npe, nc));
  TgvMriOptimizer<detail::Img<double>, detail::Ksp<double>> opt;
  opt.getParameters().reduction = prm.reduction;
  opt.getParameters().max_iter = prm.max_iter;
  opt.getParameters().alpha0 = prm.alpha0;
  opt.getParameters().alpha1 = prm.alpha1;
  opt.setOperator(op);
  opt.setNoisyData(f);
  opt.solve();
  return detail::to_colmajor(opt.getResult()->download(), detail::image_fe(remove_roos, nfe_k), npe, 1);
}

}  // namespace matlab
}  // namespace pdt
