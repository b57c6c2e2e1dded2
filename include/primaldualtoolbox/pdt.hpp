// pdt.hpp -- header-only C++ mirror of the toolbox classes over the C ABI.
//
// Same class names, template shape <InputType, OutputType>, methods and
// parameter fields as the reference's src/optimizer and src/operator headers
// (tvoptimizer.h, tgvoptimizer.h, tgvmrioptimizer.h, tvoptimizerwithop.h,
// tgvoptimizerwithop.h, operatorbase.h, mricartesianoperator.h,
// mrisamplingoperator.h), with pdt::DeviceArray<Pixel, Ndim> in place of
// iu::LinearDeviceMemory<Pixel, Ndim> and pdt::Exception in place of
// IuException.  Every call forwards to libpdt_hip.so.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../pdt_capi.h"

namespace pdt {

class Exception : public std::runtime_error {
 public:
  Exception(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != PDT_OK) throw Exception(rc, pdt_last_error());
}

template <class P> struct dtype_of;
template <> struct dtype_of<float> { static constexpr int value = PDT_F32; };
template <> struct dtype_of<double> { static constexpr int value = PDT_F64; };
template <> struct dtype_of<std::complex<float>> { static constexpr int value = PDT_C64; };
template <> struct dtype_of<std::complex<double>> { static constexpr int value = PDT_C128; };

template <class P> struct real_of { using type = P; };
template <class R> struct real_of<std::complex<R>> { using type = R; };

// iu::Size<N>: dim 0 fastest.
template <unsigned N>
struct Size : std::array<int64_t, N> {
  Size() { this->fill(0); }
  Size(std::initializer_list<int64_t> l) {
    this->fill(1);
    unsigned i = 0;
    for (int64_t v : l) (*this)[i++] = v;
  }
  int64_t numel() const {
    int64_t n = 1;
    for (auto v : *this) n *= v;
    return n;
  }
};

// iu::LinearDeviceMemory<Pixel, Ndim>.
template <class Pixel, unsigned Ndim>
class DeviceArray {
 public:
  using pixel_type = Pixel;
  static constexpr unsigned ndim = Ndim;

  explicit DeviceArray(const Size<Ndim>& s) : size_(s) {
    check(pdt_array_create(dtype_of<Pixel>::value, Ndim, s.data(), &h_));
  }
  ~DeviceArray() {
    if (h_ && own_) pdt_array_destroy(h_);
  }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;

  const Size<Ndim>& size() const { return size_; }
  void upload(const std::vector<Pixel>& host) {
    check(pdt_array_upload(h_, host.data(), host.size() * sizeof(Pixel)));
  }
  std::vector<Pixel> download() const {
    std::vector<Pixel> host(size_t(size_.numel()));
    check(pdt_array_download(h_, host.data(), host.size() * sizeof(Pixel)));
    return host;
  }
  pdt_array* handle() const { return h_; }

  // wraps a borrowed handle owned by the library (getResult)
  static std::unique_ptr<DeviceArray> borrow(pdt_array* h) {
    int dt = 0, nd = 0;
    int64_t s3[3];
    check(pdt_array_info(h, &dt, &nd, s3));
    std::unique_ptr<DeviceArray> a(new DeviceArray());
    a->h_ = h;
    a->own_ = false;
    for (unsigned i = 0; i < Ndim; ++i) a->size_[i] = s3[i];
    return a;
  }

 private:
  DeviceArray() = default;
  pdt_array* h_ = nullptr;
  bool own_ = true;
  Size<Ndim> size_;
};

// OperatorBase<InputType, OutputType> (operatorbase.h:204-507).
template <class InputType, class OutputType>
class OperatorBase {
 public:
  using complex_type = typename InputType::pixel_type;
  using real_type = typename real_of<complex_type>::type;

  virtual ~OperatorBase() {
    if (h_) pdt_op_destroy(h_);
  }
  OperatorBase(const OperatorBase&) = delete;
  void operator=(const OperatorBase&) = delete;

  // addConstant: constant order matters (coil sensitivities, then mask).
  template <class Pixel, unsigned Nd>
  void addConstant(const std::vector<Pixel>& host, const Size<Nd>& size) {
    check(pdt_op_add_constant(h_, host.data(), dtype_of<Pixel>::value, Nd, size.data()));
  }
  void forward(const InputType& src, OutputType& dst) { check(pdt_op_forward(h_, src.handle(), dst.handle())); }
  void adjoint(const OutputType& src, InputType& dst) { check(pdt_op_adjoint(h_, src.handle(), dst.handle())); }
  Size<InputType::ndim> getInputSize(const OutputType& out) {
    int nd = 0;
    int64_t s[3];
    check(pdt_op_input_size(h_, OutputType::ndim, out.size().data(), &nd, s));
    Size<InputType::ndim> r;
    for (unsigned i = 0; i < InputType::ndim; ++i) r[i] = s[i];
    return r;
  }
  Size<OutputType::ndim> getOutputSize(const InputType& in) {
    int nd = 0;
    int64_t s[3];
    check(pdt_op_output_size(h_, InputType::ndim, in.size().data(), &nd, s));
    Size<OutputType::ndim> r;
    for (unsigned i = 0; i < OutputType::ndim; ++i) r[i] = s[i];
    return r;
  }
  // adjointnessCheck: returns |<Au,p> - <u,A^*p>| / |<Au,p>|.
  double adjointnessCheck() {
    double a, b, c, d;
    check(pdt_op_adjointness(h_, &a, &b, &c, &d));
    return std::abs(std::complex<double>(a - c, b - d)) / std::abs(std::complex<double>(a, b));
  }
  std::string print() const {
    char buf[256];
    check(pdt_op_name(h_, buf, sizeof(buf)));
    return buf;
  }
  friend std::ostream& operator<<(std::ostream& os, const OperatorBase& op) { return os << op.print(); }
  pdt_op* handle() const { return h_; }

 protected:
  explicit OperatorBase(int kind) { check(pdt_op_create(kind, dtype_of<complex_type>::value, &h_)); }

 private:
  pdt_op* h_ = nullptr;
};

template <class InputType, class OutputType>
class MriCartesianOperator : public OperatorBase<InputType, OutputType> {
 public:
  MriCartesianOperator() : OperatorBase<InputType, OutputType>(PDT_OP_CARTESIAN) {}
};
template <class InputType, class OutputType>
class MriCartesianRemoveROOSOperator : public OperatorBase<InputType, OutputType> {
 public:
  MriCartesianRemoveROOSOperator() : OperatorBase<InputType, OutputType>(PDT_OP_CARTESIAN_REMOVE_ROOS) {}
};
template <class InputType, class OutputType>
class MriSamplingOperator : public OperatorBase<InputType, OutputType> {
 public:
  MriSamplingOperator() : OperatorBase<InputType, OutputType>(PDT_OP_SAMPLING) {}
};

// Parameter structs (tvoptimizer.h:32-61, tgvoptimizer.h:34-67,
// tgvmrioptimizer.h:33-66) with the reference defaults.
template <class T>
struct TvParameters {
  unsigned int check = 100;
  T lambda = 1;
  unsigned int max_iter = 1;
  double gap_tol = 0;
  unsigned int gap_every = 0;
};
template <class T>
struct TgvParameters {
  unsigned int check = 100;
  T lambda = 1;
  unsigned int max_iter = 1;
  T alpha0 = 1;
  T alpha1 = 1;
  double gap_tol = 0;
  unsigned int gap_every = 0;
  bool dual_fp16 = false;
};
template <class T>
struct TgvMriParameters {
  T alpha0 = 1;
  T alpha1 = 1;
  unsigned int check = 100;
  T reduction = 1;
  unsigned int max_iter = 1;
  double gap_tol = 0;
  unsigned int gap_every = 0;
};

namespace detail {
template <class P>
pdt_params common(int kind, const P& p) {
  pdt_params c;
  pdt_params_default(kind, &c);
  c.max_iter = p.max_iter;
  c.check = p.check;
  c.gap_tol = p.gap_tol;
  c.gap_every = p.gap_every;
  return c;
}
template <class T>
pdt_params to_c(int kind, const TvParameters<T>& p) {
  pdt_params c = common(kind, p);
  c.lambda = p.lambda;
  return c;
}
template <class T>
pdt_params to_c(int kind, const TgvParameters<T>& p) {
  pdt_params c = common(kind, p);
  c.lambda = p.lambda;
  c.alpha0 = p.alpha0;
  c.alpha1 = p.alpha1;
  c.dual_fp16 = p.dual_fp16 ? 1 : 0;
  return c;
}
template <class T>
pdt_params to_c(int kind, const TgvMriParameters<T>& p) {
  pdt_params c = common(kind, p);
  c.alpha0 = p.alpha0;
  c.alpha1 = p.alpha1;
  c.reduction = p.reduction;
  return c;
}

template <class InputType, class OutputType, int Kind, class Params>
class SolverBase {
 public:
  using pixel_type = typename InputType::pixel_type;
  using real_type = typename real_of<pixel_type>::type;

  SolverBase() { check(pdt_solver_create(Kind, dtype_of<pixel_type>::value, &h_)); }
  virtual ~SolverBase() {
    if (h_) pdt_solver_destroy(h_);
  }
  SolverBase(const SolverBase&) = delete;
  void operator=(const SolverBase&) = delete;

  void setNoisyData(const std::shared_ptr<OutputType>& f) {
    f_ = f;
    check(pdt_solver_set_noisy_data(h_, f->handle()));
  }
  void setInput0(const InputType& u0) { check(pdt_solver_set_input0(h_, u0.handle(), nullptr)); }
  // NEW: also warm-starts v (TGV solvers), v0 of size {W, H, 2}.
  void setInput0(const InputType& u0, const DeviceArray<pixel_type, 3>& v0) {
    check(pdt_solver_set_input0(h_, u0.handle(), v0.handle()));
  }
  pdt_solver* handle() const { return h_; }
  void solve(bool verbose = false) {
    pdt_params c = to_c(Kind, params_);
    check(pdt_solver_set_params(h_, &c));
    check(pdt_solver_solve(h_, verbose ? 1 : 0));
    result_.reset();
  }
  InputType* getResult() {
    pdt_array* u = nullptr;
    check(pdt_solver_get_result(h_, &u));
    result_ = InputType::borrow(u);
    return result_.get();
  }
  Params& getParameters() { return params_; }
  pdt_stats getStats() const {
    pdt_stats s;
    check(pdt_solver_get_stats(h_, &s));
    return s;
  }

 protected:
  pdt_solver* h_ = nullptr;
  Params params_;
  std::shared_ptr<OutputType> f_;
  std::unique_ptr<InputType> result_;
};

template <class InputType, class OutputType, int Kind, class Params>
class OpSolverBase : public SolverBase<InputType, OutputType, Kind, Params> {
 public:
  void setOperator(const std::shared_ptr<OperatorBase<InputType, OutputType>>& op) {
    op_ = op;
    check(pdt_solver_set_operator(this->h_, op->handle()));
  }

 private:
  std::shared_ptr<OperatorBase<InputType, OutputType>> op_;
};
}  // namespace detail

template <class InputType, class OutputType>
class TvOptimizer : public detail::SolverBase<InputType, OutputType, PDT_SOLVER_TV,
                                              TvParameters<typename real_of<typename InputType::pixel_type>::type>> {};
template <class InputType, class OutputType>
class TgvOptimizer
    : public detail::SolverBase<InputType, OutputType, PDT_SOLVER_TGV,
                                TgvParameters<typename real_of<typename InputType::pixel_type>::type>> {};
template <class InputType, class OutputType>
class TgvMriOptimizer
    : public detail::OpSolverBase<InputType, OutputType, PDT_SOLVER_TGV_MRI,
                                  TgvMriParameters<typename real_of<typename InputType::pixel_type>::type>> {};
template <class InputType, class OutputType>
class TvOptimizerWithOp
    : public detail::OpSolverBase<InputType, OutputType, PDT_SOLVER_TV_OP,
                                  TvParameters<typename real_of<typename InputType::pixel_type>::type>> {};
template <class InputType, class OutputType>
class TgvOptimizerWithOp
    : public detail::OpSolverBase<InputType, OutputType, PDT_SOLVER_TGV_OP,
                                  TgvParameters<typename real_of<typename InputType::pixel_type>::type>> {};

// NEW: learned warm start of the complex64 TGV operator solvers (no
// counterpart in the reference): a conv3x3 network 2 -> width -> ... -> 6
// predicting (u, v) from A^*f; see pdt_cnn_* in pdt_capi.h.
class WarmStartNetwork {
 public:
  explicit WarmStartNetwork(int depth = 5, int width = 32) { check(pdt_cnn_create(2, width, depth, 6, &h_)); }
  ~WarmStartNetwork() {
    if (h_) pdt_cnn_destroy(h_);
  }
  WarmStartNetwork(const WarmStartNetwork&) = delete;
  void operator=(const WarmStartNetwork&) = delete;
  size_t numParams() const {
    size_t n = 0;
    check(pdt_cnn_num_params(h_, &n));
    return n;
  }
  void setParams(const std::vector<float>& p) { check(pdt_cnn_set_params(h_, p.data(), p.size())); }
  std::vector<float> getParams() const {
    std::vector<float> p(numParams());
    check(pdt_cnn_get_params(h_, p.data(), p.size()));
    return p;
  }
  // Sets (u0, v0) of a configured TgvMriOptimizer / TgvOptimizerWithOp.
  template <class Solver>
  void warmStart(Solver& s) {
    check(pdt_solver_warm_start(s.handle(), h_));
  }
  pdt_cnn* handle() const { return h_; }

 private:
  pdt_cnn* h_ = nullptr;
};

}  // namespace pdt
