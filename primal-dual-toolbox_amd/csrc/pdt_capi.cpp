// pdt_capi.cpp -- host side of libpdt_hip.so: device arrays, operators,
// solvers and the batched iteration engine behind the C ABI of pdt_capi.h.
//
// The solver state lives in HBM for the whole solve; an iteration is 2 (TV),
// 2 (TGV) or 3 (MRI) kernel launches over persistent work lists, replayed in
// chunks through a hipGraph.  The only host/device traffic inside a solve is
// one small control/statistics copy per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "pdt_capi.h"
#include "pdt_cnn.hpp"
#include "pdt_desc.hpp"
#include "pdt_launch.hpp"
#include "pdt_tune.hpp"

using namespace pdt;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_err;

// Held by every graph capture and by the host set-up that issues synchronous
// operations on the legacy stream (engine builds, solve preparation,
// operator set-up and applications, pinned-buffer allocation and release,
// work-list uploads): HIP refuses a legacy-stream operation that would
// depend on a stream another thread is capturing, so solves on several host
// threads (pdt_solve_concurrent) never run the two at once.  The operator's
// applications, which use its own descriptor and intermediates, also run one
// at a time under it.
std::recursive_mutex g_setup_mu;
using SetupLock = std::lock_guard<std::recursive_mutex>;

struct Err {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Err{code, msg}; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::ostringstream ss;
    ss << what << ": " << hipGetErrorString(e);
    fail(PDT_ERR_HIP, ss.str());
  }
}
#define HIPC(x) hip_check((x), #x)

template <class F>
int guard(F&& f) {
  try {
    f();
    return PDT_OK;
  } catch (const Err& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return PDT_ERR_ARG;
  }
}

size_t esize(int dt) {
  switch (dt) {
    case PDT_F32: return 4;
    case PDT_C64: return 8;
    case PDT_F64: return 8;
    case PDT_C128: return 16;
  }
  fail(PDT_ERR_ARG, "unknown dtype");
}
bool is_cplx(int dt) { return dt == PDT_C64 || dt == PDT_C128; }
bool is_dbl(int dt) { return dt == PDT_F64 || dt == PDT_C128; }
int real_of(int dt) { return is_dbl(dt) ? PDT_F64 : PDT_F32; }
int cplx_of(int dt) { return is_dbl(dt) ? PDT_C128 : PDT_C64; }
const char* dt_name(int dt) {
  switch (dt) {
    case PDT_F32: return "float";
    case PDT_C64: return "float2";
    case PDT_F64: return "double";
    case PDT_C128: return "double2";
  }
  return "?";
}

// RAII device buffer
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  void alloc(size_t b) {
    release();
    if (b) HIPC(hipMalloc(&p, b));
    bytes = b;
  }
  void ensure(size_t b) {
    if (b > bytes) alloc(b);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Twiddle tables, one per (device, length, precision), kept for the process.
struct TwKey {
  int dev, n, dbl;
  bool operator<(const TwKey& o) const {
    return dev != o.dev ? dev < o.dev : n != o.n ? n < o.n : dbl < o.dbl;
  }
};
std::mutex g_tw_mu;
std::map<TwKey, void*> g_tw;

std::vector<int> factorize(int n) {
  std::vector<int> f;
  while (n % 8 == 0 && n > 8) { f.push_back(8); n /= 8; }
  for (int p : {8, 4, 2, 3, 5}) while (n % p == 0 && n >= p) { f.push_back(p); n /= p; }
  for (int p = 7; n > 1; p += 2) while (n % p == 0) { f.push_back(p); n /= p; }
  return f;
}

FftPlanD make_plan(int n, bool dbl) {
  FftPlanD pl{};
  pl.n = n;
  std::vector<int> f = n > 1 ? factorize(n) : std::vector<int>{};
  if ((int)f.size() > kMaxStages) fail(PDT_ERR_UNSUPPORTED, "FFT length has too many factors");
  for (int p : f)
    if (p > 64) fail(PDT_ERR_UNSUPPORTED, "FFT length has a prime factor above 64: " + std::to_string(n));
  pl.ns = (int)f.size();
  int pp = 1;
  for (size_t i = 0; i < f.size(); ++i) {
    pl.r[i] = f[i];
    pl.nr[i] = n / f[i];
    pl.p[i] = pp;
    pl.tws[i] = n / (pp * f[i]);
    pl.rnr[i] = 1.0f / float(pl.nr[i]);
    pl.rp[i] = 1.0f / float(pp);
    pp *= f[i];
  }
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  TwKey key{dev, n, dbl ? 1 : 0};
  auto it = g_tw.find(key);
  if (it == g_tw.end()) {
    const long double twopi = 6.283185307179586476925286766559L;
    void* d = nullptr;
    if (dbl) {
      std::vector<double> h(2 * size_t(n));
      for (int m = 0; m < n; ++m) {
        long double a = -twopi * m / n;
        h[2 * m] = double(cosl(a));
        h[2 * m + 1] = double(sinl(a));
      }
      HIPC(hipMalloc(&d, h.size() * sizeof(double)));
      HIPC(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
      std::vector<float> h(2 * size_t(n));
      for (int m = 0; m < n; ++m) {
        long double a = -twopi * m / n;
        h[2 * m] = float(cosl(a));
        h[2 * m + 1] = float(sinl(a));
      }
      HIPC(hipMalloc(&d, h.size() * sizeof(float)));
      HIPC(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    it = g_tw.emplace(key, d).first;
  }
  pl.tw = it->second;
  return pl;
}

// Largest tiles that fit the LDS of a workgroup: coils per row work item and
// columns per column group of the MRI passes.
int max_cb(int dt, int W, int nc, int C) {
  int cb = std::max(C, 1);
  while (cb > 1 && lds_rows_bytes(dt, cb, W, nc) > init_kernel_attrs()) --cb;
  return cb;
}
int max_cg(int dt, int Hk, int nc) {
  int cg = std::max(nc, 1);
  while (cg > 1 && lds_cols_bytes(dt, cg, Hk) > init_kernel_attrs()) --cg;
  return cg;
}
// Default tiles before any autotuning: 4 coils per row item, 2 columns per
// column group; temporally blocked ROF launches of 8 iterations on 32 x 16
// tiles.
constexpr int kDefaultCb = 4, kDefaultCg = 2;
constexpr int kDefaultTvBlock[3] = {32, 16, 8};
constexpr int kDefaultTgvBlock[3] = {16, 16, 4};

}  // namespace

// ---------------------------------------------------------------------------
// arrays
// ---------------------------------------------------------------------------
struct pdt_array {
  int dtype = PDT_F32;
  int ndim = 0;
  int64_t size[3] = {1, 1, 1};
  void* ptr = nullptr;
  bool owned = false;
  size_t numel() const { return size_t(size[0]) * size_t(size[1]) * size_t(size[2]); }
  size_t bytes() const { return numel() * esize(dtype); }
  ~pdt_array() {
    if (owned && ptr) (void)hipFree(ptr);
  }
};

namespace {
pdt_array* new_array(int dtype, int ndim, const int64_t* size) {
  if (ndim < 1 || ndim > 3) fail(PDT_ERR_ARG, "ndim must be 1..3");
  auto a = std::make_unique<pdt_array>();
  a->dtype = dtype;
  a->ndim = ndim;
  for (int i = 0; i < ndim; ++i) {
    if (size[i] < 0) fail(PDT_ERR_ARG, "negative size");
    a->size[i] = size[i];
  }
  (void)esize(dtype);
  if (a->bytes()) HIPC(hipMalloc(&a->ptr, a->bytes()));
  a->owned = true;
  return a.release();
}

std::string size_str(const int64_t* s, int nd) {
  std::ostringstream ss;
  ss << "(";
  for (int i = 0; i < nd; ++i) ss << (i ? ", " : "") << s[i];
  ss << ")";
  return ss.str();
}

// IU_SIZE_CHECK
void size_check(const int64_t* a, int na, const int64_t* b, int nb) {
  bool ok = na == nb;
  for (int i = 0; ok && i < na; ++i) ok = a[i] == b[i];
  if (!ok) fail(PDT_ERR_SIZE, "Size mismatch: " + size_str(a, na) + " vs " + size_str(b, nb));
}
}  // namespace

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
struct HostConst {
  int dtype, ndim;
  int64_t size[3];
  std::unique_ptr<pdt_array> dev;
  std::vector<char> host;
};

struct pdt_op {
  std::atomic<int> refs{1};
  int kind = PDT_OP_CARTESIAN;
  int dtype = PDT_C64;
  std::vector<HostConst> consts;
  // prepared geometry (valid once prepared == true)
  bool prepared = false;
  int W = 0, Hk = 0, H = 0, C = 1, y0 = 0, ncols = 0;
  DevBuf Mc, cols, colmap;
  std::vector<int> h_cols;
  FftPlanD planW{}, planH{};
  // operator-mode scratch
  DevBuf T1, T2, wpart, ctl, slice, rows, rowgroups, colitems;
  int nrows = 0, nrg = 0, ncolitems = 0, cb = 1, cg = 1;
  size_t lds_rows = 0, lds_cols = 0;
  hipStream_t stream = nullptr;
  // host source of the descriptor copy (pinned, outlives the asynchronous
  // copy) and the completion of the last application, which shares the
  // descriptor and the scratch planes with the next one
  SliceDev* h_desc = nullptr;
  hipEvent_t done = nullptr;

  int required() const { return kind == PDT_OP_SAMPLING ? 1 : 2; }
  const char* name() const {
    switch (kind) {
      case PDT_OP_CARTESIAN: return "MriCartesianOperator";
      case PDT_OP_CARTESIAN_REMOVE_ROOS: return "MriCartesianRemoveROOSOperator";
      default: return "MriSamplingOperator";
    }
  }
  std::string prefix() const { return std::string("Operator '") + name() + "': "; }

  // OperatorBase::getConstant type check (operatorbase.h:263-290)
  const HostConst& get_const(unsigned idx, int dtype_want, int ndim_want) const {
    if (idx >= consts.size()) {
      std::ostringstream ss;
      ss << prefix() << "Requested index (" << idx << ") is not valid. Constants available: " << consts.size();
      fail(PDT_ERR_CONST, ss.str());
    }
    const HostConst& c = consts[idx];
    if (c.dtype != dtype_want || c.ndim != ndim_want) {
      std::ostringstream ss;
      ss << prefix() << "Data at idx (" << idx
         << ") cannot be cast to type iu::LinearDeviceMemory<PixelType, Ndim> with PixelType="
         << dt_name(dtype_want) << " Ndim=" << ndim_want;
      fail(PDT_ERR_CONST, ss.str());
    }
    return c;
  }

  // OperatorBase::operatorCheck constant-count part (operatorbase.h:491-503)
  void check_count() const {
    if ((int)consts.size() != required()) {
      std::ostringstream ss;
      ss << prefix() << "Number of constants (" << consts.size()
         << ") does not match number of required constants (" << required() << ")";
      fail(PDT_ERR_CONST, ss.str());
    }
  }

  const HostConst& mask() const {
    return get_const(kind == PDT_OP_SAMPLING ? 0 : 1, real_of(dtype), 2);
  }
  const HostConst* sens() const {
    return kind == PDT_OP_SAMPLING ? nullptr : &get_const(0, cplx_of(dtype), 3);
  }
  const void* S_dev() const {
    const HostConst* s = sens();
    return s ? s->dev->ptr : nullptr;
  }

  void prepare() {
    SetupLock lk(g_setup_mu);
    check_count();
    if (prepared) return;
    const HostConst& m = mask();
    const HostConst* s = sens();
    W = int(m.size[0]);
    Hk = int(m.size[1]);
    if (s) {
      C = int(s->size[2]);
      const int64_t want[2] = {s->size[0], s->size[1]};
      size_check(m.size, 2, want, 2);
    } else {
      C = 1;
    }
    if (kind == PDT_OP_CARTESIAN_REMOVE_ROOS) {
      H = Hk / 2;
      y0 = int((unsigned)(float(Hk) * 0.25f + 1));  // mricartesianoperator.cu:341
      if (y0 + H > Hk) fail(PDT_ERR_SIZE, prefix() + "read-out window exceeds the k-space grid");
    } else {
      H = Hk;
      y0 = 0;
    }
    if (W > kMaxRowLen) fail(PDT_ERR_UNSUPPORTED, prefix() + "PE length above 1024 is not supported");
    // listed columns: PE lines with any nonzero mask entry
    const bool dbl = is_dbl(dtype);
    auto mval = [&](int kx, int kfe) -> double {
      size_t i = size_t(kfe) * W + kx;
      return dbl ? reinterpret_cast<const double*>(m.host.data())[i]
                 : double(reinterpret_cast<const float*>(m.host.data())[i]);
    };
    h_cols.clear();
    std::vector<int> cmap(W, -1);
    for (int kx = 0; kx < W; ++kx) {
      bool any = false;
      for (int kfe = 0; kfe < Hk && !any; ++kfe) any = mval(kx, kfe) != 0.0;
      if (any) {
        cmap[kx] = (int)h_cols.size();
        h_cols.push_back(kx);
      }
    }
    ncols = (int)h_cols.size();
    const size_t re = dbl ? 8 : 4;
    std::vector<char> mc(std::max<size_t>(1, size_t(ncols) * Hk * re));
    for (int ks = 0; ks < ncols; ++ks)
      for (int kfe = 0; kfe < Hk; ++kfe) {
        double v = mval(h_cols[ks], kfe);
        if (dbl) reinterpret_cast<double*>(mc.data())[size_t(ks) * Hk + kfe] = v;
        else reinterpret_cast<float*>(mc.data())[size_t(ks) * Hk + kfe] = float(v);
      }
    Mc.alloc(mc.size());
    HIPC(hipMemcpy(Mc.p, mc.data(), mc.size(), hipMemcpyHostToDevice));
    cols.alloc(std::max<size_t>(1, h_cols.size()) * sizeof(int));
    if (ncols) HIPC(hipMemcpy(cols.p, h_cols.data(), h_cols.size() * sizeof(int), hipMemcpyHostToDevice));
    colmap.alloc(size_t(W) * sizeof(int));
    HIPC(hipMemcpy(colmap.p, cmap.data(), cmap.size() * sizeof(int), hipMemcpyHostToDevice));
    planW = make_plan(W, dbl);
    planH = make_plan(Hk, dbl);
    prepared = true;
  }

  // LDS chunking shared by the operator mode and the solvers
  void choose_tiles(int& cb_out, int& cg_out, size_t& lr, size_t& lc) const {
    const int dt = dtype == PDT_C128 ? DT_C128 : DT_C64;
    if (lds_cols_bytes(dt, 1, Hk) > init_kernel_attrs() || lds_rows_bytes(dt, 1, W, ncols) > init_kernel_attrs())
      fail(PDT_ERR_UNSUPPORTED, prefix() + "transform length too large for the LDS passes");
    cb_out = std::min(kDefaultCb, max_cb(dt, W, ncols, C));
    cg_out = std::min(kDefaultCg, max_cg(dt, Hk, ncols));
    lr = lds_rows_bytes(dt, cb_out, W, ncols);
    lc = lds_cols_bytes(dt, cg_out, Hk);
  }

  void ensure_opmode() {
    prepare();
    if (!stream) HIPC(hipStreamCreate(&stream));
    const size_t ce = esize(cplx_of(dtype));
    T1.ensure(std::max<size_t>(1, size_t(C) * H * ncols * ce));
    T2.ensure(std::max<size_t>(1, size_t(C) * H * ncols * ce));
    if (!ctl.p) {
      ctl.alloc(CTL_WORDS * sizeof(int));
      int h[CTL_WORDS] = {1, 0, 0, 0, 0, 0, 0, 0};
      HIPC(hipMemcpy(ctl.p, h, sizeof(h), hipMemcpyHostToDevice));
    }
    choose_tiles(cb, cg, lds_rows, lds_cols);
    const int ngr = (C + cb - 1) / cb;
    wpart.ensure(size_t(ngr) * H * W * ce);
    if (!rows.p) {
      std::vector<int2> r(H), rg;
      for (int y = 0; y < H; ++y) r[y] = make_int2(0, y);
      for (int y = 0; y < H; ++y)
        for (int g = 0; g < ngr; ++g) rg.push_back(make_int2(0, y * ngr + g));
      rows.alloc(r.size() * sizeof(int2));
      HIPC(hipMemcpy(rows.p, r.data(), r.size() * sizeof(int2), hipMemcpyHostToDevice));
      nrows = H;
      rowgroups.alloc(rg.size() * sizeof(int2));
      HIPC(hipMemcpy(rowgroups.p, rg.data(), rg.size() * sizeof(int2), hipMemcpyHostToDevice));
      nrg = (int)rg.size();
      const int ng = (ncols + cg - 1) / cg;
      std::vector<int2> c;
      for (int cc = 0; cc < C; ++cc)
        for (int g = 0; g < ng; ++g) c.push_back(make_int2(0, cc * ng + g));
      colitems.alloc(std::max<size_t>(1, c.size()) * sizeof(int2));
      if (!c.empty()) HIPC(hipMemcpy(colitems.p, c.data(), c.size() * sizeof(int2), hipMemcpyHostToDevice));
      ncolitems = (int)c.size();
      slice.alloc(sizeof(SliceDev));
    }
  }

  SliceDev geometry() const {
    SliceDev d{};
    d.W = W;
    d.H = H;
    d.Hk = Hk;
    d.y0 = y0;
    d.C = C;
    d.ncols = ncols;
    d.S = S_dev();
    d.Mc = Mc.p;
    d.cols = cols.as<int>();
    d.colmap = colmap.as<int>();
    d.planW = planW;
    d.planH = planH;
    return d;
  }

  // Applies A (forward) or A^* (adjoint) from device src to device dst.
  void apply(bool forward, const void* src, void* dst, hipStream_t st) {
    SetupLock lk(g_setup_mu);
    ensure_opmode();
    SliceDev d = geometry();
    d.ngroups = (ncols + cg - 1) / cg;
    d.ngr = (C + cb - 1) / cb;
    d.ngr_a = d.ngr;
    d.T1 = T1.p;
    d.T2 = T2.p;
    d.wpart = wpart.p;
    d.op_in = src;
    d.op_out = dst;
    d.ctl = ctl.as<int>();
    if (!h_desc) HIPC(hipHostMalloc((void**)&h_desc, sizeof(SliceDev), hipHostMallocDefault));
    if (done) HIPC(hipEventSynchronize(done));  // the previous application (any stream) has finished
    else HIPC(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    *h_desc = d;
    HIPC(hipMemcpyAsync(slice.p, h_desc, sizeof(d), hipMemcpyHostToDevice, st));
    Launch L;
    L.stream = st;
    L.sl = slice.as<SliceDev>();
    L.nsl = 1;
    L.rows = rows.as<int2>();
    L.nrows = nrows;
    L.rowgroups = rowgroups.as<int2>();
    L.nrg = nrg;
    L.cols = colitems.as<int2>();
    L.ncolitems = ncolitems;
    L.cb = cb;
    L.cg = cg;
    L.lds_rows = lds_rows;
    L.lds_cols = lds_cols;
    L.fft_w = W;
    L.fft_h = Hk;
    L.split_primal = d.ngr_a > 1;
    L.primal_threads = primal_block(W);
    L.rowgroups_a = L.rowgroups;
    L.nrg_a = L.nrg;
    L.cb_a = L.cb;
    L.lds_rows_a = L.lds_rows;
    L.adj_threads = kRowThreads;
    const int dt = dtype == PDT_C128 ? DT_C128 : DT_C64;
    if (forward) {
      run_zero(st, dst, size_t(W) * Hk * C * esize(cplx_of(dtype)));
      if (ncols) run_op_forward(L, dt);
    } else {
      if (ncols) run_op_adjoint(L, dt);
      else run_zero(st, dst, size_t(W) * H * esize(cplx_of(dtype)));
    }
    HIPC(hipGetLastError());
    HIPC(hipEventRecord(done, st));
  }

  // apply() without any device allocation: the intermediates (descriptor,
  // control words, T1, T2, per-group partial adjoints) live in the per-GPU
  // scratch array, the work lists are the identity (null lists), and the
  // descriptor is written by a kernel from its by-value argument.  Needs the
  // prepared geometry (prepare() allocates) and enqueues on st without
  // waiting.  Calls that overlap on one GPU share the array and race.
  void apply_noalloc(bool forward, const void* src, void* dst, hipStream_t st) {
    int cbn = 1, cgn = 1;
    size_t lr = 0, lc = 0;
    choose_tiles(cbn, cgn, lr, lc);
    const size_t ce = esize(cplx_of(dtype));
    const int ngr = (C + cbn - 1) / cbn, ng = (ncols + cgn - 1) / cgn;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_ctl = al(sizeof(SliceDev));
    const size_t o_t1 = o_ctl + al(CTL_WORDS * sizeof(int));
    const size_t tb = al(size_t(C) * H * ncols * ce);
    const size_t o_t2 = o_t1 + tb, o_wp = o_t2 + tb;
    const size_t need = o_wp + (ngr > 1 ? al(size_t(ngr) * H * W * ce) : 0);
    if (need > kNoAllocScratchBytes) {
      std::ostringstream ss;
      ss << prefix() << "the allocation-free application needs " << need << " bytes of scratch, above its maximum of "
         << kNoAllocScratchBytes;
      fail(PDT_ERR_SIZE, ss.str());
    }
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    char* base = static_cast<char*>(noalloc_scratch(dev));
    if (!base) fail(PDT_ERR_HIP, "the scratch array of the allocation-free application is not available on this device");
    SliceDev d = geometry();
    d.ngroups = ng;
    d.ngr = ngr;
    d.ngr_a = ngr;
    d.T1 = base + o_t1;
    d.T2 = base + o_t2;
    d.wpart = ngr > 1 ? base + o_wp : nullptr;
    d.op_in = src;
    d.op_out = dst;
    d.ctl = reinterpret_cast<int*>(base + o_ctl);
    SliceDev* d_dev = reinterpret_cast<SliceDev*>(base);
    run_store_desc(st, d, d_dev, d.ctl);
    Launch L;
    L.stream = st;
    L.sl = d_dev;
    L.nsl = 1;
    L.rows = nullptr;  // identity lists: rows y, (row, coil group), (coil, column group)
    L.nrows = H;
    L.rowgroups = nullptr;
    L.nrg = H * ngr;
    L.cols = nullptr;
    L.ncolitems = C * ng;
    L.cb = cbn;
    L.cg = cgn;
    L.lds_rows = lr;
    L.lds_cols = lc;
    L.fft_w = W;
    L.fft_h = Hk;
    L.split_primal = ngr > 1;
    L.primal_threads = primal_block(W);
    L.rowgroups_a = nullptr;
    L.nrg_a = L.nrg;
    L.cb_a = cbn;
    L.lds_rows_a = lr;
    L.adj_threads = kRowThreads;
    const int dt = dtype == PDT_C128 ? DT_C128 : DT_C64;
    if (forward) {
      run_zero(st, dst, size_t(W) * Hk * C * ce);
      if (ncols) run_op_forward(L, dt);
    } else {
      if (ncols) run_op_adjoint(L, dt);
      else run_zero(st, dst, size_t(W) * H * ce);
    }
    HIPC(hipGetLastError());
  }

  ~pdt_op() {
    if (done) {
      (void)hipEventSynchronize(done);
      (void)hipEventDestroy(done);
    }
    if (h_desc) (void)hipHostFree(h_desc);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
void op_release(pdt_op* op) {
  if (op && --op->refs == 0) delete op;
}
}  // namespace

// ---------------------------------------------------------------------------
// solvers
// ---------------------------------------------------------------------------
struct pdt_solver {
  int kind = PDT_SOLVER_TV;
  int dtype = PDT_F32;
  pdt_params prm{};
  pdt_op* op = nullptr;
  // data
  std::unique_ptr<pdt_array> f;  // copy of the noisy data
  bool data_dirty = true;         // compact data / r must be rebuilt
  // state
  int W = 0, H = 0;
  bool have_u = false;
  std::unique_ptr<pdt_array> u, v;  // result handles (u: {W,H}, v: {W,H,2})
  DevBuf ub, vb, p, q, tmp;
  // second copy of the state for the temporally blocked denoising launches:
  // u, ubar, p (ROF and TGV) and v, vbar, q (TGV)
  DevBuf u2, ub2, p2, v2, vb2, q2;
  DevBuf fk, rk, T1, T2, wpart;
  // set-up buffers kept across solves: descriptor and item list of the data
  // compaction, its partial sums, A^*f, and two reduction results
  DevBuf prep_sl, prep_items, prep_part, atf, red;
  SliceDev prep_d{};  // host source of prep_sl (alive until its copy completes)
  int prep_C = 0, prep_Hk = 0;
  int Hk = 0, C = 0, ncols = 0;
  double F0 = 0, atf_norm = 0;
  int rupd = 0;  // r updates since set_noisy_data
  pdt_stats stats{};
  hipStream_t stream = nullptr;
  std::shared_ptr<void> engine;  // cached single-solver engine (work lists, graphs)

  bool is_op() const { return kind == PDT_SOLVER_TGV_MRI || kind == PDT_SOLVER_TV_OP || kind == PDT_SOLVER_TGV_OP; }
  bool is_tgv() const { return kind == PDT_SOLVER_TGV || kind == PDT_SOLVER_TGV_MRI || kind == PDT_SOLVER_TGV_OP; }
  size_t pe() const { return esize(dtype); }
  size_t pq_e() const { return (prm.dual_fp16 && kind == PDT_SOLVER_TGV && !is_dbl(dtype)) ? pe() / 2 : pe(); }

  ~pdt_solver() {
    engine.reset();
    op_release(op);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

void default_params(int kind, pdt_params* p) {
  std::memset(p, 0, sizeof(*p));
  // TvParameters / TgvParameters / TgvMriParameters constructors
  // (tvoptimizer.cu:26-29, tgvoptimizer.cu:27-30, tgvmrioptimizer.cu:28-31)
  p->lambda = 1;
  p->alpha0 = 1;
  p->alpha1 = 1;
  p->reduction = 1;
  p->max_iter = 1;
  p->check = 100;
  p->gap_tol = 0;
  p->gap_every = 0;
  p->dual_fp16 = 0;
  p->use_graph = 1;
  p->radius_u = 0;
  p->radius_v = 0;
  (void)kind;
}

// State buffers of a W x H solve.  They are kept while the shape is unchanged:
// a new input (setInput0) then overwrites them in place, so the cached
// engine's descriptors and captured graphs, which hold these addresses, stay
// valid, and no memory that a graph refers to is freed and reallocated.
void alloc_state(pdt_solver* s, int W, int H) {
  const size_t n = size_t(W) * H, e = s->pe();
  auto exact = [](DevBuf& b, size_t bytes) {
    if (b.bytes != bytes) b.alloc(bytes);
  };
  if (!s->u || s->u->size[0] != W || s->u->size[1] != H) {
    const int64_t us[2] = {W, H};
    s->u.reset(new_array(s->dtype, 2, us));
  }
  exact(s->ub, n * e);
  exact(s->p, 2 * n * s->pq_e());
  if (s->is_tgv()) {
    if (!s->v || s->v->size[0] != W || s->v->size[1] != H) {
      const int64_t vs[3] = {W, H, 2};
      s->v.reset(new_array(s->dtype, 3, vs));
    }
    exact(s->vb, 2 * n * e);
    exact(s->q, 3 * n * s->pq_e());
    if (s->kind == PDT_SOLVER_TGV) exact(s->tmp, 2 * n * e);
  }
  if (s->kind == PDT_SOLVER_TV || s->kind == PDT_SOLVER_TGV) {
    exact(s->u2, n * e);
    exact(s->ub2, n * e);
    exact(s->p2, 2 * n * s->pq_e());
  }
  if (s->kind == PDT_SOLVER_TGV) {
    exact(s->v2, 2 * n * e);
    exact(s->vb2, 2 * n * e);
    exact(s->q2, 3 * n * s->pq_e());
  }
  s->W = W;
  s->H = H;
}

// setInput0 (tgvmrioptimizer.cu:63-79): u = u_bar = input, v = p = q = 0.
void set_input0(pdt_solver* s, const void* u0_dev, const void* v0_dev, int W, int H) {
  if (!s->stream) HIPC(hipStreamCreate(&s->stream));
  alloc_state(s, W, H);
  const size_t n = size_t(W) * H, e = s->pe();
  HIPC(hipMemcpyAsync(s->u->ptr, u0_dev, n * e, hipMemcpyDeviceToDevice, s->stream));
  HIPC(hipMemcpyAsync(s->ub.p, u0_dev, n * e, hipMemcpyDeviceToDevice, s->stream));
  run_zero(s->stream, s->p.p, s->p.bytes);
  if (s->is_tgv()) {
    run_zero(s->stream, s->q.p, s->q.bytes);
    if (v0_dev) {
      HIPC(hipMemcpyAsync(s->v->ptr, v0_dev, 2 * n * e, hipMemcpyDeviceToDevice, s->stream));
      HIPC(hipMemcpyAsync(s->vb.p, v0_dev, 2 * n * e, hipMemcpyDeviceToDevice, s->stream));
    } else {
      run_zero(s->stream, s->v->ptr, 2 * n * e);
      run_zero(s->stream, s->vb.p, 2 * n * e);
    }
  }
  HIPC(hipStreamSynchronize(s->stream));
  s->have_u = true;
}

// Builds the compact data layout and r = 0 for an operator solver, and the
// quantities of the partial gap (||A^* f||, ||f||^2 off the sampled lines).
void prepare_op_data(pdt_solver* s) {
  pdt_op* op = s->op;
  op->prepare();
  const int64_t want[3] = {op->W, op->Hk, op->C};
  const int fnd = op->kind == PDT_OP_SAMPLING ? 2 : 3;
  size_check(s->f->size, s->f->ndim, want, fnd);
  if (!s->data_dirty && s->Hk == op->Hk && s->C == op->C && s->ncols == op->ncols) return;
  if (!s->stream) HIPC(hipStreamCreate(&s->stream));
  const size_t ce = s->pe();
  s->Hk = op->Hk;
  s->C = op->C;
  s->ncols = op->ncols;
  // buffers are reused while large enough: no device allocation (and no
  // synchronising free) per solve, and stable addresses for the cached engine
  s->fk.ensure(std::max<size_t>(1, size_t(op->C) * op->ncols * op->Hk * ce));
  s->rk.ensure(std::max<size_t>(1, size_t(op->C) * op->ncols * op->Hk * ce));
  s->T1.ensure(std::max<size_t>(1, size_t(op->C) * op->H * op->ncols * ce));
  s->T2.ensure(std::max<size_t>(1, size_t(op->C) * op->H * op->ncols * ce));
  // compact gather + F0
  SliceDev& d = s->prep_d;
  d = op->geometry();
  d.fk = s->fk.p;
  d.rk = s->rk.p;
  s->prep_sl.ensure(sizeof(SliceDev));
  HIPC(hipMemcpyAsync(s->prep_sl.p, &d, sizeof(d), hipMemcpyHostToDevice, s->stream));
  const int nitems = op->C * op->Hk;
  if (s->prep_C != op->C || s->prep_Hk != op->Hk) {
    std::vector<int2> items;
    for (int c = 0; c < op->C; ++c)
      for (int kfe = 0; kfe < op->Hk; ++kfe) items.push_back(make_int2(0, c * op->Hk + kfe));
    s->prep_items.ensure(items.size() * sizeof(int2));
    HIPC(hipMemcpyAsync(s->prep_items.p, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, s->stream));
    HIPC(hipStreamSynchronize(s->stream));  // the host vector goes out of scope
    s->prep_C = op->C;
    s->prep_Hk = op->Hk;
  }
  s->prep_part.ensure(size_t(nitems) * kGapTerms * sizeof(double));
  s->red.ensure(2 * sizeof(double));
  Launch L;
  L.stream = s->stream;
  L.sl = s->prep_sl.as<SliceDev>();
  L.nsl = 1;
  L.rows = s->prep_items.as<int2>();
  L.nrows = nitems;
  const int dt = is_dbl(s->dtype) ? DT_C128 : DT_C64;
  run_compact_data(L, dt, s->f->ptr, s->prep_part.as<double>());
  run_sum_strided(s->stream, s->prep_part.as<double>(), nitems, kGapTerms, s->red.as<double>());
  // ||A^* f|| (radius of the partial-gap ball); A^* f is kept for the cold start
  const size_t n = size_t(op->W) * op->H;
  s->atf.ensure(n * ce);
  op->apply(false, s->f->ptr, s->atf.p, s->stream);
  run_sumsq(s->stream, dt, s->atf.p, int64_t(n), s->red.as<double>() + 1);
  HIPC(hipGetLastError());
  double h[2];
  HIPC(hipMemcpyAsync(h, s->red.p, sizeof(h), hipMemcpyDeviceToHost, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  s->F0 = h[0];
  s->atf_norm = std::sqrt(h[1]);
  s->rupd = 0;
  s->data_dirty = false;
}

// ---------------------------------------------------------------------------
// The engine: one or many slices, persistent work lists, graph-replayed chunks
// ---------------------------------------------------------------------------
struct Engine {
  int kind = 0, dtype = 0;
  bool fp16 = false;
  bool pair_duals = true;  // TGV denoising: the dual step as 2-pixel vectors (even widths)
  int force_dpx = 0;       // dual step pixels per thread (0: cache / default)
  int force_cb = 0, force_cg = 0, force_rpi = 0, force_cb_a = 0;  // autotuner overrides (0: cache / default; cb_a -1: default)
  int force_wpc[4] = {-1, -1, -1, -1};  // occupancy caps (-1: cache; 0: none)
  int force_tb[3] = {0, 0, -1};         // blocked ROF tiles (tb[2] -1: cache / default; 0: off)
  int force_unpaired = -1;              // MRI: 1 three launches per iteration, 0 pairs where possible (-1: cache)
  bool unpaired = false;
  bool plan_only = false;                         // build(): stop after the signature
  std::string sig;
  TuneKey key{};
  std::vector<pdt_solver*> sv;
  hipStream_t stream = nullptr;
  DevBuf d_slices, d_rows, d_rg, d_rg_a, d_cols, d_part;
  // statistics (ns x ST_WORDS doubles) then control words (ns x CTL_WORDS
  // ints) in one device block, so one copy brings both to the host
  DevBuf d_cs;
  double* d_stats = nullptr;
  int* d_ctl = nullptr;
  size_t cs_bytes = 0;
  char* h_cs = nullptr;  // pinned mirror of d_cs
  int* h_ctl = nullptr;
  double* h_stats = nullptr;
  // single-slice pipelining: two pinned copies of the block (one per chunk
  // in flight) and their completion events
  char* h_cs_q[2] = {nullptr, nullptr};
  hipEvent_t ev_q[2] = {nullptr, nullptr};
  int nrows = 0, nrg = 0, nrg_a = 0, ncolitems = 0;
  Launch L;                             // every slice: gap reduction, full-size buffers
  std::vector<Launch> G;                // one launch per group of slices with equal FFT lengths
  std::vector<std::vector<int>> groups;  // slice indices of each group
  int listed_items = 0;                 // items in the current work lists
  std::vector<SliceDev> hs;
  struct GraphEntry {
    int len;
    bool gap;
    bool first;  // paired iterations: the chunk starts a solve (forward row pass first)
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;

  ~Engine() {
    SetupLock lk(g_setup_mu);
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    if (h_cs) (void)hipHostFree(h_cs);
    for (int q = 0; q < 2; ++q) {
      if (h_cs_q[q]) (void)hipHostFree(h_cs_q[q]);
      if (ev_q[q]) (void)hipEventDestroy(ev_q[q]);
    }
  }

  void build() {
    SetupLock lk(g_setup_mu);
    pdt_solver* s0 = sv[0];
    kind = s0->kind;
    dtype = s0->dtype;
    fp16 = s0->prm.dual_fp16 && kind == PDT_SOLVER_TGV && !is_dbl(dtype);
    stream = s0->stream;
    const int ns = (int)sv.size();
    const bool mri = s0->is_op();
    // tile choices (shared by the whole batch)
    int cb = 64, cg = 4, rpi = 1, cb_a = 64;
    size_t lr = 0, lc = 0, lra = 0;
    int maxW = 0, maxHk = 0;
    for (auto* s : sv) {
      maxW = std::max(maxW, s->W);
      if (mri) maxHk = std::max(maxHk, s->op->Hk);
    }
    const int dt = is_dbl(dtype) ? (is_cplx(dtype) ? DT_C128 : DT_F64) : (is_cplx(dtype) ? DT_C64 : DT_F32);
    if (mri) {
      init_kernel_attrs();
      int maxC = 1, maxNc = 1;
      for (auto* s : sv) {
        maxC = std::max(maxC, s->op->C);
        maxNc = std::max(maxNc, s->op->ncols);
      }
      key = TuneKey{kind, dt, maxW, sv[0]->H, maxHk, maxC, maxNc, ns};
      TuneVal tv = tune_lookup(key);
      for (int i = 0; i < 4; ++i) L.wpc[i] = force_wpc[i] >= 0 ? force_wpc[i] : tv.wpc[i];
      unpaired = force_unpaired >= 0 ? force_unpaired != 0 : tv.unpaired != 0;
      cb = force_cb > 0 ? force_cb : tv.cb > 0 ? tv.cb : kDefaultCb;
      cg = force_cg > 0 ? force_cg : tv.cg > 0 ? tv.cg : kDefaultCg;
      // default adjoint grouping: every coil of a row in one wide block where
      // a single slice's straight-line adjoint path fits (with the u and v
      // steps in that pass and, by default, paired iterations), else cb
      const bool fused_fits = ns == 1 && sv[0]->op->kind != PDT_OP_SAMPLING &&
                              maxC * maxW >= 8 * kRowThreads && rows_adj_fused_ok(dt, maxW, maxC, maxNc);
      const int def_cb_a = fused_fits ? maxC : cb;
      cb_a = force_cb_a > 0 ? force_cb_a : force_cb_a < 0 ? def_cb_a : tv.cb_a > 0 ? tv.cb_a : def_cb_a;
      cb = std::max(1, std::min(cb, max_cb(dt, maxW, maxNc, maxC)));
      cb_a = std::max(1, std::min(cb_a, max_cb(dt, maxW, maxNc, maxC)));
      cg = std::max(1, std::min(cg, max_cg(dt, maxHk, maxNc)));
      lr = lds_rows_bytes(dt, cb, maxW, maxNc);
      lra = lds_rows_bytes(dt, cb_a, maxW, maxNc);
      lc = lds_cols_bytes(dt, cg, maxHk);
    } else {
      key = TuneKey{kind, dt, maxW, sv[0]->H, 0, 0, 0, ns};
      TuneVal tv = tune_lookup(key);
      for (int i = 0; i < 4; ++i) L.wpc[i] = force_wpc[i] >= 0 ? force_wpc[i] : tv.wpc[i];
      rpi = force_rpi > 0 ? force_rpi : tv.rpi > 0 ? tv.rpi : 1;
      // default: pairs for single precision (8-16 B accesses); the 32-byte
      // double pairs were slower than one pixel per thread (512^2 complex
      // double: 12.2 against 9.5 us per dual step)
      const int dpx = force_dpx > 0 ? force_dpx : tv.dual_px > 0 ? tv.dual_px : is_dbl(dtype) ? 1 : 2;
      pair_duals = dpx == 2;
      // temporally blocked launches: one slice only, tiles that fit the LDS
      const int* tb = force_tb[2] >= 0 ? force_tb
                      : tv.tb[2] >= 0  ? tv.tb
                      : kind == PDT_SOLVER_TGV ? kDefaultTgvBlock
                                               : kDefaultTvBlock;
      if ((kind == PDT_SOLVER_TV || kind == PDT_SOLVER_TGV) && ns == 1 &&
          block_fits(kind, dt, tb[0], tb[1], tb[2], init_kernel_attrs())) {
        L.tb_x = tb[0];
        L.tb_y = tb[1];
        L.tb_k = tb[2];
      }
    }
    // slices, grouped by FFT lengths, and their work lists
    hs.assign(ns, SliceDev{});
    for (int i = 0; i < ns; ++i) {
      pdt_solver* s = sv[i];
      SliceDev d = mri ? s->op->geometry() : SliceDev{};
      d.W = s->W;
      d.H = s->H;
      d.kind = s->kind;
      d.dual_fp16 = fp16;
      d.theta = 1.0;
      d.alpha0 = s->prm.alpha0;
      d.alpha1 = s->prm.alpha1;
      d.alpha0_end = s->prm.alpha0 * s->prm.reduction;
      d.alpha1_end = s->prm.alpha1 * s->prm.reduction;
      d.kmax = double(s->prm.max_iter);
      d.schedule = s->kind == PDT_SOLVER_TGV_MRI;
      d.gap_tol = s->prm.gap_every ? s->prm.gap_tol : 0.0;
      switch (s->kind) {
        case PDT_SOLVER_TV:  // tvoptimizer.cu:98-100
          d.tau = d.sigma = 1.0 / std::sqrt(8.0);
          d.lambda = s->prm.lambda;
          break;
        case PDT_SOLVER_TGV:  // tgvoptimizer.cu:102-104
          d.tau = d.sigma = 1.0 / std::sqrt(12.0);
          d.lambda = s->prm.lambda;
          break;
        case PDT_SOLVER_TGV_MRI:  // tgvmrioptimizer.cu:132-133, lambda = 1 (:174)
          d.tau = 1.0 / 16.0;
          d.sigma = 1.0 / 8.0;
          d.lambda = 1.0;
          break;
        default:  // with-operator solvers (tgvoptimizerwithop.cu:123-124)
          d.tau = d.sigma = 1.0 / 8.0;
          d.lambda = s->prm.lambda;
          break;
      }
      if (s->kind == PDT_SOLVER_TV_OP) {
        d.alpha0 = d.alpha1 = 1.0;
        d.alpha0_end = d.alpha1_end = 1.0;
      }
      d.radius_u = s->prm.radius_u > 0 ? s->prm.radius_u : 0.01 * s->atf_norm;
      d.radius_v = s->prm.radius_v > 0 ? s->prm.radius_v : std::sqrt(8.0) * d.radius_u;
      d.F0 = s->F0;
      d.u = s->u->ptr;
      d.ub = s->ub.p;
      d.p = s->p.p;
      if (s->is_tgv()) {
        d.v = s->v->ptr;
        d.vb = s->vb.p;
        d.q = s->q.p;
        d.tmp = s->tmp.p;
      }
      d.u2 = s->u2.p;
      d.ub2 = s->ub2.p;
      d.p2 = s->p2.p;
      d.v2 = s->v2.p;
      d.vb2 = s->vb2.p;
      d.q2 = s->q2.p;
      d.f = s->f->ptr;
      if (mri) {
        d.fk = s->fk.p;
        d.rk = s->rk.p;
        d.T1 = s->T1.p;
        d.T2 = s->T2.p;
        d.ngroups = (d.ncols + cg - 1) / cg;
        d.ngr = (d.C + cb - 1) / cb;
        d.ngr_a = (d.C + cb_a - 1) / cb_a;
        s->wpart.ensure(size_t(d.ngr_a) * d.W * d.H * s->pe());
        d.wpart = s->wpart.p;
      }
      hs[i] = d;
    }
    groups.clear();
    if (mri) {
      std::map<std::pair<int, int>, size_t> gi;
      for (int i = 0; i < ns; ++i) {
        const auto k = std::make_pair(hs[i].W, hs[i].Hk);
        auto it = gi.find(k);
        if (it == gi.end()) {
          it = gi.emplace(k, groups.size()).first;
          groups.emplace_back();
        }
        groups[it->second].push_back(i);
      }
    } else {
      groups.emplace_back();
      for (int i = 0; i < ns; ++i) groups[0].push_back(i);
    }
    L.rpi = rpi;
    L.cb = cb;
    L.cg = cg;
    L.cb_a = cb_a;
    std::vector<int2> rows, rgs, rgs_a, cols;
    std::vector<GroupRange> gr;
    make_lists(std::vector<char>(ns, 1), rows, rgs, rgs_a, cols, gr);
    nrows = (int)rows.size();
    nrg = (int)rgs.size();
    nrg_a = (int)rgs_a.size();
    ncolitems = (int)cols.size();
    // signature of everything the kernels and graphs depend on (the control
    // and statistics words are owned by the engine and added below)
    sig.assign(reinterpret_cast<const char*>(hs.data()), hs.size() * sizeof(SliceDev));
    const int tiles[14] = {cb,       cg,       rpi,      int(sv[0]->prm.gap_every), cb_a,   L.wpc[0],
                           L.wpc[1], L.wpc[2], L.wpc[3], L.tb_x,                    L.tb_y, L.tb_k,
                           int(pair_duals), int(unpaired)};
    sig.append(reinterpret_cast<const char*>(tiles), sizeof(tiles));
    if (plan_only) return;
    cs_bytes = size_t(ns) * (ST_WORDS * sizeof(double) + CTL_WORDS * sizeof(int));
    d_cs.alloc(cs_bytes);
    d_stats = d_cs.as<double>();
    d_ctl = reinterpret_cast<int*>(d_stats + size_t(ns) * ST_WORDS);
    HIPC(hipHostMalloc((void**)&h_cs, cs_bytes, hipHostMallocDefault));
    h_stats = reinterpret_cast<double*>(h_cs);
    h_ctl = reinterpret_cast<int*>(h_stats + size_t(ns) * ST_WORDS);
    for (int i = 0; i < ns; ++i) {
      hs[i].ctl = d_ctl + size_t(i) * CTL_WORDS;
      hs[i].stats = d_stats + size_t(i) * ST_WORDS;
    }
    d_slices.alloc(hs.size() * sizeof(SliceDev));
    d_rows.alloc(std::max<size_t>(1, rows.size()) * sizeof(int2));
    d_rg.alloc(std::max<size_t>(1, rgs.size()) * sizeof(int2));
    d_rg_a.alloc(std::max<size_t>(1, rgs_a.size()) * sizeof(int2));
    d_cols.alloc(std::max<size_t>(1, cols.size()) * sizeof(int2));
    // (a single slice's dual-step workgroups write after its row items)
    const size_t rg_items = rgs.size() + (mri && ns == 1 ? size_t(dual_groups(hs[0])) : 0);
    // (and its v-step workgroups after its rows)
    const size_t row_items = rows.size() + (mri && ns == 1 ? size_t(vstep_groups(hs[0])) : 0);
    const size_t nitem = std::max<size_t>(1, std::max({row_items, rg_items, rgs_a.size(), cols.size()}));
    d_part.alloc(4 * nitem * kGapTerms * sizeof(double));
    // (a paired gap iteration leaves the forward row items' partial sums,
    // zero in a single-slice solve, unwritten: they start at zero)
    HIPC(hipMemset(d_part.p, 0, 4 * nitem * kGapTerms * sizeof(double)));
    L.stream = stream;
    L.sl = d_slices.as<SliceDev>();
    L.nsl = ns;
    L.rows = d_rows.as<int2>();
    L.nrows = nrows;
    L.rowgroups = d_rg.as<int2>();
    L.nrg = nrg;
    L.rowgroups_a = d_rg_a.as<int2>();
    L.nrg_a = nrg_a;
    L.cols = d_cols.as<int2>();
    L.ncolitems = ncolitems;
    for (int k = 0; k < 4; ++k) L.part[k] = d_part.as<double>() + size_t(k) * nitem * kGapTerms;
    L.gap_rows = L.part[2];
    L.lds_rows = lr;
    L.lds_rows_a = lra;
    L.lds_cols = lc;
    upload_lists(rows, rgs, rgs_a, cols, gr);
  }

  // Item ranges of one slice group in the work lists.
  struct GroupRange {
    int g;  // index into groups
    int row0, nrow, rg0, nrg, rga0, nrga, col0, ncol;
  };

  // Work lists of the slices with use[i] != 0, group after group: row items
  // (slice, first row), row-group items of the forward (slice, row * ngr +
  // coil group) and adjoint (slice, row * ngr_a + coil group) row passes and
  // column items (slice, coil * ngroups + column group).  Sets the item
  // ranges of hs; slices left out get empty ranges.
  void make_lists(const std::vector<char>& use, std::vector<int2>& rows, std::vector<int2>& rgs,
                  std::vector<int2>& rgs_a, std::vector<int2>& cols, std::vector<GroupRange>& gr) {
    const bool mri = sv[0]->is_op();
    const int step = mri ? 1 : L.rpi;
    for (size_t g = 0; g < groups.size(); ++g) {
      GroupRange r{int(g), (int)rows.size(), 0, (int)rgs.size(), 0, (int)rgs_a.size(), 0, (int)cols.size(), 0};
      for (int i : groups[g]) {
        SliceDev& d = hs[i];
        d.row_beg = d.row_end = d.rg_beg = d.rg_end = d.col_beg = d.col_end = 0;
        if (!use[i]) continue;
        d.row_beg = (int)rows.size();
        for (int y = 0; y < d.H; y += step) rows.push_back(make_int2(i, y));
        d.row_end = (int)rows.size();
        d.rg_beg = (int)rgs.size();
        if (mri)
          for (int y = 0; y < d.H; ++y)
            for (int q = 0; q < d.ngr; ++q) rgs.push_back(make_int2(i, y * d.ngr + q));
        d.rg_end = (int)rgs.size();
        if (mri)
          for (int y = 0; y < d.H; ++y)
            for (int q = 0; q < d.ngr_a; ++q) rgs_a.push_back(make_int2(i, y * d.ngr_a + q));
        d.col_beg = (int)cols.size();
        if (mri)
          for (int c = 0; c < d.C; ++c)
            for (int q = 0; q < d.ngroups; ++q) cols.push_back(make_int2(i, c * d.ngroups + q));
        d.col_end = (int)cols.size();
      }
      r.nrow = (int)rows.size() - r.row0;
      r.nrg = (int)rgs.size() - r.rg0;
      r.nrga = (int)rgs_a.size() - r.rga0;
      r.ncol = (int)cols.size() - r.col0;
      if (r.nrow > 0) gr.push_back(r);
    }
  }

  // Copies slices and lists to the device (the stream is idle) and derives
  // one launch per non-empty group: its sub-ranges of the lists and of the
  // partial-sum buffers, its FFT lengths and LDS sizes.
  // Workgroups of k_rows_fwd that take a single slice's dual step off its
  // row items: two pixels per thread.
  static int dual_groups(const SliceDev& d) { return (d.W * d.H + 2 * kRowThreads - 1) / (2 * kRowThreads); }
  // Workgroups of k_rows_adj that take a single TGV slice's v step off its
  // wide row items: one pixel per thread.
  static int vstep_groups(const SliceDev& d) { return (d.W * d.H + kRowAdjThreadsWide - 1) / kRowAdjThreadsWide; }

  void upload_lists(const std::vector<int2>& rows, const std::vector<int2>& rgs, const std::vector<int2>& rgs_a,
                    const std::vector<int2>& cols, const std::vector<GroupRange>& gr) {
    auto identity = [](const std::vector<int2>& v, int step) {
      for (size_t i = 0; i < v.size(); ++i)
        if (v[i].x != 0 || v[i].y != int(i) * step) return false;
      return true;
    };
    const bool one = sv.size() == 1;
    const bool dense = one && (sv[0]->is_op() ? identity(rgs, 1) && identity(rgs_a, 1) && identity(cols, 1)
                                              : identity(rows, L.rpi));
    // a dense single MRI slice hands its dual step to separate workgroups,
    // whose gap terms follow the row items'
    const int ndual = dense && sv[0]->is_op() ? dual_groups(hs[0]) : 0;
    if (one && sv[0]->is_op()) hs[0].rg_end = hs[0].rg_beg + (int)rgs.size() + ndual;
    // and its v step, where the wide adjoint blocks would otherwise do it
    // before their FFT: the same conditions as the kernel's (1024-thread
    // adjoint blocks, a compile-time FFT length, two threads per pixel of a
    // row, one adjoint coil group)
    const bool tgv = kind == PDT_SOLVER_TGV_MRI || kind == PDT_SOLVER_TGV_OP;
    const int nvstep = dense && sv[0]->is_op() && tgv && L.cb_a * hs[0].W >= 8 * kRowThreads &&
                               fft_len_compiled(dtk(), hs[0].W) && 2 * hs[0].W <= kRowAdjThreadsWide &&
                               hs[0].ngr_a == 1
                           ? vstep_groups(hs[0])
                           : 0;
    if (one && sv[0]->is_op()) hs[0].row_end = hs[0].row_beg + (int)rows.size() + nvstep;
    HIPC(hipMemcpy(d_slices.p, hs.data(), hs.size() * sizeof(SliceDev), hipMemcpyHostToDevice));
    if (!rows.empty()) HIPC(hipMemcpy(d_rows.p, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!rgs.empty()) HIPC(hipMemcpy(d_rg.p, rgs.data(), rgs.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!rgs_a.empty()) HIPC(hipMemcpy(d_rg_a.p, rgs_a.data(), rgs_a.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!cols.empty()) HIPC(hipMemcpy(d_cols.p, cols.data(), cols.size() * sizeof(int2), hipMemcpyHostToDevice));
    const bool mri = sv[0]->is_op();
    const int dt = dtk();
    G.clear();
    listed_items = 0;
    for (const GroupRange& r : gr) {
      Launch x = L;
      x.rows = L.rows + r.row0;
      x.nrows = r.nrow;
      x.rowgroups = L.rowgroups + r.rg0;
      x.nrg = r.nrg;
      x.rowgroups_a = L.rowgroups_a + r.rga0;
      x.nrg_a = r.nrga;
      x.cols = L.cols + r.col0;
      x.ncolitems = r.ncol;
      if (mri) {
        // rows_fwd writes per row-group item, cols per column item, primal per row item
        x.part[0] = L.part[0] + size_t(r.rg0) * kGapTerms;
        x.part[1] = L.part[1] + size_t(r.col0) * kGapTerms;
        x.part[2] = L.part[2] + size_t(r.row0) * kGapTerms;
        const SliceDev& d0 = hs[groups[r.g][0]];
        int nc = 1;
        for (int i : groups[r.g]) nc = std::max(nc, hs[i].ncols);
        x.fft_w = d0.W;
        x.fft_h = d0.Hk;
        x.primal_threads = primal_block(d0.W);
        x.split_primal = false;
        for (int i : groups[r.g]) x.split_primal = x.split_primal || hs[i].ngr_a > 1;
        x.lds_rows = lds_rows_bytes(dt, x.cb, d0.W, nc);
        x.lds_cols = lds_cols_bytes(dt, x.cg, d0.Hk);
        // wide blocks for the adjoint pass when an item holds many coil rows
        x.adj_threads = x.cb_a * d0.W >= 8 * kRowThreads ? kRowAdjThreadsWide : kRowThreads;
        x.lds_rows_a = lds_rows_bytes(dt, x.cb_a, d0.W, nc, x.adj_threads != kRowAdjThreadsWide);
      } else {
        // one thread per pixel of an item where the workgroup size allows
        int maxW = 1;
        for (int i : groups[r.g]) maxW = std::max(maxW, hs[i].W);
        const int px = maxW * L.rpi;
        x.den_threads = px > 512 ? 1024 : px > 256 ? 512 : kBlock;
        // the TGV dual step as 2-pixel vectors where every width is even
        bool even = true;
        for (int i : groups[r.g]) even = even && hs[i].W % 2 == 0;
        x.den_pair = kind == PDT_SOLVER_TGV && even && pair_duals;
      }
      x.dense = dense;
      x.ndual = ndual;
      x.nvstep = nvstep;
      if (mri) {
        const SliceDev& d0 = hs[groups[r.g][0]];
        x.adj_fused = dense && x.adj_threads == kRowAdjThreadsWide && d0.ngr_a == 1 && d0.S != nullptr &&
                      rows_adj_fused_ok(dt, d0.W, d0.C, d0.ncols) && (!tgv || nvstep > 0);
        x.pair = x.adj_fused && ndual > 0 && dt == DT_C64 && !unpaired;
      }
      if (!(dense && (kind == PDT_SOLVER_TV || kind == PDT_SOLVER_TGV))) x.tb_x = x.tb_y = x.tb_k = 0;
      x.tb_w = hs[groups[r.g][0]].W;
      x.tb_h = hs[groups[r.g][0]].H;
      listed_items += r.nrow + r.nrg + r.nrga + r.ncol;
      G.push_back(x);
    }
  }

  // Drops the items of finished slices from the work lists once they make
  // up more than a sixth of them (batched solves whose slices stop at
  // different iterations); the captured graphs are rebuilt on next use.
  void compact_lists() {
    const int ns = (int)sv.size();
    if (ns < 2) return;
    std::vector<char> use(ns, 0);
    int items = 0;
    for (int i = 0; i < ns; ++i) {
      use[i] = h_ctl[size_t(i) * CTL_WORDS + CTL_ACTIVE] != 0;
      if (use[i]) {
        const SliceDev& d = hs[i];
        const int step = sv[0]->is_op() ? 1 : L.rpi;
        items += (d.H + step - 1) / step + (sv[0]->is_op() ? d.H * (d.ngr + d.ngr_a) + d.C * d.ngroups : 0);
      }
    }
    if (items == 0 || 6 * items > 5 * listed_items) return;
    SetupLock lk(g_setup_mu);
    std::vector<int2> rows, rgs, rgs_a, cols;
    std::vector<GroupRange> gr;
    make_lists(use, rows, rgs, rgs_a, cols, gr);
    upload_lists(rows, rgs, rgs_a, cols, gr);
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }

  // One stage (fused kernel) of an iteration over every group.
  void stage(int st, bool gap) {
    for (const Launch& x : G) run_stage(x, kind, dtk(), fp16, gap, st);
  }
  void iteration(bool gap) {
    for (int st = 0; st < num_stages(kind); ++st) stage(st, gap);
  }

  // Algorithmic HBM bytes of one launch of a stage, summed over the slices.
  double stage_bytes(int stage) const {
    double tot = 0;
    for (size_t i = 0; i < sv.size(); ++i) {
      const pdt_solver* s = sv[i];
      const double N = double(s->W) * s->H, pe = double(s->pe()), pq = double(s->pq_e());
      if (kind == PDT_SOLVER_TV) {
        tot += stage == 0 ? (pe + 4 * pq) * N : (2 * pe + pe + 2 * pq + pe) * N;
      } else if (kind == PDT_SOLVER_TGV) {
        tot += stage == 0 ? (10 * pe + 5 * pq) * N : (3 * pe + 10 * pq) * N;
      } else {
        const SliceDev& d = hs[i];
        const double C = d.C, nc = d.ncols, Hk = d.Hk, H = d.H;
        const double S = s->op->kind == PDT_OP_SAMPLING ? 0.0 : C * N * pe;  // sensitivities on the image window
        const double T = C * H * nc * pe;                                   // T1 or T2
        const bool tgv = kind != PDT_SOLVER_TV_OP;
        const double Wp = double(d.ngr_a) * N * pe;  // per-coil-group partial adjoints
        const double Pr = (tgv ? 14 * pe : 5 * pe) * N;  // primal step
        if (stage == 0) tot += (tgv ? 13 * pe : 5 * pe) * N + S + T;
        else if (stage == 1) tot += 2 * T + 3 * C * nc * Hk * pe + nc * Hk * (pe / 2);
        else if (stage == 2) tot += T + S + (d.ngr_a > 1 ? Wp : Pr);  // one coil group: primal step in the same kernel
        else tot += d.ngr_a > 1 ? Wp + Pr : 0.0;
      }
    }
    return tot;
  }

  int dtk() const {
    return is_dbl(dtype) ? (is_cplx(dtype) ? DT_C128 : DT_F64) : (is_cplx(dtype) ? DT_C64 : DT_F32);
  }

  void reset_ctl() {
    const int ns = (int)sv.size();
    for (int i = 0; i < ns; ++i) {
      int* c = h_ctl + size_t(i) * CTL_WORDS;
      std::fill(c, c + CTL_WORDS, 0);
      c[CTL_ACTIVE] = 1;
      c[CTL_ITER] = 0;
      c[CTL_RUPD] = sv[i]->rupd;
      c[CTL_DONE_AT] = -1;
    }
    std::fill(h_stats, h_stats + size_t(ns) * ST_WORDS, 0.0);
    HIPC(hipMemcpyAsync(d_cs.p, h_cs, cs_bytes, hipMemcpyHostToDevice, stream));
  }

  bool blocked() const { return G.size() == 1 && G[0].tb_k > 0; }
  bool paired() const { return G.size() == 1 && G[0].pair; }

  // The launches of one iteration without the gap evaluation, as timing and
  // profiling take them: the two of a paired iteration (the column pass with
  // the dual step, the adjoint row pass with the next forward row FFTs) or
  // the stages.
  int iter_launches() const { return paired() ? 2 : num_stages(kind); }
  // before timing the launches alone: the T1 the first column pass reads
  void iter_prologue() {
    if (paired()) stage(0, false);
  }
  void iter_launch(int st) {
    if (paired()) run_pair_stage(G[0], kind, dtk(), st, false);
    else stage(st, false);
  }
  std::string iter_launch_name(int st) const {
    if (paired()) return st == 0 ? "k_cols_dual" : "k_rows_adj_next";
    return stage_name(kind, st);
  }
  // algorithmic bytes: the column pass plus the dual step's planes, the
  // adjoint row pass plus the T1 it writes
  double iter_launch_bytes(int st) const {
    if (!paired()) return stage_bytes(st);
    const pdt_solver* s = sv[0];
    const SliceDev& d = hs[0];
    const double N = double(s->W) * s->H, pe = double(s->pe());
    const bool tgv = kind != PDT_SOLVER_TV_OP;
    return st == 0 ? stage_bytes(1) + (tgv ? 13 * pe : 5 * pe) * N : stage_bytes(2) + double(d.C) * d.H * d.ncols * pe;
  }

  // Enqueues `len` iterations; the last one evaluates the gap when asked.
  // first: the iterations start a solve (paired iterations take the T1 of
  // the first one from the forward row pass, later chunks from the last
  // adjoint row pass of the chunk before).
  // With temporally blocked ROF launches, the iterations before the gap
  // iteration run as pairs of launches of tb_k iterations each (the state
  // goes to the second copy and back), the rest one iteration at a time.
  void enqueue(int len, bool gap, bool first = true) {
    const int dt = dtk();
    if (blocked()) {
      const int k = G[0].tb_k;
      int plain = len - (gap ? 1 : 0);
      for (; plain >= 2 * k; plain -= 2 * k) {
        run_block(G[0], kind, dt, fp16, 0);
        run_block(G[0], kind, dt, fp16, 1);
      }
      for (int i = 0; i < plain; ++i) iteration(false);
      if (gap) {
        iteration(true);
        run_gap_reduce(L, kind, dt);
      }
      return;
    }
    if (paired()) {
      // every iteration in a pair of launches, the gap iteration included;
      // each adjoint row pass writes the T1 of the next iteration, so only
      // the first iteration of a solve takes its T1 from the forward row pass
      // (which also does its dual step)
      for (int i = 0; i < len; ++i) {
        const bool g = gap && i == len - 1;
        if (i == 0 && first) {
          stage(0, g);
          stage(1, g);
        } else {
          run_pair_stage(G[0], kind, dt, 0, g);
        }
        run_pair_stage(G[0], kind, dt, 1, g);
        if (g) run_gap_reduce(L, kind, dt);
      }
      return;
    }
    for (int i = 0; i < len; ++i) {
      const bool g = gap && i == len - 1;
      iteration(g);
      if (g) run_gap_reduce(L, kind, dt);
    }
  }

  // Statistics and control words of every slice to pinned host memory laid
  // out like d_cs, in one copy (stream-ordered after the chunk enqueued before).
  void copy_ctl(char* host) { HIPC(hipMemcpyAsync(host, d_cs.p, cs_bytes, hipMemcpyDeviceToHost, stream)); }

  void launch_chunk(int len, bool gap, bool use_graph, bool first) {
    first = first && paired();  // only paired iterations start differently
    if (!use_graph) {
      enqueue(len, gap, first);
      HIPC(hipGetLastError());
      return;
    }
    for (auto& g : graphs)
      if (g.len == len && g.gap == gap && g.first == first) {
        HIPC(hipGraphLaunch(g.exec, stream));
        return;
      }
    SetupLock lk(g_setup_mu);
    hipGraph_t graph;
    HIPC(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    enqueue(len, gap, first);
    HIPC(hipStreamEndCapture(stream, &graph));
    hipGraphExec_t exec;
    HIPC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    graphs.push_back({len, gap, first, exec});
    HIPC(hipGraphLaunch(exec, stream));
  }

  void run(int verbose) {
    const int ns = (int)sv.size();
    const unsigned gap_every = sv[0]->prm.gap_every;
    const bool gap = gap_every > 0;
    std::vector<unsigned> limit(ns);
    unsigned total = 0;
    for (int i = 0; i < ns; ++i) {
      limit[i] = sv[i]->prm.max_iter + 1;  // the reference runs k = 0..max_iter
      total = std::max(total, limit[i]);
    }
    // Graphs hold at most kChunk iterations (a long graph's launch latency is
    // not hidden); the gap is evaluated at the multiples of gap_every.
    constexpr unsigned kChunk = 50;
    unsigned chunk = std::min<unsigned>(gap ? gap_every : total, kChunk);
    // blocked ROF launches come in pairs of tb_k iterations: chunks of a
    // multiple of 2 tb_k leave no iterations to the one-iteration kernels
    if (blocked()) {
      const unsigned pair = 2u * unsigned(G[0].tb_k);
      chunk = std::max(pair, chunk - chunk % pair);
    }
    const bool use_graph = sv[0]->prm.use_graph != 0;
    std::vector<double> stop_ms(ns, -1.0);  // batched solves: time from the start to each slice's stop
    reset_ctl();
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, stream));
    auto boundary = [&](unsigned k) {
      unsigned b = std::min(total, (k / chunk + 1) * chunk);
      if (gap) b = std::min(b, (k / gap_every + 1) * gap_every);
      for (int i = 0; i < ns; ++i)
        if (limit[i] > k) b = std::min(b, limit[i]);
      return b;
    };
    auto report = [&]() {
      if (!verbose) return;
      for (int i = 0; i < std::min(ns, 4); ++i) {
        const double* st = h_stats + size_t(i) * ST_WORDS;
        std::printf("[pdt] slice %d iter=%d P=%.9g D=%.9g rel_gap=%.3e\n", i, h_ctl[i * CTL_WORDS + CTL_ITER],
                    st[ST_P], st[ST_D], st[ST_REL]);
      }
      std::fflush(stdout);
    };
    if (ns == 1) {
      // One slice: its iteration limit is the end of the loop, so no host
      // decision is needed before the next chunk.  One chunk stays queued
      // behind the one whose control words the host is reading; once the
      // slice has stopped, the queued chunk's kernels return at once.  When
      // the last two gap values extrapolate (geometrically) to convergence
      // at the chunk in flight, nothing is queued behind it: a chunk of
      // returning launches costs more than the host's wait for the verdict.
      const double tol = gap ? sv[0]->prm.gap_tol : 0.0;
      double last_rel = 0, prev_rel = 0;
      int seen_ngap = 0;
      auto converging = [&]() {
        if (tol <= 0 || seen_ngap < 2 || !(prev_rel > 0)) return false;
        return last_rel * std::min(1.0, last_rel / prev_rel) <= 1.02 * tol;
      };
      for (int q = 0; q < 2; ++q) {
        SetupLock lk(g_setup_mu);
        if (!h_cs_q[q]) HIPC(hipHostMalloc((void**)&h_cs_q[q], cs_bytes, hipHostMallocDefault));
        if (!ev_q[q]) HIPC(hipEventCreateWithFlags(&ev_q[q], hipEventDisableTiming));
      }
      int queued[2], nq = 0, head = 0, next = 0;
      unsigned k = 0;
      bool live = true;
      while (true) {
        while (live && k < total && nq < (converging() ? 1 : 2)) {
          const unsigned b = boundary(k);
          launch_chunk(int(b - k), gap && (b % gap_every) == 0, use_graph, k == 0);
          const int q = next;
          next ^= 1;
          copy_ctl(h_cs_q[q]);
          HIPC(hipEventRecord(ev_q[q], stream));
          queued[(head + nq) % 2] = q;
          ++nq;
          k = b;
        }
        if (nq == 0) break;
        const int q = queued[head];
        head = (head + 1) % 2;
        --nq;
        HIPC(hipEventSynchronize(ev_q[q]));
        std::memcpy(h_cs, h_cs_q[q], cs_bytes);
        report();
        if (h_ctl[CTL_NGAP] > seen_ngap) {
          seen_ngap = h_ctl[CTL_NGAP];
          prev_rel = last_rel;
          last_rel = h_stats[ST_REL];
        }
        if (!h_ctl[CTL_ACTIVE]) live = false;
      }
    } else {
      unsigned k = 0;
      const auto t0 = std::chrono::steady_clock::now();
      while (k < total) {
        const unsigned b = boundary(k);
        launch_chunk(int(b - k), gap && (b % gap_every) == 0, use_graph, k == 0);
        copy_ctl(h_cs);
        HIPC(hipStreamSynchronize(stream));
        k = b;
        // the end of the chunk in which a slice stopped (at its gap, seen in
        // the copied control words, or at its limit, below) is its time to stop
        const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (int i = 0; i < ns; ++i)
          if (stop_ms[i] < 0 && (!h_ctl[size_t(i) * CTL_WORDS + CTL_ACTIVE] || limit[i] <= k)) stop_ms[i] = el;
        // a slice stops at its own iteration limit (host-driven, so that no
        // kernel ever reads a control word that another block of the same
        // launch writes)
        bool changed = false, any = false;
        for (int i = 0; i < ns; ++i) {
          int* c = h_ctl + size_t(i) * CTL_WORDS;
          if (c[CTL_ACTIVE] && limit[i] <= k) {
            c[CTL_ACTIVE] = 0;
            changed = true;
          }
          any = any || c[CTL_ACTIVE];
        }
        report();
        if (!any) break;
        if (changed)
          HIPC(hipMemcpyAsync(d_ctl, h_ctl, size_t(ns) * CTL_WORDS * sizeof(int), hipMemcpyHostToDevice, stream));
        compact_lists();
      }
    }
    HIPC(hipEventRecord(e1, stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int i = 0; i < ns; ++i) {
      pdt_solver* s = sv[i];
      const int* c = h_ctl + size_t(i) * CTL_WORDS;
      const double* st = h_stats + size_t(i) * ST_WORDS;
      s->stats.iters = unsigned(c[CTL_ITER]);
      s->stats.n_gap = unsigned(c[CTL_NGAP]);
      s->stats.primal = st[ST_P];
      s->stats.dual = st[ST_D];
      s->stats.gap = st[ST_GAP];
      s->stats.rel_gap = st[ST_REL];
      s->stats.ms = ms;
      s->stats.ms_to_stop = stop_ms[i] >= 0 ? stop_ms[i] : ms;
      s->stats.converged = c[CTL_DONE_AT] >= 0;
      s->rupd = c[CTL_RUPD];
    }
  }
};

void check_solvable(pdt_solver* s) {
  if (s->is_op() && !s->op) fail(PDT_ERR_STATE, "No operator is set!");  // tgvmrioptimizer.cu:96-102
  if (!s->f) fail(PDT_ERR_STATE, "No noisy data is set!");              // :104-110
  if (s->is_op() && !is_cplx(s->dtype)) fail(PDT_ERR_UNSUPPORTED, "operator solvers need complex images");
}

void prepare_solve(pdt_solver* s, int verbose) {
  SetupLock lk(g_setup_mu);
  check_solvable(s);
  if (!s->stream) HIPC(hipStreamCreate(&s->stream));
  if (s->is_op()) {
    if (s->op->dtype != s->dtype) fail(PDT_ERR_ARG, "operator and solver precision differ");
    prepare_op_data(s);
    if (!s->have_u) {
      // Initialization with A^T(f) (tgvmrioptimizer.cu:112-118)
      if (verbose) std::printf("Initialization with A^T(f)\n");
      const size_t n = size_t(s->op->W) * s->op->H;
      s->atf.ensure(n * s->pe());
      s->op->apply(false, s->f->ptr, s->atf.p, s->stream);
      set_input0(s, s->atf.p, nullptr, s->op->W, s->op->H);  // stream-ordered after the apply
    } else {
      const int64_t want[2] = {s->op->W, s->op->H};
      size_check(s->u->size, 2, want, 2);
    }
  } else {
    if (s->f->ndim != 2) fail(PDT_ERR_SIZE, "denoising data must be 2D");
    if (!s->have_u) {
      // Initialization with f (tvoptimizer.cu:87-92)
      if (verbose) std::printf("Initialization with f\n");
      set_input0(s, s->f->ptr, nullptr, int(s->f->size[0]), int(s->f->size[1]));
    } else {
      size_check(s->u->size, 2, s->f->size, 2);
    }
  }
  if (s->is_tgv() && s->kind == PDT_SOLVER_TGV && !s->tmp.p) s->tmp.alloc(2 * size_t(s->W) * s->H * s->pe());
}

// What one engine (one set of launches) requires of the solvers it runs
// together: distinct handles, one kernel family (kind, precision), one gap
// schedule, and one storage size of the dual fields p, q.  The engine
// instantiates its kernels from the first solver, so a solver whose p, q were
// sized otherwise would be written out of bounds.  Checked on the host before
// any state is touched or any kernel is launched.
void check_batch(const std::vector<pdt_solver*>& sv) {
  if (sv.empty()) fail(PDT_ERR_ARG, "empty batch");
  for (size_t i = 0; i < sv.size(); ++i) {
    if (!sv[i]) fail(PDT_ERR_ARG, "null solver in batch");
    for (size_t j = 0; j < i; ++j)
      if (sv[j] == sv[i]) fail(PDT_ERR_ARG, "a solver appears twice in the batch");
  }
  const pdt_solver* s0 = sv[0];
  for (size_t i = 1; i < sv.size(); ++i) {
    const pdt_solver* s = sv[i];
    if (s->kind != s0->kind || s->dtype != s0->dtype) fail(PDT_ERR_ARG, "batched solvers must share kind and dtype");
    if (s->prm.gap_every != s0->prm.gap_every) fail(PDT_ERR_ARG, "batched solvers must share gap_every");
    if (s->pq_e() != s0->pq_e()) fail(PDT_ERR_ARG, "batched solvers must share dual_fp16");
  }
}

void solve_many(std::vector<pdt_solver*>& sv, int verbose) {
  check_batch(sv);
  for (auto* s : sv) prepare_solve(s, verbose);
  if (sv.size() == 1) {
    // reuse the engine of the previous solve when nothing it depends on
    // changed: its work lists and instantiated graphs stay valid (a replaced
    // engine is released, and the new one built, under the set-up lock)
    pdt_solver* s = sv[0];
    std::shared_ptr<Engine> cached;
    {
      SetupLock lk(g_setup_mu);
      Engine probe;
      probe.sv = sv;
      probe.plan_only = true;
      probe.build();
      cached = std::static_pointer_cast<Engine>(s->engine);
      if (!cached || cached->sig != probe.sig) {
        cached.reset();
        s->engine.reset();
        auto e = std::make_shared<Engine>();
        e->sv = sv;
        e->build();
        s->engine = e;
        cached = e;
      }
    }
    cached->run(verbose);
    return;
  }
  auto e = std::make_unique<Engine>();
  e->sv = sv;
  e->build();
  e->run(verbose);
  SetupLock lk(g_setup_mu);  // its graphs and device buffers are released under the lock
  e.reset();
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* pdt_last_error(void) { return g_err.c_str(); }
int pdt_version(void) { return 100; }

int pdt_set_device(int device) {
  return guard([&] { HIPC(hipSetDevice(device)); });
}

int pdt_array_create(int dtype, int ndim, const int64_t* size, pdt_array** out) {
  return guard([&] {
    if (!out || !size) fail(PDT_ERR_ARG, "null argument");
    *out = new_array(dtype, ndim, size);
  });
}

int pdt_array_wrap(int dtype, int ndim, const int64_t* size, void* device_ptr, pdt_array** out) {
  return guard([&] {
    if (!out || !size) fail(PDT_ERR_ARG, "null argument");
    if (ndim < 1 || ndim > 3) fail(PDT_ERR_ARG, "ndim must be 1..3");
    auto a = std::make_unique<pdt_array>();
    a->dtype = dtype;
    a->ndim = ndim;
    for (int i = 0; i < ndim; ++i) a->size[i] = size[i];
    (void)esize(dtype);
    a->ptr = device_ptr;
    a->owned = false;
    *out = a.release();
  });
}

int pdt_array_upload(pdt_array* a, const void* host, size_t bytes) {
  return guard([&] {
    if (!a || !host) fail(PDT_ERR_ARG, "null argument");
    if (bytes != a->bytes()) fail(PDT_ERR_SIZE, "upload size does not match the array");
    if (bytes) HIPC(hipMemcpy(a->ptr, host, bytes, hipMemcpyHostToDevice));
  });
}

int pdt_array_download(const pdt_array* a, void* host, size_t bytes) {
  return guard([&] {
    if (!a || !host) fail(PDT_ERR_ARG, "null argument");
    if (bytes != a->bytes()) fail(PDT_ERR_SIZE, "download size does not match the array");
    if (bytes) HIPC(hipMemcpy(host, a->ptr, bytes, hipMemcpyDeviceToHost));
  });
}

int pdt_array_info(const pdt_array* a, int* dtype, int* ndim, int64_t* size3) {
  return guard([&] {
    if (!a) fail(PDT_ERR_ARG, "null array");
    if (dtype) *dtype = a->dtype;
    if (ndim) *ndim = a->ndim;
    if (size3)
      for (int i = 0; i < 3; ++i) size3[i] = a->size[i];
  });
}

void* pdt_array_device_ptr(pdt_array* a) { return a ? a->ptr : nullptr; }

int pdt_array_destroy(pdt_array* a) {
  return guard([&] { delete a; });
}

// ---- operators ----
int pdt_op_create(int kind, int dtype, pdt_op** out) {
  return guard([&] {
    if (!out) fail(PDT_ERR_ARG, "null argument");
    if (kind < PDT_OP_CARTESIAN || kind > PDT_OP_SAMPLING) fail(PDT_ERR_ARG, "unknown operator kind");
    if (dtype != PDT_C64 && dtype != PDT_C128) fail(PDT_ERR_ARG, "operators are complex (PDT_C64 / PDT_C128)");
    auto op = std::make_unique<pdt_op>();
    op->kind = kind;
    op->dtype = dtype;
    *out = op.release();
  });
}

int pdt_op_add_constant(pdt_op* op, const void* host, int dtype, int ndim, const int64_t* size) {
  return guard([&] {
    if (!op || !host || !size) fail(PDT_ERR_ARG, "null argument");
    // OperatorBase::addConstant (operatorbase.h:238-257)
    if ((int)op->consts.size() >= op->required())
      fail(PDT_ERR_CONST, op->prefix() + "Too many InputTypeConstants. Could not add new constant.'");
    HostConst c;
    c.dtype = dtype;
    c.ndim = ndim;
    for (int i = 0; i < 3; ++i) c.size[i] = i < ndim ? size[i] : 1;
    c.dev.reset(new_array(dtype, ndim, size));
    const size_t b = c.dev->bytes();
    c.host.assign(static_cast<const char*>(host), static_cast<const char*>(host) + b);
    if (b) HIPC(hipMemcpy(c.dev->ptr, host, b, hipMemcpyHostToDevice));
    op->consts.push_back(std::move(c));
    op->prepared = false;
  });
}

int pdt_op_num_constants(const pdt_op* op, int* have, int* required) {
  return guard([&] {
    if (!op) fail(PDT_ERR_ARG, "null operator");
    if (have) *have = (int)op->consts.size();
    if (required) *required = op->required();
  });
}

namespace {
// sizeCheck of the three operators (mricartesianoperator.cu:173-180,310-318,
// mrisamplingoperator.cu:39-44), in = image array, out = k-space array.
void op_size_check(pdt_op* op, const pdt_array* in, const pdt_array* out) {
  op->check_count();
  const HostConst& m = op->mask();
  if (op->kind == PDT_OP_SAMPLING) {
    size_check(out->size, out->ndim, m.size, 2);
    size_check(in->size, in->ndim, m.size, 2);
    return;
  }
  const HostConst* s = op->sens();
  size_check(out->size, out->ndim, s->size, 3);
  if (op->kind == PDT_OP_CARTESIAN) {
    size_check(in->size, in->ndim, m.size, 2);
    const int64_t o2[2] = {out->size[0], out->size[1]};
    size_check(in->size, in->ndim, o2, 2);
  } else {
    int64_t sz[2] = {in->size[0], in->size[1] * 2};
    if (in->ndim != 2) fail(PDT_ERR_SIZE, "image must be 2D");
    size_check(sz, 2, m.size, 2);
  }
  if (in->dtype != op->dtype || out->dtype != op->dtype) fail(PDT_ERR_ARG, "array dtype does not match the operator");
}
}  // namespace

int pdt_op_forward(pdt_op* op, const pdt_array* src, pdt_array* dst) {
  return guard([&] {
    if (!op || !src || !dst) fail(PDT_ERR_ARG, "null argument");
    op_size_check(op, src, dst);
    op->prepare();
    op->apply(true, src->ptr, dst->ptr, nullptr);
    HIPC(hipStreamSynchronize(nullptr));
  });
}

int pdt_op_adjoint(pdt_op* op, const pdt_array* src, pdt_array* dst) {
  return guard([&] {
    if (!op || !src || !dst) fail(PDT_ERR_ARG, "null argument");
    op_size_check(op, dst, src);
    op->prepare();
    op->apply(false, src->ptr, dst->ptr, nullptr);
    HIPC(hipStreamSynchronize(nullptr));
  });
}

int pdt_op_apply_noalloc(pdt_op* op, int forward, const pdt_array* src, pdt_array* dst, void* stream) {
  return guard([&] {
    if (!op || !src || !dst) fail(PDT_ERR_ARG, "null argument");
    if (forward) op_size_check(op, src, dst);
    else op_size_check(op, dst, src);
    if (!op->prepared)
      fail(PDT_ERR_STATE, op->prefix() +
                              "not prepared: the first pdt_op_forward, pdt_op_adjoint or pdt_op_adjointness "
                              "prepares its geometry");
    op->apply_noalloc(forward != 0, src->ptr, dst->ptr, static_cast<hipStream_t>(stream));
  });
}

int pdt_op_input_size(pdt_op* op, int out_ndim, const int64_t* out_size, int* in_ndim, int64_t* in_size) {
  return guard([&] {
    if (!op || !out_size || !in_ndim || !in_size) fail(PDT_ERR_ARG, "null argument");
    (void)out_ndim;
    *in_ndim = 2;
    in_size[0] = out_size[0];
    in_size[1] = op->kind == PDT_OP_CARTESIAN_REMOVE_ROOS ? out_size[1] / 2 : out_size[1];
  });
}

int pdt_op_output_size(pdt_op* op, int in_ndim, const int64_t* in_size, int* out_ndim, int64_t* out_size) {
  return guard([&] {
    if (!op || !out_ndim || !out_size) fail(PDT_ERR_ARG, "null argument");
    (void)in_ndim;
    (void)in_size;
    if (op->kind == PDT_OP_SAMPLING) {
      const HostConst& m = op->get_const(0, real_of(op->dtype), 2);
      *out_ndim = 2;
      out_size[0] = m.size[0];
      out_size[1] = m.size[1];
    } else {
      const HostConst& s = op->get_const(0, cplx_of(op->dtype), 3);
      *out_ndim = 3;
      for (int i = 0; i < 3; ++i) out_size[i] = s.size[i];
    }
  });
}

int pdt_op_adjointness(pdt_op* op, double* lre, double* lim, double* rre, double* rim) {
  return guard([&] {
    if (!op) fail(PDT_ERR_ARG, "null operator");
    op->prepare();
    // iu::random::fillRandomFloatingNumbers (iurandom.h:120-136): U[0,1)
    // pairs from std::default_random_engine seeded with a global seed that
    // advances by the element count after every fill.
    static unsigned seed = 42;
    const bool dbl = is_dbl(op->dtype);
    const size_t nu = size_t(op->W) * op->H, np = size_t(op->W) * op->Hk * op->C;
    auto fill = [&](size_t n) {
      std::vector<double> v(2 * n);
      std::default_random_engine gen(seed);
      std::uniform_real_distribution<> dis(0, 1);
      for (size_t i = 0; i < n; ++i) {
        v[2 * i] = dis(gen);
        v[2 * i + 1] = dis(gen);
      }
      seed += unsigned(n);
      return v;
    };
    std::vector<double> hu = fill(nu), hp = fill(np);
    const size_t ce = esize(op->dtype);
    DevBuf du(nu * ce), dp(np * ce), dAu(np * ce), dAtp(nu * ce);
    auto up = [&](DevBuf& d, const std::vector<double>& h) {
      if (dbl) {
        HIPC(hipMemcpy(d.p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
      } else {
        std::vector<float> f(h.begin(), h.end());
        HIPC(hipMemcpy(d.p, f.data(), f.size() * 4, hipMemcpyHostToDevice));
      }
    };
    auto down = [&](DevBuf& d, size_t n) {
      std::vector<double> h(2 * n);
      if (dbl) {
        HIPC(hipMemcpy(h.data(), d.p, h.size() * 8, hipMemcpyDeviceToHost));
      } else {
        std::vector<float> f(2 * n);
        HIPC(hipMemcpy(f.data(), d.p, f.size() * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); ++i) h[i] = f[i];
      }
      return h;
    };
    up(du, hu);
    up(dp, hp);
    op->apply(true, du.p, dAu.p, nullptr);
    op->apply(false, dp.p, dAtp.p, nullptr);
    HIPC(hipStreamSynchronize(nullptr));
    std::vector<double> Au = down(dAu, np), Atp = down(dAtp, nu);
    // <a, b> = sum conj(a) b
    auto dot = [](const std::vector<double>& a, const std::vector<double>& b, double& re, double& im) {
      re = im = 0;
      for (size_t i = 0; i < a.size() / 2; ++i) {
        re += a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
        im += a[2 * i] * b[2 * i + 1] - a[2 * i + 1] * b[2 * i];
      }
    };
    double a, b, c, d;
    dot(Au, hp, a, b);
    dot(hu, Atp, c, d);
    if (lre) *lre = a;
    if (lim) *lim = b;
    if (rre) *rre = c;
    if (rim) *rim = d;
  });
}

int pdt_op_name(const pdt_op* op, char* buf, size_t len) {
  return guard([&] {
    if (!op || !buf || !len) fail(PDT_ERR_ARG, "null argument");
    std::ostringstream ss;
    ss << op->name() << ": Added constants: (" << op->consts.size() << "/" << op->required() << ")";
    std::snprintf(buf, len, "%s", ss.str().c_str());
  });
}

int pdt_op_destroy(pdt_op* op) {
  return guard([&] {
    SetupLock lk(g_setup_mu);
    op_release(op);
  });
}

// ---- solvers ----
int pdt_solver_create(int kind, int dtype, pdt_solver** out) {
  return guard([&] {
    if (!out) fail(PDT_ERR_ARG, "null argument");
    if (kind < PDT_SOLVER_TV || kind > PDT_SOLVER_TGV_OP) fail(PDT_ERR_ARG, "unknown solver kind");
    (void)esize(dtype);
    auto s = std::make_unique<pdt_solver>();
    s->kind = kind;
    s->dtype = dtype;
    default_params(kind, &s->prm);
    if (s->is_op() && !is_cplx(dtype)) fail(PDT_ERR_UNSUPPORTED, "operator solvers need complex images");
    *out = s.release();
  });
}

int pdt_params_default(int kind, pdt_params* p) {
  return guard([&] {
    if (!p) fail(PDT_ERR_ARG, "null argument");
    default_params(kind, p);
  });
}

int pdt_solver_get_params(const pdt_solver* s, pdt_params* p) {
  return guard([&] {
    if (!s || !p) fail(PDT_ERR_ARG, "null argument");
    *p = s->prm;
  });
}

int pdt_solver_set_params(pdt_solver* s, const pdt_params* p) {
  return guard([&] {
    if (!s || !p) fail(PDT_ERR_ARG, "null argument");
    // validated before anything is assigned: a refused call changes nothing
    // (p and q keep the storage size they were allocated with)
    const bool fp16_changed = (p->dual_fp16 != 0) != (s->prm.dual_fp16 != 0);
    if (fp16_changed && s->have_u) fail(PDT_ERR_STATE, "dual_fp16 must be set before the first solve");
    s->prm = *p;
  });
}

int pdt_solver_set_operator(pdt_solver* s, pdt_op* op) {
  return guard([&] {
    if (!s || !op) fail(PDT_ERR_ARG, "null argument");
    if (!s->is_op()) fail(PDT_ERR_ARG, "this solver takes no operator");
    op->refs++;
    op_release(s->op);
    s->op = op;
    s->data_dirty = true;
  });
}

int pdt_solver_set_noisy_data(pdt_solver* s, const pdt_array* f) {
  return guard([&] {
    if (!s || !f) fail(PDT_ERR_ARG, "null argument");
    if (f->dtype != s->dtype) fail(PDT_ERR_ARG, "data dtype does not match the solver");
    // same shape as the held copy: overwrite it (no reallocation, and the
    // address the cached engine holds stays valid)
    bool same = s->f && s->f->dtype == f->dtype && s->f->ndim == f->ndim;
    for (int k = 0; same && k < f->ndim; ++k) same = s->f->size[k] == f->size[k];
    if (!same) s->f.reset(new_array(f->dtype, f->ndim, f->size));
    if (f->bytes()) HIPC(hipMemcpy(s->f->ptr, f->ptr, f->bytes(), hipMemcpyDeviceToDevice));
    s->data_dirty = true;  // r = 0 (tgvmrioptimizer.cu:88-90)
  });
}

int pdt_solver_set_input0(pdt_solver* s, const pdt_array* u0, const pdt_array* v0) {
  return guard([&] {
    if (!s || !u0) fail(PDT_ERR_ARG, "null argument");
    if (u0->dtype != s->dtype) fail(PDT_ERR_ARG, "input dtype does not match the solver");
    if (u0->ndim != 2) fail(PDT_ERR_SIZE, "input must be 2D");
    if (v0) {
      if (!s->is_tgv()) fail(PDT_ERR_ARG, "v0 is only taken by the TGV solvers");
      const int64_t want[3] = {u0->size[0], u0->size[1], 2};
      size_check(v0->size, v0->ndim, want, 3);
      if (v0->dtype != s->dtype) fail(PDT_ERR_ARG, "v0 dtype does not match the solver");
    }
    set_input0(s, u0->ptr, v0 ? v0->ptr : nullptr, int(u0->size[0]), int(u0->size[1]));
  });
}

int pdt_solver_solve(pdt_solver* s, int verbose) {
  return guard([&] {
    if (!s) fail(PDT_ERR_ARG, "null solver");
    std::vector<pdt_solver*> v{s};
    solve_many(v, verbose);
  });
}

int pdt_solver_get_result(pdt_solver* s, pdt_array** u) {
  return guard([&] {
    if (!s || !u) fail(PDT_ERR_ARG, "null argument");
    if (!s->have_u) fail(PDT_ERR_STATE, "There is no result yet!");
    *u = s->u.get();
  });
}

int pdt_solver_get_v(pdt_solver* s, pdt_array** v) {
  return guard([&] {
    if (!s || !v) fail(PDT_ERR_ARG, "null argument");
    if (!s->have_u) fail(PDT_ERR_STATE, "There is no result yet!");
    if (!s->is_tgv()) fail(PDT_ERR_ARG, "this solver has no v");
    *v = s->v.get();
  });
}

int pdt_solver_get_stats(const pdt_solver* s, pdt_stats* st) {
  return guard([&] {
    if (!s || !st) fail(PDT_ERR_ARG, "null argument");
    *st = s->stats;
  });
}

int pdt_solver_destroy(pdt_solver* s) {
  return guard([&] {
    SetupLock lk(g_setup_mu);  // its engine's graphs and its device buffers
    delete s;
  });
}

// The solvers' iterates (u, ubar, v, vbar, p, q and the MRI dual r), copied
// aside while the autotuner or the kernel profile runs iterations on them
// and put back when it ends: a solve after either starts where it would
// have without it.
class StateGuard {
 public:
  explicit StateGuard(const std::vector<pdt_solver*>& sv) {
    HIPC(hipDeviceSynchronize());
    for (auto* s : sv) {
      if (s->u) keep(s->u->ptr, s->u->bytes());
      if (s->v) keep(s->v->ptr, s->v->bytes());
      keep(s->ub.p, s->ub.bytes);
      keep(s->vb.p, s->vb.bytes);
      keep(s->p.p, s->p.bytes);
      keep(s->q.p, s->q.bytes);
      keep(s->rk.p, s->rk.bytes);
    }
  }
  ~StateGuard() {
    (void)hipDeviceSynchronize();
    for (auto& k : saved_) (void)hipMemcpy(k.first, k.second.p, k.second.bytes, hipMemcpyDeviceToDevice);
  }
  StateGuard(const StateGuard&) = delete;
  StateGuard& operator=(const StateGuard&) = delete;

 private:
  void keep(void* ptr, size_t bytes) {
    if (!ptr || !bytes) return;
    DevBuf b(bytes);
    HIPC(hipMemcpy(b.p, ptr, bytes, hipMemcpyDeviceToDevice));
    saved_.emplace_back(ptr, std::move(b));
  }
  std::vector<std::pair<void*, DevBuf>> saved_;
};

namespace {
// Per-stage duration and algorithmic bytes of one launch, over a solver or a
// batch (pdt_solver_profile, pdt_batch_profile).
void profile_engine(std::vector<pdt_solver*> sv, int reps, pdt_kernel_prof* out, int cap, int* n) {
  if (!out || !n || reps <= 0) fail(PDT_ERR_ARG, "bad argument");
  check_batch(sv);
  for (auto* s : sv) prepare_solve(s, 0);
  StateGuard keep_state(sv);
  {
    Engine e;
    e.sv = sv;
    e.build();
    e.reset_ctl();
    if (e.blocked()) {
      // temporally blocked launches, as a solve runs them: a graph of reps
      // pairs of launches timed with an event pair; per launch (tb_k
      // iterations) each plane read or written once: ROF reads u, ubar, p,
      // f and writes u, ubar, p; TGV reads u, v, p, q, f and writes u, ubar,
      // v, vbar, p, q
      hipEvent_t e0, e1;
      HIPC(hipEventCreate(&e0));
      HIPC(hipEventCreate(&e1));
      hipGraph_t graph;
      hipGraphExec_t ge;
      {
        SetupLock lk(g_setup_mu);
        HIPC(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < reps; ++r) {
          run_block(e.G[0], e.kind, e.dtk(), e.fp16, 0);
          run_block(e.G[0], e.kind, e.dtk(), e.fp16, 1);
        }
        HIPC(hipStreamEndCapture(e.stream, &graph));
        HIPC(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
      }
      HIPC(hipGraphLaunch(ge, e.stream));
      HIPC(hipEventRecord(e0, e.stream));
      HIPC(hipGraphLaunch(ge, e.stream));
      HIPC(hipEventRecord(e1, e.stream));
      HIPC(hipEventSynchronize(e1));
      (void)hipGraphExecDestroy(ge);
      float loop_ms = 0;
      HIPC(hipEventElapsedTime(&loop_ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      double bytes = 0;
      for (auto* s : sv) {
        const double N = double(s->W) * s->H, pe = double(s->pe()), pq = double(s->pq_e());
        bytes += e.kind == PDT_SOLVER_TGV ? (10 * pe + 10 * pq) * N : (5 * pe + 4 * pq) * N;
      }
      *n = 0;
      if (cap > 0) {
        std::snprintf(out[0].name, sizeof(out[0].name), e.kind == PDT_SOLVER_TGV ? "k_tgv_block" : "k_tv_block");
        out[0].ms = double(loop_ms) / (2.0 * reps);
        out[0].bytes = bytes;
        *n = 1;
      }
      return;
    }
    // The iteration as a solve runs it: a hipGraph of reps iterations, timed
    // with an event pair on the solver stream (t_it per iteration).  Each
    // stage's share of it comes from event pairs around its direct launches
    // in iteration order (e_st); an event pair adds a fixed dispatch delay to
    // every launch, so the stage duration reported is t_it * e_st / sum(e).
    // (A paired engine's iterations are taken as the two launches of the
    // steady state, each reading the T1 the one before wrote.)
    const int ns = e.iter_launches();
    e.iter_prologue();
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    hipGraph_t graph;
    hipGraphExec_t ge;
    {
      SetupLock lk(g_setup_mu);
      HIPC(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
      for (int r = 0; r < reps; ++r)
        for (int st = 0; st < ns; ++st) e.iter_launch(st);
      HIPC(hipStreamEndCapture(e.stream, &graph));
      HIPC(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
    }
    HIPC(hipGraphLaunch(ge, e.stream));  // warm-up replay
    HIPC(hipEventRecord(e0, e.stream));
    HIPC(hipGraphLaunch(ge, e.stream));
    HIPC(hipEventRecord(e1, e.stream));
    HIPC(hipEventSynchronize(e1));
    (void)hipGraphExecDestroy(ge);
    float loop_ms = 0;
    HIPC(hipEventElapsedTime(&loop_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<hipEvent_t> ev(size_t(2) * ns * reps);
    for (auto& x : ev) HIPC(hipEventCreate(&x));
    for (int r = 0; r < reps; ++r)
      for (int st = 0; st < ns; ++st) {
        HIPC(hipEventRecord(ev[(size_t(r) * ns + st) * 2], e.stream));
        e.iter_launch(st);
        HIPC(hipEventRecord(ev[(size_t(r) * ns + st) * 2 + 1], e.stream));
      }
    HIPC(hipStreamSynchronize(e.stream));
    HIPC(hipGetLastError());
    std::vector<double> share(ns, 0.0);
    double sum = 0;
    for (int st = 0; st < ns; ++st) {
      if (e.iter_launch_bytes(st) <= 0) continue;  // a stage without launches (primal step inside k_rows_adj)
      for (int r = 0; r < reps; ++r) {
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, ev[(size_t(r) * ns + st) * 2], ev[(size_t(r) * ns + st) * 2 + 1]));
        share[st] += ms;
      }
      sum += share[st];
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    const double t_it = double(loop_ms) / reps;
    int k = 0;
    for (int st = 0; st < ns && k < cap; ++st) {
      const double bytes = e.iter_launch_bytes(st);
      if (bytes <= 0) continue;
      std::snprintf(out[k].name, sizeof(out[k].name), "%s", e.iter_launch_name(st).c_str());
      out[k].ms = sum > 0 ? t_it * share[st] / sum : 0.0;
      out[k].bytes = bytes;
      ++k;
    }
    *n = k;
  }
}
}  // namespace

int pdt_solver_profile(pdt_solver* s, int reps, pdt_kernel_prof* out, int cap, int* n) {
  return guard([&] {
    if (!s) fail(PDT_ERR_ARG, "null solver");
    profile_engine({s}, reps, out, cap, n);
  });
}

int pdt_batch_profile(pdt_solver** solvers, int count, int reps, pdt_kernel_prof* out, int cap, int* n) {
  return guard([&] {
    if (!solvers || count <= 0) fail(PDT_ERR_ARG, "empty batch");
    profile_engine(std::vector<pdt_solver*>(solvers, solvers + count), reps, out, cap, n);
  });
}

// helpers of the C ABI below, with C++ linkage
extern "C++" {
namespace {
// Device time of each stage kernel of an iteration (ms per launch), each
// stage timed over `reps` back-to-back launches after one untimed launch.
std::vector<double> time_stages(Engine& e, int reps) {
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  std::vector<double> out;
  e.iter_prologue();
  for (int st = 0; st < e.iter_launches(); ++st) {
    e.iter_launch(st);
    HIPC(hipEventRecord(e0, e.stream));
    for (int r = 0; r < reps; ++r) e.iter_launch(st);
    HIPC(hipEventRecord(e1, e.stream));
    HIPC(hipEventSynchronize(e1));
    HIPC(hipGetLastError());
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    out.push_back(double(ms) / reps);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return out;
}
double time_iteration(Engine& e, int reps) {
  if (e.blocked()) {
    // pairs of temporally blocked launches, per iteration
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    run_block(e.G[0], e.kind, e.dtk(), e.fp16, 0);
    run_block(e.G[0], e.kind, e.dtk(), e.fp16, 1);
    HIPC(hipEventRecord(e0, e.stream));
    for (int r = 0; r < reps; ++r) {
      run_block(e.G[0], e.kind, e.dtk(), e.fp16, 0);
      run_block(e.G[0], e.kind, e.dtk(), e.fp16, 1);
    }
    HIPC(hipEventRecord(e1, e.stream));
    HIPC(hipEventSynchronize(e1));
    HIPC(hipGetLastError());
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return double(ms) / (2.0 * reps * e.G[0].tb_k);
  }
  double tot = 0;
  for (double t : time_stages(e, reps)) tot += t;
  return tot;
}
}  // namespace
}  // extern "C++"

int pdt_autotune(pdt_solver** solvers, int count, int reps, pdt_tiles* out, double* out_ms) {
  return guard([&] {
    if (!solvers || count <= 0 || reps <= 0) fail(PDT_ERR_ARG, "bad argument");
    std::vector<pdt_solver*> sv(solvers, solvers + count);
    check_batch(sv);
    for (auto* s : sv) prepare_solve(s, 0);
    StateGuard keep_state(sv);
    const bool mri = sv[0]->is_op();
    // Phase 1, tile shapes without occupancy caps: coils per forward row
    // item, columns per column item, rows per denoising item, coils per
    // adjoint row item (the forward choice, or every coil of the row, which
    // also fuses the u step into that pass); for a single ROF or TGV
    // denoising slice also temporally blocked launches (interior tile,
    // iterations per launch), compared per iteration
    struct Cand {
      std::array<int, 4> t;
      std::array<int, 3> tb;
      int dpx = 0;
      int unp = 0;  // MRI: three launches per iteration where two would do
    };
    std::vector<Cand> cand;
    if (mri) {
      int maxC = 1;
      for (auto* s : sv) maxC = std::max(maxC, s->op->C);
      std::vector<int> cbs{1, 2, 3, 4, 5, 8, (maxC + 1) / 2, maxC};
      std::sort(cbs.begin(), cbs.end());
      cbs.erase(std::unique(cbs.begin(), cbs.end()), cbs.end());
      for (int cb : cbs)
        if (cb <= maxC)
          for (int cg : {1, 2, 4}) {
            cand.push_back({{cb, cg, 0, cb}, {0, 0, 0}, 0});
            if (cb != maxC) cand.push_back({{cb, cg, 0, maxC}, {0, 0, 0}, 0});
            // every coil in the adjoint row blocks: also the three-launch iteration
            cand.push_back({{cb, cg, 0, maxC}, {0, 0, 0}, 0, 1});
          }
    } else {
      for (int rpi : {1, 2, 4, 8})
        for (int dpx : sv[0]->kind == PDT_SOLVER_TGV ? std::vector<int>{1, 2} : std::vector<int>{0})
          cand.push_back({{0, 0, rpi, 0}, {0, 0, 0}, dpx});
      if (sv[0]->kind == PDT_SOLVER_TV && sv.size() == 1)
        for (const auto& tb : std::vector<std::array<int, 3>>{
                 {16, 16, 4}, {16, 16, 6}, {32, 16, 4}, {32, 16, 8}, {32, 32, 8}, {64, 32, 8}, {32, 32, 12},
                 {64, 64, 8}, {64, 64, 16}})
          cand.push_back({{0, 0, 1, 0}, tb, 0});
      // TGV holds 8 planes per tile pixel in LDS: smaller tiles and halos
      if (sv[0]->kind == PDT_SOLVER_TGV && sv.size() == 1)
        for (const auto& tb : std::vector<std::array<int, 3>>{
                 {16, 8, 2}, {16, 16, 2}, {16, 16, 4}, {32, 8, 2}, {32, 16, 2}, {32, 16, 4}, {32, 32, 2}, {32, 32, 4},
                 {32, 32, 6}, {48, 32, 4}, {64, 16, 4}, {64, 32, 3}})
          cand.push_back({{0, 0, 1, 0}, tb, 0});
    }
    auto engine = [&](const Cand& c, const int (&wpc)[4]) {
      auto e = std::make_unique<Engine>();
      e->sv = sv;
      e->force_cb = c.t[0];
      e->force_cg = c.t[1];
      e->force_rpi = c.t[2];
      e->force_cb_a = c.t[3];
      for (int i = 0; i < 3; ++i) e->force_tb[i] = c.tb[i];
      e->force_dpx = c.dpx;
      e->force_unpaired = c.unp;
      for (int i = 0; i < 4; ++i) e->force_wpc[i] = wpc[i];
      e->build();
      e->reset_ctl();
      return e;
    };
    const int none[4] = {0, 0, 0, 0};
    double best = 1e300;
    Cand bc{{0, 0, 0, 0}, {0, 0, 0}, 0, 0};
    std::vector<double> best_st;
    bool best_paired = false;
    TuneKey key{};
    for (const auto& c : cand) {
      auto e = engine(c, none);
      if (c.tb[2] > 0 && !e->blocked()) continue;  // tile does not fit
      const std::vector<double> st = e->blocked() ? std::vector<double>{time_iteration(*e, reps)} : time_stages(*e, reps);
      double ms = 0;
      for (double t : st) ms += t;
      key = e->key;
      if (ms < best) {
        best = ms;
        bc = {{e->L.cb, e->L.cg, e->L.rpi, e->L.cb_a}, {e->G[0].tb_x, e->G[0].tb_y, e->G[0].tb_k}, c.dpx, c.unp};
        best_st = st;
        best_paired = e->paired();
      }
    }
    // Phase 2, waves per CU: with the chosen tiles, a cap on the workgroups
    // per CU of each pass in turn (its kernel timed alone); a cap is kept
    // only when it beats no cap by 2 %, so timing noise keeps the default.
    // The denoising stages share one cap (their total is compared); the
    // blocked ROF launches take none.
    int wpc[4] = {0, 0, 0, 0};
    std::vector<std::pair<int, std::vector<int>>> axes;  // (cap slot, stages it governs)
    // (a paired iteration: the column and adjoint row launches, the forward
    // row pass running once per chunk only)
    if (mri && best_paired) axes = {{WPC_COLS, {0}}, {WPC_ADJ, {1}}};
    else if (mri) axes = {{WPC_FWD, {0}}, {WPC_COLS, {1}}, {WPC_ADJ, {2}}};
    else if (bc.tb[2] == 0) axes = {{WPC_DEN, {0, 1}}};
    const std::vector<int> caps = mri ? std::vector<int>{1, 2, 3, 4, 6} : std::vector<int>{3, 4, 6};
    for (const auto& ax : axes) {
      auto cost = [&](const std::vector<double>& st) {
        double t = 0;
        for (int i : ax.second) t += st[size_t(i)];
        return t;
      };
      double base = cost(best_st);
      for (int cap : caps) {
        int w[4];
        for (int i = 0; i < 4; ++i) w[i] = wpc[i];
        w[ax.first] = cap;
        auto e = engine(bc, w);
        const double t = cost(time_stages(*e, reps));
        if (t < 0.98 * base) {
          base = t;
          wpc[ax.first] = cap;
        }
      }
    }
    TuneVal tv = mri ? TuneVal{bc.t[0], bc.t[1], 0, bc.t[3]} : TuneVal{0, 0, bc.t[2], 0};
    for (int i = 0; i < 4; ++i) tv.wpc[i] = wpc[i];
    for (int i = 0; i < 3; ++i) tv.tb[i] = bc.tb[i];
    tv.dual_px = bc.dpx;
    tv.unpaired = bc.unp;
    tune_store(key, tv);
    double ms = 0;
    {
      auto e = engine(bc, wpc);
      ms = time_iteration(*e, reps);
    }
    if (out) {
      out->cb = bc.t[0];
      out->cg = bc.t[1];
      out->rpi = bc.t[2];
      out->cb_a = bc.t[3];
      for (int i = 0; i < 4; ++i) out->wpc[i] = wpc[i];
      out->tb_x = bc.tb[0];
      out->tb_y = bc.tb[1];
      out->tb_k = bc.tb[2];
      out->dual_px = bc.dpx;
      out->unpaired = bc.unp;
    }
    if (out_ms) *out_ms = ms;
  });
}

int pdt_set_tiles(pdt_solver** solvers, int count, const pdt_tiles* t) {
  return guard([&] {
    if (!solvers || count <= 0 || !t) fail(PDT_ERR_ARG, "bad argument");
    std::vector<pdt_solver*> sv(solvers, solvers + count);
    check_batch(sv);
    for (auto* s : sv) prepare_solve(s, 0);
    Engine e;
    e.sv = sv;
    e.plan_only = true;
    // zero or negative: the default, not a cached choice
    e.force_cb = t->cb > 0 ? t->cb : kDefaultCb;
    e.force_cg = t->cg > 0 ? t->cg : kDefaultCg;
    e.force_rpi = t->rpi > 0 ? t->rpi : 1;
    e.force_cb_a = t->cb_a > 0 ? t->cb_a : -1;  // (-1: the engine's default grouping)
    for (int i = 0; i < 4; ++i) e.force_wpc[i] = std::max(0, std::min(t->wpc[i], 6));
    e.force_tb[0] = t->tb_x;
    e.force_tb[1] = t->tb_y;
    e.force_tb[2] = t->tb_k > 0 ? t->tb_k : 0;
    e.force_dpx = t->dual_px == 1 || t->dual_px == 2 ? t->dual_px : 0;
    e.force_unpaired = t->unpaired != 0 ? 1 : 0;
    e.build();
    const bool mri = sv[0]->is_op();
    TuneVal tv = mri ? TuneVal{e.L.cb, e.L.cg, 0, e.L.cb_a} : TuneVal{0, 0, e.L.rpi, 0};
    for (int i = 0; i < 4; ++i) tv.wpc[i] = e.L.wpc[i];
    tv.tb[0] = e.L.tb_x;
    tv.tb[1] = e.L.tb_y;
    tv.tb[2] = e.L.tb_k;
    tv.dual_px = e.force_dpx;
    tv.unpaired = e.unpaired ? 1 : 0;
    tune_store(e.key, tv);
  });
}

int pdt_batch_solve(pdt_solver** solvers, int count, int verbose) {
  return guard([&] {
    if (!solvers || count <= 0) fail(PDT_ERR_ARG, "empty batch");
    std::vector<pdt_solver*> v(solvers, solvers + count);
    solve_many(v, verbose);
  });
}

int pdt_solve_concurrent(pdt_solver** solvers, int count, int streams, int verbose) {
  return guard([&] {
    if (!solvers || count <= 0) fail(PDT_ERR_ARG, "empty batch");
    if (streams < 1 || streams > PDT_MAX_CONCURRENT) fail(PDT_ERR_ARG, "streams must be in [1, 16]");
    std::vector<pdt_solver*> v(solvers, solvers + count);
    for (int i = 0; i < count; ++i) {
      if (!v[i]) fail(PDT_ERR_ARG, "null solver in batch");
      for (int j = 0; j < i; ++j)
        if (v[j] == v[i]) fail(PDT_ERR_ARG, "a solver appears twice in the batch");
    }
    // Workers (the calling thread and streams - 1 more) take the next
    // unsolved solver in order and solve it alone on its own stream; the
    // first failure stops the others from taking more and is reported here.
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::mutex mu;
    int err_code = PDT_OK;
    std::string err_msg;
    auto work = [&] {
      const int rc = guard([&] {
        HIPC(hipSetDevice(dev));
        for (;;) {
          if (failed.load()) return;
          const int i = next.fetch_add(1);
          if (i >= count) return;
          std::vector<pdt_solver*> one{v[i]};
          solve_many(one, verbose);
        }
      });
      if (rc != PDT_OK) {
        failed.store(true);
        std::lock_guard<std::mutex> lk(mu);
        if (err_code == PDT_OK) {
          err_code = rc;
          err_msg = g_err;
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < std::min(streams, count); ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (err_code != PDT_OK) fail(err_code, err_msg);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned warm start
// ---------------------------------------------------------------------------
struct pdt_cnn {
  int nl = 0;
  std::vector<int> ch;         // nl + 1 channel counts
  std::vector<size_t> off;     // offset of layer l's W in the parameter vector; b follows it
  size_t np = 0;
  DevBuf w, g, m, v;           // fp32 parameters, gradient, Adam moments
  std::vector<DevBuf> wk, wt;  // packed bf16 weights (forward, backward-data)
  int64_t step = 0;
  int64_t cap = 0;             // pixels the activation buffers hold
  DevBuf xb, y, dza, dzb, part, lpart;
  std::vector<DevBuf> act;     // hidden activations (bf16), nl - 1 of them
  DevBuf ws_x, ws_u, ws_v, ws_ss;  // warm-start input, (u0, v0) and sum of |A^*f|^2
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;  // solver stream <-> network stream ordering
  ~pdt_cnn() {
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (stream) (void)hipStreamDestroy(stream);
  }

  size_t wsize(int l) const { return size_t(9) * ch[l] * ch[l + 1]; }
  const float* W(int l) const { return w.as<float>() + off[l]; }
  const float* b(int l) const { return w.as<float>() + off[l] + wsize(l); }

  void ensure(int64_t P) {
    if (P <= cap) return;
    xb.alloc(size_t(P) * ch[0] * 2);
    for (int l = 0; l + 1 < nl; ++l) act[l].alloc(size_t(P) * ch[l + 1] * 2);
    y.alloc(size_t(P) * ch[nl] * 4);
    int cmax = 1;
    for (int c : ch) cmax = std::max(cmax, c);
    dza.alloc(size_t(P) * cmax * 2);
    dzb.alloc(size_t(P) * cmax * 2);
    size_t pf = 0;
    for (int l = 0; l < nl; ++l) pf = std::max(pf, cnn_wgrad_part_floats(ch[l], ch[l + 1], 1024));
    part.alloc(pf * 4);
    lpart.alloc(1024 * sizeof(double));
    cap = P;
  }

  void pack() {
    for (int l = 0; l < nl; ++l) cnn_pack_weights(stream, W(l), ch[l], ch[l + 1], wk[l].p, l > 0 ? wt[l].p : nullptr);
  }

  // y = net(x) for P = n h w pixels (x: fp32 NHWC on the device)
  void forward(const float* x, int n, int h, int wd) {
    const int64_t P = int64_t(n) * h * wd;
    ensure(P);
    cnn_to_bf16(stream, x, xb.p, P * ch[0]);
    const void* in = xb.p;
    for (int l = 0; l < nl; ++l) {
      const bool last = l + 1 == nl;
      void* out = last ? y.p : act[l].p;
      cnn_conv_forward(stream, ch[l], ch[l + 1], !last, in, wk[l].p, b(l), out, n, h, wd);
      in = out;
    }
    HIPC(hipGetLastError());
  }

  // loss and gradient (into g) of the last forward against t
  double backward(const float* t, int n, int h, int wd) {
    const int64_t P = int64_t(n) * h * wd, M = P * ch[nl];
    const int nb = cnn_mse_blocks(M);
    void* dz = dza.p;
    void* dzn = dzb.p;
    cnn_mse(stream, y.as<float>(), t, M, dz, lpart.as<double>(), nb);
    const int nblk = cnn_wgrad_blocks(n, h, wd);
    for (int l = nl - 1; l >= 0; --l) {
      const void* in = l == 0 ? xb.p : act[l - 1].p;
      cnn_conv_wgrad(stream, ch[l], ch[l + 1], in, dz, part.as<float>(), nblk, g.as<float>() + off[l], n, h, wd);
      if (l > 0) {
        cnn_conv_backward_data(stream, ch[l], ch[l + 1], dz, wt[l].p, act[l - 1].p, dzn, n, h, wd);
        std::swap(dz, dzn);
      }
    }
    HIPC(hipGetLastError());
    std::vector<double> lp(nb);
    HIPC(hipMemcpyAsync(lp.data(), lpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPC(hipStreamSynchronize(stream));
    double ss = 0;
    for (double x : lp) ss += x;
    return ss / (2.0 * double(M));
  }

  void adam(const pdt_adam& o) {
    ++step;
    const double c1 = 1.0 - std::pow(o.beta1, double(step)), c2 = 1.0 - std::pow(o.beta2, double(step));
    cnn_adam(stream, w.as<float>(), g.as<float>(), m.as<float>(), v.as<float>(), int64_t(np), float(o.lr),
             float(o.beta1), float(o.beta2), float(o.eps), float(c1), float(c2));
    pack();
    HIPC(hipGetLastError());
  }
};

namespace {
const float* cnn_tensor(const pdt_array* a, int64_t want, const char* what) {
  if (!a) fail(PDT_ERR_ARG, std::string("null ") + what);
  if (a->dtype != PDT_F32) fail(PDT_ERR_ARG, std::string(what) + " must be PDT_F32");
  if (int64_t(a->numel()) != want)
    fail(PDT_ERR_SIZE, std::string(what) + " holds " + std::to_string(a->numel()) + " floats, expected " +
                           std::to_string(want));
  return static_cast<const float*>(a->ptr);
}
void cnn_dims(int n, int h, int w) {
  if (n <= 0 || h <= 0 || w <= 0) fail(PDT_ERR_ARG, "n, h, w must be positive");
}
}  // namespace

extern "C" {

int pdt_cnn_create(int cin, int width, int depth, int cout, pdt_cnn** out) {
  return guard([&] {
    if (!out) fail(PDT_ERR_ARG, "null output");
    if (depth < 2) fail(PDT_ERR_ARG, "depth must be at least 2");
    auto net = std::make_unique<pdt_cnn>();
    net->nl = depth;
    net->ch.assign(depth + 1, width);
    net->ch[0] = cin;
    net->ch[depth] = cout;
    for (int l = 0; l < depth; ++l)
      if (!cnn_layer_supported(net->ch[l], net->ch[l + 1]))
        fail(PDT_ERR_UNSUPPORTED, "no kernels for a " + std::to_string(net->ch[l]) + " -> " +
                                      std::to_string(net->ch[l + 1]) + " conv layer (supported: 2-32-...-32-6)");
    for (int l = 0; l < depth; ++l) {
      net->off.push_back(net->np);
      net->np += net->wsize(l) + size_t(net->ch[l + 1]);
    }
    HIPC(hipStreamCreate(&net->stream));
    net->w.alloc(net->np * 4);
    net->g.alloc(net->np * 4);
    net->m.alloc(net->np * 4);
    net->v.alloc(net->np * 4);
    net->wk.resize(depth);
    net->wt.resize(depth);
    net->act.resize(depth - 1);
    for (int l = 0; l < depth; ++l) {
      net->wk[l].alloc(cnn_wk_elems(net->ch[l], net->ch[l + 1]) * 2);
      if (l > 0) net->wt[l].alloc(cnn_wt_elems(net->ch[l], net->ch[l + 1]) * 2);
    }
    run_zero(net->stream, net->w.p, net->np * 4);
    run_zero(net->stream, net->m.p, net->np * 4);
    run_zero(net->stream, net->v.p, net->np * 4);
    net->pack();
    HIPC(hipStreamSynchronize(net->stream));
    *out = net.release();
  });
}

int pdt_cnn_destroy(pdt_cnn* net) {
  return guard([&] { delete net; });
}

int pdt_cnn_num_params(const pdt_cnn* net, size_t* n) {
  return guard([&] {
    if (!net || !n) fail(PDT_ERR_ARG, "null argument");
    *n = net->np;
  });
}

int pdt_cnn_set_params(pdt_cnn* net, const float* host, size_t n) {
  return guard([&] {
    if (!net || !host) fail(PDT_ERR_ARG, "null argument");
    if (n != net->np) fail(PDT_ERR_SIZE, "expected " + std::to_string(net->np) + " parameters, got " + std::to_string(n));
    HIPC(hipMemcpy(net->w.p, host, n * 4, hipMemcpyHostToDevice));
    run_zero(net->stream, net->m.p, n * 4);
    run_zero(net->stream, net->v.p, n * 4);
    net->step = 0;
    net->pack();
    HIPC(hipStreamSynchronize(net->stream));
  });
}

int pdt_cnn_get_params(const pdt_cnn* net, float* host, size_t n) {
  return guard([&] {
    if (!net || !host) fail(PDT_ERR_ARG, "null argument");
    if (n != net->np) fail(PDT_ERR_SIZE, "expected " + std::to_string(net->np) + " parameters, got " + std::to_string(n));
    HIPC(hipStreamSynchronize(net->stream));
    HIPC(hipMemcpy(host, net->w.p, n * 4, hipMemcpyDeviceToHost));
  });
}

int pdt_cnn_forward(pdt_cnn* net, const pdt_array* x, pdt_array* y, int n, int h, int w) {
  return guard([&] {
    if (!net) fail(PDT_ERR_ARG, "null network");
    cnn_dims(n, h, w);
    const int64_t P = int64_t(n) * h * w;
    const float* xp = cnn_tensor(x, P * net->ch[0], "x");
    float* yp = const_cast<float*>(cnn_tensor(y, P * net->ch[net->nl], "y"));
    net->forward(xp, n, h, w);
    HIPC(hipMemcpyAsync(yp, net->y.p, size_t(P) * net->ch[net->nl] * 4, hipMemcpyDeviceToDevice, net->stream));
    HIPC(hipStreamSynchronize(net->stream));
  });
}

int pdt_cnn_loss_grad(pdt_cnn* net, const pdt_array* x, const pdt_array* t, int n, int h, int w, double* loss,
                      float* grad_host, size_t n_grad) {
  return guard([&] {
    if (!net) fail(PDT_ERR_ARG, "null network");
    cnn_dims(n, h, w);
    const int64_t P = int64_t(n) * h * w;
    const float* xp = cnn_tensor(x, P * net->ch[0], "x");
    const float* tp = cnn_tensor(t, P * net->ch[net->nl], "t");
    if (grad_host && n_grad != net->np) fail(PDT_ERR_SIZE, "gradient buffer must hold every parameter");
    net->forward(xp, n, h, w);
    const double l = net->backward(tp, n, h, w);
    if (loss) *loss = l;
    if (grad_host) HIPC(hipMemcpy(grad_host, net->g.p, net->np * 4, hipMemcpyDeviceToHost));
  });
}

int pdt_cnn_adam_step(pdt_cnn* net, const pdt_adam* opt) {
  return guard([&] {
    if (!net || !opt) fail(PDT_ERR_ARG, "null argument");
    net->adam(*opt);
    HIPC(hipStreamSynchronize(net->stream));
  });
}

int pdt_cnn_train_step(pdt_cnn* net, const pdt_array* x, const pdt_array* t, int n, int h, int w,
                       const pdt_adam* opt, double* loss) {
  return guard([&] {
    if (!net || !opt) fail(PDT_ERR_ARG, "null argument");
    cnn_dims(n, h, w);
    const int64_t P = int64_t(n) * h * w;
    const float* xp = cnn_tensor(x, P * net->ch[0], "x");
    const float* tp = cnn_tensor(t, P * net->ch[net->nl], "t");
    net->forward(xp, n, h, w);
    const double l = net->backward(tp, n, h, w);
    if (loss) *loss = l;
    net->adam(*opt);
    HIPC(hipStreamSynchronize(net->stream));
  });
}

int pdt_solver_warm_start(pdt_solver* s, pdt_cnn* net) {
  return guard([&] {
    if (!s || !net) fail(PDT_ERR_ARG, "null argument");
    if (s->kind != PDT_SOLVER_TGV_MRI && s->kind != PDT_SOLVER_TGV_OP)
      fail(PDT_ERR_UNSUPPORTED, "the warm start predicts (u, v) of the TGV operator solvers");
    if (s->dtype != PDT_C64) fail(PDT_ERR_UNSUPPORTED, "the warm start runs on complex64 solvers");
    if (net->ch[0] != 2 || net->ch[net->nl] != 6) fail(PDT_ERR_ARG, "the warm-start network maps 2 to 6 channels");
    prepare_solve(s, 0);
    const int W = s->op->W, H = s->op->H;
    const int64_t n = int64_t(W) * H;
    // A^*f and its sum of squares on the solver stream; the scale RMS(A^*f)
    // is taken on the device by the pack and unpack kernels (no host trip)
    s->atf.ensure(size_t(n) * 8);
    net->ws_x.ensure(size_t(n) * 2 * 4);
    net->ws_u.ensure(size_t(n) * 8);
    net->ws_v.ensure(size_t(n) * 16);
    net->ws_ss.ensure(sizeof(double));
    if (!net->ev_in) HIPC(hipEventCreateWithFlags(&net->ev_in, hipEventDisableTiming));
    if (!net->ev_out) HIPC(hipEventCreateWithFlags(&net->ev_out, hipEventDisableTiming));
    s->op->apply(false, s->f->ptr, s->atf.p, s->stream);
    run_sumsq(s->stream, DT_C64, s->atf.p, n, net->ws_ss.as<double>());
    cnn_ws_pack(s->stream, s->atf.p, n, net->ws_ss.as<double>(), net->ws_x.as<float>());
    HIPC(hipEventRecord(net->ev_in, s->stream));
    HIPC(hipStreamWaitEvent(net->stream, net->ev_in, 0));
    net->forward(net->ws_x.as<float>(), 1, H, W);
    cnn_ws_unpack(net->stream, net->ws_x.as<float>(), net->y.as<float>(), n, net->ws_ss.as<double>(), net->ws_u.p,
                  net->ws_v.p);
    HIPC(hipGetLastError());
    HIPC(hipEventRecord(net->ev_out, net->stream));
    HIPC(hipStreamWaitEvent(s->stream, net->ev_out, 0));
    set_input0(s, net->ws_u.p, net->ws_v.p, W, H);  // on the solver stream; synchronises
  });
}

}  // extern "C"
