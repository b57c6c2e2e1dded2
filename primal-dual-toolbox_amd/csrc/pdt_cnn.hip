// pdt_cnn.hip -- the warm-start network: conv3x3 layers (zero padding,
// stride 1) with ReLU between them, as implicit GEMMs on the bf16 MFMA
// (v_mfma_f32_16x16x32_bf16, fp32 accumulation), its backward pass, an MSE
// loss and Adam.
//
// Layout: activations NHWC (channel fastest), bf16 between layers, fp32 at
// the output.  Master parameters fp32, per layer W[tap][co][ci] (tap =
// 3 ky + kx) then b[co].  For the MFMA the weights are repacked to bf16
// Wk[co][k], k = tap * CIN + ci (zero-padded to 32 in k and 16 in co), and
// for the backward-data pass to Wt[ci][k'], k' = (8 - tap) * COUT + co.
//
// MFMA operand layout (v_mfma_f32_16x16x32_bf16, wave64): lane l holds
// A[l % 16][(l / 16) * 8 + i] and B[(l / 16) * 8 + i][l % 16], i < 8, and
// receives C[(l / 16) * 4 + i][l % 16], i < 4.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pdt_cnn.hpp"

namespace pdt {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kConvThreads = 256;  // 4 waves, 64 pixels each
constexpr int kPixPerBlock = 256;

__device__ __forceinline__ f4 mfma(const bf8& a, const bf8& b, const f4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int kpad(int cin) { return ((9 * cin + 31) / 32) * 32; }

// Epilogues of conv_kernel.
enum { EPI_RELU_BF16 = 0, EPI_F32 = 1, EPI_MASK_BF16 = 2 };

// out[p][co] = epi(sum_{tap, ci} Wk[co][tap * CIN + ci] * in[p + off(tap)][ci])
// for the pixels p of n images of h x w.  CIN = 32 takes one tap per k-step
// (16-byte channel loads); other CIN gather k elements one by one.
template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(kConvThreads) void conv_kernel(const __bf16* __restrict__ in, const __bf16* __restrict__ wk,
                                                            const float* __restrict__ bias, void* out,
                                                            const __bf16* __restrict__ mask, int n, int h, int w) {
  constexpr int KP = (9 * CIN + 31) / 32 * 32;  // padded k
  constexpr int NTN = (COUT + 15) / 16;         // 16-wide output-channel tiles
  constexpr int COP = NTN * 16;
  __shared__ float stage[4][64][COP + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t P = int64_t(n) * h * w;
  const int64_t pw = int64_t(blockIdx.x) * kPixPerBlock + wave * 64;  // first pixel of this wave
  f4 acc[4][NTN];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < NTN; ++nt) acc[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
  // pixel coordinates of the 4 A rows this lane loads (row l % 16 of each M tile)
  int py[4], px[4];
  int64_t pimg[4];
  bool pin[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const int64_t p = pw + mt * 16 + (lane & 15);
    pin[mt] = p < P;
    const int64_t pp = pin[mt] ? p : 0;
    const int64_t img = pp / (int64_t(h) * w);
    const int r = int(pp - img * int64_t(h) * w);
    py[mt] = r / w;
    px[mt] = r - py[mt] * w;
    pimg[mt] = img * int64_t(h) * w;
  }
  const int kq = (lane >> 4) * 8;  // this lane's first k within a 32-wide k-step
#pragma unroll 1
  for (int ks = 0; ks < KP / 32; ++ks) {
    bf8 b[NTN];
#pragma unroll
    for (int nt = 0; nt < NTN; ++nt) {
      const int co = nt * 16 + (lane & 15);
      b[nt] = *reinterpret_cast<const bf8*>(wk + size_t(co) * KP + ks * 32 + kq);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      bf8 a;
      if constexpr (CIN == 32) {
        const int tap = ks, dy = tap / 3 - 1, dx = tap % 3 - 1;
        const int yy = py[mt] + dy, xx = px[mt] + dx;
        if (pin[mt] && yy >= 0 && yy < h && xx >= 0 && xx < w)
          a = *reinterpret_cast<const bf8*>(in + (pimg[mt] + int64_t(yy) * w + xx) * 32 + kq);
        else
          for (int i = 0; i < 8; ++i) a[i] = __bf16(0.f);
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int k = ks * 32 + kq + i;
          float v = 0.f;
          if (k < 9 * CIN && pin[mt]) {
            const int tap = k / CIN, ci = k - tap * CIN;
            const int yy = py[mt] + tap / 3 - 1, xx = px[mt] + tap % 3 - 1;
            if (yy >= 0 && yy < h && xx >= 0 && xx < w) v = float(in[(pimg[mt] + int64_t(yy) * w + xx) * CIN + ci]);
          }
          a[i] = __bf16(v);
        }
      }
#pragma unroll
      for (int nt = 0; nt < NTN; ++nt) acc[mt][nt] = mfma(a, b[nt], acc[mt][nt]);
    }
  }
  // C tiles -> LDS [pixel][channel] -> epilogue with contiguous channel runs
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < NTN; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) stage[wave][mt * 16 + (lane >> 4) * 4 + i][nt * 16 + (lane & 15)] = acc[mt][nt][i];
  __syncthreads();
  for (int e = lane; e < 64 * COUT; e += 64) {
    const int pl = e / COUT, co = e - pl * COUT;
    const int64_t p = pw + pl;
    if (p >= P) continue;
    const float v = stage[wave][pl][co];
    if constexpr (EPI == EPI_RELU_BF16) {
      static_cast<__bf16*>(out)[p * COUT + co] = __bf16(fmaxf(v + bias[co], 0.f));
    } else if constexpr (EPI == EPI_F32) {
      static_cast<float*>(out)[p * COUT + co] = v + bias[co];
    } else {
      const bool on = float(mask[p * COUT + co]) > 0.f;
      static_cast<__bf16*>(out)[p * COUT + co] = __bf16(on ? v : 0.f);
    }
  }
}

// Weight gradient partials: for the strips (image, row, 64 columns) of this
// block, dWk[co][k] += sum_p dz[p][co] in[p + off(tap)][ci] and
// db[co] += sum_p dz[p][co]; one partial per block, reduced in order later.
constexpr int kStrip = 64;
template <int CIN, int COUT>
__global__ __launch_bounds__(kConvThreads) void wgrad_kernel(const __bf16* __restrict__ in, const __bf16* __restrict__ dz,
                                                             float* part, int n, int h, int w) {
  constexpr int KP = (9 * CIN + 31) / 32 * 32;
  constexpr int NTN = KP / 16;            // k tiles (GEMM columns)
  constexpr int MTN = (COUT + 15) / 16;   // co tiles (GEMM rows)
  constexpr int TILES = MTN * NTN;
  constexpr int TPW = (TILES + 3) / 4;    // tiles per wave
  __shared__ __bf16 sin[3][kStrip + 2][CIN];
  __shared__ __bf16 sdz[kStrip][COUT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sx = (w + kStrip - 1) / kStrip;
  const int64_t nstrips = int64_t(n) * h * sx;
  f4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  float db = 0.f;  // thread co < COUT of wave 0: bias gradient
  for (int64_t s = blockIdx.x; s < nstrips; s += gridDim.x) {
    const int img = int(s / (int64_t(h) * sx));
    const int rem = int(s - int64_t(img) * h * sx);
    const int y = rem / sx, x0 = (rem - y * sx) * kStrip;
    const int64_t ibase = int64_t(img) * h * w;
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * (kStrip + 2) * CIN; e += kConvThreads) {
      const int r = e / ((kStrip + 2) * CIN), rr = e - r * (kStrip + 2) * CIN;
      const int c = rr / CIN, ci = rr - c * CIN;
      const int yy = y + r - 1, xx = x0 + c - 1;
      sin[r][c][ci] = (yy >= 0 && yy < h && xx >= 0 && xx < w) ? in[(ibase + int64_t(yy) * w + xx) * CIN + ci] : __bf16(0.f);
    }
    for (int e = threadIdx.x; e < kStrip * COUT; e += kConvThreads) {
      const int c = e / COUT, co = e - c * COUT;
      const int xx = x0 + c;
      sdz[c][co] = xx < w ? dz[(ibase + int64_t(y) * w + xx) * COUT + co] : __bf16(0.f);
    }
    __syncthreads();
    if (wave == 0 && lane < COUT)
      for (int c = 0; c < kStrip; ++c) db += float(sdz[c][lane]);
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int tile = wave * TPW + t;
      if (tile >= TILES) break;
      const int mt = tile / NTN, nt = tile - mt * NTN;
      const int co = mt * 16 + (lane & 15);
      const int k = nt * 16 + (lane & 15);
      const int tap = k / CIN, ci = k - tap * CIN;
      const bool kin = k < 9 * CIN;
#pragma unroll
      for (int ks = 0; ks < kStrip / 32; ++ks) {
        bf8 a, b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int c = ks * 32 + (lane >> 4) * 8 + i;  // pixel within the strip
          a[i] = co < COUT ? sdz[c][co] : __bf16(0.f);
          b[i] = kin ? sin[tap / 3][c + tap % 3][ci] : __bf16(0.f);
        }
        acc[t] = mfma(a, b, acc[t]);
      }
    }
  }
  // partial of this block: [co][k] fp32, then db
  float* pb = part + size_t(blockIdx.x) * (size_t(MTN) * 16 * KP + COUT);
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = wave * TPW + t;
    if (tile >= TILES) break;
    const int mt = tile / NTN, nt = tile - mt * NTN;
#pragma unroll
    for (int i = 0; i < 4; ++i) pb[size_t(mt * 16 + (lane >> 4) * 4 + i) * KP + nt * 16 + (lane & 15)] = acc[t][i];
  }
  if (wave == 0 && lane < COUT) pb[size_t(MTN) * 16 * KP + lane] = db;
}

// Sums the per-block partials into the master-layout gradient dW[tap][co][ci]
// and db[co].  A workgroup takes kRedOut consecutive outputs; its kRedGroups
// thread groups each sum every kRedGroups-th block partial, in block order
// (a group reads kRedOut neighbouring floats of a partial row at a time),
// and the group sums are added in group order: a fixed summation order, so
// the gradient is reproducible run to run.
constexpr int kRedOut = 64, kRedGroups = 16;

__global__ __launch_bounds__(kRedOut * kRedGroups) void wgrad_reduce(const float* part, int nblk, int cin, int cout,
                                                                     float* g) {
  __shared__ double red[kRedGroups][kRedOut];
  const int KP = kpad(cin), MT = (cout + 15) / 16 * 16;
  const int nw = 9 * cout * cin;
  const size_t stride = size_t(MT) * KP + cout;
  const int o = threadIdx.x % kRedOut, grp = threadIdx.x / kRedOut;
  const int e = blockIdx.x * kRedOut + o;
  double s = 0;
  if (e < nw + cout) {
    size_t src;
    if (e < nw) {
      const int tap = e / (cout * cin), r = e - tap * cout * cin, co = r / cin, ci = r - co * cin;
      src = size_t(co) * KP + tap * cin + ci;
    } else {
      src = size_t(MT) * KP + (e - nw);
    }
    for (int b = grp; b < nblk; b += kRedGroups) s += part[size_t(b) * stride + src];
  }
  red[grp][o] = s;
  __syncthreads();
  if (grp == 0 && e < nw + cout) {
    double t = 0;
    for (int j = 0; j < kRedGroups; ++j) t += red[j][o];
    g[e] = float(t);
  }
}

// Master fp32 W[tap][co][ci] -> bf16 Wk[co][tap * cin + ci] (padded) and
// Wt[ci][(8 - tap) * cout + co] (padded).
__global__ __launch_bounds__(256) void pack_weights(const float* W, int cin, int cout, __bf16* wk, __bf16* wt) {
  const int KP = kpad(cin), COP = (cout + 15) / 16 * 16;
  const int KPt = kpad(cout), CIP = (cin + 15) / 16 * 16;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < COP * KP) {
    const int co = e / KP, k = e - co * KP;
    float v = 0.f;
    if (co < cout && k < 9 * cin) {
      const int tap = k / cin, ci = k - tap * cin;
      v = W[(tap * cout + co) * cin + ci];
    }
    wk[e] = __bf16(v);
  }
  if (wt && e < CIP * KPt) {
    const int ci = e / KPt, k = e - ci * KPt;
    float v = 0.f;
    if (ci < cin && k < 9 * cout) {
      const int tp = k / cout, co = k - tp * cout;
      v = W[((8 - tp) * cout + co) * cin + ci];
    }
    wt[e] = __bf16(v);
  }
}

__global__ __launch_bounds__(256) void to_bf16(const float* x, __bf16* y, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __bf16(x[i]);
}

// Loss 1/(2M) sum (y - t)^2 over M = P * c elements: per-block fp64 partial
// sums, and dz = (y - t) / M in bf16.
__global__ __launch_bounds__(256) void mse_kernel(const float* y, const float* t, int64_t m, __bf16* dz,
                                                  double* part) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x) {
    const float d = y[i] - t[i];
    s += double(d) * d;
    dz[i] = __bf16(d / float(m));
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba, Algorithm 1) on fp32 master parameters.
__global__ __launch_bounds__(256) void adam_kernel(float* w, const float* g, float* m, float* v, int64_t n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Warm-start input: x[p] = (re, im) of u0[p] / scale (complex64 planar ->
// NHWC fp32, 2 channels), and per-block partial sums of |u0|^2.
// Scale of the warm-start network input: the RMS of A^*f from its sum of
// squares ss (n pixels), 1 for an all-zero image.
__device__ __forceinline__ double ws_scale(const double* ss, int64_t n) {
  return ss[0] > 0 ? sqrt(ss[0] / double(n)) : 1.0;
}
__global__ __launch_bounds__(256) void ws_pack(const float2* u0, int64_t n, const double* ss, float* x) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float inv_scale = float(1.0 / ws_scale(ss, n));
  const float2 z = u0[i];
  x[2 * i] = z.x * inv_scale;
  x[2 * i + 1] = z.y * inv_scale;
}
__global__ __launch_bounds__(256) void ws_unpack(const float* x, const float* y, int64_t n, const double* ss,
                                                 float2* u, float2* v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float scale = float(ws_scale(ss, n));
  const float* yi = y + 6 * i;
  u[i] = make_float2(scale * (x[2 * i] + yi[0]), scale * (x[2 * i + 1] + yi[1]));
  v[i] = make_float2(scale * yi[2], scale * yi[3]);
  v[n + i] = make_float2(scale * yi[4], scale * yi[5]);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int blocks_for(int64_t pixels) { return int((pixels + kPixPerBlock - 1) / kPixPerBlock); }

bool cnn_layer_supported(int cin, int cout) {
  return (cin == 2 && cout == 32) || (cin == 32 && cout == 32) || (cin == 32 && cout == 6);
}

void cnn_conv_forward(hipStream_t st, int cin, int cout, bool relu, const void* in, const void* wk, const float* b,
                      void* out, int n, int h, int w) {
  const int g = blocks_for(int64_t(n) * h * w);
  const __bf16* x = static_cast<const __bf16*>(in);
  const __bf16* k = static_cast<const __bf16*>(wk);
  if (cin == 2 && cout == 32 && relu)
    hipLaunchKernelGGL((conv_kernel<2, 32, EPI_RELU_BF16>), dim3(g), dim3(kConvThreads), 0, st, x, k, b, out, nullptr, n, h, w);
  else if (cin == 32 && cout == 32 && relu)
    hipLaunchKernelGGL((conv_kernel<32, 32, EPI_RELU_BF16>), dim3(g), dim3(kConvThreads), 0, st, x, k, b, out, nullptr, n, h, w);
  else if (cin == 32 && cout == 6 && !relu)
    hipLaunchKernelGGL((conv_kernel<32, 6, EPI_F32>), dim3(g), dim3(kConvThreads), 0, st, x, k, b, out, nullptr, n, h, w);
}

void cnn_conv_backward_data(hipStream_t st, int cin, int cout, const void* dz, const void* wt, const void* act,
                            void* din, int n, int h, int w) {
  // input gradient of a layer cin -> cout is a conv cout -> cin with Wt,
  // masked by the ReLU of the layer input
  const int g = blocks_for(int64_t(n) * h * w);
  const __bf16* z = static_cast<const __bf16*>(dz);
  const __bf16* k = static_cast<const __bf16*>(wt);
  const __bf16* a = static_cast<const __bf16*>(act);
  if (cin == 32 && cout == 32)
    hipLaunchKernelGGL((conv_kernel<32, 32, EPI_MASK_BF16>), dim3(g), dim3(kConvThreads), 0, st, z, k, nullptr, din, a, n, h, w);
  else if (cin == 32 && cout == 6)
    hipLaunchKernelGGL((conv_kernel<6, 32, EPI_MASK_BF16>), dim3(g), dim3(kConvThreads), 0, st, z, k, nullptr, din, a, n, h, w);
}

int cnn_wgrad_blocks(int n, int h, int w) {
  const int64_t strips = int64_t(n) * h * ((w + kStrip - 1) / kStrip);
  return int(strips < 1024 ? strips : 1024);
}

size_t cnn_wgrad_part_floats(int cin, int cout, int nblk) {
  return size_t(nblk) * (size_t((cout + 15) / 16 * 16) * ((9 * cin + 31) / 32 * 32) + cout);
}

void cnn_conv_wgrad(hipStream_t st, int cin, int cout, const void* in, const void* dz, float* part, int nblk,
                    float* grad, int n, int h, int w) {
  const __bf16* x = static_cast<const __bf16*>(in);
  const __bf16* z = static_cast<const __bf16*>(dz);
  if (cin == 2 && cout == 32)
    hipLaunchKernelGGL((wgrad_kernel<2, 32>), dim3(nblk), dim3(kConvThreads), 0, st, x, z, part, n, h, w);
  else if (cin == 32 && cout == 32)
    hipLaunchKernelGGL((wgrad_kernel<32, 32>), dim3(nblk), dim3(kConvThreads), 0, st, x, z, part, n, h, w);
  else if (cin == 32 && cout == 6)
    hipLaunchKernelGGL((wgrad_kernel<32, 6>), dim3(nblk), dim3(kConvThreads), 0, st, x, z, part, n, h, w);
  const int ne = 9 * cin * cout + cout;
  hipLaunchKernelGGL(wgrad_reduce, dim3((ne + kRedOut - 1) / kRedOut), dim3(kRedOut * kRedGroups), 0, st, part, nblk,
                     cin, cout, grad);
}

size_t cnn_wk_elems(int cin, int cout) { return size_t((cout + 15) / 16 * 16) * ((9 * cin + 31) / 32 * 32); }
size_t cnn_wt_elems(int cin, int cout) { return size_t((cin + 15) / 16 * 16) * ((9 * cout + 31) / 32 * 32); }

void cnn_pack_weights(hipStream_t st, const float* W, int cin, int cout, void* wk, void* wt) {
  const size_t ne = cnn_wk_elems(cin, cout) > cnn_wt_elems(cin, cout) ? cnn_wk_elems(cin, cout) : cnn_wt_elems(cin, cout);
  hipLaunchKernelGGL(pack_weights, dim3(int((ne + 255) / 256)), dim3(256), 0, st, W, cin, cout,
                     static_cast<__bf16*>(wk), static_cast<__bf16*>(wt));
}

void cnn_to_bf16(hipStream_t st, const float* x, void* y, int64_t n) {
  hipLaunchKernelGGL(to_bf16, dim3(int((n + 255) / 256)), dim3(256), 0, st, x, static_cast<__bf16*>(y), n);
}

int cnn_mse_blocks(int64_t m) {
  const int64_t b = (m + 255) / 256;
  return int(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

void cnn_mse(hipStream_t st, const float* y, const float* t, int64_t m, void* dz, double* part, int nblk) {
  hipLaunchKernelGGL(mse_kernel, dim3(nblk), dim3(256), 0, st, y, t, m, static_cast<__bf16*>(dz), part);
}

void cnn_adam(hipStream_t st, float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
              float eps, float c1, float c2) {
  hipLaunchKernelGGL(adam_kernel, dim3(int((n + 255) / 256)), dim3(256), 0, st, w, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

void cnn_ws_pack(hipStream_t st, const void* u0, int64_t n, const double* ss, float* x) {
  hipLaunchKernelGGL(ws_pack, dim3(int((n + 255) / 256)), dim3(256), 0, st, static_cast<const float2*>(u0), n, ss,
                     x);
}

void cnn_ws_unpack(hipStream_t st, const float* x, const float* y, int64_t n, const double* ss, void* u, void* v) {
  hipLaunchKernelGGL(ws_unpack, dim3(int((n + 255) / 256)), dim3(256), 0, st, x, y, n, ss,
                     static_cast<float2*>(u), static_cast<float2*>(v));
}

}  // namespace pdt
