// pdt_cnn.hpp -- host entry points into the warm-start network kernels
// (pdt_cnn.hip).  Activations are NHWC; bf16 buffers are passed as void*.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pdt {

// Layer shapes with kernels: 2 -> 32, 32 -> 32, 32 -> 6.
bool cnn_layer_supported(int cin, int cout);
// Packed bf16 weight sizes (elements) of a layer, forward and backward-data.
size_t cnn_wk_elems(int cin, int cout);
size_t cnn_wt_elems(int cin, int cout);
// fp32 master W[tap][co][ci] -> packed bf16 copies (wt may be null).
void cnn_pack_weights(hipStream_t st, const float* W, int cin, int cout, void* wk, void* wt);
void cnn_to_bf16(hipStream_t st, const float* x, void* y, int64_t n);
// out = conv(in) + b, then ReLU and bf16 (relu) or fp32 (last layer).
void cnn_conv_forward(hipStream_t st, int cin, int cout, bool relu, const void* in, const void* wk, const float* b,
                      void* out, int n, int h, int w);
// din = (conv^T dz) * [act > 0], bf16.
void cnn_conv_backward_data(hipStream_t st, int cin, int cout, const void* dz, const void* wt, const void* act,
                            void* din, int n, int h, int w);
// grad (master layout W then b) of one layer; part holds cnn_wgrad_part_floats.
int cnn_wgrad_blocks(int n, int h, int w);
size_t cnn_wgrad_part_floats(int cin, int cout, int nblk);
void cnn_conv_wgrad(hipStream_t st, int cin, int cout, const void* in, const void* dz, float* part, int nblk,
                    float* grad, int n, int h, int w);
// MSE 1/(2m) sum (y - t)^2: fp64 per-block partials and dz = (y - t) / m (bf16).
int cnn_mse_blocks(int64_t m);
void cnn_mse(hipStream_t st, const float* y, const float* t, int64_t m, void* dz, double* part, int nblk);
// One Adam update; c1 = 1 - b1^t, c2 = 1 - b2^t.
void cnn_adam(hipStream_t st, float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
              float eps, float c1, float c2);
// Warm start from a complex64 image u0 (n pixels); ss points to the sum of
// |u0|^2 on the device, from which the input scale RMS(u0) is taken.
void cnn_ws_pack(hipStream_t st, const void* u0, int64_t n, const double* ss, float* x);
void cnn_ws_unpack(hipStream_t st, const float* x, const float* y, int64_t n, const double* ss, void* u, void* v);

}  // namespace pdt
