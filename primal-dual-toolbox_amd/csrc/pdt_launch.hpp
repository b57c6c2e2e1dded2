// pdt_launch.hpp -- host entry points into the HIP kernels (pdt_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "pdt_desc.hpp"

namespace pdt {

// Element types (match PDT_F32 / PDT_C64 / PDT_F64 / PDT_C128).
enum DType { DT_F32 = 0, DT_C64 = 1, DT_F64 = 2, DT_C128 = 3 };

struct Launch {
  hipStream_t stream = nullptr;
  const SliceDev* sl = nullptr;  // device array of slices
  int nsl = 0;
  const int2* rows = nullptr;    // (slice, first row) items
  int nrows = 0;
  const int2* rowgroups = nullptr;  // (slice, row * ngr + coil group) items (MRI row FFT passes)
  int nrg = 0;
  int rpi = 1;                   // rows per item (denoising kernels)
  const int2* cols = nullptr;    // (slice, coil * ngroups + group) items
  int ncolitems = 0;
  double* part[4] = {nullptr, nullptr, nullptr, nullptr};
  int cb = 1;                    // coils per row work item (coil group size)
  int cg = 1;                    // columns per group in the column kernel
  size_t lds_rows = 0;           // dynamic LDS bytes of the row kernels
  size_t lds_cols = 0;           // dynamic LDS bytes of the column kernel
  int fft_w = 0;                 // row FFT length shared by every slice of the launch (0: mixed)
  int fft_h = 0;                 // column FFT length shared by every slice (0: mixed)
  double* gap_rows = nullptr;    // row-item partial sums indexed by global row item (primal in k_rows_adj)
  bool split_primal = true;      // some slice has several coil groups: k_mri_primal completes the step
  int primal_threads = 256;      // block size of k_mri_primal: rpi * W rounded up to 64, at most 1024
  // adjoint row pass: its own coil groups (slice, row * ngr_a + group), LDS
  // bytes and block size (kRowThreads or kRowAdjThreadsWide)
  const int2* rowgroups_a = nullptr;
  int nrg_a = 0;
  int cb_a = 1;
  size_t lds_rows_a = 0;
  int adj_threads = 256;
  // single-slice engine: every list is (0, i) for i < n ((0, i rpi) for
  // the denoising rows), so the kernels take their item from the block
  // index (no list load)
  bool dense = false;
  // single dense slice: workgroups of k_rows_fwd that do the dual step alone
  // after the row items (0: the row items share it out)
  int ndual = 0;
  // single dense slice, TGV, wide adjoint blocks: workgroups of k_rows_adj
  // that do the v step alone after the row items (0: the row items do it)
  int nvstep = 0;
  // single dense slice, wide adjoint blocks, one coil group per row: the
  // straight-line adjoint row path (rows_adj_fused_ok; no v step in the row
  // blocks)
  bool adj_fused = false;
  // with adj_fused and ndual > 0 (single precision): an iteration without
  // the gap evaluation may run as two launches (run_pair_stage) instead of
  // three, the dual step in the column pass and the next forward row FFTs in
  // the adjoint row pass
  bool pair = false;
  // workgroup size of the denoising half-iterations (256, 512 or 1024): the
  // smallest that covers rpi rows of the widest slice in one pass, at most 1024
  int den_threads = 256;
  // TGV denoising dual step with two pixels per thread (every slice of the
  // launch has an even width)
  bool den_pair = false;
  // occupancy caps, workgroups per CU (0: none) of the forward row pass, the
  // column pass, the adjoint row pass and the denoising half-iterations: a
  // launch asks for enough LDS that no more than that many workgroups fit
  // the LDS of a CU (the autotuner's waves-per-CU axis)
  int wpc[4] = {0, 0, 0, 0};
  // temporally blocked denoising launches (one dense ROF or TGV slice):
  // interior tile tb_x x tb_y pixels, tb_k iterations per launch (0: off)
  int tb_x = 0, tb_y = 0, tb_k = 0;
  int tb_w = 0, tb_h = 0;  // the slice's image size (grid of tiles)
};
enum { WPC_FWD = 0, WPC_COLS = 1, WPC_ADJ = 2, WPC_DEN = 3 };

// Dynamic LDS bytes of a launch that needs `need` bytes and is capped at
// wpc workgroups per CU (0: no cap): at least `need`, raised to the share of
// the CU's LDS that leaves room for no more than wpc workgroups.  Caps that
// would need more than `limit` bytes (the dynamic LDS the kernel may ask
// for) are not applied.
size_t lds_occupancy(size_t need, int wpc, size_t limit);

// Block size of k_mri_primal covering rows of width W in one pass.
inline int primal_block(int W) { return W >= 1024 ? 1024 : ((W + 63) / 64) * 64; }

// One full iteration of the solver kind on every active slice.
void run_iteration(const Launch& L, int kind, int dtype, bool fp16, bool gap);
// One stage (fused kernel) of an iteration; stages run in order 0..n-1.
int num_stages(int kind);
const char* stage_name(int kind, int stage);
void run_stage(const Launch& L, int kind, int dtype, bool fp16, bool gap, int stage);
// One launch of a paired MRI iteration (L.pair): which 0 = the column pass
// with the dual step (reads the T1 of the current ubar), 1 = the adjoint row
// pass with the primal step that also writes the T1 of the new ubar; gap:
// the gap-evaluation variants (partial sums where the three passes put them).
// The first iteration of a solve starts from stages 0 and 1 instead of 0.
void run_pair_stage(const Launch& L, int kind, int dtype, int which, bool gap);
// tb_k ROF or TGV (kind) iterations of a single dense slice in one launch
// (tiles held in LDS with a tb_k-pixel halo), reading the state (ROF: u,
// ubar, p; TGV: u, v, p, q) and writing its second copy with ubar (and vbar)
// (dir = 0) or the reverse (dir = 1); fp16: TGV duals stored as half.
void run_block(const Launch& L, int kind, int dtype, bool fp16, int dir);
// LDS bytes of that launch, and whether a tile shape fits (the LDS a
// workgroup may take, and for TGV the pixels its threads can own).
size_t block_lds(int kind, int dtype, int tx, int ty, int k);
bool block_fits(int kind, int dtype, int tx, int ty, int k, size_t lds_limit);
// Gap evaluation from the partial sums of the last iteration (one block/slice).
void run_gap_reduce(const Launch& L, int kind, int dtype);
// Operator application in operator mode (op_in -> op_out of every slice).
void run_op_forward(const Launch& L, int dtype);
void run_op_adjoint(const Launch& L, int dtype);
// Gathers full-layout data {W, Hk, C} into the compact [c][ks][kfe] layout,
// zeroes r, and writes per-block partial sums of |f|^2 over columns without
// samples into f0_part (nblocks entries, returned).
int run_compact_data(const Launch& L, int dtype, const void* f_full, double* f0_part);
// Writes slice descriptor d to dst and opens its control words ctl (a
// kernel: nothing is copied from the host).
void run_store_desc(hipStream_t st, const SliceDev& d, SliceDev* dst, int* ctl);

// Bytes of the per-GPU scratch array of the allocation-free operator
// application, and the largest device index it serves.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
// Device address of that array on `device`, which must be the current
// device (looked up once per device); null on failure.
void* noalloc_scratch(int device);

// Fills n elements of the given size (bytes per element 4, 8 or 16) with 0.
void run_zero(hipStream_t st, void* ptr, size_t bytes);
// out[0] = sum of a[i * stride], i < n (one block, fixed order).
void run_sum_strided(hipStream_t st, const double* a, int n, int stride, double* out);
// out[0] = sum of |z_i|^2 over n complex pixels (DT_C64 or DT_C128).
void run_sumsq(hipStream_t st, int dtype, const void* z, int64_t n, double* out);

// True when the row/column kernels of this precision have a compile-time
// FFT plan for length n (the kernels fall back to the runtime plan otherwise).
bool fft_len_compiled(int dtype, int n);
// Whether the straight-line adjoint row path fits a slice with rows of W
// pixels, C coils (all in one coil group) and nc listed PE columns.
bool rows_adj_fused_ok(int dtype, int W, int C, int nc);

// Allows the MRI kernels the full LDS of a workgroup (call before launches);
// returns the dynamic LDS bytes a workgroup may use.
size_t init_kernel_attrs();

// LDS bytes needed by the row / column kernels.
// (padded: the FFT buffers with one pad slot per 8 entries; the wide adjoint
// row pass uses them unpadded)
size_t lds_rows_bytes(int dtype, int cb, int W, int nc, bool padded = true);
size_t lds_cols_bytes(int dtype, int cg, int Hk);

}  // namespace pdt
