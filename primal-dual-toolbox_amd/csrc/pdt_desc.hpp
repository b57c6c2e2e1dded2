// pdt_desc.hpp -- descriptors shared by the host engine and the HIP kernels.
//
// One "slice" is one reconstruction problem (one image, its data, its duals).
// A single solve is a batch of one slice; a batched solve packs many slices of
// uneven size and coil count into the same launches.  Kernels receive an array
// of slice descriptors plus a work list of (slice, index) items, so every
// launch is a persistent work-list over the whole batch.
#pragma once

#include <cstdint>

namespace pdt {

constexpr int kBlock = 256;       // threads per workgroup of the stencil kernels (4 wave64)
constexpr int kRowThreads = 256;  // threads per workgroup of the MRI row FFT passes
constexpr int kRowAdjThreadsWide = 1024;  // adjoint row pass with many coils per item
constexpr int kColThreads = 256;  // threads per workgroup of the MRI column pass
constexpr int kPrefetch = 8;       // row-pass elements per thread held in registers ahead of use
constexpr int kColPrefetchCols = 2;  // column groups up to this width prefetch their prox operands
constexpr int kMaxStages = 14;    // FFT stages per transform
constexpr int kGapTerms = 8;      // partial-sum slots per work item
constexpr int kMaxRowLen = 4096;  // W limit of the MRI row kernels
constexpr size_t kLdsMax = 160 * 1024;  // LDS per workgroup on gfx950
// LDS elements of an FFT buffer of n entries: one pad slot per 8 entries
// (the compile-time-planned kernels address element e at e + e / 8).
constexpr int lds_pad(int n) { return n + n / 8 + 1; }

// Solver kinds (match PDT_SOLVER_* of pdt_capi.h).
enum Kind { K_TV = 0, K_TGV = 1, K_TGV_MRI = 2, K_TV_OP = 3, K_TGV_OP = 4 };

// Mixed-radix Stockham plan of one transform length n.
// Per stage s: radix r, nr = n / r, p = product of the earlier radices,
// tws = n / (p r) (twiddle stride), and float reciprocals of nr and p for
// division-free index arithmetic.
struct FftPlanD {
  const void* tw;  // Cx<R>[n]: exp(-2 pi i m / n)
  int n;
  int ns;
  int r[kMaxStages];
  int nr[kMaxStages];
  int p[kMaxStages];
  int tws[kMaxStages];
  float rnr[kMaxStages];
  float rp[kMaxStages];
};

// Control words of a slice (device memory).
enum Ctl { CTL_ACTIVE = 0, CTL_ITER = 1, CTL_RUPD = 2, CTL_DONE_AT = 3, CTL_NGAP = 4, CTL_WORDS = 8 };
// Statistics doubles of a slice (device memory), written by the gap reduction.
enum Stat { ST_P = 0, ST_D = 1, ST_GAP = 2, ST_REL = 3, ST_C0 = 4, ST_C1 = 5, ST_C2 = 6, ST_WORDS = 8 };

// Field order: what the iteration kernels read first (control words, device
// pointers, sizes, scalar parameters) sits in the first few cache lines, so
// a block's scalar loads of its descriptor touch few lines; the FFT plans,
// read in full only by the runtime-planned kernels, come last.
struct SliceDev {
  // control and results
  int* ctl;
  double* stats;
  // k-space (operator solvers)
  const int* cols;    // ks -> kx
  const int* colmap;  // kx -> ks or -1
  void* fk;           // compact data [c][ks][kfe]
  void* rk;           // compact dual r [c][ks][kfe]
  const void* S;      // coil sensitivities {W, Hk, C}; null for the sampling op
  const void* Mc;     // compact mask [ks][kfe] (real)
  void* T1;           // row-pass output [c][y][ks] (row-major per coil, as T2)
  void* T2;           // column-pass output [c][y][ks]
  void* wpart;        // per-coil-group partial adjoints [g][y][x]
  const void* op_in;  // operator mode input
  void* op_out;       // operator mode output
  // image fields, planes of W*H pixels
  void* u;
  void* ub;
  void* v;   // 2 planes (x, y)
  void* vb;  // 2 planes
  void* p;   // 2 planes
  void* q;   // 3 planes (xx, yy, xy)
  const void* f;  // denoising data
  void* tmp;      // 2 planes: TGV dual-restoration scratch
  // the second copy of the state that the temporally blocked denoising
  // launches write and read in turn (ping-pong): u, ubar, p (2 planes) and,
  // for TGV, v, vbar (2 planes each) and q (3 planes)
  void* u2;
  void* ub2;
  void* p2;
  void* v2;
  void* vb2;
  void* q2;
  // geometry
  int W, H;       // image {W (x, contiguous), H (y)}
  int Hk, y0, C;  // k-space FE length, FE offset of the image window, coils
  int ncols;      // listed (sampled) PE columns
  int ngroups;    // column groups of the column pass (per coil)
  int ngr;        // coil groups of the forward row pass
  int ngr_a;      // coil groups of the adjoint row pass (planes of wpart; 1: u step fused)
  int kind;
  // scalar parameters
  double tau, sigma, theta, lambda;
  double alpha0, alpha1;          // alpha at k = 0
  double alpha0_end, alpha1_end;  // alpha * reduction (TGV-MRI schedule)
  double kmax;                    // max_iter as a real, for the schedule
  int schedule;                   // 1: TGV-MRI alpha continuation
  int dual_fp16;
  double radius_u, radius_v;      // partial-gap ball radii
  double F0;                      // ||f||^2 on columns without samples
  double gap_tol;
  // partial-sum ranges of this slice in the work lists: image rows,
  // (row, coil group) items and column items
  int row_beg, row_end, rg_beg, rg_end, col_beg, col_end;
  FftPlanD planW, planH;
};

}  // namespace pdt
