// pdt_tune.hpp -- tile-shape cache of the autotuner.
//
// Tiles are keyed by the problem shape; values are the coils per LDS chunk
// of the forward (cb) and adjoint (cb_a) row kernels, the columns per group
// (cg) of the column kernel, the rows per item (rpi) of the denoising
// kernels and the occupancy caps (workgroups per CU) of the forward row,
// column, adjoint row and denoising passes.  A value of 0 means "use the
// built-in default" (no cap, for the occupancy, and two launches per MRI
// iteration where possible).  The cache is process-wide and, when the environment
// variable PDT_TUNE_CACHE names a file, loaded from / saved to that file, so a
// tuned choice is deterministic across runs.
#pragma once

#include <string>

namespace pdt {

struct TuneKey {
  int kind, dt, W, H, Hk, C, nc, ns;
  bool operator<(const TuneKey& o) const;
  std::string str() const;
};

struct TuneVal {
  int cb = 0, cg = 0, rpi = 0, cb_a = 0;
  int wpc[4] = {0, 0, 0, 0};  // workgroups per CU of the passes (0: no cap)
  int tb[3] = {0, 0, -1};     // temporally blocked ROF tiles x, y, iterations (-1: default, 0: off)
  int dual_px = 0;            // TGV denoising dual step: pixels per thread (0: default)
  int unpaired = 0;           // MRI: three launches per iteration even where two would do
};

TuneVal tune_lookup(const TuneKey& k);
void tune_store(const TuneKey& k, const TuneVal& v);

}  // namespace pdt
