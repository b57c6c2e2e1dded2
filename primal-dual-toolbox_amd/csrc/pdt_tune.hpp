// pdt_tune.hpp -- tile-shape cache of the autotuner.
//
// Tiles are keyed by the problem shape; values are the coils per LDS chunk
// of the forward (cb) and adjoint (cb_a) row kernels, the columns per group
// (cg) of the column kernel and the rows per item (rpi) of the denoising
// kernels.  A value of 0 means "use
// the built-in default".  The cache is process-wide and, when the environment
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
};

TuneVal tune_lookup(const TuneKey& k);
void tune_store(const TuneKey& k, const TuneVal& v);

}  // namespace pdt
