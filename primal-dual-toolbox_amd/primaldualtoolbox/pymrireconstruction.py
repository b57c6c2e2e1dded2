"""Cartesian MRI reconstruction (the reference's ``pymrireconstruction``
module, python/pyMriReconstruction.cpp:112-268; the radial gpuNUFFT operators
are not part of this package)."""
from ._core import (MriCartesianOperator, MriCartesianRemoveROOSOperator,  # noqa: F401
                    MriSamplingOperator, TgvMriOptimizer_2c2c, TgvMriOptimizer_2c3c,
                    TgvMriParameters, TgvOptimizer_2c2c, TgvOptimizer_2c3c, TgvParameters,
                    TvOptimizer_2c2c, TvOptimizerWithOp_2c3c, TvParameters, autotune, batch_profile, batch_solve,
                    set_tiles, solve_concurrent)
