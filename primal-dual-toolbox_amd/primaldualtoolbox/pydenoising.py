"""TV / TGV denoising (the reference's ``pydenoising`` module,
python/pyDenoising.cpp:39-94)."""
from ._core import (TgvOptimizer_2c, TgvOptimizer_2f, TgvParameters, TvOptimizer_2c,  # noqa: F401
                    TvOptimizer_2f, TvParameters, autotune, batch_solve, set_tiles, solve_concurrent)
