"""hestonexotics_amd -- MI355X-native Heston QE Monte-Carlo pricing (gfx950 HIP behind a C ABI).

The product is libhexo.so (hestonexotics_amd/csrc); this package is its Python binding.
"""
from .api import (ASIAN, ASIAN_GEO_CV, CV_REGRESSION, DISCOUNT, DRIFT, EUROPEAN, FP32, FP32_COMPENSATED, FP64, GEO_ASIAN,
                  MARTINGALE, TRAPEZOID, Batch, Context, HexoError, Request, SwiftCall, SwiftChain, asian_weights,
                  cv_optimal, swift_grid, time_grid, validate)

__all__ = ["ASIAN", "ASIAN_GEO_CV", "CV_REGRESSION", "DISCOUNT", "DRIFT", "EUROPEAN", "FP32", "FP32_COMPENSATED", "FP64",
           "GEO_ASIAN", "MARTINGALE", "TRAPEZOID", "Batch", "Context", "HexoError", "Request", "SwiftCall",
           "SwiftChain", "asian_weights", "cv_optimal", "swift_grid", "time_grid", "validate"]
