"""fmm2d_amd -- MI355X-native 2-D fast multipole methods.

Public names mirror python/fmm2dpy/__init__.py:1 of flatironinstitute/fmm2d,
so ``import fmm2d_amd as fmm2dpy`` is the drop-in switch.
"""
from .fmm2d import (hfmm2d, rfmm2d, lfmm2d, cfmm2d, bhfmm2d, stfmm2d, h2ddir, r2ddir,
                    l2ddir, c2ddir, bh2ddir, st2ddir, comperr, Output)
from .fmm2d import fmm2d_batch  # many FMMs per call; not in the reference
from .fmm2d import hfmm2d_mps  # helmholtz/hfmm2d_mps.f (no Python wrapper in the reference)

__version__ = "0.1.0"
