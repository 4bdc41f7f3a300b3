"""ofp -- MI355X frontier transfer-matrix solver for corner-to-corner simple
paths on an (n+1)x(n+1) grid (OEIS A007764), host side.

The compute path is libofp.so (hand-written HIP kernels for gfx950 behind the
C ABI in include/ofp.h); this package binds it and carries the host logic
around it.
"""
from ._lib import (  # noqa: F401
    EXPORTS, OFP_MAX_N, Context, Job, OfpError, Plan, group_plan, height_cap, load, moduli_groups, state_counts,
    version,
)
