"""FP64 host oracle for the PME reciprocal solve and the direct-space Ewald
sum (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package.  It is the checker; the product path in ``gpu_recip_amd``
never imports it.
"""
from .oracle import (  # noqa: F401
    OracleParams,
    INDEX_F32,
    INDEX_F64,
    NYQUIST_REFERENCE,
    NYQUIST_EXACT,
    CCELEC,
    build,
    lib,
    index_f32,
    index_f64,
    bspline,
    bsp_mod,
    spread,
    fft1d,
    fft3d_r2c,
    fft3d_c2r,
    scalar_sum,
    gather,
    self_energy,
    solve,
    ewald_recip,
    ewald_recip9,
    direct_tiles,
    direct_ewald,
    pair14,
)
