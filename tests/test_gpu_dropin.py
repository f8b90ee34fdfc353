"""The C++ drop-in (include/pme_recip.hpp) running the reference's gpu_recip
test4/test6 sequence with its tolerances (gpu_recip.cu:68-522) through
CudaPMERecip<float|int|long long, float, float2> (CudaPMERecip.cu:2584-2600;
<int, float, float2> is the production instantiation, CudaDomdecRecip.h:12).
GPU only."""
import os
import subprocess

import numpy as np
import pytest

from gpu_recip_amd import build as B

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_cpp_dropin_reference_sequence(tmp_path):
    assert os.path.exists(B.DROPIN_BIN), "build the drop-in test first (gpu_recip_amd.build.build_dropin_test)"
    d = np.load(os.path.join(GOLD, "recip_system_23558.npz"))
    x = tmp_path / "xyzq.f32"
    f4 = tmp_path / "f4.f64"
    f6 = tmp_path / "f6.f64"
    np.ascontiguousarray(d["xyzq"], np.float32).tofile(x)
    np.ascontiguousarray(d["force_recip_4"], np.float64).tofile(f4)
    np.ascontiguousarray(d["force_recip_6"], np.float64).tofile(f6)
    r = subprocess.run([B.DROPIN_BIN, str(x), str(f4), str(f6)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "ALL OK" in r.stdout
    # three checks (energy, virial, forces) for each of the six instantiations
    # (AT = float, int, long long; BOX, COLUMN and SLAB handles; orders 4 and 6)
    assert r.stdout.count(" OK (tolerance") == 18
    assert "FAILED" not in r.stdout
