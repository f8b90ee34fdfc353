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


def test_cpp_dropin_direct_space(tmp_path):
    """CudaPMEDirectForce<long long, float> of the C++ header over a water box
    with its exclusions and a few 1-4 pairs, against the FP64 oracle."""
    import oracle as O
    from gpu_recip_amd import nlist, systems

    assert os.path.exists(B.DIRECT_DROPIN_BIN), "build the drop-in tests first (gpu_recip_amd.build.build_dropin_test)"
    box = (25.0, 26.0, 27.0)
    kappa, roff, e14fac = 0.34, 9.0, 0.5
    x = systems.triplets(500, box, seed=5)
    x[:, :3] = np.mod(x[:, :3].astype(np.float64), box).astype(np.float32)
    water = nlist.water_exclusions(len(x))
    in14 = np.array([[0, 4], [7, 11], [30, 40], [100, 250]], np.int64)
    tl = nlist.build(x, box, roff, excl_pairs=np.concatenate([water, in14]))
    inv = tl.sorted_index()
    l_ex = nlist.ex14_list(x, box, water)
    l_in = nlist.ex14_list(x, box, in14)
    od = O.direct_ewald(x, box, kappa, roff, excl=np.concatenate([water, in14]))
    oi = O.pair14(x, box, kappa, e14fac, l_in)
    oe = O.pair14(x, box, kappa, 0.0, l_ex)
    f = od["force"] + oi["force"] + oe["force"]
    fs = np.zeros((tl.npad, 3))
    fs[inv] = f
    prm = [*box, kappa, roff, e14fac, od["energy"] + oi["energy"], oe["energy"], 1e-5, 1e-5]
    np.asarray(prm, np.float64).tofile(tmp_path / "params.f64")
    np.ascontiguousarray(tl.xyzq, np.float32).tofile(tmp_path / "xyzq.f32")
    np.ascontiguousarray(tl.ientry, np.int32).tofile(tmp_path / "ientry.i32")
    np.ascontiguousarray(tl.tile_indj, np.int32).tofile(tmp_path / "tile_indj.i32")
    np.ascontiguousarray(tl.tile_excl, np.uint32).tofile(tmp_path / "tile_excl.u32")
    for name, l in (("in14", l_in), ("ex14", l_ex)):
        s = np.stack([inv[l[:, 0]], inv[l[:, 1]], l[:, 2]], 1)
        np.ascontiguousarray(s, np.int32).tofile(tmp_path / f"{name}.i32")
    fs.tofile(tmp_path / "force.f64")
    r = subprocess.run([B.DIRECT_DROPIN_BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "ALL OK" in r.stdout
    assert r.stdout.count(" OK (tolerance") == 7
