"""Convert the reference's own test data for the PME reciprocal path into compact
fixtures under tests/golden/.  Run in a container where /root/reference exists:

    python tests/golden/make_fixtures.py

Sources (data files the reference's tests hold; values copied verbatim):
  test_data/xyzq.txt           23,558 atoms x (x, y, z, q)       gpu_recip.cu:100 (XYZQ loader XYZQ.cu:131-168)
  test_data/force_recip_4.txt  order-4 recip gradient              gpu_recip.cu:93, compared :286
  test_data/force_recip_6.txt  order-6 recip gradient              gpu_recip.cu:338, compared :515
  test_data/dataFFT{in,out}*.txt  forward R2C goldens              test_fbfft.cu:203-207, 274-278
Energy/virial constants and tolerances of gpu_recip.cu:71-84 and :301-314 are
written to recip_constants.json.
"""
import json
import os

import numpy as np

REF = "/root/reference/test_data"
OUT = os.path.dirname(os.path.abspath(__file__))


def load(name, ncol):
    a = np.loadtxt(os.path.join(REF, name), dtype=np.float64)
    return a.reshape(-1, ncol)


def main():
    xyzq = load("xyzq.txt", 4)
    # The reference reads the text with iostream into float (XYZQ.cu:141-150)
    np.savez_compressed(os.path.join(OUT, "recip_system_23558.npz"),
                        xyzq=xyzq.astype(np.float32),
                        force_recip_4=load("force_recip_4.txt", 3),
                        force_recip_6=load("force_recip_6.txt", 3))
    fft = {}
    fft["in64"] = load("dataFFTin64.txt", 1).ravel()
    o = load("dataFFTout64.txt", 2)
    fft["out64"] = o[:, 0] + 1j * o[:, 1]
    fft["in64x2"] = load("dataFFTin64x2.txt", 1).ravel()
    o = load("dataFFTout64x2.txt", 2)
    fft["out64x2"] = o[:, 0] + 1j * o[:, 1]
    fft["in64x64"] = load("dataFFTin64x64.txt", 1).ravel()
    o = load("dataFFTout64x64.txt", 2)
    fft["out64x64"] = o[:, 0] + 1j * o[:, 1]
    fft["in4x4"] = load("dataFFTin4x4.txt", 1).ravel()
    o = load("dataFFTout4x4.txt", 2)
    fft["out4x4"] = o[:, 0] + 1j * o[:, 1]
    np.savez_compressed(os.path.join(OUT, "fft_goldens.npz"), **fft)

    const = {
        "source": "gpu_recip.cu:71-84 (order 4), :301-314 (order 6); tolerances :146,157,286,373,384,515",
        "box": [62.23, 62.23, 62.23],
        "kappa": 0.32,
        "nfft": [64, 64, 64],
        "ncoord": 23558,
        "order4": {
            "energy": 0.5939730270e03,
            "virial": [-0.5282341105e03, 0.5596133994e01, -0.6254453157e01,
                       0.5596133994e01, -0.5054581254e03, -0.7652312605e-01,
                       -0.6254453157e01, -0.7652312605e-01, -0.5575246246e03],
            "tol_energy": 2.0e-3, "tol_virial": 4.0e-3, "tol_force": 1.0e-3,
        },
        "order6": {
            "energy": 0.5951144996e03,
            "virial": [-0.5314776507e03, 0.5670623182e01, -0.6175456643e01,
                       0.5670623182e01, -0.5086150173e03, -0.1026251692e00,
                       -0.6175456643e01, -0.1026251692e00, -0.5606459148e03],
            "tol_energy": 1.3e-3, "tol_virial": 2.0e-3, "tol_force": 3.6e-4,
        },
    }
    with open(os.path.join(OUT, "recip_constants.json"), "w") as f:
        json.dump(const, f, indent=1)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
