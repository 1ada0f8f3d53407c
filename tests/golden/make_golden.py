"""Write the golden fixtures under tests/golden/ from the reference's own
test data.  Run in the development container (reads /root/reference):

    python tests/golden/make_golden.py

Sources (data only):
  splinepractice/kernel.csv          B-spline kernels 0/1/2 at t = -3:.01:3
  splinepractice/interpolate1d.csv   1-D value / d1 / d2 at t = -30:.01:44
                                     (inputs: renderer/samples/spline_tests.cpp:8-11)
  splinepractice/interpolate2d.csv   2-D value and 5 derivatives on a 25x25 lattice
                                     (inputs: spline_tests.cpp:35-38, 45-97)
  splinepractice/coefficients3d.csv  prefiltered 10^3 coefficients
  splinepractice/interpolate3d.csv   7 derivative orders at 576 points
                                     (inputs: spline_tests.cpp:100-134)
  renderer/images/Disk_501.pfm       inscribed-disk emitter texture
  renderings/besselriffile.vol       the 101^3 grid of n that renderings/testing_splines.sh
                                     renders (rifgridFile); it does not vary along x, so one
                                     y-z slice and the header reproduce it exactly
"""
import json
import os
import re

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_csv_values(path):
    vals = []
    with open(path) as f:
        for line in f:
            for tok in line.replace(",", " ").split():
                vals.append(float(tok))
    return np.array(vals)


def main():
    sp = os.path.join(REF, "splinepractice")
    out = {}
    # kernels: rows "t, value" for d = 0, 1, 2 (601 rows each)
    k = np.loadtxt(os.path.join(sp, "kernel.csv"), delimiter=",")
    out["kernel_t"] = k[:601, 0]
    out["kernel_v"] = k[:, 1].reshape(3, 601)
    # 1-D: 3 blocks of 7401 values (value, d1, d2), subsampled every 10th
    v1 = read_csv_values(os.path.join(sp, "interpolate1d.csv"))
    n1 = len(v1) // 3
    t1 = -30.0 + 0.01 * np.arange(n1)
    sel = np.arange(0, n1, 10)
    out["i1d_t"] = t1[sel]
    out["i1d_v"] = v1.reshape(3, n1)[:, sel]
    out["i1d_data"] = np.array([1, 2, 4, 2, 3, 1, 1, 2, 3, 1, -2, 1, 2, 3], float)
    out["i1d_xmin"] = np.sqrt(13.0)
    out["i1d_xmax"] = np.sqrt(803.0)
    # 2-D: 6 blocks of 25 rows x 25 values; rows = x, columns = y
    v2 = read_csv_values(os.path.join(sp, "interpolate2d.csv"))
    out["i2d_v"] = v2.reshape(6, 25, 25)
    src = open(os.path.join(REF, "renderer/samples/spline_tests.cpp")).read()
    m2 = re.search(r"Float x\[\] = \{(-7, -14[^}]*)\}", src)
    out["i2d_data"] = np.array([float(t) for t in m2.group(1).split(",")])
    out["i2d_N"] = np.array([20, 15])
    # 3-D input (1000 values) and outputs
    m3 = re.search(r"Float x\[\] = \{(0\.417022[^}]*)\}", src)
    out["i3d_data"] = np.array([float(t) for t in m3.group(1).split(",")])
    out["c3d"] = read_csv_values(os.path.join(sp, "coefficients3d.csv")).reshape(100, 10)
    out["i3d_v"] = read_csv_values(os.path.join(sp, "interpolate3d.csv")).reshape(576, 7)
    np.savez_compressed(os.path.join(HERE, "spline_golden.npz"), **out)

    # disk texture: count and rule check
    raw = open(os.path.join(REF, "renderer/images/Disk_501.pfm"), "rb").read()
    n = 501 * 501
    px = np.frombuffer(raw[len(raw) - 8 * n:], dtype="<f8").reshape(501, 501)[::-1]
    lit = px > 0
    rows = [[int(np.argmax(r)), int(len(r) - np.argmax(r[::-1]))] if r.any() else [0, 0] for r in lit]
    json.dump({"shape": [501, 501], "lit": int(lit.sum()), "values": sorted(set(np.unique(px).tolist())),
               "row_ranges": rows}, open(os.path.join(HERE, "disk_501.json"), "w"))
    # VOL grid: header, x-invariance check, the y-z slice at x index 50
    raw = open(os.path.join(REF, "renderings/besselriffile.vol"), "rb").read()
    assert raw[:3] == b"VOL" and raw[3] == 3
    N = np.frombuffer(raw[8:20], "<i4")
    lo = np.frombuffer(raw[24:36], "<f4")
    hi = np.frombuffer(raw[36:48], "<f4")
    grid = np.frombuffer(raw[48:], "<f4").reshape(N[2], N[1], N[0])
    xvar = float(np.abs(grid - grid[:, :, :1]).max())
    np.savez_compressed(os.path.join(HERE, "besselriffile_slice.npz"), N=N, xmin=lo, xmax=hi,
                        slice_zy=grid[:, :, N[0] // 2].copy(), x_variation=np.float64(xvar))
    print("wrote", sorted(out), "disk_501.json and besselriffile_slice.npz")


if __name__ == "__main__":
    main()
