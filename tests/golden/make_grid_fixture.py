"""Writes test_case_grid_y.json: the y-grid stored in the reference's
test_case/grid.h5 (dataset /grids/y, 194 float64 = Ny+1 for Ny = 193).

h5py is not installed here, so the file is read as bytes: the dataset is
stored contiguously, its dataspace dimension (194) sits at byte 1864 and its
data at byte 2432 (found by inspecting the file).  Run once, in a container
that has the reference tree; the output is the committed fixture.
"""
import json
import os
import struct
import sys

import numpy as np

SRC = "/root/reference/test_case/grid.h5"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "test_case_grid_y.json")


def main():
    b = open(SRC, "rb").read()
    n = struct.unpack("<Q", b[1864:1872])[0]
    if n != 194:
        sys.exit(f"unexpected dataspace size {n}")
    y = np.frombuffer(b[2432:2432 + 8 * n], dtype="<f8")
    json.dump(
        {
            "source": "test_case/grid.h5:/grids/y",
            "grid_def": {"Nx": 256, "Ny": 193, "Nz": 16, "Nth": 1},
            "y": [float(v) for v in y],
        },
        open(OUT, "w"),
        indent=1,
    )
    print(f"wrote {n} values to {OUT}")


if __name__ == "__main__":
    main()
