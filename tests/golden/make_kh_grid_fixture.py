"""Writes tests/golden/kh_grid_y.txt: the 66 GY values stored in the
reference's KH_test/grid.h5 (dataset /grids/y, for/hdf5.f:547-709 reads it).

Run in the build container only (the reference tree is not on the GPU box):
    python tests/golden/make_kh_grid_fixture.py

h5py is not installed, so the file is read as bytes: the dataset is stored
contiguously as 66 little-endian float64 at the end of the file.  The script
checks the HDF5 signature, that the block is strictly increasing and that it
is antisymmetric (the KH grid is symmetric about y = 0), then writes the
values with 17 significant digits.
"""
import os
import sys

import numpy as np

SRC = "/root/reference/diablo_2.0/KH_test/grid.h5"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kh_grid_y.txt")
N = 66


def main():
    raw = open(SRC, "rb").read()
    if raw[:8] != b"\x89HDF\r\n\x1a\n":
        sys.exit("not an HDF5 file")
    y = np.frombuffer(raw[len(raw) - 8 * N:], dtype="<f8")
    if not (np.all(np.diff(y) > 0) and np.allclose(y, -y[::-1], rtol=0, atol=1e-12)):
        sys.exit("unexpected dataset layout")
    with open(OUT, "w") as f:
        f.write("# GY(1..66) from KH_test/grid.h5 /grids/y (NY=65, LY=20, GRID_TYPE=2, CS=1.5)\n")
        for v in y:
            f.write(f"{v:.17g}\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
