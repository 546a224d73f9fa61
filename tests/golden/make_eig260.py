"""Writes tests/golden/eig260.npz from the reference's shipped eigendecomposition
files source_beta_to_alpha/{d,v,vi}{0..3}260.dat (data files, read as text).

Run once in a container that has the reference mounted:
    python tests/golden/make_eig260.py /root/reference
Stored per boundary id k (0..3):
    d{k}      all 260 eigenvalues, in file order
    vrow{k}   rows ROWS of V (every column), V as A(r, c) = file row r, column c
    virow{k}  rows of V^-1 for the eigen-indices COLS (every column)
"""
import os
import sys

import numpy as np

N = 260
ROWS = np.array([0, 1, 2, 63, 129, 200, 258, 259])
COLS = np.array([0, 1, 7, 64, 130, 201, 258, 259])


def main(root):
    src = os.path.join(root, "source_beta_to_alpha")
    out = {"rows": ROWS, "cols": COLS}
    for k in range(4):
        d = np.loadtxt(os.path.join(src, f"d{k}{N}.dat")).reshape(-1)
        v = np.loadtxt(os.path.join(src, f"v{k}{N}.dat")).reshape(N, N)
        vi = np.loadtxt(os.path.join(src, f"vi{k}{N}.dat")).reshape(N, N)
        out[f"d{k}"] = d
        out[f"vrow{k}"] = v[ROWS, :]
        out[f"virow{k}"] = vi[COLS, :]
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "eig260.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
