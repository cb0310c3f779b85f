"""Writes exact_diag_256cube.npz: the exact diagonal of A^-1 for the 7-point
operator (6, 1, 1) with zero Dirichlet boundary on the 256^3 grid, at
  * the 4096 strided cells of ref_256cube.npz (cell c = 4096 m, m < 4096:
    i1 = 0, the boundary face, every 16th row, every plane), and
  * 4096 cells drawn uniformly over the grid (numpy default_rng(2024)),
computed by the oracle's oracle_exact_diag_inv (the separable eigen-expansion
diag(A^-1)_i = sum_k prod_d (2/(n+1)) sin^2(pi k_d i_d/(n+1)) / lambda_k,
pinned there against dense inverses and known values).  diag(A^-1) is at
least 0.1855 (the corner value) at every cell, so |dD_i| / diag(A^-1)_i is a
well-posed per-cell relative error for any two estimates D of it.

Usage: python tests/golden/make_exact_diag.py   (about 4 minutes on 2 cores)
"""
import concurrent.futures
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE)))
import oracle_lib  # noqa: E402

SHAPE = (256, 256, 256)


def cells():
    n = SHAPE[0] * SHAPE[1] * SHAPE[2]
    strided = np.arange(0, n, n // 4096, dtype=np.int64)[:4096]
    rnd = np.sort(np.random.default_rng(2024).choice(n, 4096, replace=False).astype(np.int64))
    return strided, rnd


def exact(idx):
    return oracle_lib.exact_diag_inv(SHAPE, idx)


def main():
    strided, rnd = cells()
    chunks = np.array_split(np.concatenate([strided, rnd]), 16)
    with concurrent.futures.ProcessPoolExecutor(2) as ex:
        vals = np.concatenate(list(ex.map(exact, chunks)))
    np.savez_compressed(os.path.join(HERE, "exact_diag_256cube.npz"), idx_strided=strided,
                        diag_strided=vals[:4096], idx_random=rnd, diag_random=vals[4096:])
    print("min %.6f max %.6f" % (vals.min(), vals.max()))


if __name__ == "__main__":
    main()
