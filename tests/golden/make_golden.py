"""Writes the golden fixtures of this directory (run from the repo root:
python tests/golden/make_golden.py).

Every expected output here is computed without the oracle:
  * six_by_six      -- the reference's built-in example (spmv_cuda/main.cu:441-474)
                       with x = 1; y is its known answer.
  * laplace_ramp    -- 7-point Laplacian on a 5x4x3 grid, x_i = i; y from the grid
                       geometry (6 x_i minus the neighbours that exist), not from CSR.
  * int_skewed      -- integer values and x; leading, trailing and runs of empty
                       rows, a row longer than a wave of tiles; y = dense int64 product.
  * tall_dirty      -- 5000 x 64, rows of length 0 or 1 (almost every tile holds an
                       empty row); y = dense int64 product.
For the last three the tile partition tile_row[p] = max{r : row_ptr[r] <= min(pP, nnz)}
is also stored for P in {256, 512, 1024}, found by a linear scan.
All values are small integers, so every summation order gives the exact y in
both float64 and float32.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def csr_from_lengths(lengths, n, rng):
    m = len(lengths)
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lengths)
    col = np.concatenate([np.sort(rng.choice(n, L, replace=False)) if L else
                          np.zeros(0, np.int64) for L in lengths]).astype(np.int64)
    val = rng.integers(-9, 10, int(rp[-1]))
    val[val == 0] = 1
    return rp, col, val


def dense_product(m, n, rp, col, val, x):
    D = np.zeros((m, n), np.int64)
    for i in range(m):
        for j in range(rp[i], rp[i + 1]):
            D[i, col[j]] += val[j]
    return D @ np.asarray(x, np.int64)


def partitions(rp):
    nnz = int(rp[-1])
    m = len(rp) - 1
    out = {}
    for P in (256, 512, 1024):
        NP = (nnz + P - 1) // P
        tr = []
        for p in range(NP + 1):
            key = min(p * P, nnz)
            r = 0
            while r + 1 <= m and rp[r + 1] <= key:
                r += 1
            tr.append(r)
        out[str(P)] = tr
    return out


def fixture(name, m, n, rp, col, val, x, y, note, with_partition=True):
    d = {"name": name, "note": note, "m": int(m), "n": int(n),
         "row_ptr": [int(v) for v in rp], "col": [int(v) for v in col],
         "val": [float(v) for v in val], "x": [float(v) for v in x],
         "y": [float(v) for v in y]}
    if with_partition:
        d["tile_row"] = partitions(rp)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(d, f, separators=(",", ":"))


def main():
    # spmv_cuda/main.cu:441-474
    rp = [0, 3, 9, 11, 11, 12, 15]
    col = [0, 2, 5, 0, 1, 2, 3, 4, 5, 2, 4, 4, 2, 3, 4]
    val = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15]
    fixture("six_by_six", 6, 6, rp, col, val, [1] * 6, [6, 39, 21, 0, 12, 42],
            "reference example, spmv_cuda/main.cu:441-474, x = 1", with_partition=False)

    nx, ny, nz = 5, 4, 3
    m = nx * ny * nz
    rp, col, val, y = [0], [], [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = i + nx * (j + ny * k)
                nb = []
                if k > 0: nb.append(r - nx * ny)
                if j > 0: nb.append(r - nx)
                if i > 0: nb.append(r - 1)
                entries = sorted([(c, -1.0) for c in nb] + [(r, 6.0)] +
                                 [(c, -1.0) for c in
                                  ([r + 1] if i < nx - 1 else []) +
                                  ([r + nx] if j < ny - 1 else []) +
                                  ([r + nx * ny] if k < nz - 1 else [])])
                col += [c for c, _ in entries]
                val += [v for _, v in entries]
                rp.append(len(col))
                # x_c = c: y_r = 6 r - sum of the neighbour indices that exist
                y.append(6 * r - sum(c for c, v in entries if c != r))
    fixture("laplace_ramp", m, m, rp, col, val, list(range(m)), y,
            "7-point Laplacian 5x4x3, x_i = i, y from the grid geometry")

    rng = np.random.default_rng(20240601)
    n = 2000
    L = np.concatenate([np.zeros(40, int), rng.integers(1, 12, 300), np.zeros(7, int),
                        [1800], rng.integers(0, 3, 200), np.zeros(90, int),
                        rng.integers(20, 300, 12), rng.integers(0, 6, 400), np.zeros(25, int)])
    rp, col, val = csr_from_lengths(L, n, rng)
    x = rng.integers(-9, 10, n)
    y = dense_product(len(L), n, rp, col, val, x)
    fixture("int_skewed", len(L), n, rp, col, val, x, y,
            "integer data: empty runs, a 1800-long row, trailing empties; y = dense int64 product")

    L = (rng.random(5000) < 0.3).astype(int)
    rp, col, val = csr_from_lengths(L, 64, rng)
    x = rng.integers(-9, 10, 64)
    y = dense_product(5000, 64, rp, col, val, x)
    fixture("tall_dirty", 5000, 64, rp, col, val, x, y,
            "5000 x 64, 70% empty rows; y = dense int64 product")


if __name__ == "__main__":
    main()
