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
  * mtx/*.mtx       -- Matrix Market files (general / symmetric / hermitian /
                       skew-symmetric; real / integer / pattern; comments, blank
                       lines, unsorted rows, duplicates) written from explicit
                       entry lists, with the dense matrix those entries define
                       (mirroring rules applied here, not by the reader) in
                       mtx/expected.json, and files the reader must reject.
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


def write_mtx_fixtures():
    d = os.path.join(HERE, "mtx")
    os.makedirs(d, exist_ok=True)
    expected = {}

    def emit(name, m, n, field, sym, entries, extra_lines=()):
        with open(os.path.join(d, name), "w") as f:
            f.write(f"%%MatrixMarket matrix coordinate {field} {sym}\n")
            f.write("% written by make_golden.py\n%\n")
            f.write(f"{m} {n} {len(entries)}\n")
            for k, (i, j, v) in enumerate(entries):
                if k in extra_lines:
                    f.write("\n% a comment between entries\n")
                f.write(f"{i} {j}\n" if field == "pattern" else
                        f"{i} {j} {int(v) if field == 'integer' else v!r}\n")
        D = np.zeros((m, n))
        for i, j, v in entries:
            v = 1.0 if field == "pattern" else float(v)
            D[i - 1, j - 1] += v
            if sym in ("symmetric", "hermitian") and i != j:
                D[j - 1, i - 1] += v
            if sym == "skew-symmetric":
                D[j - 1, i - 1] -= v
        expected[name] = {"m": m, "n": n, "dense": D.tolist(),
                          "entries_in_file": len(entries)}

    emit("general_real.mtx", 4, 5, "real", "general",
         [(1, 3, 0.5), (1, 1, -2.0), (3, 5, 1.25), (2, 2, 3.0), (3, 1, 4.5), (1, 3, 0.25),
          (4, 4, -1.0)], extra_lines=(3,))
    emit("symmetric_integer.mtx", 5, 5, "integer", "symmetric",
         [(1, 1, 4), (2, 1, -1), (3, 2, -1), (3, 3, 4), (5, 1, 7), (5, 4, 2), (4, 4, 1)])
    emit("hermitian_real.mtx", 3, 3, "real", "hermitian",
         [(1, 1, 2.0), (2, 1, 0.5), (3, 2, -0.75)])
    emit("pattern_general.mtx", 3, 4, "pattern", "general",
         [(1, 4, 0), (3, 1, 0), (1, 2, 0), (2, 3, 0)])
    emit("skew_real.mtx", 4, 4, "real", "skew-symmetric",
         [(2, 1, 1.5), (3, 1, -2.0), (4, 3, 0.5)])
    emit("empty_rows.mtx", 6, 6, "real", "general", [(2, 2, 1.0), (6, 1, -3.0)])
    with open(os.path.join(d, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1)
    # files the reader must reject, with the code it must return
    bad = {
        "complex.mtx": ("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1.0 0.0\n",
                        -1),
        "array.mtx": ("%%MatrixMarket matrix array real general\n2 2\n1.0\n2.0\n3.0\n4.0\n",
                      -3),
        "bad_index.mtx": ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", -4),
        "truncated.mtx": ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n", -4),
        "no_banner.mtx": ("2 2 1\n1 1 1.0\n", -3),
        "skew_diagonal.mtx": ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n"
                              "1 1 1.0\n", -4),
    }
    for name, (txt, code) in bad.items():
        with open(os.path.join(d, name), "w") as f:
            f.write(txt)
    with open(os.path.join(d, "reject.json"), "w") as f:
        json.dump({k: v[1] for k, v in bad.items()}, f, indent=1)


if __name__ == "__main__":
    main()
    write_mtx_fixtures()
