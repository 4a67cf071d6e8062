"""Adversarial CSR inputs for the segmented sum (test helper).

Shapes follow what the reference's tiling makes hard (spmv_cuda/kernels_cuda.h
and bhsparse_spmv_cuda.h:184-222): no nonzeros, fewer than one tile, exact
multiples of the tile, rows starting exactly on tile and wave boundaries, one
row spanning many tiles and waves, empty rows leading / trailing / in runs
across tile boundaries (a dirty last tile), all rows of length 1, and
rectangular matrices.  Every case is built from a row-length list with seeded
distinct sorted columns.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Case:
    name: str
    m: int
    n: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    x: np.ndarray

    @property
    def nnz(self):
        return int(self.row_ptr[-1])


def from_lengths(name, lengths, n=None, seed=0, integer=False):
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, np.int64)
    m = len(lengths)
    if n is None:
        n = max(int(lengths.max()) if m else 1, m, 1)
    lengths = np.minimum(lengths, n)
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lengths)
    nnz = int(rp[-1])
    col = np.empty(nnz, np.int32)
    for i in range(m):
        L = int(lengths[i])
        if L:
            col[rp[i]:rp[i + 1]] = np.sort(rng.choice(n, L, replace=False))
    if integer:
        val = rng.integers(-9, 10, nnz).astype(np.float64)
        x = rng.integers(-9, 10, n).astype(np.float64)
    else:
        val = rng.uniform(-1, 1, nnz)
        x = rng.uniform(-1, 1, n)
    return Case(name, m, n, rp.astype(np.int32), col, val, x)


def six_by_six():
    """The reference's built-in example (spmv_cuda/main.cu:441-474), x = 1."""
    rp = np.array([0, 3, 9, 11, 11, 12, 15], np.int32)
    col = np.array([0, 2, 5, 0, 1, 2, 3, 4, 5, 2, 4, 4, 2, 3, 4], np.int32)
    val = np.arange(1, 16, dtype=np.float64)
    return Case("6x6 main.cu:441", 6, 6, rp, col, val, np.ones(6))


SIX_BY_SIX_Y = np.array([6, 39, 21, 0, 12, 42], np.float64)


def adversarial(P, T, seed=0, integer=False):
    """Cases aimed at a tile of P nonzeros and T tiles per wave."""
    rng = np.random.default_rng(seed + 12345)
    cs = []

    def add(name, lengths, n=None):
        cs.append(from_lengths(name, lengths, n, seed=seed + len(cs), integer=integer))

    cs.append(six_by_six())
    add("m=1 nnz=0", [0])
    add("m=5 nnz=0", [0] * 5)
    add("m=1 one entry", [1])
    add("nnz < P", [3, 0, 5, 2, 0, 7])
    add("nnz == P", [P // 4] * 4)
    add("nnz == 3P", [P // 2] * 6)
    # rows starting exactly at tile and wave boundaries
    add("rows of length P", [P] * (2 * T + 1), n=P + 7)
    add("rows of length P*T (wave aligned)", [P * T] * 3, n=P * T + 3)
    add("rows of length P/2 then 1", [P // 2, P // 2, 1, P - 1, P, 1])
    # a long row across many tiles and waves
    add("one long row", [5, 3 * P * T + 17, 4], n=3 * P * T + 40)
    add("long row at start", [4 * P * T + 1, 2, 2], n=4 * P * T + 10)
    add("long row at end", [2, 2, 4 * P * T - 3], n=4 * P * T + 10)
    add("two long rows", [2 * P * T + 5, 2 * P * T - 5, 1], n=2 * P * T + 10)
    # empty rows
    add("leading empty rows", [0] * 70 + [3] * 50)
    add("trailing empty rows", [3] * 50 + [0] * 70)
    add("leading+trailing empty", [0] * 3 + [7] * (P // 2) + [0] * 5)
    L = [2] * (P // 2 - 1) + [0] * 130 + [2] * P   # empty run across a tile boundary
    add("empty run across tile boundary", L)
    add("empty rows at tile boundary", [P // 2, 0, 0, P // 2, 0, P, 0, 1])
    add("dirty last tile", [5] * (P // 5) + [1, 0, 2, 0, 0])
    add("huge empty run", [3] + [0] * 5000 + [4] + [0] * 300 + [2])
    # runs of empty rows around 64-row multiples (the chunk parts and the
    # 256-row chunk), a run far longer than a tile inside it, and one reaching
    # the last row
    add("empty runs of 63/64/65/127/128/129", [2] + [0] * 63 + [3] + [0] * 64 + [1] + [0] * 65 +
        [2] + [0] * 127 + [1] + [0] * 128 + [3] + [0] * 129 + [2])
    add("empty runs of 191/192/193/255/256/257", [1] + [0] * 191 + [2] + [0] * 192 + [1] +
        [0] * 193 + [3] + [0] * 255 + [1] + [0] * 256 + [2] + [0] * 257 + [1])
    # tiles of 128-256 and 256+ short rows with a few empty ones (the chunk
    # covers 256 rows; longer tiles walk the rest)
    add("short rows with sparse empties", ([1, 1, 2, 0] * 300 + [1] * 100)[:1300])
    # about 190 rows per tile (so the 256-row chunk is chosen), 5% of them empty
    L = P / 190.0
    nr = 190 * (3 * T + 2)
    lens = np.floor(L) + (rng.random(nr) < L - np.floor(L))
    lens[rng.random(nr) < 0.05] = 0
    add("about 190 rows per tile", lens.astype(np.int64))
    add("100k empty rows inside a tile", [5] * 40 + [0] * 100000 + [3] * 30 + [0] * 700 + [1] * 9,
        n=512)
    add("empty run to the last row", [4] * 20 + [0] * 20000, n=256)
    add("all rows length 1", [1] * (3 * P + 11))
    add("alternating 0/1", [1, 0] * (2 * P))
    add("rectangular n > m", rng.integers(0, 40, 200), n=1000)
    add("rectangular n < m", rng.integers(0, 8, 600), n=9)
    # random mixtures
    for i in range(3):
        lens = rng.integers(0, 33, 700 + 300 * i)
        add(f"uniform[0,32] #{i}", lens, n=2000)
    pl = np.minimum((rng.pareto(1.0, 1500) + 1).astype(np.int64), 4 * P)
    pl[rng.random(1500) < 0.1] = 0
    add("power-law with 10% empty", pl, n=5000)
    return cs


def stack_panels(c, npan):
    """The column-panel copy of a case (the layout of BHS_VARIANT_SEGSUM_PANELS):
    one stacked row per nonempty (panel, row) pair, panel-major, holding row
    i's entries whose column lies in panel p = col // ceil(n / npan), columns in
    order.  Returns the stacked case and idx (npan x m, -1 where empty)."""
    pw = (c.n + npan - 1) // npan
    rows = np.repeat(np.arange(c.m, dtype=np.int64), np.diff(c.row_ptr))
    key = (c.col.astype(np.int64) // pw) * c.m + rows
    order = np.argsort(key, kind="stable")
    cnt = np.bincount(key, minlength=npan * c.m)
    nz = np.flatnonzero(cnt)
    idx = np.full(npan * c.m, -1, np.int64)
    idx[nz] = np.arange(len(nz))
    rp = np.zeros(len(nz) + 1, np.int64)
    rp[1:] = np.cumsum(cnt[nz])
    st = Case(c.name + f" panels={npan}", len(nz), c.n, rp.astype(np.int32), c.col[order],
              c.val[order], c.x)
    return st, idx.reshape(npan, c.m)
