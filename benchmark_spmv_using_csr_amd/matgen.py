"""Seeded synthetic CSR workloads (wrappers over libbhs_host.so).

Each recipe is fully determined by its seed: row lengths, columns (distinct,
sorted per row) and values come from per-row counter-based streams, values and
x are uniform in [-1, 1] in FP64; FP32 runs use the same numbers rounded.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import Csr, host_lib


@dataclass
class CSR:
    m: int
    n: int
    row_ptr: np.ndarray   # int32, m+1
    col: np.ndarray       # int32, nnz
    val: np.ndarray       # float64 (or float32 after astype)
    name: str = ""

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1]) if len(self.row_ptr) else 0

    def astype(self, dtype) -> "CSR":
        return CSR(self.m, self.n, self.row_ptr, self.col, self.val.astype(dtype), self.name)

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.row_ptr)

    def dense(self) -> np.ndarray:
        D = np.zeros((self.m, self.n), np.float64)
        for i in range(self.m):
            for j in range(self.row_ptr[i], self.row_ptr[i + 1]):
                D[i, self.col[j]] += self.val[j]
        return D


def _take(ptr, name) -> CSR:
    if not ptr:
        raise ValueError(f"generator {name} rejected its arguments")
    A = ptr.contents
    m, n, nnz = A.m, A.n, A.nnz
    rp = np.ctypeslib.as_array(A.row_ptr, shape=(m + 1,)).copy()
    col = np.ctypeslib.as_array(A.col, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(A.val, shape=(max(nnz, 1),))[:nnz].copy()
    host_lib().bhs_csr_free(ptr)
    return CSR(m, n, rp, col, val, name)


MM_SORT_COLUMNS = 1
MM_MERGE_DUPLICATES = 2


def read_mtx(path, sort=True, merge=True) -> CSR:
    """Matrix Market coordinate file -> CSR (bhs_mm_read, include/bhs_mmio.h):
    pattern -> 1.0, symmetric/hermitian mirrored, skew-symmetric mirrored with
    the sign flipped.  sort / merge: sort every row by column / add duplicate
    entries; both False keeps the reference's layout (main.cu:581-733)."""
    from ._lib import BhsError
    flags = (MM_SORT_COLUMNS if sort else 0) | (MM_MERGE_DUPLICATES if merge else 0)
    ptr = C.POINTER(Csr)()
    rc = host_lib().bhs_mm_read(str(path).encode(), flags, C.byref(ptr))
    if rc != 0:
        raise BhsError(rc, "bhs_mm_read", str(path))
    return _take(ptr, str(path))


def write_mtx(A: CSR, path, field="real", symmetry="general") -> None:
    """Writes A as a Matrix Market coordinate file (every stored entry, 1-based;
    with symmetry != general the caller passes the lower triangle)."""
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {symmetry}\n")
        f.write(f"{A.m} {A.n} {A.nnz}\n")
        rows = np.repeat(np.arange(A.m), np.diff(A.row_ptr))
        for r, c, v in zip(rows.tolist(), A.col.tolist(), A.val.tolist()):
            if field == "pattern":
                f.write(f"{r + 1} {c + 1}\n")
            elif field == "integer":
                f.write(f"{r + 1} {c + 1} {int(v)}\n")
            else:
                f.write(f"{r + 1} {c + 1} {v!r}\n")     # shortest repr: round-trips


def vector(seed: int, n: int) -> np.ndarray:
    x = np.empty(max(n, 1), np.float64)
    if n > 0:
        host_lib().bhs_gen_vector(C.c_ulonglong(seed), n, x)
    return x[:n]


def uniform_rows(seed, m, n, lmin, lmax) -> CSR:
    return _take(host_lib().bhs_gen_uniform_rows(seed, m, n, lmin, lmax),
                 f"uniform[{lmin},{lmax}] m={m}")


def power_law(seed, m, n, alpha=2.0, kmax=100000, empty_frac=0.1, target_nnz=0) -> CSR:
    return _take(host_lib().bhs_gen_power_law(seed, m, n, alpha, kmax, empty_frac, target_nnz),
                 f"powerlaw a={alpha} m={m}")


def laplacian7(nx, ny, nz) -> CSR:
    return _take(host_lib().bhs_gen_laplacian7(nx, ny, nz), f"laplace7 {nx}x{ny}x{nz}")


def block3(seed, m, blocks_per_row=9, bandwidth=5000) -> CSR:
    return _take(host_lib().bhs_gen_block3(seed, m, blocks_per_row, bandwidth),
                 f"block3 m={m}")


def banded(seed, m, n, bandwidth) -> CSR:
    return _take(host_lib().bhs_gen_banded(seed, m, n, bandwidth), f"banded bw={bandwidth} m={m}")


# ---- the benchmark workloads ----------------------------------------------

def uniform_65536(seed: int = 1) -> CSR:
    """m = n = 65,536, row lengths U{0..32} (~3% empty, ~1.07M nnz)."""
    A = uniform_rows(seed, 65536, 65536, 0, 32)
    A.name = "uniform-rows m=65536 len~U{0..32}"
    return A


def power_law_1m(seed: int = 2) -> CSR:
    """m = n = 2^20, P(k) ~ k^-2 on [1, 1e5], 10% rows empty, ~10M nnz."""
    A = power_law(seed, 1 << 20, 1 << 20, 2.0, 100000, 0.1, 10_000_000)
    A.name = "power-law m=2^20 a=2 10%-empty nnz~10M"
    return A


def laplacian_160() -> CSR:
    A = laplacian7(160, 160, 160)
    A.name = "laplace7 160^3"
    return A


def block3_2m(seed: int = 3) -> CSR:
    A = block3(seed, 2_000_000, 9, 5000)
    A.name = "block3x3 m=2e6 9 blocks/row bw=5000"
    return A


WORKLOADS = ("power-law", "uniform", "laplace", "block3")


def workload(name: str):
    """One of the benchmark matrices by name, with its seeded x vector."""
    if name == "power-law":
        return power_law_1m(), vector(102, 1 << 20)
    if name == "uniform":
        return uniform_65536(), vector(101, 65536)
    if name == "laplace":
        A = laplacian_160()
        return A, vector(103, A.n)
    if name == "block3":
        A = block3_2m()
        return A, vector(104, A.n)
    raise ValueError(f"unknown workload {name!r}; one of {WORKLOADS}")


# ---- tuner and batch families ---------------------------------------------

def _rng(seed):
    return np.random.default_rng(np.random.SeedSequence([seed, 0x5EED]))


def family_matrix(seed: int, m_lo: float, m_hi: float, family: int | None = None) -> CSR:
    """One matrix of the four mixed families: m log-uniform in [m_lo, m_hi],
    family = seed mod 4 unless given:
    0 uniform rows in [0, L], L in [4, 64]; 1 power law alpha in [1.5, 3] with
    a fraction in [0, 0.2] of the rows empty; 2 banded, bandwidth 3..201;
    3 3x3 block rows."""
    g = _rng(seed)
    m = int(10 ** g.uniform(np.log10(m_lo), np.log10(m_hi)))
    fam = seed % 4 if family is None else family
    if fam == 0:
        L = int(g.integers(4, 65))
        A = uniform_rows(seed, m, m, 0, L)
    elif fam == 1:
        a = float(g.uniform(1.5, 3.0))
        ef = float(g.uniform(0.0, 0.2))
        A = power_law(seed, m, m, a, min(100000, m), ef, 0)
    elif fam == 2:
        bw = int(g.integers(1, 101)) * 2 + 1
        A = banded(seed, m, m, bw)
    else:
        A = block3(seed, max(m, 3), 9, int(g.integers(30, 5001)))
    A.name = f"fam{fam} seed={seed} m={A.m}"
    return A
