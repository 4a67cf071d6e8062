/*
 * bhs_mmio.h -- Matrix Market input to CSR (libbhs_host.so): the reference's
 * only real input path (spmv_cuda/main.cu:581-733 with the vendored NIST
 * mmio.h:249-362), host code, no GPU.
 *
 * Accepted: "%%MatrixMarket matrix coordinate <field> <symmetry>" with field
 * real | double | integer | pattern (pattern entries read as 1.0, integers as
 * doubles) and symmetry general | symmetric | hermitian | skew-symmetric.
 * Symmetric and hermitian files get every off-diagonal entry mirrored, as the
 * reference does (main.cu:653-710); skew-symmetric files get the mirror
 * negated (the reference does not mirror them).  Indices in the file are
 * 1-based; the CSR is 0-based with 32-bit indices.
 *
 * Return codes are those of bhsparse_spmv.h: BHSPARSE_UNSUPPORTED_VALUE_TYPE
 * for complex files (main.cu:594-598), BHSPARSE_INVALID_ARGUMENT for a file
 * that cannot be opened or is not a coordinate Matrix Market file,
 * BHSPARSE_INVALID_MATRIX for malformed or out-of-range entries.
 */
#ifndef BHS_MMIO_H
#define BHS_MMIO_H

#ifdef __cplusplus
extern "C" {
#endif

/* A CSR matrix owned by the library (free with bhs_csr_free). */
typedef struct {
    int m, n, nnz;
    int *row_ptr;
    int *col;
    double *val;
} bhs_csr;

/* flags for bhs_mm_read */
#define BHS_MM_SORT_COLUMNS      1   /* sort every row by column (stable) */
#define BHS_MM_MERGE_DUPLICATES  2   /* add up entries with the same (row, col); needs SORT */
/* flags = 0 keeps the reference's layout: entries of a row in file order
 * (mirrored entries where the reference inserts them), duplicates kept. */

/* Reads `path` into *out (on success; *out is untouched otherwise). */
int bhs_mm_read(const char *path, int flags, bhs_csr **out);
void bhs_csr_free(bhs_csr *A);

#ifdef __cplusplus
}
#endif
#endif /* BHS_MMIO_H */
