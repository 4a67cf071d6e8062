/*
 * bhs_vendor.h -- vendor CSR SpMV baseline: rocSPARSE on the same matrix, as
 * the reference times cuSPARSE beside its own kernel (call_cusparse_ref,
 * spmv_cuda/main.cu:49-183: non-transposed, general, base 0, alpha = 1,
 * beta = 0).  A comparison column only, never part of the bhsparse_spmv
 * path; it lives in its own library (libbhs_vendor.so, linked with
 * librocsparse) so the drop-in library carries no vendor dependency.
 *
 * Return codes are those of bhsparse_spmv.h.
 */
#ifndef BHS_VENDOR_H
#define BHS_VENDOR_H

#ifdef __cplusplus
extern "C" {
#endif

/* rocSPARSE CSR algorithms (rocsparse_v2_spmv). */
#define BHS_VENDOR_ALG_ADAPTIVE  0   /* rocsparse_spmv_alg_csr_adaptive */
#define BHS_VENDOR_ALG_ROWSPLIT  1   /* rocsparse_spmv_alg_csr_rowsplit */
#define BHS_VENDOR_ALG_LRB       2   /* rocsparse_spmv_alg_csr_lrb */
#define BHS_VENDOR_ALG_NNZSPLIT  3   /* rocsparse_spmv_alg_csr_nnzsplit */

/* Uploads A and x to `device`, runs rocSPARSE's analysis stage for `alg`
 * (its host time in *analysis_ms), then `warmup` untimed and `reps` timed
 * y = alpha*A*x calls (HIP events around each compute stage; if flush_bytes
 * > 0 a buffer of that size is written before every call, outside the
 * events).  ms[i] per timed call; y_out (m values, may be NULL) gets the
 * last result.  Frees everything before returning. */
int bhs_vendor_d_csrmv_bench(int device, int m, int n, int nnz, const int *row_ptr,
                             const int *col, const double *val, const double *x, double alpha,
                             int alg, int warmup, int reps, long long flush_bytes, double *ms,
                             double *y_out, double *analysis_ms);
int bhs_vendor_s_csrmv_bench(int device, int m, int n, int nnz, const int *row_ptr,
                             const int *col, const float *val, const float *x, float alpha,
                             int alg, int warmup, int reps, long long flush_bytes, double *ms,
                             float *y_out, double *analysis_ms);
/* rocSPARSE's version as major*100000 + minor*100 + patch. */
int bhs_vendor_version(void);

#ifdef __cplusplus
}
#endif
#endif /* BHS_VENDOR_H */
