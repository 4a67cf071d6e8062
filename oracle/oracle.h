/*
 * oracle.h -- CPU oracle for the speculative segmented-sum CSR SpMV.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library calls this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker (or as the timed CPU baseline), never as the thing
 * measured on the GPU.
 *
 * Every function cites the reference file:line it restates
 * (weifengliu-ssslab/Benchmark_SpMV_using_CSR).
 */
#ifndef BHS_ORACLE_H
#define BHS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Serial CSR row loop in FP64: y[i] = sum_j x[col[j]] * val[j], left to right.
 * Restates spmv_cuda/main.cu:755-761 (there in value_type; here always FP64). */
void oracle_spmv_serial_f64(int m, const int *row_ptr, const int *col,
                            const double *val, const double *x, double *y);

/* Same loop in value_type = float (main.cu:755-761 with USE_DOUBLE=0). */
void oracle_spmv_serial_f32(int m, const int *row_ptr, const int *col,
                            const float *val, const float *x, float *y);

/* Row-parallel OpenMP loop (main.cu:377-393); nthreads<=0 uses all cores.
 * Returns the thread count used. */
int oracle_spmv_omp_f64(int m, const int *row_ptr, const int *col,
                        const double *val, const double *x, double *y, int nthreads);
int oracle_spmv_omp_f32(int m, const int *row_ptr, const int *col,
                        const float *val, const float *x, float *y, int nthreads);

/* Times `reps` calls of the serial (nthreads==1) or OpenMP loop; returns the
 * median seconds per call and writes the thread count used. */
double oracle_time_spmv_f64(int m, const int *row_ptr, const int *col,
                            const double *val, const double *x, double *y,
                            int nthreads, int reps, int *threads_used);
double oracle_time_spmv_f32(int m, const int *row_ptr, const int *col,
                            const float *val, const float *x, float *y,
                            int nthreads, int reps, int *threads_used);

/* upper_bound(row_ptr[0..size), key): first index with row_ptr[idx] > key.
 * Restates spmv_cuda/utils_cuda.h:6-34 (binary_search_right_boundary_kernel). */
int oracle_upper_bound(const int *row_ptr, int size, int key);

/* rowOf(k) = upper_bound(row_ptr[0..m], min(k, nnz)) - 1; tile_row[p] = rowOf(p*P)
 * for p in [0, NP], NP = ceil(nnz / P).  Restates kernels_cuda.h:39-49. */
void oracle_partition(int m, int nnz, const int *row_ptr, int P, int *tile_row);

/* Per-tile classification for a tile size P and T tiles per wave (thread bunch):
 *   n_fast  tiles with r0 == r1 (fast track, kernels_cuda.h:79)
 *   n_dirty general tiles with an empty row in [r0, min(r1, m-1)]
 *           (kernels_cuda.h:119-145)
 *   n_sync  waves whose first row starts before the wave's first element
 *           (those rows need a cross-wave partial; the reference routes every
 *           bunch's first row through the synchronizer, kernels_cuda.h:98-102)
 *   n_lead_empty rows before rowOf(0) (never covered by a tile). */
typedef struct {
    long long n_tiles, n_waves, n_fast, n_dirty, n_sync, n_lead_empty;
} oracle_tile_stats;
void oracle_tile_stats_compute(int m, int nnz, const int *row_ptr, int P, int T,
                               oracle_tile_stats *out);

/* The reference's speculative segmented sum restated on the CPU in FP64, for a
 * thread bunch of W lanes, SEG_H = S items per lane and STEP = T tiles per
 * bunch.  Follows spmv_cuda/kernels_cuda.h:7-269 (tiles, fast track with
 * xor-butterfly, bit flags, lane-sequential segments, serial spanning walk,
 * speculative emission, synchronizer, dirty record) and the host calibration
 * and fix-up of spmv_cuda/bhsparse_spmv_cuda.h:166-222, with two corrections:
 *   - y is cleared before the call (beta = 0), so results do not depend on
 *     y's previous contents (the reference is only right on the first call);
 *   - the fix-up walk stops at row m-1 and only adds the carried partial to
 *     row z when row z has elements in the dirty tile (the reference reads
 *     row_ptr[m+1] when z == m).
 * y has m entries.  Writes the counts it saw into stats (may be NULL).
 * Returns 0, or -1 on bad arguments. */
typedef struct {
    long long n_tiles, n_bunches, n_fast, n_dirty;
} oracle_ref_stats;
int oracle_segsum_ref(int W, int S, int T, int m, int nnz, const int *row_ptr,
                      const int *col, const double *val, const double *x,
                      double *y, oracle_ref_stats *stats);

/* Per-row error of y against y_ref: |y_i - yref_i| / sum_j |a_ij x_j| (FP64),
 * 0/0 -> 0, nonzero/0 -> +inf.  err may be NULL; returns the maximum. */
double oracle_rel_error_f64(int m, const int *row_ptr, const int *col,
                            const double *val, const double *x,
                            const double *y, const double *y_ref, double *err);

/* The tuner's MLP n_in -> h -> h -> n_out (ReLU) in FP64: parameters laid out
 * W1[h*n_in], b1[h], W2[h*h], b2[h], W3[n_out*h], b3[n_out]; inputs
 * standardised with norm = {mean[n_in], std[n_in]}.  Returns the masked mean
 * squared error sum(M (O - T)^2) / sum(M) over X[N*n_in]; writes the outputs O
 * (N*n_out) and the gradient w.r.t. the parameters when the pointers are non-NULL.
 * No reference counterpart (the tuner is new).  Pinned by central finite
 * differences of the loss over every parameter and by a numpy matrix-product
 * forward (tests/test_oracle.py). */
double oracle_mlp_loss_grad(int n_in, int h, int n_out, const double *params,
                            const double *norm, const double *X, const double *T,
                            const double *M, int N, double *out, double *grad);

/* Planted mistakes, for the tests that show the oracle's checks catch them.
 * Every other value (0 included) leaves the oracle as written.  Process-wide;
 * a test sets it and restores 0. */
enum {
    ORACLE_MUT_NONE = 0,
    ORACLE_MUT_SERIAL_DROP_LAST = 1,   /* serial loop skips each row's last term */
    ORACLE_MUT_SERIAL_SIGN = 2,        /* first term of each row subtracted */
    ORACLE_MUT_SERIAL_F32_ACC = 3,     /* accumulates in float */
    ORACLE_MUT_SERIAL_TRANSPOSE = 4,   /* computes A^T x */
    ORACLE_MUT_LOWER_BOUND = 5,        /* upper_bound becomes lower_bound */
    ORACLE_MUT_NO_CARRY = 6,           /* segsum_ref: tile carry not added to item 0 */
    ORACLE_MUT_NO_SPAN = 7,            /* segsum_ref: lane-spanning partials dropped */
    ORACLE_MUT_NO_CALIBRATION = 8,     /* segsum_ref: synchronizer partials dropped */
    ORACLE_MUT_NO_FIXUP = 9,           /* segsum_ref: dirty tiles not de-speculated */
    ORACLE_MUT_MLP_NO_MASK = 10,       /* mlp_loss_grad: mask dropped from dLoss/dOut */
    ORACLE_MUT_MLP_RELU_GRAD = 11,     /* mlp_loss_grad: ReLU derivative ignored */
    ORACLE_MUT_MLP_TRANSPOSE = 12      /* mlp_loss_grad: W2 transposed in backprop */
};
void oracle_set_mutation(int kind);

#ifdef __cplusplus
}
#endif
#endif
