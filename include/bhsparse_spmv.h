/*
 * bhsparse_spmv.h -- C ABI of the MI355X (gfx950) speculative segmented-sum
 * CSR SpMV, y = alpha * A * x.
 *
 * Drop-in for the operator class of weifengliu-ssslab/Benchmark_SpMV_using_CSR:
 *   class bhsparse_spmv_cuda        spmv_cuda/bhsparse_spmv_cuda.h:7-55
 *   class bhsparse_spmv_opencl      spmv_opencl_amd/bhsparse_spmv_opencl.h:7-19
 * Every entry point exists twice: bhs_s_* computes in float (the reference's
 * USE_DOUBLE=0 build, spmv_cuda/common.h:27-31) and bhs_d_* in double
 * (USE_DOUBLE=1).  Plain pointers and sizes only; all indices are 32-bit
 * (0-based CSR, as in the reference).  Functions return BHSPARSE_SUCCESS (0)
 * or a negative code and never exit the process.
 *
 * A header-only C++ class with the reference's method names sits in
 * bhsparse_spmv.hpp.
 */
#ifndef BHSPARSE_SPMV_H
#define BHSPARSE_SPMV_H

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes.  0, -1 and -2 keep the reference's meaning
 * (spmv_cuda/common.h:24-25, spmv_opencl_amd/common.h:20-21). */
#define BHSPARSE_SUCCESS                 0
#define BHSPARSE_UNSUPPORTED_VALUE_TYPE -1
#define BHSPARSE_UNSUPPORTED_DEVICE     -2
#define BHSPARSE_INVALID_ARGUMENT       -3   /* null pointer, negative size, bad config */
#define BHSPARSE_INVALID_MATRIX         -4   /* row_ptr not monotone, col out of range, nnz too large */
#define BHSPARSE_OUT_OF_MEMORY          -5
#define BHSPARSE_HIP_ERROR              -6   /* a HIP runtime call failed; see bhs_last_error() */
#define BHSPARSE_NOT_INITIALIZED        -7   /* call order violated (e.g. spmv before initData) */

/* Kernel variants a handle can run (bhs_*_set_config). */
#define BHS_VARIANT_SEGSUM  0   /* speculative segmented sum (kernels_cuda.h:7-269) */
#define BHS_VARIANT_CSRVEC  1   /* csr-vector (kernels_cuda.h:272-342, CUSP v0.4.0) */
/* The segmented sum over a column-panel copy of A (no reference counterpart):
 * row i's entries in column panel p = [p*ceil(n/tpv), (p+1)*ceil(n/tpv)) form
 * one row of a stacked matrix (panel-major, nonempty (panel, row) pairs only),
 * so the x gathers of consecutive tiles stay in one panel of x (one L2's worth
 * on a random-column matrix); a last kernel adds the tpv panel partials of
 * every row in panel order.  tpv = panels (2, 4 or 8).  Built once per panel
 * count from the device copy of A (host pass, O(nnz)) and kept with the handle. */
#define BHS_VARIANT_SEGSUM_PANELS 2

/* Where init_data puts row_ptr, col, val, x and y (bhs_*_set_memory_mode).
 * DEVICE: hipMalloc (default).  MANAGED: hipMallocManaged, the analogue of the
 * reference's USE_SVM_ALWAYS=1 build (spmv_cuda/Makefile:6,
 * bhsparse_spmv_cuda.h:250-263), where A and x live in unified memory; needs a
 * device with managed memory (hipDeviceAttributeManagedMemory), else
 * init_data returns BHSPARSE_UNSUPPORTED_DEVICE.  HOST_MAPPED: pinned host
 * memory mapped into the device (hipHostMalloc), one pointer valid on host and
 * device, which every device that can map host memory supports; kernels read
 * A and x over the host link.  The tile partition, calibration buffers and
 * panel copy stay in device memory in every mode. */
#define BHS_MEMORY_DEVICE      0
#define BHS_MEMORY_MANAGED     1
#define BHS_MEMORY_HOST_MAPPED 2

typedef struct bhs_handle_opaque *bhs_handle_t;
/* The tuner's network (a device-resident MLP); see "learned tuner" below. */
typedef struct bhs_mlp_opaque *bhs_mlp_t;

/* Counters of one segmented-sum launch (tiles of P = 64*S nonzeros, T tiles
 * per wave).  Filled by bhs_*_get_stats. */
typedef struct {
    long long n_tiles;       /* NP = ceil(nnz / P) */
    long long n_waves;       /* ceil(NP / T) */
    long long n_fast;        /* tiles inside one row (fast track) */
    long long n_dirty;       /* tiles holding an empty row (speculation fixed up) */
    long long n_sync;        /* waves that sent a boundary-row partial to calibration */
    long long n_lead_empty;  /* empty rows before the first nonzero */
    long long chunk_rows;    /* rows whose row_ptr each tile loads ahead: 128, or 256 when
                                at least 5% of the tiles span 128-255 rows (0: csr-vector) */
} bhs_spmv_stats;

/* Tile shape / variant.  S = items per lane (4, 8, 16, 32), waves_per_block
 * in [1, 8], tiles_per_wave >= 1; tpv = csr-vector lanes per row in
 * {0 = rule of main.cu:238-287 extended to 64, 2, 4, 8, 16, 32, 64} for
 * BHS_VARIANT_CSRVEC, column panels in {2, 4, 8} for BHS_VARIANT_SEGSUM_PANELS,
 * ignored for BHS_VARIANT_SEGSUM. */
typedef struct {
    int variant;
    int items_per_lane;
    int waves_per_block;
    int tiles_per_wave;
    int tpv;
} bhs_config;

/* A batch of independent matrices (bhs_*_batch_*). */
typedef struct bhs_batch_opaque *bhs_batch_t;

/* Schedule of a batch: tile shape of the segmented sum, waves per block,
 * persistent grid size (0 = about 16 resident waves per CU), item order
 * (0 = matrix order, 1 = most work first) and x staging: 0, or (fp64 with
 * items_per_lane 4 only) 2 or 4 -- matrices with an even number of
 * right-hand sides load each tile's x values into LDS, x_stage 16-byte
 * chunks (2 right-hand sides each) per nonzero at a time, one cache request
 * per nonzero and pass instead of one per pair; the arithmetic, and y, are
 * the same.  Adds 1 KB + 4 KB * x_stage of LDS per wave. */
typedef struct {
    int items_per_lane;
    int tiles_per_wave;
    int waves_per_block;
    int grid_blocks;
    int order;
    int x_stage;
} bhs_batch_schedule;

/* ---- library-wide ------------------------------------------------------ */
const char *bhs_error_string(int code);
/* Text of the last error raised on this thread (HIP message included). */
const char *bhs_last_error(void);
/* hipDeviceSynchronize on the current device. */
int bhs_device_synchronize(void);
/* Number of visible devices (0 when none). */
int bhs_device_count(void);
/* Device copy bandwidth probe (no reference counterpart): copies `bytes`
 * between two device buffers `reps` times with a 16-byte-per-lane copy kernel
 * and writes each copy's HIP-event time in ms to ms[i].  It gives the box's
 * measured stream-copy ceiling, and under rocprofv3 --pmc the FETCH_SIZE /
 * WRITE_SIZE reading of a kernel whose traffic is known exactly. */
int bhs_stream_copy(int device, long long bytes, int reps, double *ms);

/* ---- fp32 build (value_type = float) ---------------------------------- */
/* new bhsparse_spmv_cuda() -- main.cu:772 */
int bhs_s_create(bhs_handle_t *h);
/* init_platform(): select the device, create the handle's stream
 * (bhsparse_spmv_cuda.h:61-75; the reference always uses device 0). */
int bhs_s_init_platform(bhs_handle_t h, int device);
/* prepare_mem(m, n, nnzA, csrRowPtrA, csrColIdxA, csrValA, x, y)
 * (bhsparse_spmv_cuda.h:238-351): validates and copies A and x to the device
 * and partitions the nonzeros into tiles.  y is borrowed, as in the reference
 * (:248): get_y() writes m values into it later.  Its contents are never read
 * (beta = 0). */
int bhs_s_init_data(bhs_handle_t h, int m, int n, int nnz, const int *row_ptr,
                    const int *col, const float *val, const float *x, float *y);
/* Replace x (n values) without re-uploading A. No reference counterpart
 * (the reference re-runs prepare_mem). */
int bhs_s_set_x(bhs_handle_t h, const float *x);
/* Memory mode used by the next init_data (USE_SVM_ALWAYS, a compile-time
 * switch in the reference, spmv_cuda/Makefile:6).  Rejected while data loaded in the
 * other mode is held (free_mem first).  BHSPARSE_UNSUPPORTED_DEVICE from
 * init_data if the device has no managed memory. */
int bhs_s_set_memory_mode(bhs_handle_t h, int mode);
/* Mode of the loaded data, or the mode the next init_data will use. */
int bhs_s_get_memory_mode(bhs_handle_t h, int *mode);
/* run_benchmark() (bhsparse_spmv_cuda.h:121-228), plus alpha, which the
 * reference drops (utils_cuda.h:133).  Blocking, like the reference. */
int bhs_s_spmv(bhs_handle_t h, float alpha);
/* The same work enqueued on the handle's stream without waiting. */
int bhs_s_spmv_async(bhs_handle_t h, float alpha);
/* The handle's HIP stream (a hipStream_t), on which bhs_s_spmv_async enqueues.
 * No reference counterpart.  Once init_data (and set_config, if used) have
 * run, an spmv_async call enqueues kernel launches only -- no memset, copy or
 * host step -- so a caller may capture it into a hipGraph and replay it; the
 * graph reads the handle's current x on every replay.  The graph holds the
 * buffers of the tile shape and layout it was captured with: capture again
 * after set_config, init_data or free_mem. */
int bhs_s_get_stream(bhs_handle_t h, void **stream);
/* get_y() (bhsparse_spmv_cuda.h:230-236): copies y into the pointer given to
 * init_data. */
int bhs_s_get_y(bhs_handle_t h);
/* Copies y into `out` (m values) instead of the borrowed pointer. */
int bhs_s_get_y_to(bhs_handle_t h, float *out);
/* free_mem() (bhsparse_spmv_cuda.h:84-118); safe to call twice. */
int bhs_s_free_mem(bhs_handle_t h);
/* free_platform() (bhsparse_spmv_cuda.h:77-82). */
int bhs_s_free_platform(bhs_handle_t h);
/* delete: frees everything the handle still owns. */
int bhs_s_destroy(bhs_handle_t h);
/* Tile shape and variant (compile-time THREADBUNCH/SEG_H/STEP in the
 * reference, spmv_cuda/common.h:39-48).  Re-partitions if the tile size changes. */
int bhs_s_set_config(bhs_handle_t h, const bhs_config *cfg);
int bhs_s_get_config(bhs_handle_t h, bhs_config *cfg);
/* Debug: copies tile_row[0..NP] (first row of every tile boundary) into
 * tile_row (capacity cap) and NP into *np. */
int bhs_s_get_partition(bhs_handle_t h, int *tile_row, int cap, int *np);
/* Debug: runs one counting launch of the segmented sum and reports its counters. */
int bhs_s_get_stats(bhs_handle_t h, bhs_spmv_stats *out);
/* Timing: `warmup` untimed calls, then `reps` calls each bracketed by two HIP
 * events on the handle's stream (nothing between the call's kernels).  If
 * flush_bytes > 0, a buffer of that size is written before every call (outside
 * the events) so caches start cold.  ms_total[i] = whole call; ms_main[i] =
 * main kernel only, from a second pass of calls with events attached to that
 * kernel's dispatch (both may be NULL). */
int bhs_s_bench(bhs_handle_t h, float alpha, int warmup, int reps, long long flush_bytes,
                double *ms_total, double *ms_main);
/* GPU reduction over row_ptr/col: 8 row-length features (m, nnz, mean, var,
 * max, empty fraction, bandwidth estimate, precision flag) used by the tuner. */
int bhs_s_row_features(bhs_handle_t h, double *feat8);
/* Data-movement floor (no reference counterpart): times a kernel that streams
 * val and col and gathers x exactly like the main kernel (same tile loads, same
 * gathers) but does no segmentation and writes no y.  ms[i] per rep; with
 * flush_bytes > 0 caches are flushed before each rep.  The ratio of this time
 * to bhs_s_bench's says how close the SpMV is to what the matrix's own memory
 * traffic allows on the device. */
int bhs_s_bench_floor(bhs_handle_t h, int reps, long long flush_bytes, double *ms);
/* Timing calibration (no reference counterpart): ms[i] = an empty interval
 * between the two events bhs_s_bench uses for ms_main, after the same flush.
 * Those events cost a few microseconds of stream time; ms_main minus the
 * median of this is the kernel's own duration, as a kernel trace reports it. */
int bhs_s_bench_event_overhead(bhs_handle_t h, int reps, long long flush_bytes, double *ms);
/* Exhaustive autotuner (no reference counterpart; the reference fixes its tile
 * shape at compile time, spmv_cuda/common.h:39-48): times every candidate
 * (median of `reps` calls, caches flushed when flush_bytes > 0), writes us[i]
 * (-1 when a candidate cannot run), installs the fastest and returns its index. */
int bhs_s_autotune(bhs_handle_t h, const bhs_config *cands, int n_cands, int reps,
                   long long flush_bytes, double *us, int *best);
/* Learned choice: the 8 row statistics are reduced on the GPU, transformed,
 * run through `net` on the GPU, and the class with the lowest predicted cost
 * among those that can run is installed (*chosen = its index). */
int bhs_s_apply_tuner(bhs_handle_t h, bhs_mlp_t net, const bhs_config *classes, int n_classes,
                      int *chosen);
/* Learned shortlist (no reference counterpart): the topk classes that can run
 * with the lowest predicted cost are timed as bhs_s_autotune times them
 * (median of `reps` calls, caches flushed when flush_bytes > 0) and the
 * fastest is installed (*chosen = its index in classes).  topk = 1 is
 * bhs_s_apply_tuner; topk = n_classes is the exhaustive autotuner. */
int bhs_s_apply_tuner_topk(bhs_handle_t h, bhs_mlp_t net, const bhs_config *classes,
                           int n_classes, int topk, int reps, long long flush_bytes, int *chosen);

/* ---- fp64 build (value_type = double) --------------------------------- */
int bhs_d_create(bhs_handle_t *h);
int bhs_d_init_platform(bhs_handle_t h, int device);
int bhs_d_init_data(bhs_handle_t h, int m, int n, int nnz, const int *row_ptr,
                    const int *col, const double *val, const double *x, double *y);
int bhs_d_set_x(bhs_handle_t h, const double *x);
int bhs_d_set_memory_mode(bhs_handle_t h, int mode);
int bhs_d_get_memory_mode(bhs_handle_t h, int *mode);
int bhs_d_spmv(bhs_handle_t h, double alpha);
int bhs_d_spmv_async(bhs_handle_t h, double alpha);
int bhs_d_get_stream(bhs_handle_t h, void **stream);
int bhs_d_get_y(bhs_handle_t h);
int bhs_d_get_y_to(bhs_handle_t h, double *out);
int bhs_d_free_mem(bhs_handle_t h);
int bhs_d_free_platform(bhs_handle_t h);
int bhs_d_destroy(bhs_handle_t h);
int bhs_d_set_config(bhs_handle_t h, const bhs_config *cfg);
int bhs_d_get_config(bhs_handle_t h, bhs_config *cfg);
int bhs_d_get_partition(bhs_handle_t h, int *tile_row, int cap, int *np);
int bhs_d_get_stats(bhs_handle_t h, bhs_spmv_stats *out);
int bhs_d_bench(bhs_handle_t h, double alpha, int warmup, int reps, long long flush_bytes,
                double *ms_total, double *ms_main);
int bhs_d_row_features(bhs_handle_t h, double *feat8);
int bhs_d_bench_floor(bhs_handle_t h, int reps, long long flush_bytes, double *ms);
int bhs_d_bench_event_overhead(bhs_handle_t h, int reps, long long flush_bytes, double *ms);
int bhs_d_autotune(bhs_handle_t h, const bhs_config *cands, int n_cands, int reps,
                   long long flush_bytes, double *us, int *best);
int bhs_d_apply_tuner(bhs_handle_t h, bhs_mlp_t net, const bhs_config *classes, int n_classes,
                      int *chosen);
int bhs_d_apply_tuner_topk(bhs_handle_t h, bhs_mlp_t net, const bhs_config *classes,
                           int n_classes, int topk, int reps, long long flush_bytes, int *chosen);

/* ---- allocation-free SpMV ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): y = alpha * A * x over the caller's DEVICE arrays
 * (row_ptr[m+1], col[nnz], val[nnz], x[n], y[m]; col and val 16-byte
 * aligned), enqueued on `stream` (a hipStream_t, NULL = the null stream) and
 * not waited for.  It runs the kernels of a handle's default tile (segmented
 * sum, 8 items per lane, 4 waves per block, 4 tiles per wave, plain layout),
 * so y equals bhs_*_spmv's with that tile bit for bit, with the tile
 * partition and the cross-wave partials in a static device scratch array of
 * 8 MiB per device (held until the process ends).  It calls no hipMalloc,
 * hipFree, memset or copy.  A matrix needing more scratch than that
 * ((nnz/512 + 1) * 4 bytes of partition plus (nnz/2048) * (4 + sizeof(T))
 * bytes of partials, about 600M nonzeros in fp64) is refused with
 * BHSPARSE_INVALID_ARGUMENT.  The caller guarantees a valid CSR (row_ptr
 * monotone from 0 to nnz, 0 <= col < n): device data is not checked.
 * Every call on a device shares its scratch: calls that overlap on one device
 * (two host threads, or two streams) race and must be serialised. */
int bhs_s_spmv_noalloc(int m, int n, int nnz, const int *row_ptr, const int *col,
                       const float *val, const float *x, float *y, float alpha, void *stream);
int bhs_d_spmv_noalloc(int m, int n, int nnz, const int *row_ptr, const int *col,
                       const double *val, const double *x, double *y, double alpha, void *stream);

/* ---- batched solves (no reference counterpart: the reference runs one
 * matrix and one right-hand side per process, spmv_cuda/main.cu:813-816) ---- */
/* Copies n_mats CSR matrices (validated like init_data); matrix i gets k[i]
 * (1..16) right-hand sides.  Default schedule {8, 4, 4, 0, 1}. */
int bhs_d_batch_create(bhs_batch_t *b, int device, int n_mats, const int *m, const int *n,
                       const int *nnz, const int *const *row_ptr, const int *const *col,
                       const double *const *val, const int *k);
/* X[i] is n[i] x k[i], row-major. */
int bhs_d_batch_set_x(bhs_batch_t b, const double *const *X);
int bhs_d_batch_set_schedule(bhs_batch_t b, const bhs_batch_schedule *s);
int bhs_d_batch_get_schedule(bhs_batch_t b, bhs_batch_schedule *s);
/* Y_i = alpha * A_i * X_i for every matrix: one hipGraph replay of
 * [queue reset] -> [persistent kernel over a global (matrix, tile range) queue]
 * -> [calibration].  Blocking. */
int bhs_d_batch_run(bhs_batch_t b, double alpha);
/* The baseline: one plain SpMV (main kernel + calibration) per matrix per
 * right-hand side, back to back on one stream, on contiguous vectors. */
int bhs_d_batch_run_plain(bhs_batch_t b, double alpha);
/* Y[i] (m[i] x k[i], row-major) from the batched (plain = 0) or baseline run. */
int bhs_d_batch_get_y(bhs_batch_t b, double *const *Y, int plain);
/* HIP-event times (ms) of `reps` runs: plain = 0 the batched graph, 1 the
 * baseline's launches, 2 the baseline's launches replayed as one hipGraph
 * (its kernels without the per-launch overhead). */
int bhs_d_batch_bench(bhs_batch_t b, double alpha, int plain, int warmup, int reps, double *ms);
/* Times every candidate schedule (median of reps), installs the fastest. */
int bhs_d_batch_autotune(bhs_batch_t b, const bhs_batch_schedule *cands, int n_cands, int reps,
                         double *ms, int *best);
int bhs_d_batch_destroy(bhs_batch_t b);
/* The same in float. */
int bhs_s_batch_create(bhs_batch_t *b, int device, int n_mats, const int *m, const int *n,
                       const int *nnz, const int *const *row_ptr, const int *const *col,
                       const float *const *val, const int *k);
int bhs_s_batch_set_x(bhs_batch_t b, const float *const *X);
int bhs_s_batch_set_schedule(bhs_batch_t b, const bhs_batch_schedule *s);
int bhs_s_batch_get_schedule(bhs_batch_t b, bhs_batch_schedule *s);
int bhs_s_batch_run(bhs_batch_t b, float alpha);
int bhs_s_batch_run_plain(bhs_batch_t b, float alpha);
int bhs_s_batch_get_y(bhs_batch_t b, float *const *Y, int plain);
int bhs_s_batch_bench(bhs_batch_t b, float alpha, int plain, int warmup, int reps, double *ms);
int bhs_s_batch_autotune(bhs_batch_t b, const bhs_batch_schedule *cands, int n_cands, int reps,
                         double *ms, int *best);
int bhs_s_batch_destroy(bhs_batch_t b);

/* ---- learned tuner (no reference counterpart) -------------------------- */
/* An MLP n_in(=8) -> hidden -> hidden -> n_out with ReLU, trained on the GPU
 * by the library's own forward/backward/Adam kernels.  Outputs are predicted
 * log2(time / best time) per class; the tuner takes the argmin.
 * Parameters are one float array: W1[h*8], b1[h], W2[h*h], b2[h], W3[k*h], b3[k];
 * inputs are standardised as (x - mean) / std inside the network. */
long long bhs_mlp_param_count(int n_in, int n_hidden, int n_out);
/* raw = {m, nnz, mean len, var len, max len, empty fraction, mean column span,
 * 1 if fp64}: the transform the tuner applies before standardising (logs of
 * the sizes and lengths, the span as a fraction of m). */
int bhs_tuner_transform(const double *raw8, float *out8);
int bhs_mlp_create(bhs_mlp_t *net, int device, int n_in, int n_hidden, int n_out);
int bhs_mlp_destroy(bhs_mlp_t net);
int bhs_mlp_set_params(bhs_mlp_t net, const float *params, long long n, const float *norm_mean,
                       const float *norm_std);
int bhs_mlp_get_params(bhs_mlp_t net, float *params, long long n);
/* out[n_rows * n_out] for transformed features x[n_rows * 8]. */
int bhs_mlp_forward(bhs_mlp_t net, const float *x, int n_rows, float *out);
/* Masked mean squared error against target and its gradient w.r.t. the parameters. */
int bhs_mlp_loss_grad(bhs_mlp_t net, const float *x, const float *target, const float *mask,
                      int n_rows, double *loss, float *grad);
/* `epochs` full-batch Adam steps on the device; loss_hist[epochs] (may be NULL). */
int bhs_mlp_train(bhs_mlp_t net, const float *x, const float *target, const float *mask,
                  int n_rows, int epochs, float lr, float weight_decay, double *loss_hist);
/* The same forward/backward kernels instantiated in double, for gradient checks;
 * norm = {mean[8], std[8]}. */
int bhs_mlp_loss_grad_f64(int device, int n_in, int n_hidden, int n_out, const double *params,
                          const double *norm, const double *x, const double *target,
                          const double *mask, int n_rows, double *loss, double *grad);

#ifdef __cplusplus
}
#endif
#endif /* BHSPARSE_SPMV_H */
