/*
 * oracle.c -- CPU oracle for the speculative segmented-sum CSR SpMV.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  The shipped library never links or calls it.  Built with
 * -ffp-contract=off so every sum below rounds exactly as written.
 *
 * Citations are into weifengliu-ssslab/Benchmark_SpMV_using_CSR.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Planted-mistake switch (tests only): 0 = the oracle as written. */
static int g_mutation = 0;
void oracle_set_mutation(int kind) { g_mutation = kind; }

/* spmv_cuda/main.cu:755-761 -- the serial loop, accumulating in FP64. */
void oracle_spmv_serial_f64(int m, const int *row_ptr, const int *col,
                            const double *val, const double *x, double *y)
{
    if (g_mutation >= ORACLE_MUT_SERIAL_DROP_LAST && g_mutation <= ORACLE_MUT_SERIAL_TRANSPOSE) {
        for (int i = 0; i < m; i++) y[i] = 0;
        for (int i = 0; i < m; i++) {
            double sum = 0;
            float sum32 = 0;
            for (int j = row_ptr[i]; j < row_ptr[i + 1]; j++) {
                switch (g_mutation) {
                case ORACLE_MUT_SERIAL_DROP_LAST:
                    if (j + 1 < row_ptr[i + 1]) sum += x[col[j]] * val[j];
                    break;
                case ORACLE_MUT_SERIAL_SIGN:
                    sum += (j == row_ptr[i] ? -1.0 : 1.0) * x[col[j]] * val[j];
                    break;
                case ORACLE_MUT_SERIAL_F32_ACC:
                    sum32 += (float)(x[col[j]] * val[j]);
                    break;
                case ORACLE_MUT_SERIAL_TRANSPOSE:      /* x[i] read as if n == m */
                    if (col[j] < m && i < m) y[col[j]] += x[i] * val[j];
                    break;
                }
            }
            if (g_mutation == ORACLE_MUT_SERIAL_F32_ACC) y[i] = sum32;
            else if (g_mutation != ORACLE_MUT_SERIAL_TRANSPOSE) y[i] = sum;
        }
        return;
    }
    for (int i = 0; i < m; i++) {
        double sum = 0;
        for (int j = row_ptr[i]; j < row_ptr[i + 1]; j++)
            sum += x[col[j]] * val[j];
        y[i] = sum;
    }
}

/* spmv_cuda/main.cu:755-761 with value_type = float. */
void oracle_spmv_serial_f32(int m, const int *row_ptr, const int *col,
                            const float *val, const float *x, float *y)
{
    for (int i = 0; i < m; i++) {
        float sum = 0;
        for (int j = row_ptr[i]; j < row_ptr[i + 1]; j++)
            sum += x[col[j]] * val[j];
        y[i] = sum;
    }
}

static int set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* spmv_cuda/main.cu:377-393 -- the OpenMP row-parallel loop (the reference
 * pins 4 threads at main.cu:370; here the caller chooses). */
int oracle_spmv_omp_f64(int m, const int *row_ptr, const int *col,
                        const double *val, const double *x, double *y, int nthreads)
{
    int t = set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; i++) {
        double sum = 0;
        for (int j = row_ptr[i]; j < row_ptr[i + 1]; j++)
            sum += x[col[j]] * val[j];
        y[i] = sum;
    }
    return t;
}

int oracle_spmv_omp_f32(int m, const int *row_ptr, const int *col,
                        const float *val, const float *x, float *y, int nthreads)
{
    int t = set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; i++) {
        float sum = 0;
        for (int j = row_ptr[i]; j < row_ptr[i + 1]; j++)
            sum += x[col[j]] * val[j];
        y[i] = sum;
    }
    return t;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int cmp_dbl(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

double oracle_time_spmv_f64(int m, const int *row_ptr, const int *col,
                            const double *val, const double *x, double *y,
                            int nthreads, int reps, int *threads_used)
{
    if (reps < 1) reps = 1;
    double *t = (double *)malloc(sizeof(double) * (size_t)reps);
    int used = 1;
    for (int r = 0; r < reps; r++) {
        double t0 = now_s();
        if (nthreads == 1) oracle_spmv_serial_f64(m, row_ptr, col, val, x, y);
        else used = oracle_spmv_omp_f64(m, row_ptr, col, val, x, y, nthreads);
        t[r] = now_s() - t0;
    }
    qsort(t, (size_t)reps, sizeof(double), cmp_dbl);
    double med = t[reps / 2];
    free(t);
    if (threads_used) *threads_used = used;
    return med;
}

double oracle_time_spmv_f32(int m, const int *row_ptr, const int *col,
                            const float *val, const float *x, float *y,
                            int nthreads, int reps, int *threads_used)
{
    if (reps < 1) reps = 1;
    double *t = (double *)malloc(sizeof(double) * (size_t)reps);
    int used = 1;
    for (int r = 0; r < reps; r++) {
        double t0 = now_s();
        if (nthreads == 1) oracle_spmv_serial_f32(m, row_ptr, col, val, x, y);
        else used = oracle_spmv_omp_f32(m, row_ptr, col, val, x, y, nthreads);
        t[r] = now_s() - t0;
    }
    qsort(t, (size_t)reps, sizeof(double), cmp_dbl);
    double med = t[reps / 2];
    free(t);
    if (threads_used) *threads_used = used;
    return med;
}

/* spmv_cuda/utils_cuda.h:6-34: bisection returning the first index whose
 * row_ptr value exceeds key (std::upper_bound over size entries). */
static int upper_bound_impl(const int *row_ptr, int size, int key, int lower)
{
    int start = 0, stop = size - 1;
    while (stop >= start) {
        int median = start + (stop - start) / 2;
        if (lower ? key > row_ptr[median] : key >= row_ptr[median]) start = median + 1;
        else stop = median - 1;
    }
    return start;
}

int oracle_upper_bound(const int *row_ptr, int size, int key)
{
    return upper_bound_impl(row_ptr, size, key, g_mutation == ORACLE_MUT_LOWER_BOUND);
}

/* Row lookups inside the restatement never take the planted mistake (a wrong
 * row there could index outside y); the mistake reaches the partition and
 * upper_bound entry points only. */
static int row_of(const int *row_ptr, int m, int nnz, long long k)
{
    int key = k > nnz ? nnz : (int)k;          /* clamp, kernels_cuda.h:47 */
    return upper_bound_impl(row_ptr, m + 1, key, 0) - 1;
}

/* spmv_cuda/kernels_cuda.h:39-49 evaluated for every tile boundary. */
void oracle_partition(int m, int nnz, const int *row_ptr, int P, int *tile_row)
{
    long long NP = ((long long)nnz + P - 1) / P;
    for (long long p = 0; p <= NP; p++) {
        long long k = p * (long long)P;
        int key = k > nnz ? nnz : (int)k;
        tile_row[p] = oracle_upper_bound(row_ptr, m + 1, key) - 1;
    }
}

void oracle_tile_stats_compute(int m, int nnz, const int *row_ptr, int P, int T,
                               oracle_tile_stats *o)
{
    memset(o, 0, sizeof(*o));
    long long NP = ((long long)nnz + P - 1) / P;
    long long NW = (NP + T - 1) / T;
    o->n_tiles = NP;
    o->n_waves = NW;
    int r_first0 = row_of(row_ptr, m, nnz, 0);
    o->n_lead_empty = r_first0;
    for (long long p = 0; p < NP; p++) {
        long long B = p * P;
        int r0 = row_of(row_ptr, m, nnz, B);
        int r1 = row_of(row_ptr, m, nnz, B + P);
        if (r0 == r1) { o->n_fast++; continue; }
        int stop = r1 == m ? m - 1 : r1;
        for (int r = r0; r <= stop; r++)
            if (row_ptr[r] == row_ptr[r + 1]) { o->n_dirty++; break; }
    }
    for (long long w = 0; w < NW; w++) {
        long long E0 = w * T * (long long)P;
        int rf = row_of(row_ptr, m, nnz, E0);
        if ((long long)row_ptr[rf] < E0) o->n_sync++;
    }
}

/* ------------------------------------------------------------------------ */
/* Reference speculative segmented sum (kernels_cuda.h:7-269 +              */
/* bhsparse_spmv_cuda.h:166-222), FP64, with y cleared first and the fix-up */
/* kept inside [0, m).                                                      */
/* ------------------------------------------------------------------------ */
int oracle_segsum_ref(int W, int S, int T, int m, int nnz, const int *row_ptr,
                      const int *col, const double *val, const double *x,
                      double *y, oracle_ref_stats *stats)
{
    if (W < 1 || S < 1 || T < 1 || m < 0 || nnz < 0 || S > 32) return -1;
    const int P = W * S;
    const long long NP = ((long long)nnz + P - 1) / P;       /* bhsparse_spmv_cuda.h:326-327 */
    const long long NB = (NP + T - 1) / T;                   /* :328 */
    oracle_ref_stats st;
    memset(&st, 0, sizeof(st));
    st.n_tiles = NP;
    st.n_bunches = NB;

    for (int i = 0; i < m; i++) y[i] = 0;                    /* beta = 0 (correction) */

    int *sync_idx = (int *)calloc((size_t)(NB > 0 ? NB : 1), sizeof(int));
    double *sync_val = (double *)calloc((size_t)(NB > 0 ? NB : 1), sizeof(double));
    /* dirty record (start row, stop row, tile end); the reference keeps the
     * first two (kernels_cuda.h:264-266) */
    long long *spec = (long long *)malloc(sizeof(long long) * 3 * (size_t)(NP > 0 ? NP : 1));
    long long n_spec = 0;

    double *prod = (double *)malloc(sizeof(double) * (size_t)P);
    unsigned *bf = (unsigned *)malloc(sizeof(unsigned) * (size_t)W);
    double *slot = (double *)malloc(sizeof(double) * (size_t)W * (size_t)S);
    int *lstart = (int *)malloc(sizeof(int) * (size_t)W);
    int *lstop = (int *)malloc(sizeof(int) * (size_t)W);
    int *present = (int *)malloc(sizeof(int) * (size_t)(W + 1));
    int *segn = (int *)malloc(sizeof(int) * (size_t)W);
    double *lead_next = (double *)malloc(sizeof(double) * (size_t)W);
    double *lane = (double *)malloc(sizeof(double) * (size_t)W);

    for (long long b = 0; b < NB; b++) {
        const long long off = b * T * (long long)P;                        /* kernels_cuda.h:39 */
        const int first = row_of(row_ptr, m, nnz, off);                    /* tbunch_start, :53 */
        double tail = 0;                                                   /* s_tail, :58 */
        int bs_stop = first;
        for (int i = 0; i < T; i++) {
            if (i + b * T >= NP) break;                                    /* :67 */
            const long long B = off + (long long)i * P;
            const int bs_start = bs_stop;
            bs_stop = row_of(row_ptr, m, nnz, B + P);                      /* :76 */
            int dirty = 0;
            if (bs_start == bs_stop) {                                     /* fast track :79-112 */
                st.n_fast++;
                for (int l = 0; l < W; l++) {                              /* lane-strided sum :81-89 */
                    double fsum = 0;
                    for (int j = 0; j < S; j++) {
                        long long e = B + (long long)j * W + l;
                        fsum += e < nnz ? val[e] * x[col[e]] : 0;
                    }
                    lane[l] = fsum;
                }
                for (int mask = W / 2; mask > 0; mask >>= 1) {             /* sum_32_shfl :69-78 */
                    for (int l = 0; l < W; l++) lead_next[l] = lane[l] + lane[l ^ mask];
                    memcpy(lane, lead_next, sizeof(double) * (size_t)W);
                }
                double fsum = (row_ptr[bs_start] == B) ? lane[0] : (tail + lane[0]);   /* :96 */
                if (bs_start == first) { sync_idx[b] = bs_start; sync_val[b] = fsum; }
                else y[bs_start] = fsum;                                   /* :98-106 */
                tail = fsum;                                               /* :108 */
                continue;
            }
            /* step 2: bit flags (:114-145); bit (S-1-k) of lane word marks item k
             * (the reference uses bit 31-k of a 32-bit word). */
            for (int l = 0; l < W; l++) bf[l] = 0;
            const int stop = bs_stop == m ? m - 1 : bs_stop;               /* :119 */
            for (int r = bs_start; r <= stop; r++) {
                if (row_ptr[r] != row_ptr[r + 1]) {
                    long long o = row_ptr[r] - B;
                    if (o >= 0) {
                        long long slice = o / S;
                        if (slice < W) bf[slice] |= 1u << (unsigned)(o % S);   /* clamp :132-133 */
                    }
                } else {
                    dirty = 1;
                }
            }
            /* step 3: products, lane l owns items [lS, lS+S) (:150-162) */
            for (int k = 0; k < P; k++) {
                long long e = B + k;
                prod[k] = e < nnz ? val[e] * x[col[e]] : 0;
            }
            /* step 3-1: carry into item 0 unless it starts a row (:164-172) */
            if (!(bf[0] & 1u) && g_mutation != ORACLE_MUT_NO_CARRY) prod[0] = prod[0] + tail;
            bf[0] |= 1u;
            /* step 4-0: lane-sequential segments (:174-203) */
            for (int l = 0; l < W; l++) {
                const double *pl = prod + (size_t)l * S;
                double *sl = slot + (size_t)l * S;
                int bit = bf[l] & 1u;
                lstart[l] = !bit;
                lstop[l] = 0;
                double sum = pl[0];
                present[l] = (l == 0) | bit;
                for (int j = 1; j < S; j++) {
                    bit = (bf[l] >> j) & 1u;
                    if (bit) { sl[lstop[l]] = sum; sum = 0; lstop[l]++; }
                    sum += pl[j];
                    present[l] |= bit;
                }
                sl[lstop[l]] = sum;
                int sn = lstop[l] - lstart[l] + present[l];
                segn[l] = sn > 0 ? sn : 0;
            }
            present[W] = 1;                                                /* sentinel :57 */
            /* step 4: exclusive scan of segment counts (:205-207) */
            int acc = 0;
            for (int l = 0; l < W; l++) { int t = segn[l]; segn[l] = acc; acc += t; }
            /* step 4: leading partial of lane l+1, shifted down one lane (:210-214) */
            for (int l = 0; l < W; l++)
                lead_next[l] = (l == W - 1) ? 0 : lstart[l + 1] * slot[(size_t)(l + 1) * S];
            /* step 4-1/4-2: serial spanning walk, then add to the trailing slot (:218-235) */
            for (int l = 0; l < W; l++) {
                double tmp = 0;
                if (present[l] && g_mutation != ORACLE_MUT_NO_SPAN) {
                    tmp = lead_next[l];
                    int id = l + 1;
                    while (!present[id]) { tmp += lead_next[id]; id++; }
                }
                slot[(size_t)l * S + lstop[l]] += tmp;
            }
            /* steps 5/6: speculative emission (:238-257) */
            for (int l = 0; l < W; l++) {
                int loc = bs_start + segn[l];
                for (int j = lstart[l]; j <= lstop[l]; j++) {
                    double v = slot[(size_t)l * S + j];
                    if (loc == first) { sync_idx[b] = first; sync_val[b] = v; }
                    else if (loc < m) y[loc] = v;
                    if (loc == bs_stop) tail = v;
                    loc++;
                }
            }
            /* step 7: dirty record (:259-267) */
            if (dirty) {
                st.n_dirty++;
                tail = 0;
                spec[3 * n_spec] = bs_start;
                spec[3 * n_spec + 1] = bs_stop;
                spec[3 * n_spec + 2] = B + P;
                n_spec++;
            }
        }
    }

    /* step 8: calibration (bhsparse_spmv_cuda.h:170-173) */
    if (g_mutation != ORACLE_MUT_NO_CALIBRATION)
        for (long long b = 0; b < NB; b++) y[sync_idx[b]] += sync_val[b];

    /* step 9: fix-up of dirty tiles (bhsparse_spmv_cuda.h:184-222) */
    if (n_spec && g_mutation != ORACLE_MUT_NO_FIXUP) {
        double *t = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
        memcpy(t, y, sizeof(double) * (size_t)m);
        for (long long i = 0; i < n_spec; i++) {
            const int a = (int)spec[3 * i], z = (int)spec[3 * i + 1];
            const long long tend = spec[3 * i + 2];
            int q = a;
            const int zz = z < m ? z : m - 1;
            for (int r = a; r <= zz; r++) {
                if (row_ptr[r] != row_ptr[r + 1]) {
                    if (r == z) {
                        /* add the partial only when row z was the dirty tile's
                         * trailing segment, i.e. it starts before the tile's end */
                        if ((long long)row_ptr[z] < tend) { y[z] += t[q]; t[z] += t[q]; }
                        q++;
                    } else {
                        y[r] = t[q++];
                    }
                } else {
                    y[r] = 0;
                }
            }
        }
        free(t);
    }

    free(sync_idx); free(sync_val); free(spec); free(prod); free(bf); free(slot);
    free(lstart); free(lstop); free(present); free(segn); free(lead_next); free(lane);
    if (stats) *stats = st;
    return 0;
}

double oracle_rel_error_f64(int m, const int *row_ptr, const int *col,
                            const double *val, const double *x,
                            const double *y, const double *y_ref, double *err)
{
    double emax = 0;
    for (int i = 0; i < m; i++) {
        double den = 0;
        for (int j = row_ptr[i]; j < row_ptr[i + 1]; j++) den += fabs(val[j] * x[col[j]]);
        double num = fabs(y[i] - y_ref[i]);
        double e;
        if (num == 0) e = 0;
        else if (den == 0) e = INFINITY;
        else e = num / den;
        if (y[i] != y[i]) e = INFINITY;     /* NaN never passes */
        if (err) err[i] = e;
        if (e > emax) emax = e;
    }
    return emax;
}

/* Textbook backpropagation for the tuner's MLP (FP64). */
double oracle_mlp_loss_grad(int n_in, int h, int n_out, const double *params,
                            const double *norm, const double *X, const double *T,
                            const double *M, int N, double *out, double *grad)
{
    const double *W1 = params, *b1 = W1 + (size_t)h * n_in, *W2 = b1 + h,
                 *b2 = W2 + (size_t)h * h, *W3 = b2 + h, *b3 = W3 + (size_t)n_out * h;
    const size_t P = (size_t)h * n_in + h + (size_t)h * h + h + (size_t)n_out * h + n_out;
    double *xs = malloc(sizeof(double) * (size_t)n_in);
    double *a1 = malloc(sizeof(double) * (size_t)h), *a2 = malloc(sizeof(double) * (size_t)h);
    double *o = malloc(sizeof(double) * (size_t)n_out), *d3 = malloc(sizeof(double) * (size_t)n_out);
    double *d2 = malloc(sizeof(double) * (size_t)h), *d1 = malloc(sizeof(double) * (size_t)h);
    double msum = 0, loss = 0;
    for (size_t i = 0; i < (size_t)N * n_out; i++) msum += M[i];
    if (msum <= 0) msum = 1;
    if (grad) for (size_t i = 0; i < P; i++) grad[i] = 0;
    double *gW1 = grad, *gb1 = grad ? gW1 + (size_t)h * n_in : NULL, *gW2 = grad ? gb1 + h : NULL,
           *gb2 = grad ? gW2 + (size_t)h * h : NULL, *gW3 = grad ? gb2 + h : NULL,
           *gb3 = grad ? gW3 + (size_t)n_out * h : NULL;
    for (int n = 0; n < N; n++) {
        for (int i = 0; i < n_in; i++) xs[i] = (X[(size_t)n * n_in + i] - norm[i]) / norm[n_in + i];
        for (int j = 0; j < h; j++) {
            double a = b1[j];
            for (int i = 0; i < n_in; i++) a += xs[i] * W1[(size_t)j * n_in + i];
            a1[j] = a > 0 ? a : 0;
        }
        for (int j = 0; j < h; j++) {
            double a = b2[j];
            for (int i = 0; i < h; i++) a += a1[i] * W2[(size_t)j * h + i];
            a2[j] = a > 0 ? a : 0;
        }
        for (int k = 0; k < n_out; k++) {
            double a = b3[k];
            for (int i = 0; i < h; i++) a += a2[i] * W3[(size_t)k * h + i];
            o[k] = a;
            if (out) out[(size_t)n * n_out + k] = a;
            const double mk = M[(size_t)n * n_out + k], d = a - T[(size_t)n * n_out + k];
            loss += mk * d * d;
            /* planted mistake: the mask dropped from dLoss/dOut */
            d3[k] = 2 * (g_mutation == ORACLE_MUT_MLP_NO_MASK ? 1.0 : mk) * d / msum;
        }
        if (!grad) continue;
        for (int k = 0; k < n_out; k++) {
            gb3[k] += d3[k];
            for (int i = 0; i < h; i++) gW3[(size_t)k * h + i] += d3[k] * a2[i];
        }
        for (int i = 0; i < h; i++) {
            double a = 0;
            for (int k = 0; k < n_out; k++) a += d3[k] * W3[(size_t)k * h + i];
            /* planted mistake: the ReLU derivative ignored */
            d2[i] = (a2[i] > 0 || g_mutation == ORACLE_MUT_MLP_RELU_GRAD) ? a : 0;
        }
        for (int j = 0; j < h; j++) {
            gb2[j] += d2[j];
            for (int i = 0; i < h; i++) gW2[(size_t)j * h + i] += d2[j] * a1[i];
        }
        for (int i = 0; i < h; i++) {
            double a = 0;
            /* planted mistake: W2 transposed in the backward pass */
            for (int j = 0; j < h; j++)
                a += d2[j] * (g_mutation == ORACLE_MUT_MLP_TRANSPOSE ? W2[(size_t)i * h + j]
                                                                      : W2[(size_t)j * h + i]);
            d1[i] = a1[i] > 0 ? a : 0;
        }
        for (int j = 0; j < h; j++) {
            gb1[j] += d1[j];
            for (int i = 0; i < n_in; i++) gW1[(size_t)j * n_in + i] += d1[j] * xs[i];
        }
    }
    free(xs); free(a1); free(a2); free(o); free(d3); free(d2); free(d1);
    return loss / msum;
}
