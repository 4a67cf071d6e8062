// bhs_tuner.h -- the learned tile-shape tuner: feature transform (shared by
// host and device), MLP object, and the kernels' launch helpers.
//
// No reference counterpart: the reference fixes the tile shape at compile time
// (THREADBUNCH / SEG_H / STEP, spmv_cuda/common.h:39-48) and picks the
// csr-vector width with a nnz/m rule (main.cu:238-287).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace bhs {

constexpr int kTunerFeatures = 8;
constexpr int kMaxHidden = 128;
constexpr int kMaxClasses = 64;

// raw = {m, nnz, mean row length, row-length variance, max row length,
// empty-row fraction, mean column span of nonempty rows, 1 if fp64 else 0}
// -> scale-free inputs of comparable range.  The span enters relative to m
// (the x window a row touches, as a fraction of x for a square matrix): it
// separates clustered columns (stencils, bands) from scattered ones whatever
// the matrix size.
__host__ __device__ inline void tuner_transform(const double *raw, float *f)
{
    f[0] = (float)(log10(raw[0] > 1 ? raw[0] : 1.0) / 7.0);
    f[1] = (float)(log10(raw[1] + 1.0) / 9.0);
    f[2] = (float)(log2(raw[2] + 1.0) / 8.0);
    f[3] = (float)(log2(sqrt(raw[3] > 0 ? raw[3] : 0.0) + 1.0) / 10.0);
    f[4] = (float)(log2(raw[4] + 1.0) / 17.0);
    f[5] = (float)raw[5];
    f[6] = (float)(log2((raw[6] + 1.0) / ((raw[0] > 1 ? raw[0] : 1.0) + 1.0)) / 23.0);
    f[7] = (float)raw[7];
}

// Raw features from the integer sums of row_stats_kernel
// (s[0] = sum len, s[1] = sum len^2, s[2] = empty rows, s[3] = sum span,
// s[4] = max len).
__host__ __device__ inline void tuner_raw(const unsigned long long *s, int m, int nnz,
                                          int is_double, double *raw)
{
    const double dm = m, mean = m > 0 ? (double)s[0] / dm : 0.0;
    const double var = m > 0 ? (double)s[1] / dm - mean * mean : 0.0;
    const double nonempty = dm - (double)s[2];
    raw[0] = dm;
    raw[1] = nnz;
    raw[2] = mean;
    raw[3] = var > 0 ? var : 0.0;
    raw[4] = (double)s[4];
    raw[5] = m > 0 ? (double)s[2] / dm : 0.0;
    raw[6] = nonempty > 0 ? (double)s[3] / nonempty : 0.0;
    raw[7] = is_double ? 1.0 : 0.0;
}

// Device-resident multilayer perceptron n_in -> hidden -> hidden -> n_out with
// ReLU, inputs standardised by (x - mean) / std.  Parameters are one float
// array: W1 [h x n_in], b1 [h], W2 [h x h], b2 [h], W3 [n_out x h], b3 [n_out].
struct Mlp {
    int device = 0;
    int n_in = 0, h = 0, n_out = 0;
    long long n_params = 0;
    float *d_params = nullptr;
    float *d_norm = nullptr;      // mean[n_in], std[n_in]
    hipStream_t stream = nullptr;
};

// Launches the fused inference kernel (defined in bhs_mlp.hip): one block per
// matrix; stats[8 * mat] from row_stats_kernel, mn[2 * mat] = {m, nnz}.
hipError_t launch_tuner_infer(const unsigned long long *stats, const int *mn, const int *is_double,
                              int n_mat, const Mlp &net, int K, const int *valid, int *choice,
                              float *scores, hipStream_t stream);

inline long long mlp_param_count(int n_in, int h, int n_out)
{
    return (long long)h * n_in + h + (long long)h * h + h + (long long)n_out * h + n_out;
}

} // namespace bhs
