// bhs_vendor.hip -- rocSPARSE CSR SpMV timed on the same matrix as the
// library's own kernel (the reference's cuSPARSE column, call_cusparse_ref,
// spmv_cuda/main.cu:49-183).  Not part of the bhsparse_spmv path.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <chrono>
#include <cstdio>
#include <cstring>

#include "../../include/bhs_vendor.h"
#include "../../include/bhsparse_spmv.h"

namespace {

struct Buffers {
    int *rp = nullptr, *col = nullptr;
    void *val = nullptr, *x = nullptr, *y = nullptr, *work = nullptr, *flush = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rocsparse_handle handle = nullptr;
    rocsparse_spmv_descr sd = nullptr;
    rocsparse_const_spmat_descr A = nullptr;
    rocsparse_const_dnvec_descr X = nullptr;
    rocsparse_dnvec_descr Y = nullptr;

    ~Buffers()
    {
        if (Y) rocsparse_destroy_dnvec_descr(Y);
        if (X) rocsparse_destroy_dnvec_descr(X);
        if (A) rocsparse_destroy_spmat_descr(A);
        if (sd) rocsparse_destroy_spmv_descr(sd);
        if (handle) rocsparse_destroy_handle(handle);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (void *p : {(void *)rp, (void *)col, val, x, y, work, flush})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

#define VH(call)                                                                                  \
    do {                                                                                          \
        if ((call) != hipSuccess) return BHSPARSE_HIP_ERROR;                                      \
    } while (0)
#define VS(call)                                                                                  \
    do {                                                                                          \
        if ((call) != rocsparse_status_success) return BHSPARSE_HIP_ERROR;                        \
    } while (0)

rocsparse_spmv_alg alg_of(int a)
{
    switch (a) {
    case BHS_VENDOR_ALG_ROWSPLIT: return rocsparse_spmv_alg_csr_rowsplit;
    case BHS_VENDOR_ALG_LRB: return rocsparse_spmv_alg_csr_lrb;
    case BHS_VENDOR_ALG_NNZSPLIT: return rocsparse_spmv_alg_csr_nnzsplit;
    default: return rocsparse_spmv_alg_csr_adaptive;
    }
}

template <typename T>
int csrmv_bench(int device, int m, int n, int nnz, const int *row_ptr, const int *col,
                const T *val, const T *x, T alpha, int alg, int warmup, int reps,
                long long flush_bytes, double *ms, T *y_out, double *analysis_ms)
{
    if (m < 1 || n < 1 || nnz < 0 || !row_ptr || (nnz > 0 && (!col || !val)) || !x || reps < 0 ||
        warmup < 0 || alg < 0 || alg > 3 || (reps > 0 && !ms))
        return BHSPARSE_INVALID_ARGUMENT;
    const rocsparse_datatype dt =
        sizeof(T) == 8 ? rocsparse_datatype_f64_r : rocsparse_datatype_f32_r;
    Buffers b;
    VH(hipSetDevice(device));
    VH(hipStreamCreate(&b.stream));
    // timed the same way as the library's own calls: events without the
    // system-scope fence of a default event
    VH(hipEventCreateWithFlags(&b.e0, hipEventDisableSystemFence));
    VH(hipEventCreateWithFlags(&b.e1, hipEventDisableSystemFence));
    VH(hipMalloc((void **)&b.rp, sizeof(int) * (size_t)(m + 1)));
    VH(hipMalloc((void **)&b.col, sizeof(int) * (size_t)(nnz > 0 ? nnz : 1)));
    VH(hipMalloc(&b.val, sizeof(T) * (size_t)(nnz > 0 ? nnz : 1)));
    VH(hipMalloc(&b.x, sizeof(T) * (size_t)n));
    VH(hipMalloc(&b.y, sizeof(T) * (size_t)m));
    VH(hipMemcpy(b.rp, row_ptr, sizeof(int) * (size_t)(m + 1), hipMemcpyHostToDevice));
    if (nnz > 0) {
        VH(hipMemcpy(b.col, col, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
        VH(hipMemcpy(b.val, val, sizeof(T) * (size_t)nnz, hipMemcpyHostToDevice));
    }
    VH(hipMemcpy(b.x, x, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
    VH(hipMemset(b.y, 0, sizeof(T) * (size_t)m));
    if (flush_bytes > 0) VH(hipMalloc(&b.flush, (size_t)flush_bytes));

    VS(rocsparse_create_handle(&b.handle));
    VS(rocsparse_set_stream(b.handle, b.stream));
    VS(rocsparse_create_const_csr_descr(&b.A, m, n, nnz, b.rp, b.col, b.val,
                                        rocsparse_indextype_i32, rocsparse_indextype_i32,
                                        rocsparse_index_base_zero, dt));
    VS(rocsparse_create_const_dnvec_descr(&b.X, n, b.x, dt));
    VS(rocsparse_create_dnvec_descr(&b.Y, m, b.y, dt));
    VS(rocsparse_create_spmv_descr(&b.sd));
    const rocsparse_spmv_alg a = alg_of(alg);
    const rocsparse_operation op = rocsparse_operation_none;
    VS(rocsparse_spmv_set_input(b.handle, b.sd, rocsparse_spmv_input_alg, &a, sizeof(a), nullptr));
    VS(rocsparse_spmv_set_input(b.handle, b.sd, rocsparse_spmv_input_operation, &op, sizeof(op),
                                nullptr));
    VS(rocsparse_spmv_set_input(b.handle, b.sd, rocsparse_spmv_input_scalar_datatype, &dt,
                                sizeof(dt), nullptr));
    VS(rocsparse_spmv_set_input(b.handle, b.sd, rocsparse_spmv_input_compute_datatype, &dt,
                                sizeof(dt), nullptr));

    const T beta = T(0);
    size_t bytes_a = 0, bytes_c = 0;
    VS(rocsparse_v2_spmv_buffer_size(b.handle, b.sd, b.A, b.X, b.Y,
                                     rocsparse_v2_spmv_stage_analysis, &bytes_a, nullptr));
    VS(rocsparse_v2_spmv_buffer_size(b.handle, b.sd, b.A, b.X, b.Y,
                                     rocsparse_v2_spmv_stage_compute, &bytes_c, nullptr));
    const size_t wbytes = bytes_a > bytes_c ? bytes_a : bytes_c;
    if (wbytes > 0) VH(hipMalloc(&b.work, wbytes));
    VH(hipDeviceSynchronize());
    const auto t0 = std::chrono::steady_clock::now();
    VS(rocsparse_v2_spmv(b.handle, b.sd, &alpha, b.A, b.X, &beta, b.Y,
                         rocsparse_v2_spmv_stage_analysis, bytes_a, b.work, nullptr));
    VH(hipStreamSynchronize(b.stream));
    if (analysis_ms)
        *analysis_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count();

    for (int i = 0; i < warmup + reps; ++i) {
        if (b.flush) VH(hipMemsetAsync(b.flush, i & 0xff, (size_t)flush_bytes, b.stream));
        VH(hipEventRecord(b.e0, b.stream));
        VS(rocsparse_v2_spmv(b.handle, b.sd, &alpha, b.A, b.X, &beta, b.Y,
                             rocsparse_v2_spmv_stage_compute, bytes_c, b.work, nullptr));
        VH(hipEventRecord(b.e1, b.stream));
        VH(hipEventSynchronize(b.e1));
        if (i >= warmup) {
            float t = 0.f;
            VH(hipEventElapsedTime(&t, b.e0, b.e1));
            ms[i - warmup] = t;
        }
    }
    if (y_out) VH(hipMemcpy(y_out, b.y, sizeof(T) * (size_t)m, hipMemcpyDeviceToHost));
    return BHSPARSE_SUCCESS;
}

} // namespace

extern "C" {

int bhs_vendor_d_csrmv_bench(int device, int m, int n, int nnz, const int *row_ptr,
                             const int *col, const double *val, const double *x, double alpha,
                             int alg, int warmup, int reps, long long flush_bytes, double *ms,
                             double *y_out, double *analysis_ms)
{
    return csrmv_bench<double>(device, m, n, nnz, row_ptr, col, val, x, alpha, alg, warmup, reps,
                               flush_bytes, ms, y_out, analysis_ms);
}

int bhs_vendor_s_csrmv_bench(int device, int m, int n, int nnz, const int *row_ptr,
                             const int *col, const float *val, const float *x, float alpha,
                             int alg, int warmup, int reps, long long flush_bytes, double *ms,
                             float *y_out, double *analysis_ms)
{
    return csrmv_bench<float>(device, m, n, nnz, row_ptr, col, val, x, alpha, alg, warmup, reps,
                              flush_bytes, ms, y_out, analysis_ms);
}

int bhs_vendor_version(void)
{
    return ROCSPARSE_VERSION_MAJOR * 100000 + ROCSPARSE_VERSION_MINOR * 100 +
           ROCSPARSE_VERSION_PATCH;
}

} // extern "C"
