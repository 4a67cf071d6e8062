// bhs_api.hip -- the C ABI of include/bhsparse_spmv.h: device selection,
// validated upload of A and x, per-matrix tile partition, the spmv launch
// sequence, timing and debug queries.
//
// Reference counterparts: bhsparse_spmv_cuda (spmv_cuda/bhsparse_spmv_cuda.h)
// and bhsparse_spmv_opencl (spmv_opencl_amd/bhsparse_spmv_opencl.h).  Unlike
// them, every buffer is plain device memory (hipMalloc) unless the handle is
// put in managed mode, nothing is computed on the host per call, and a call
// never depends on y's previous contents.
#include <hip/hip_ext.h>
#include <hipcub/hipcub.hpp>

#include "bhs_kernels.h"
#include "bhs_tuner.h"
#include "../../include/bhsparse_spmv.h"

#include <algorithm>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <vector>

// Device scratch of the allocation-free entry points (bhs_*_spmv_noalloc):
// the tile partition and the waves' boundary-row partials, at most
// kScratchBytes.  One copy per device, held until the process ends; host code
// reaches it only through the address scratch_ptr() looks up.
constexpr size_t kScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

// Events used only for timing are created without the system-scope fence
// that a default event performs when it is recorded (a cache writeback and
// invalidation inside every timed interval: 4.3 us for an empty interval
// between two default events, 3.2 us without it).  The results a timed call
// leaves are read only after a stream synchronisation.
inline hipError_t create_timing_event(hipEvent_t *e)
{
    return hipEventCreateWithFlags(e, hipEventDisableSystemFence);
}

#define BHS_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return set_error(e_ == hipErrorOutOfMemory ? BHSPARSE_OUT_OF_MEMORY          \
                                                       : BHSPARSE_HIP_ERROR,             \
                             std::string(#call) + ": " + hipGetErrorString(e_));         \
    } while (0)

} // namespace

namespace bhs {
int mlp_set_error(int code, const std::string &msg) { return set_error(code, msg); }
} // namespace bhs

namespace {

// The largest tile (64 lanes x 32 items) -- arrays are padded to it.
constexpr int kMaxTile = 64 * 32;
// nnz limit keeps every tile index, tile boundary and padded index in int.
constexpr long long kMaxNnz = 0x7fffffffLL - 4 * kMaxTile;

enum ValueKind { kFloat = 1, kDouble = 2 };

template <typename T>
struct Impl {
    int device = -1;
    hipStream_t stream = nullptr;
    bool platform = false;

    bool data = false;
    int m = 0, n = 0, nnz = 0;
    long long nnz_pad = 0;
    int *d_row_ptr = nullptr, *d_col = nullptr;
    T *d_val = nullptr, *d_x = nullptr, *d_y = nullptr;
    T *h_y = nullptr;   // borrowed, as _h_y in bhsparse_spmv_cuda.h:248
    // BHS_MEMORY_MANAGED / BHS_MEMORY_HOST_MAPPED: row_ptr, col, val, x and y
    // come from hipMallocManaged / pinned host memory mapped into the device
    // (the USE_SVM_ALWAYS=1 build, bhsparse_spmv_cuda.h:250-263); data_mode is
    // the mode the loaded data was allocated in
    int mem_mode = BHS_MEMORY_DEVICE, data_mode = BHS_MEMORY_DEVICE;

    // the fixed default tile (the tuner's "no tuner" baseline class)
    bhs_config cfg{BHS_VARIANT_SEGSUM, 8, 4, 4, 0};

    // column-panel copy of A (BHS_VARIANT_SEGSUM_PANELS): one stacked row per
    // nonempty (panel, row) pair, panel-major.  d_pmask[i] bit p: row i has
    // entries in panel p; d_pbase[g*pan_n+p]: stacked rows before (panel p,
    // row 64g).  pan_n = 0 when not built
    int pan_n = 0, pan_rows = 0;
    int *d_prow_ptr = nullptr, *d_pcol = nullptr, *d_pbase = nullptr;
    unsigned char *d_pmask = nullptr;
    T *d_pval = nullptr, *d_ystack = nullptr;
    // host copies for the chain records: stacked row_ptr, the row of A each
    // stacked row belongs to, and the first stacked row of every panel
    std::vector<int> h_prow_ptr, h_prow_of, h_pan_start;
    // chain records of the combine (one per combine wave, then the overflow),
    // rebuilt with every analysis of the panel layout
    bhs::ChainRec *d_chains = nullptr;
    long long chain_cap = 0;

    // analysis for the current tile shape and layout (an_layout = panels, 0 = A)
    int an_P = 0, an_TPW = 0, an_layout = -1;
    int NP = 0, NW = 0, lead_empty = 0;
    int chunk_parts = 2;   // row chunk of the segmented sum: 2 or 4 parts of 64 rows
    int *d_tile_row = nullptr;
    int tile_cap = 0;
    int *d_sync_row = nullptr;
    T *d_sync_val = nullptr;
    int sync_cap = 0;

    unsigned long long *d_counters = nullptr;   // 8 x u64 scratch for stats/features
    T *d_floor = nullptr;                       // per-lane partials of the floor probe
    long long floor_cap = 0;
    int *d_tuner_io = nullptr;                  // tuner inference inputs/outputs (ints)
    float *d_tuner_scores = nullptr;            // tuner's predicted cost per class
    void *d_flush = nullptr;
    long long flush_cap = 0;
    // ev[0]/ev[2]: recorded around a whole call; ev[1]/ev[3]: attached to the
    // main kernel's dispatch (hipExtLaunchKernelGGL), its exact duration
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t max_lds = 65536;   // per-block LDS limit of the device
};

struct HandleBox {
    int kind;
    void *impl;
};

template <typename T> constexpr int kind_of();
template <> constexpr int kind_of<float>() { return kFloat; }
template <> constexpr int kind_of<double>() { return kDouble; }

template <typename T>
int get_impl(bhs_handle_t h, Impl<T> **out)
{
    if (!h) return set_error(BHSPARSE_INVALID_ARGUMENT, "null handle");
    HandleBox *b = reinterpret_cast<HandleBox *>(h);
    if (b->kind != kind_of<T>())
        return set_error(BHSPARSE_UNSUPPORTED_VALUE_TYPE,
                         "handle was created for the other precision");
    *out = static_cast<Impl<T> *>(b->impl);
    return BHSPARSE_SUCCESS;
}

template <typename T>
void free_ptr(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <typename T>
int create(bhs_handle_t *h)
{
    if (!h) return set_error(BHSPARSE_INVALID_ARGUMENT, "null handle pointer");
    HandleBox *b = new (std::nothrow) HandleBox;
    Impl<T> *im = new (std::nothrow) Impl<T>;
    if (!b || !im) {
        delete b;
        delete im;
        return set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
    }
    b->kind = kind_of<T>();
    b->impl = im;
    *h = reinterpret_cast<bhs_handle_t>(b);
    return BHSPARSE_SUCCESS;
}

int scratch_ptr(void **out);

template <typename T>
int init_platform(bhs_handle_t h, int device)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return set_error(BHSPARSE_UNSUPPORTED_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count)
        return set_error(BHSPARSE_UNSUPPORTED_DEVICE, "device index out of range");
    hipDeviceProp_t prop;
    BHS_HIP(hipGetDeviceProperties(&prop, device));
    if (prop.warpSize != 64)
        return set_error(BHSPARSE_UNSUPPORTED_DEVICE, "wave64 device required");
    if (im->platform && im->device != device)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "handle already bound to another device");
    BHS_HIP(hipSetDevice(device));
    im->max_lds = std::min<size_t>(prop.sharedMemPerBlock, 65536);
    if (!im->platform) {
        BHS_HIP(hipStreamCreateWithFlags(&im->stream, hipStreamNonBlocking));
        for (auto &e : im->ev) BHS_HIP(create_timing_event(&e));
        BHS_HIP(hipMalloc(&im->d_counters, 8 * sizeof(unsigned long long)));
    }
    im->device = device;
    im->platform = true;
    void *sp = nullptr;              // the device's scratch address, looked up once
    return scratch_ptr(&sp);
}

// Frees an array allocated by alloc_data in `mode`.
template <typename U>
void free_data(U *&p, int mode)
{
    if (p) {
        if (mode == BHS_MEMORY_HOST_MAPPED)
            (void)hipHostFree((void *)p);
        else
            (void)hipFree((void *)p);
    }
    p = nullptr;
}

template <typename T>
void release_data(Impl<T> *im)
{
    if (im->platform) (void)hipSetDevice(im->device);
    if (im->stream) (void)hipStreamSynchronize(im->stream);
    free_data(im->d_row_ptr, im->data_mode);
    free_data(im->d_col, im->data_mode);
    free_data(im->d_val, im->data_mode);
    free_data(im->d_x, im->data_mode);
    free_data(im->d_y, im->data_mode);
    free_ptr(im->d_tile_row);
    free_ptr(im->d_sync_row);
    free_ptr(im->d_sync_val);
    free_ptr(im->d_floor);
    free_ptr(im->d_prow_ptr);
    free_ptr(im->d_pcol);
    free_ptr(im->d_pval);
    free_ptr(im->d_ystack);
    free_ptr(im->d_pmask);
    free_ptr(im->d_pbase);
    free_ptr(im->d_chains);
    im->chain_cap = 0;
    im->h_prow_ptr = {};
    im->h_prow_of = {};
    im->h_pan_start = {};
    im->pan_n = im->pan_rows = 0;
    im->floor_cap = 0;
    im->tile_cap = im->sync_cap = 0;
    im->an_P = im->an_TPW = 0;
    im->an_layout = -1;
    im->NP = im->NW = im->lead_empty = 0;
    im->h_y = nullptr;
    im->data = false;
    im->data_mode = BHS_MEMORY_DEVICE;
    im->m = im->n = im->nnz = 0;
}

// Allocation of the caller-visible arrays (A, x, y) in the data's memory mode
// (data_mode).  Host-mapped memory must be usable by kernels at the address
// the host sees (one pointer for both, as in unified memory).
template <typename T, typename U>
hipError_t alloc_data(Impl<T> *im, U **p, size_t bytes)
{
    if (im->data_mode == BHS_MEMORY_MANAGED)
        return hipMallocManaged(reinterpret_cast<void **>(p), bytes, hipMemAttachGlobal);
    if (im->data_mode == BHS_MEMORY_HOST_MAPPED) {
        hipError_t e = hipHostMalloc(reinterpret_cast<void **>(p), bytes, hipHostMallocMapped);
        if (e != hipSuccess) return e;
        void *dp = nullptr;
        e = hipHostGetDevicePointer(&dp, (void *)*p, 0);
        if (e == hipSuccess && dp != (void *)*p) e = hipErrorInvalidValue;
        if (e != hipSuccess) {
            (void)hipHostFree((void *)*p);
            *p = nullptr;
        }
        return e;
    }
    return hipMalloc(reinterpret_cast<void **>(p), bytes);
}

inline bool is_segsum(int variant)
{
    return variant == BHS_VARIANT_SEGSUM || variant == BHS_VARIANT_SEGSUM_PANELS;
}

int check_config(const bhs_config &c)
{
    if (c.variant != BHS_VARIANT_SEGSUM && c.variant != BHS_VARIANT_CSRVEC &&
        c.variant != BHS_VARIANT_SEGSUM_PANELS)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "unknown variant");
    if (c.variant == BHS_VARIANT_SEGSUM_PANELS && c.tpv != 2 && c.tpv != 4 && c.tpv != 8)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "column panels (tpv) must be 2, 4 or 8");
    const int S = c.items_per_lane;
    if (S != 4 && S != 8 && S != 16 && S != 32)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "items_per_lane must be 4, 8, 16 or 32");
    if (c.waves_per_block < 1 || c.waves_per_block > 8)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "waves_per_block must be in [1, 8]");
    if (c.tiles_per_wave < 1 || c.tiles_per_wave > 4096)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "tiles_per_wave must be in [1, 4096]");
    const int t = c.tpv;
    if (t != 0 && t != 2 && t != 4 && t != 8 && t != 16 && t != 32 && t != 64)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "tpv must be 0, 2, 4, 8, 16, 32 or 64");
    return BHSPARSE_SUCCESS;
}

// The matrix the segmented sum runs over: A itself, or its column-panel copy.
template <typename T>
struct Layout {
    int rows;
    const int *row_ptr, *col;
    const T *val;
    T *y;
};

template <typename T>
Layout<T> active_layout(Impl<T> *im)
{
    if (im->cfg.variant == BHS_VARIANT_SEGSUM_PANELS)
        return {im->pan_rows, im->d_prow_ptr, im->d_pcol, im->d_pval, im->d_ystack};
    return {im->m, im->d_row_ptr, im->d_col, im->d_val, im->d_y};
}

// Column-panel copy of A, built on the device from its device copy: per row
// the entries in every panel (panel_count_kernel), exclusive sums over the
// (panel, row) pairs in panel-major order of the nonempty flags (the stacked
// row of a pair) and of the counts (its first entry), then the entries
// scattered in order (panel_scatter_kernel).  Only nonempty (panel, row) pairs
// become stacked rows, so the stacked matrix has no empty rows; columns keep
// their order inside every one.  The host keeps the stacked row_ptr, the row
// of A of every stacked row and each panel's first stacked row, for the chain
// records of the combine.
struct IsPositive {
    __host__ __device__ int operator()(int v) const { return v > 0 ? 1 : 0; }
};

template <typename T, int NPAN>
int build_panels_n(Impl<T> *im)
{
    const int m = im->m, n = im->n, nnz = im->nnz;
    const int pw = std::max((n + NPAN - 1) / NPAN, 1);
    const long long PM = (long long)NPAN * m;
    const long long G = ((long long)m + 63) / 64;
    const int threads = 256;
    int *d_cnt = nullptr, *d_rows = nullptr, *d_nzs = nullptr, *d_owner = nullptr;
    void *d_tmp = nullptr;
    // frees the work arrays on every return path
    struct Work {
        int **a[4];
        void **t;
        ~Work()
        {
            for (int **p : a)
                if (*p) (void)hipFree(*p);
            if (*t) (void)hipFree(*t);
        }
    } work{{&d_cnt, &d_rows, &d_nzs, &d_owner}, &d_tmp};
    BHS_HIP(hipMalloc(&d_cnt, sizeof(int) * (size_t)PM));
    BHS_HIP(hipMalloc(&d_rows, sizeof(int) * (size_t)PM));
    BHS_HIP(hipMalloc(&d_nzs, sizeof(int) * (size_t)PM));
    BHS_HIP(hipMalloc(&im->d_pmask, std::max<size_t>((size_t)m, 1)));
    BHS_HIP(hipMalloc(&im->d_pbase, sizeof(int) * (size_t)std::max<long long>(G * NPAN, 1)));
    const unsigned row_blocks = (unsigned)(((long long)m + threads / 64 - 1) / (threads / 64));
    hipLaunchKernelGGL(bhs::panel_count_kernel<NPAN>, dim3(row_blocks),
                       dim3(threads), 0, im->stream, (const int *)im->d_row_ptr,
                       (const int *)im->d_col, m, pw, d_cnt, im->d_pmask);
    BHS_HIP(hipGetLastError());
    hipcub::TransformInputIterator<int, IsPositive, const int *> flags(d_cnt, IsPositive());
    size_t tb1 = 0, tb2 = 0;
    BHS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb1, flags, d_rows, (int)PM, im->stream));
    BHS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, (const int *)d_cnt, d_nzs, (int)PM,
                                             im->stream));
    size_t tb = std::max(tb1, tb2);
    BHS_HIP(hipMalloc(&d_tmp, std::max<size_t>(tb, 16)));
    BHS_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp, tb, flags, d_rows, (int)PM, im->stream));
    tb = std::max(tb1, tb2);
    BHS_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp, tb, (const int *)d_cnt, d_nzs, (int)PM,
                                             im->stream));
    int last[2] = {0, 0};
    BHS_HIP(hipMemcpyAsync(&last[0], d_rows + PM - 1, sizeof(int), hipMemcpyDeviceToHost, im->stream));
    BHS_HIP(hipMemcpyAsync(&last[1], d_cnt + PM - 1, sizeof(int), hipMemcpyDeviceToHost, im->stream));
    std::vector<int> pan_start((size_t)NPAN + 1, 0);
    for (int p = 0; p < NPAN; ++p)
        BHS_HIP(hipMemcpyAsync(&pan_start[(size_t)p], d_rows + (long long)p * m, sizeof(int),
                               hipMemcpyDeviceToHost, im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    const int M = last[0] + (last[1] > 0 ? 1 : 0);
    pan_start[(size_t)NPAN] = M;
    BHS_HIP(hipMalloc(&im->d_prow_ptr, sizeof(int) * ((size_t)M + 1)));
    BHS_HIP(hipMalloc(&d_owner, sizeof(int) * (size_t)std::max(M, 1)));
    BHS_HIP(hipMalloc(&im->d_pcol, sizeof(int) * (size_t)im->nnz_pad));
    BHS_HIP(hipMalloc(&im->d_pval, sizeof(T) * (size_t)im->nnz_pad));
    BHS_HIP(hipMalloc(&im->d_ystack, sizeof(T) * (size_t)std::max(M, 1)));
    BHS_HIP(hipMemsetAsync(im->d_pcol, 0, sizeof(int) * (size_t)im->nnz_pad, im->stream));
    BHS_HIP(hipMemsetAsync(im->d_pval, 0, sizeof(T) * (size_t)im->nnz_pad, im->stream));
    hipLaunchKernelGGL(bhs::panel_index_kernel<NPAN>, dim3((unsigned)((PM + threads - 1) / threads)),
                       dim3(threads), 0, im->stream, (const int *)d_cnt, (const int *)d_rows,
                       (const int *)d_nzs, m, im->d_prow_ptr, d_owner, im->d_pbase);
    hipLaunchKernelGGL(bhs::fill_kernel<int>, dim3(1), dim3(64), 0, im->stream, im->d_prow_ptr + M,
                       1LL, nnz);
    hipLaunchKernelGGL((bhs::panel_scatter_kernel<T, NPAN>), dim3(row_blocks), dim3(threads), 0,
                       im->stream,
                       (const int *)im->d_row_ptr, (const int *)im->d_col, (const T *)im->d_val, m,
                       pw, (const int *)d_cnt, (const int *)d_nzs, im->d_pcol, im->d_pval);
    BHS_HIP(hipGetLastError());
    std::vector<int> prp((size_t)M + 1), prow_of((size_t)M);
    BHS_HIP(hipMemcpyAsync(prp.data(), im->d_prow_ptr, sizeof(int) * ((size_t)M + 1),
                           hipMemcpyDeviceToHost, im->stream));
    if (M > 0)
        BHS_HIP(hipMemcpyAsync(prow_of.data(), d_owner, sizeof(int) * (size_t)M,
                               hipMemcpyDeviceToHost, im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    im->h_prow_ptr = std::move(prp);
    im->h_prow_of = std::move(prow_of);
    im->h_pan_start = std::move(pan_start);
    im->pan_n = NPAN;
    im->pan_rows = M;
    return BHSPARSE_SUCCESS;
}

template <typename T>
int build_panels(Impl<T> *im, int npan)
{
    if (im->pan_n == npan) return BHSPARSE_SUCCESS;
    free_ptr(im->d_prow_ptr);
    free_ptr(im->d_pcol);
    free_ptr(im->d_pval);
    free_ptr(im->d_ystack);
    free_ptr(im->d_pmask);
    free_ptr(im->d_pbase);
    im->h_prow_ptr = {};
    im->h_prow_of = {};
    im->h_pan_start = {};
    im->pan_n = im->pan_rows = 0;
    if ((long long)npan * im->m > 0x7fffffffLL)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "too many rows for column panels");
    switch (npan) {
    case 2: return build_panels_n<T, 2>(im);
    case 4: return build_panels_n<T, 4>(im);
    default: return build_panels_n<T, 8>(im);
    }
}

// Chain records of the panel combine for waves of PW stacked nonzeros.  Wave
// w > 0 leaves a partial for stacked row R when R = upper_bound(row_ptr,
// w*PW) - 1 began before w*PW (the segmented sum's `cont` test on its first
// tile; the stacked matrix has no empty rows, so R is the row holding nonzero
// w*PW); consecutive such waves with the same R form one chain.
template <typename T>
int build_chains(Impl<T> *im, long long PW, long long NW)
{
    const std::vector<int> &rp = im->h_prow_ptr;
    const long long groups = 64LL * bhs::kCombineGroups;
    const long long n_cw = ((long long)im->m + groups - 1) / groups;
    constexpr int K = bhs::kChainSlots;
    // K slots per combine wave, then the records linked from its last slot
    std::vector<bhs::ChainRec> rec((size_t)std::max<long long>(n_cw, 1) * K,
                                   bhs::ChainRec{0, 0, 0, -1});
    std::vector<unsigned char> used((size_t)std::max<long long>(n_cw, 1), 0);
    auto add = [&](int R, int first, int count) {
        const int i = im->h_prow_of[(size_t)R];
        const int pan = (int)(std::upper_bound(im->h_pan_start.begin(), im->h_pan_start.end(), R) -
                              im->h_pan_start.begin()) - 1;
        const long long cw = i / groups;
        bhs::ChainRec c{first, count, (int)(i % groups) << 4 | pan, -1};
        const size_t last = (size_t)(cw * K + K - 1);
        if (used[(size_t)cw] < K) {
            rec[(size_t)(cw * K + used[(size_t)cw]++)] = c;
        } else {                              // beyond the slots: first in the linked list
            c.next = rec[last].next;
            rec.push_back(c);
            rec[last].next = (int)rec.size() - 1;
        }
    };
    int R_open = -1, first = 0, count = 0;
    for (long long w = 1; w < NW; ++w) {
        const long long s = w * PW;
        const int R = (int)(std::upper_bound(rp.begin(), rp.end(), (int)s) - rp.begin()) - 1;
        const bool sync = rp[(size_t)R] < s;
        if (sync && R == R_open && first + count == w) {
            ++count;
            continue;
        }
        if (count > 0) add(R_open, first, count);
        R_open = -1;
        count = 0;
        if (sync) {
            R_open = R;
            first = (int)w;
            count = 1;
        }
    }
    if (count > 0) add(R_open, first, count);
    if ((long long)rec.size() > 0x7fffffffLL)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "too many chain records");
    if ((long long)rec.size() > im->chain_cap) {
        free_ptr(im->d_chains);
        BHS_HIP(hipMalloc(&im->d_chains, sizeof(bhs::ChainRec) * rec.size()));
        im->chain_cap = (long long)rec.size();
    }
    BHS_HIP(hipMemcpyAsync(im->d_chains, rec.data(), sizeof(bhs::ChainRec) * rec.size(),
                           hipMemcpyHostToDevice, im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    return BHSPARSE_SUCCESS;
}

// Partition for the configured tile size and layout, and sizes of the
// per-wave buffers.
template <typename T>
int ensure_analysis(Impl<T> *im)
{
    const int P = 64 * im->cfg.items_per_lane;
    const int TPW = im->cfg.tiles_per_wave;
    const int layout = im->cfg.variant == BHS_VARIANT_SEGSUM_PANELS ? im->cfg.tpv : 0;
    if (im->an_P == P && im->an_TPW == TPW && im->an_layout == layout) return BHSPARSE_SUCCESS;
    if (layout > 0 && im->m > 0)
        if (int rc = build_panels(im, layout)) return rc;
    const Layout<T> L = active_layout(im);
    const long long NP = ((long long)im->nnz + P - 1) / P;
    const long long NW = (NP + TPW - 1) / TPW;
    if (NP + 1 > im->tile_cap) {
        free_ptr(im->d_tile_row);
        BHS_HIP(hipMalloc(&im->d_tile_row, sizeof(int) * (size_t)(NP + 1)));
        im->tile_cap = (int)(NP + 1);
    }
    if (NW > im->sync_cap) {
        free_ptr(im->d_sync_row);
        free_ptr(im->d_sync_val);
        BHS_HIP(hipMalloc(&im->d_sync_row, sizeof(int) * (size_t)std::max<long long>(NW, 1)));
        BHS_HIP(hipMalloc(&im->d_sync_val, sizeof(T) * (size_t)std::max<long long>(NW, 1)));
        im->sync_cap = (int)std::max<long long>(NW, 1);
    }
    const int threads = 256;
    const int blocks = (int)((NP + 1 + threads - 1) / threads);
    if (L.rows > 0)
        hipLaunchKernelGGL(bhs::partition_kernel, dim3(blocks), dim3(threads), 0, im->stream,
                           L.row_ptr, L.rows, im->nnz, P, (int)NP, im->d_tile_row);
    BHS_HIP(hipGetLastError());
    // tiles spanning 128-255 rows: with enough of them (5% of the tiles) the
    // 256-row chunk spares their row_ptr loop, otherwise the 128-row chunk
    // costs less per tile
    unsigned long long mid_tiles = 0;
    if (L.rows > 0 && NP > 0) {
        launch_fill(im, im->d_counters, 1);
        hipLaunchKernelGGL(bhs::tile_span_count_kernel, dim3((unsigned)((NP + threads - 1) / threads)),
                           dim3(threads), 0, im->stream, (const int *)im->d_tile_row, (int)NP,
                           2 * 64, 4 * 64, im->d_counters);
        BHS_HIP(hipGetLastError());
        BHS_HIP(hipMemcpyAsync(&mid_tiles, im->d_counters, sizeof(mid_tiles),
                               hipMemcpyDeviceToHost, im->stream));
    }
    if (layout > 0 && im->m > 0)
        if (int rc = build_chains(im, (long long)P * TPW, NW)) return rc;
    int first = 0;
    BHS_HIP(hipMemcpyAsync(&first, im->d_tile_row, sizeof(int), hipMemcpyDeviceToHost, im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    im->NP = (int)NP;
    im->NW = (int)NW;
    im->lead_empty = L.rows > 0 ? first : 0;
    im->chunk_parts = (NP > 0 && mid_tiles * 20 >= (unsigned long long)NP) ? 4 : 2;
    im->an_P = P;
    im->an_TPW = TPW;
    im->an_layout = layout;
    return BHSPARSE_SUCCESS;
}

template <typename T>
int init_data(bhs_handle_t h, int m, int n, int nnz, const int *row_ptr, const int *col,
              const T *val, const T *x, T *y)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->platform) return set_error(BHSPARSE_NOT_INITIALIZED, "init_platform first");
    if (m < 0 || n < 0 || nnz < 0) return set_error(BHSPARSE_INVALID_ARGUMENT, "negative size");
    if (!row_ptr) return set_error(BHSPARSE_INVALID_ARGUMENT, "null row_ptr");
    if (nnz > 0 && (!col || !val)) return set_error(BHSPARSE_INVALID_ARGUMENT, "null col/val");
    if (n > 0 && !x) return set_error(BHSPARSE_INVALID_ARGUMENT, "null x");
    if ((long long)nnz > kMaxNnz) return set_error(BHSPARSE_INVALID_MATRIX, "nnz too large");
    if (row_ptr[0] != 0 || row_ptr[m] != nnz)
        return set_error(BHSPARSE_INVALID_MATRIX, "row_ptr[0] must be 0 and row_ptr[m] == nnz");
    for (int i = 0; i < m; i++)
        if (row_ptr[i + 1] < row_ptr[i])
            return set_error(BHSPARSE_INVALID_MATRIX, "row_ptr is not monotone");
    for (int j = 0; j < nnz; j++)
        if (col[j] < 0 || col[j] >= n)
            return set_error(BHSPARSE_INVALID_MATRIX, "column index out of range");

    release_data(im);
    BHS_HIP(hipSetDevice(im->device));
    if (im->mem_mode == BHS_MEMORY_MANAGED) {
        int managed = 0;
        BHS_HIP(hipDeviceGetAttribute(&managed, hipDeviceAttributeManagedMemory, im->device));
        if (!managed)
            return set_error(BHSPARSE_UNSUPPORTED_DEVICE,
                             "device has no managed memory (hipDeviceAttributeManagedMemory = 0)");
    } else if (im->mem_mode == BHS_MEMORY_HOST_MAPPED) {
        int can_map = 0;
        BHS_HIP(hipDeviceGetAttribute(&can_map, hipDeviceAttributeCanMapHostMemory, im->device));
        if (!can_map)
            return set_error(BHSPARSE_UNSUPPORTED_DEVICE, "device cannot map host memory");
    }
    im->m = m;
    im->n = n;
    im->nnz = nnz;
    im->nnz_pad = ((long long)nnz + kMaxTile - 1) / kMaxTile * kMaxTile + kMaxTile;
    im->data_mode = im->mem_mode;
    BHS_HIP(alloc_data(im, &im->d_row_ptr, sizeof(int) * ((size_t)m + 1)));
    BHS_HIP(alloc_data(im, &im->d_col, sizeof(int) * (size_t)im->nnz_pad));
    BHS_HIP(alloc_data(im, &im->d_val, sizeof(T) * (size_t)im->nnz_pad));
    BHS_HIP(alloc_data(im, &im->d_x, sizeof(T) * (size_t)std::max(n, 1)));
    BHS_HIP(alloc_data(im, &im->d_y, sizeof(T) * (size_t)std::max(m, 1)));
    BHS_HIP(hipMemcpyAsync(im->d_row_ptr, row_ptr, sizeof(int) * ((size_t)m + 1),
                           hipMemcpyHostToDevice, im->stream));
    BHS_HIP(hipMemsetAsync(im->d_col, 0, sizeof(int) * (size_t)im->nnz_pad, im->stream));
    BHS_HIP(hipMemsetAsync(im->d_val, 0, sizeof(T) * (size_t)im->nnz_pad, im->stream));
    BHS_HIP(hipMemsetAsync(im->d_x, 0, sizeof(T) * (size_t)std::max(n, 1), im->stream));
    if (nnz > 0) {
        BHS_HIP(hipMemcpyAsync(im->d_col, col, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice,
                               im->stream));
        BHS_HIP(hipMemcpyAsync(im->d_val, val, sizeof(T) * (size_t)nnz, hipMemcpyHostToDevice,
                               im->stream));
    }
    if (n > 0)
        BHS_HIP(hipMemcpyAsync(im->d_x, x, sizeof(T) * (size_t)n, hipMemcpyHostToDevice,
                               im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    im->h_y = y;
    im->data = true;
    return ensure_analysis(im);
}

template <typename T>
int set_x(bhs_handle_t h, const T *x)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    if (im->n > 0 && !x) return set_error(BHSPARSE_INVALID_ARGUMENT, "null x");
    BHS_HIP(hipSetDevice(im->device));
    if (im->n > 0)
        BHS_HIP(hipMemcpyAsync(im->d_x, x, sizeof(T) * (size_t)im->n, hipMemcpyHostToDevice,
                               im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    return BHSPARSE_SUCCESS;
}

// Memory mode for the next init_data (the reference's compile-time
// USE_SVM_ALWAYS switch, spmv_cuda/Makefile:6).  Changing it while data is loaded is an
// error: free_mem first.
template <typename T>
int set_memory_mode(bhs_handle_t h, int mode)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (mode != BHS_MEMORY_DEVICE && mode != BHS_MEMORY_MANAGED &&
        mode != BHS_MEMORY_HOST_MAPPED)
        return set_error(BHSPARSE_INVALID_ARGUMENT,
                         "memory mode must be device (0), managed (1) or host-mapped (2)");
    if (im->data && mode != im->data_mode)
        return set_error(BHSPARSE_INVALID_ARGUMENT,
                         "data is loaded in the other memory mode; free_mem first");
    im->mem_mode = mode;
    return BHSPARSE_SUCCESS;
}

template <typename T>
int get_memory_mode(bhs_handle_t h, int *mode)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!mode) return set_error(BHSPARSE_INVALID_ARGUMENT, "null mode");
    *mode = im->data ? im->data_mode : im->mem_mode;
    return BHSPARSE_SUCCESS;
}

template <typename T, int S, bool STATS, int CP>
void launch_segsum_s(Impl<T> *im, T alpha, unsigned long long *stats, hipEvent_t e0, hipEvent_t e1)
{
    const Layout<T> L = active_layout(im);
    bhs::SegsumParams<T> q;
    q.m = L.rows;
    q.nnz = im->nnz;
    q.NP = im->NP;
    q.TPW = im->cfg.tiles_per_wave;
    q.NW = im->NW;
    q.lead_empty = im->lead_empty;
    q.wpb = im->cfg.waves_per_block;
    q.wave_lds_bytes = (256 + 64 * S * (int)sizeof(T) + 15) / 16 * 16;
    q.alpha = alpha;
    // x is read ahead into each XCD's L2 when it takes at most 1 MB there
    const long long x_bytes = (long long)im->n * (long long)sizeof(T);
    q.x_lines = x_bytes <= (1LL << 20) ? (x_bytes + 127) / 128 : 0;
    const int blocks = (im->NW + q.wpb - 1) / q.wpb;
    const size_t lds = (size_t)q.wpb * q.wave_lds_bytes;
    // the stacked matrix of the column-panel layout has no empty rows
    auto kern = im->cfg.variant == BHS_VARIANT_SEGSUM_PANELS
                    ? bhs::segsum_kernel<T, S, STATS, CP, false, true>
                    : bhs::segsum_kernel<T, S, STATS, CP, false, false>;
    if (e0 && e1)
        hipExtLaunchKernelGGL(kern, dim3(blocks), dim3(64 * q.wpb), (std::uint32_t)lds, im->stream,
                              e0, e1, 0u, L.row_ptr, L.col, L.val, (const T *)im->d_x, L.y,
                              (const int *)im->d_tile_row, im->d_sync_row, im->d_sync_val, stats,
                              q);
    else
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * q.wpb), lds, im->stream, L.row_ptr, L.col,
                           L.val, (const T *)im->d_x, L.y, (const int *)im->d_tile_row,
                           im->d_sync_row, im->d_sync_val, stats, q);
}

template <typename T, bool STATS, int CP>
void launch_segsum_cp(Impl<T> *im, T alpha, unsigned long long *stats, hipEvent_t e0,
                      hipEvent_t e1)
{
    switch (im->cfg.items_per_lane) {
    case 4: launch_segsum_s<T, 4, STATS, CP>(im, alpha, stats, e0, e1); break;
    case 8: launch_segsum_s<T, 8, STATS, CP>(im, alpha, stats, e0, e1); break;
    case 16: launch_segsum_s<T, 16, STATS, CP>(im, alpha, stats, e0, e1); break;
    default: launch_segsum_s<T, 32, STATS, CP>(im, alpha, stats, e0, e1); break;
    }
}

// The row chunk covers 128 rows (2 parts) or 256 (4 parts), as the analysis
// of the tile partition chose.
template <typename T, bool STATS>
void launch_segsum(Impl<T> *im, T alpha, unsigned long long *stats, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr)
{
    if (im->chunk_parts == 4) launch_segsum_cp<T, STATS, 4>(im, alpha, stats, e0, e1);
    else launch_segsum_cp<T, STATS, 2>(im, alpha, stats, e0, e1);
}

// The column-panel layout is calibrated inside its combine.
template <typename T>
void launch_calibrate(Impl<T> *im)
{
    if (im->NW <= 1) return;   // wave 0 never has a boundary partial
    if (im->cfg.variant == BHS_VARIANT_SEGSUM_PANELS) return;
    const int threads = 256;
    hipLaunchKernelGGL(bhs::calibrate_kernel<T>, dim3((im->NW + threads - 1) / threads),
                       dim3(threads), 0, im->stream, im->d_sync_row, im->d_sync_val, im->NW,
                       active_layout(im).y);
}

template <typename T>
void launch_combine(Impl<T> *im)
{
    if (im->cfg.variant != BHS_VARIANT_SEGSUM_PANELS) return;
    constexpr int kGroups = bhs::kCombineGroups;   // 64-row groups per wave
    const int threads = 256;
    const long long rows_per_block = 64LL * kGroups * (threads / 64);
    const long long blocks = ((long long)im->m + rows_per_block - 1) / rows_per_block;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(threads), 0, im->stream,
                           (const T *)im->d_ystack, (const unsigned char *)im->d_pmask,
                           (const int *)im->d_pbase, im->m, im->d_y,
                           (const bhs::ChainRec *)im->d_chains, (const T *)im->d_sync_val);
    };
    switch (im->pan_n) {
    case 2: go(bhs::combine_panels_kernel<T, 2, kGroups>); break;
    case 4: go(bhs::combine_panels_kernel<T, 4, kGroups>); break;
    default: go(bhs::combine_panels_kernel<T, 8, kGroups>); break;
    }
}

// TPV rule of main.cu:238-287 / csr_vector_opencl.h:212-230 (by floor(nnz/m)).
inline int tpv_rule(int m, int nnz)
{
    const int r = m > 0 ? nnz / m : 0;
    if (r <= 2) return 2;
    if (r <= 4) return 4;
    if (r <= 8) return 8;
    if (r <= 16) return 16;
    if (r <= 32) return 32;
    return 64;
}

template <typename T, int TPV>
void launch_csrvec_t(Impl<T> *im, T alpha, hipEvent_t e0, hipEvent_t e1)
{
    const int threads = 256;
    const long long vpb = threads / TPV;
    long long blocks = ((long long)im->m + vpb - 1) / vpb;
    blocks = std::max<long long>(1, std::min<long long>(blocks, 256LL * 32));
    // x is read ahead into each XCD's L2 when it takes at most 1 MB there
    const long long x_bytes = (long long)im->n * (long long)sizeof(T);
    const long long x_lines = x_bytes <= (1LL << 20) ? (x_bytes + 127) / 128 : 0;
    if (e0 && e1)
        hipExtLaunchKernelGGL((bhs::csrvec_kernel<T, TPV>), dim3((unsigned)blocks), dim3(threads),
                              0u, im->stream, e0, e1, 0u, im->m, (const int *)im->d_row_ptr,
                              (const int *)im->d_col, (const T *)im->d_val, (const T *)im->d_x,
                              im->d_y, alpha, x_lines);
    else
        hipLaunchKernelGGL((bhs::csrvec_kernel<T, TPV>), dim3((unsigned)blocks), dim3(threads), 0,
                           im->stream, im->m, im->d_row_ptr, im->d_col, im->d_val, im->d_x,
                           im->d_y, alpha, x_lines);
}

template <typename T>
void launch_csrvec(Impl<T> *im, T alpha, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    const int tpv = im->cfg.tpv ? im->cfg.tpv : tpv_rule(im->m, im->nnz);
    switch (tpv) {
    case 2: launch_csrvec_t<T, 2>(im, alpha, e0, e1); break;
    case 4: launch_csrvec_t<T, 4>(im, alpha, e0, e1); break;
    case 8: launch_csrvec_t<T, 8>(im, alpha, e0, e1); break;
    case 16: launch_csrvec_t<T, 16>(im, alpha, e0, e1); break;
    case 32: launch_csrvec_t<T, 32>(im, alpha, e0, e1); break;
    default: launch_csrvec_t<T, 64>(im, alpha, e0, e1); break;
    }
}

template <typename T, typename U>
void launch_fill(Impl<T> *im, U *p, long long n)
{
    if (n <= 0) return;
    const int threads = 256;
    const long long blocks = std::min<long long>((n + threads - 1) / threads, 4096);
    hipLaunchKernelGGL(bhs::fill_kernel<U>, dim3((unsigned)blocks), dim3(threads), 0, im->stream,
                       p, n, U(0));
}

template <typename T>
size_t segsum_lds_bytes(const bhs_config &c)
{
    const size_t per_wave = (256 + 64 * (size_t)c.items_per_lane * sizeof(T) + 15) / 16 * 16;
    return per_wave * (size_t)c.waves_per_block;
}

// Enqueues one y = alpha*A*x.  If main_start and main_done are given they are
// attached to the main kernel's dispatch and time exactly that kernel (they
// are recorded around the fill when there is no kernel to time).
template <typename T>
int enqueue_spmv(Impl<T> *im, T alpha, hipEvent_t main_start = nullptr,
                 hipEvent_t main_done = nullptr)
{
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    if (int rc = ensure_analysis(im)) return rc;
    const bool timed = main_start && main_done;
    if (im->m == 0) {
        if (timed) {
            BHS_HIP(hipEventRecord(main_start, im->stream));
            BHS_HIP(hipEventRecord(main_done, im->stream));
        }
        return BHSPARSE_SUCCESS;
    }
    if (im->cfg.variant == BHS_VARIANT_CSRVEC) {
        launch_csrvec(im, alpha, main_start, main_done);
        BHS_HIP(hipGetLastError());
        return BHSPARSE_SUCCESS;
    }
    if (im->NP == 0) {                        // no nonzeros: y = 0
        if (timed) BHS_HIP(hipEventRecord(main_start, im->stream));
        launch_fill(im, im->d_y, im->m);
        BHS_HIP(hipGetLastError());
        if (timed) BHS_HIP(hipEventRecord(main_done, im->stream));
        return BHSPARSE_SUCCESS;
    }
    if (segsum_lds_bytes<T>(im->cfg) > im->max_lds)
        return set_error(BHSPARSE_INVALID_ARGUMENT,
                         "items_per_lane x waves_per_block needs more LDS than a block may use");
    launch_segsum<T, false>(im, alpha, nullptr, main_start, main_done);
    BHS_HIP(hipGetLastError());
    launch_calibrate(im);
    BHS_HIP(hipGetLastError());
    launch_combine(im);
    BHS_HIP(hipGetLastError());
    return BHSPARSE_SUCCESS;
}

template <typename T>
int spmv_async(bhs_handle_t h, T alpha)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->platform) return set_error(BHSPARSE_NOT_INITIALIZED, "init_platform first");
    BHS_HIP(hipSetDevice(im->device));
    return enqueue_spmv(im, alpha);
}

template <typename T>
int get_stream(bhs_handle_t h, void **stream)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!stream) return set_error(BHSPARSE_INVALID_ARGUMENT, "null stream pointer");
    if (!im->platform) return set_error(BHSPARSE_NOT_INITIALIZED, "init_platform first");
    *stream = reinterpret_cast<void *>(im->stream);
    return BHSPARSE_SUCCESS;
}

template <typename T>
int spmv(bhs_handle_t h, T alpha)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (int rc = spmv_async<T>(h, alpha)) return rc;
    BHS_HIP(hipStreamSynchronize(im->stream));
    return BHSPARSE_SUCCESS;
}

template <typename T>
int get_y_to(bhs_handle_t h, T *out)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    if (im->m > 0 && !out) return set_error(BHSPARSE_INVALID_ARGUMENT, "null y");
    BHS_HIP(hipSetDevice(im->device));
    if (im->m > 0)
        BHS_HIP(hipMemcpyAsync(out, im->d_y, sizeof(T) * (size_t)im->m, hipMemcpyDeviceToHost,
                               im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    return BHSPARSE_SUCCESS;
}

template <typename T>
int get_y(bhs_handle_t h)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    return get_y_to<T>(h, im->h_y);
}

template <typename T>
int free_mem(bhs_handle_t h)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    release_data(im);
    if (im->d_flush) {
        (void)hipFree(im->d_flush);
        im->d_flush = nullptr;
        im->flush_cap = 0;
    }
    return BHSPARSE_SUCCESS;
}

template <typename T>
int free_platform(bhs_handle_t h)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->platform) return BHSPARSE_SUCCESS;
    release_data(im);
    (void)hipSetDevice(im->device);
    if (im->d_flush) (void)hipFree(im->d_flush);
    im->d_flush = nullptr;
    im->flush_cap = 0;
    free_ptr(im->d_counters);
    free_ptr(im->d_tuner_io);
    free_ptr(im->d_tuner_scores);
    for (auto &e : im->ev)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    if (im->stream) (void)hipStreamDestroy(im->stream);
    im->stream = nullptr;
    im->platform = false;
    return BHSPARSE_SUCCESS;
}

template <typename T>
int destroy(bhs_handle_t h)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    free_platform<T>(h);
    delete im;
    delete reinterpret_cast<HandleBox *>(h);
    return BHSPARSE_SUCCESS;
}

template <typename T>
int set_config(bhs_handle_t h, const bhs_config *c)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!c) return set_error(BHSPARSE_INVALID_ARGUMENT, "null config");
    if (int rc = check_config(*c)) return rc;
    im->cfg = *c;
    if (im->data) {
        BHS_HIP(hipSetDevice(im->device));
        return ensure_analysis(im);
    }
    return BHSPARSE_SUCCESS;
}

template <typename T>
int get_config(bhs_handle_t h, bhs_config *c)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!c) return set_error(BHSPARSE_INVALID_ARGUMENT, "null config");
    *c = im->cfg;
    return BHSPARSE_SUCCESS;
}

template <typename T>
int get_partition(bhs_handle_t h, int *tile_row, int cap, int *np)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    BHS_HIP(hipSetDevice(im->device));
    if (int rc = ensure_analysis(im)) return rc;
    if (np) *np = im->NP;
    if (tile_row) {
        if (cap < im->NP + 1) return set_error(BHSPARSE_INVALID_ARGUMENT, "capacity < NP+1");
        BHS_HIP(hipMemcpyAsync(tile_row, im->d_tile_row, sizeof(int) * ((size_t)im->NP + 1),
                               hipMemcpyDeviceToHost, im->stream));
        BHS_HIP(hipStreamSynchronize(im->stream));
    }
    return BHSPARSE_SUCCESS;
}

template <typename T>
int get_stats(bhs_handle_t h, bhs_spmv_stats *out)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!out) return set_error(BHSPARSE_INVALID_ARGUMENT, "null stats");
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    BHS_HIP(hipSetDevice(im->device));
    if (int rc = ensure_analysis(im)) return rc;
    std::memset(out, 0, sizeof(*out));
    out->n_tiles = im->NP;
    out->n_waves = im->NW;
    out->n_lead_empty = im->lead_empty;
    out->chunk_rows = is_segsum(im->cfg.variant) ? 64LL * im->chunk_parts : 0;
    if (im->NP == 0 || im->m == 0) return BHSPARSE_SUCCESS;
    if (segsum_lds_bytes<T>(im->cfg) > im->max_lds)
        return set_error(BHSPARSE_INVALID_ARGUMENT,
                         "items_per_lane x waves_per_block needs more LDS than a block may use");
    launch_fill(im, im->d_counters, 8);
    launch_segsum<T, true>(im, T(1), im->d_counters);
    BHS_HIP(hipGetLastError());
    launch_calibrate(im);
    BHS_HIP(hipGetLastError());
    unsigned long long c[3];
    BHS_HIP(hipMemcpyAsync(c, im->d_counters, sizeof(c), hipMemcpyDeviceToHost, im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    out->n_fast = (long long)c[0];
    out->n_dirty = (long long)c[1];
    out->n_sync = (long long)c[2];
    return BHSPARSE_SUCCESS;
}

template <typename T>
int bench(bhs_handle_t h, T alpha, int warmup, int reps, long long flush_bytes, double *ms_total,
          double *ms_main)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    if (warmup < 0 || reps < 0 || flush_bytes < 0)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "negative count");
    BHS_HIP(hipSetDevice(im->device));
    const long long n16 = flush_bytes / 16;
    if (n16 > 0 && flush_bytes > im->flush_cap) {
        if (im->d_flush) (void)hipFree(im->d_flush);
        im->d_flush = nullptr;
        im->flush_cap = 0;
        BHS_HIP(hipMalloc(&im->d_flush, (size_t)n16 * 16));
        im->flush_cap = n16 * 16;
    }
    auto flush = [&]() {
        if (n16 <= 0) return;
        hipLaunchKernelGGL(bhs::flush_kernel, dim3(4096), dim3(256), 0, im->stream,
                           reinterpret_cast<bhs::u32x4 *>(im->d_flush), n16, 0u);
    };
    for (int i = 0; i < warmup; ++i) {
        flush();
        if (int rc = enqueue_spmv(im, alpha)) return rc;
    }
    BHS_HIP(hipStreamSynchronize(im->stream));
    // whole calls: two events around each, nothing between its kernels
    for (int i = 0; ms_total && i < reps; ++i) {
        flush();
        BHS_HIP(hipEventRecord(im->ev[0], im->stream));
        if (int rc = enqueue_spmv(im, alpha)) return rc;
        BHS_HIP(hipEventRecord(im->ev[2], im->stream));
        BHS_HIP(hipEventSynchronize(im->ev[2]));
        float a = 0;
        BHS_HIP(hipEventElapsedTime(&a, im->ev[0], im->ev[2]));
        ms_total[i] = a;
    }
    // the main kernel alone: events attached to its dispatch, separate calls
    for (int i = 0; ms_main && i < reps; ++i) {
        flush();
        if (int rc = enqueue_spmv(im, alpha, im->ev[1], im->ev[3])) return rc;
        BHS_HIP(hipEventSynchronize(im->ev[3]));
        BHS_HIP(hipStreamSynchronize(im->stream));
        float b = 0;
        BHS_HIP(hipEventElapsedTime(&b, im->ev[1], im->ev[3]));
        ms_main[i] = b;
    }
    return BHSPARSE_SUCCESS;
}

template <typename T, int S>
void launch_floor_s(Impl<T> *im, hipEvent_t e0, hipEvent_t e1)
{
    const long long threads = (long long)im->NP * 64;
    const Layout<T> L = active_layout(im);
    if (e0 && e1)
        hipExtLaunchKernelGGL((bhs::floor_kernel<T, S>), dim3((unsigned)((threads + 255) / 256)),
                              dim3(256), 0u, im->stream, e0, e1, 0u, L.col, L.val,
                              (const T *)im->d_x, im->NP, im->d_floor);
    else
        hipLaunchKernelGGL((bhs::floor_kernel<T, S>), dim3((unsigned)((threads + 255) / 256)),
                           dim3(256), 0, im->stream, L.col, L.val, im->d_x, im->NP, im->d_floor);
}

// The fixed cost of the timing events themselves: an interval between the two
// event slots that bench() attaches to the main kernel, with nothing in it,
// after the same cache flush.  Subtracting its median from ms_main gives the
// kernel's own duration (what a kernel trace reports).
template <typename T>
int bench_event_overhead(bhs_handle_t h, int reps, long long flush_bytes, double *ms)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->platform) return set_error(BHSPARSE_NOT_INITIALIZED, "init_platform first");
    if (reps < 1 || !ms || flush_bytes < 0)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "bad reps/output");
    BHS_HIP(hipSetDevice(im->device));
    const long long n16 = flush_bytes / 16;
    if (n16 > 0 && flush_bytes > im->flush_cap) {
        if (im->d_flush) (void)hipFree(im->d_flush);
        im->d_flush = nullptr;
        im->flush_cap = 0;
        BHS_HIP(hipMalloc(&im->d_flush, (size_t)n16 * 16));
        im->flush_cap = n16 * 16;
    }
    for (int i = 0; i < reps; ++i) {
        if (n16 > 0)
            hipLaunchKernelGGL(bhs::flush_kernel, dim3(4096), dim3(256), 0, im->stream,
                               reinterpret_cast<bhs::u32x4 *>(im->d_flush), n16, 0u);
        BHS_HIP(hipEventRecord(im->ev[1], im->stream));
        BHS_HIP(hipEventRecord(im->ev[3], im->stream));
        BHS_HIP(hipEventSynchronize(im->ev[3]));
        float t = 0;
        BHS_HIP(hipEventElapsedTime(&t, im->ev[1], im->ev[3]));
        ms[i] = t;
    }
    return BHSPARSE_SUCCESS;
}

template <typename T>
int bench_floor(bhs_handle_t h, int reps, long long flush_bytes, double *ms)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    if (reps < 1 || !ms || flush_bytes < 0)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "bad reps/output");
    BHS_HIP(hipSetDevice(im->device));
    if (int rc = ensure_analysis(im)) return rc;
    if (im->NP == 0) {
        for (int i = 0; i < reps; ++i) ms[i] = 0;
        return BHSPARSE_SUCCESS;
    }
    const long long need = (long long)im->NP * 64;
    if (need > im->floor_cap) {
        free_ptr(im->d_floor);
        BHS_HIP(hipMalloc(&im->d_floor, sizeof(T) * (size_t)need));
        im->floor_cap = need;
    }
    const long long n16 = flush_bytes / 16;
    if (n16 > 0 && flush_bytes > im->flush_cap) {
        if (im->d_flush) (void)hipFree(im->d_flush);
        im->d_flush = nullptr;
        im->flush_cap = 0;
        BHS_HIP(hipMalloc(&im->d_flush, (size_t)n16 * 16));
        im->flush_cap = n16 * 16;
    }
    auto launch = [&](hipEvent_t e0, hipEvent_t e1) {
        switch (im->cfg.items_per_lane) {
        case 4: launch_floor_s<T, 4>(im, e0, e1); break;
        case 8: launch_floor_s<T, 8>(im, e0, e1); break;
        case 16: launch_floor_s<T, 16>(im, e0, e1); break;
        default: launch_floor_s<T, 32>(im, e0, e1); break;
        }
    };
    launch(nullptr, nullptr);
    BHS_HIP(hipGetLastError());
    for (int i = 0; i < reps; ++i) {
        if (n16 > 0)
            hipLaunchKernelGGL(bhs::flush_kernel, dim3(4096), dim3(256), 0, im->stream,
                               reinterpret_cast<bhs::u32x4 *>(im->d_flush), n16, 0u);
        launch(im->ev[1], im->ev[3]);
        BHS_HIP(hipGetLastError());
        BHS_HIP(hipEventSynchronize(im->ev[3]));
        float t = 0;
        BHS_HIP(hipEventElapsedTime(&t, im->ev[1], im->ev[3]));
        ms[i] = t;
    }
    return BHSPARSE_SUCCESS;
}

// Median time (us) of every candidate configuration; the fastest is installed.
// Candidates that cannot run (LDS limit) get -1.
template <typename T>
int autotune(bhs_handle_t h, const bhs_config *cands, int n, int reps, long long flush_bytes,
             double *us, int *best)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    if (!cands || n < 1 || reps < 1 || !us) return set_error(BHSPARSE_INVALID_ARGUMENT, "bad candidates");
    for (int i = 0; i < n; ++i)
        if (int rc = check_config(cands[i])) return rc;
    const bhs_config keep = im->cfg;
    std::vector<double> t(reps);
    int bi = -1;
    for (int i = 0; i < n; ++i) {
        us[i] = -1;
        if (is_segsum(cands[i].variant) && segsum_lds_bytes<T>(cands[i]) > im->max_lds)
            continue;
        im->cfg = cands[i];
        if (int rc = bench<T>(h, T(1), 2, reps, flush_bytes, t.data(), nullptr)) {
            im->cfg = keep;
            return rc;
        }
        std::vector<double> s = t;
        std::nth_element(s.begin(), s.begin() + reps / 2, s.end());
        us[i] = s[reps / 2] * 1e3;
        if (bi < 0 || us[i] < us[bi]) bi = i;
    }
    if (bi < 0) {
        im->cfg = keep;
        return set_error(BHSPARSE_INVALID_ARGUMENT, "no runnable candidate");
    }
    im->cfg = cands[bi];
    if (best) *best = bi;
    BHS_HIP(hipSetDevice(im->device));
    return ensure_analysis(im);
}

// Learned prediction: row statistics (GPU reduction) -> features -> MLP, all
// on the device.  *choice = argmin of the predicted cost over the classes that
// can run; if scores is given, it receives every class's predicted cost and
// runnable[k] whether class k can run on this handle.
template <typename T>
int tuner_predict(Impl<T> *im, bhs_mlp_t net, const bhs_config *classes, int K, int *choice,
                  std::vector<float> *scores, std::vector<int> *runnable)
{
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    if (!net || !classes) return set_error(BHSPARSE_INVALID_ARGUMENT, "null network/classes");
    const bhs::Mlp *mlp = reinterpret_cast<const bhs::Mlp *>(net);
    if (mlp->n_out != K || K > bhs::kMaxClasses)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "class count does not match the network");
    if (mlp->device != im->device)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "network lives on another device");
    BHS_HIP(hipSetDevice(im->device));
    std::vector<int> io(4 + K);
    io[0] = im->m;
    io[1] = im->nnz;
    io[2] = sizeof(T) == 8;
    io[3] = -1;
    for (int k = 0; k < K; ++k) {
        const bool ok = check_config(classes[k]) == BHSPARSE_SUCCESS &&
                        !(is_segsum(classes[k].variant) &&
                          segsum_lds_bytes<T>(classes[k]) > im->max_lds);
        io[4 + k] = ok;
    }
    if (!im->d_tuner_io) BHS_HIP(hipMalloc(&im->d_tuner_io, sizeof(int) * (4 + bhs::kMaxClasses)));
    if (scores && !im->d_tuner_scores)
        BHS_HIP(hipMalloc(&im->d_tuner_scores, sizeof(float) * bhs::kMaxClasses));
    BHS_HIP(hipMemcpyAsync(im->d_tuner_io, io.data(), sizeof(int) * io.size(),
                           hipMemcpyHostToDevice, im->stream));
    launch_fill(im, im->d_counters, 8);
    if (im->m > 0) {
        const int blocks = (int)std::min<long long>(((long long)im->m + 255) / 256, 2048);
        hipLaunchKernelGGL(bhs::row_stats_kernel, dim3(blocks), dim3(256), 0, im->stream, im->m,
                           im->d_row_ptr, im->d_col, im->d_counters);
    }
    BHS_HIP(bhs::launch_tuner_infer(im->d_counters, im->d_tuner_io, im->d_tuner_io + 2, 1, *mlp, K,
                                    im->d_tuner_io + 4, im->d_tuner_io + 3,
                                    scores ? im->d_tuner_scores : nullptr, im->stream));
    int c = -1;
    BHS_HIP(hipMemcpyAsync(&c, im->d_tuner_io + 3, sizeof(int), hipMemcpyDeviceToHost, im->stream));
    if (scores) {
        scores->resize(K);
        BHS_HIP(hipMemcpyAsync(scores->data(), im->d_tuner_scores, sizeof(float) * K,
                               hipMemcpyDeviceToHost, im->stream));
    }
    BHS_HIP(hipStreamSynchronize(im->stream));
    if (c < 0 || c >= K) return set_error(BHSPARSE_INVALID_ARGUMENT, "no runnable class");
    if (runnable) runnable->assign(io.begin() + 4, io.end());
    *choice = c;
    return BHSPARSE_SUCCESS;
}

// Learned choice: installs the class with the lowest predicted cost.
template <typename T>
int apply_tuner(bhs_handle_t h, bhs_mlp_t net, const bhs_config *classes, int K, int *chosen)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    int c = -1;
    if (int rc = tuner_predict(im, net, classes, K, &c, nullptr, nullptr)) return rc;
    im->cfg = classes[c];
    if (chosen) *chosen = c;
    return ensure_analysis(im);
}

// Learned shortlist: the topk runnable classes with the lowest predicted cost
// are timed as the exhaustive autotuner times them, and the fastest is
// installed.  topk = 1 is apply_tuner.
template <typename T>
int apply_tuner_topk(bhs_handle_t h, bhs_mlp_t net, const bhs_config *classes, int K, int topk,
                     int reps, long long flush_bytes, int *chosen)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (topk < 1 || reps < 1)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "topk and reps must be at least 1");
    if (topk == 1) return apply_tuner<T>(h, net, classes, K, chosen);
    int c = -1;
    std::vector<float> sc;
    std::vector<int> ok;
    if (int rc = tuner_predict(im, net, classes, K, &c, &sc, &ok)) return rc;
    std::vector<int> order;
    for (int k = 0; k < K; ++k)
        if (ok[k]) order.push_back(k);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sc[a] < sc[b]; });
    if ((int)order.size() > topk) order.resize(topk);
    std::vector<bhs_config> cands;
    for (int k : order) cands.push_back(classes[k]);
    std::vector<double> us(cands.size());
    int bi = -1;
    if (int rc = autotune<T>(h, cands.data(), (int)cands.size(), reps, flush_bytes, us.data(), &bi))
        return rc;
    if (chosen) *chosen = order[bi];
    return BHSPARSE_SUCCESS;
}

template <typename T>
int row_features(bhs_handle_t h, double *f)
{
    Impl<T> *im;
    if (int rc = get_impl(h, &im)) return rc;
    if (!f) return set_error(BHSPARSE_INVALID_ARGUMENT, "null output");
    if (!im->data) return set_error(BHSPARSE_NOT_INITIALIZED, "init_data first");
    BHS_HIP(hipSetDevice(im->device));
    launch_fill(im, im->d_counters, 8);
    if (im->m > 0) {
        const int threads = 256;
        const int blocks = (int)std::min<long long>(((long long)im->m + threads - 1) / threads, 2048);
        hipLaunchKernelGGL(bhs::row_stats_kernel, dim3(blocks), dim3(threads), 0, im->stream,
                           im->m, im->d_row_ptr, im->d_col, im->d_counters);
        BHS_HIP(hipGetLastError());
    }
    unsigned long long c[5];
    BHS_HIP(hipMemcpyAsync(c, im->d_counters, sizeof(c), hipMemcpyDeviceToHost, im->stream));
    BHS_HIP(hipStreamSynchronize(im->stream));
    const double m = im->m, nnz = im->nnz;
    const double mean = m > 0 ? (double)c[0] / m : 0.0;
    const double var = m > 0 ? (double)c[1] / m - mean * mean : 0.0;
    const double nonempty = m - (double)c[2];
    f[0] = m;
    f[1] = nnz;
    f[2] = mean;
    f[3] = var > 0 ? var : 0.0;
    f[4] = (double)c[4];
    f[5] = m > 0 ? (double)c[2] / m : 0.0;
    f[6] = nonempty > 0 ? (double)c[3] / nonempty : 0.0;
    f[7] = sizeof(T) == 8 ? 1.0 : 0.0;
    return BHSPARSE_SUCCESS;
}

#include "bhs_batch.inc"

// ---- y = alpha*A*x with no allocation ------------------------------------
// Device scratch of the allocation-free entry point: the tile partition and
// the waves' boundary-row partials, at most kScratchBytes.  One copy per
// device, held until the process ends.  Host code never uses the array's
// name as an address: scratch_ptr() looks the device's copy up once per
// device (hipGetSymbolAddress, with that device current) and keeps it.
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kScratchDevices] = {};

// The current device's scratch (looked up on its first use on that device).
int scratch_ptr(void **out)
{
    int dev = -1;
    BHS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return set_error(BHSPARSE_UNSUPPORTED_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        BHS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(BHSPARSE_HIP_ERROR, "hipGetSymbolAddress(scratch) gave null");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return BHSPARSE_SUCCESS;
}

// The same kernels as a handle's default tile (segmented sum, S = 8, 4 waves
// per block, 4 tiles per wave, plain layout): partition, segmented sum with
// checked loads of the last tile (the caller's col/val are not padded),
// calibration.  Enqueued on `stream`; nothing is allocated, copied to the
// host or waited for, so a caller may capture it in a graph.
template <typename T>
int spmv_noalloc(int m, int n, int nnz, const int *row_ptr, const int *col, const T *val,
                 const T *x, T *y, T alpha, void *stream_v)
{
    constexpr int S = 8, WPB = 4, TPW = 4, P = 64 * S;
    if (m < 0 || n < 0 || nnz < 0) return set_error(BHSPARSE_INVALID_ARGUMENT, "negative size");
    if ((long long)nnz > kMaxNnz) return set_error(BHSPARSE_INVALID_MATRIX, "nnz too large");
    if (m == 0) return BHSPARSE_SUCCESS;
    if (!row_ptr || !y || (nnz > 0 && (!col || !val || !x)))
        return set_error(BHSPARSE_INVALID_ARGUMENT, "null device pointer");
    if (nnz > 0 && n == 0) return set_error(BHSPARSE_INVALID_MATRIX, "nonzeros in a matrix with n = 0");
    if ((reinterpret_cast<uintptr_t>(col) | reinterpret_cast<uintptr_t>(val)) & 15)
        return set_error(BHSPARSE_INVALID_ARGUMENT, "col and val must be 16-byte aligned");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream_v);
    const long long NP = ((long long)nnz + P - 1) / P;
    const long long NW = (NP + TPW - 1) / TPW;
    // scratch: tile_row[NP+1] | sync_row[NW] | sync_val[NW], each 256-byte aligned
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t o_sr = up(sizeof(int) * (size_t)(NP + 1));
    const size_t o_sv = o_sr + up(sizeof(int) * (size_t)std::max<long long>(NW, 1));
    const size_t need = o_sv + sizeof(T) * (size_t)std::max<long long>(NW, 1);
    if (need > kScratchBytes)
        return set_error(BHSPARSE_INVALID_ARGUMENT,
                         "matrix needs " + std::to_string(need) + " bytes of scratch, more than the " +
                             std::to_string(kScratchBytes) + " the allocation-free entry point has");
    void *base = nullptr;
    if (int rc = scratch_ptr(&base)) return rc;
    unsigned char *sb = static_cast<unsigned char *>(base);
    int *tile_row = reinterpret_cast<int *>(sb);
    int *sync_row = reinterpret_cast<int *>(sb + o_sr);
    T *sync_val = reinterpret_cast<T *>(sb + o_sv);
    const int threads = 256;
    if (NP == 0) {                                    // no nonzeros: y = 0
        const long long blocks = std::min<long long>((m + threads - 1) / threads, 4096);
        hipLaunchKernelGGL(bhs::fill_kernel<T>, dim3((unsigned)blocks), dim3(threads), 0, st, y,
                           (long long)m, T(0));
        BHS_HIP(hipGetLastError());
        return BHSPARSE_SUCCESS;
    }
    hipLaunchKernelGGL(bhs::partition_kernel, dim3((unsigned)((NP + 1 + threads - 1) / threads)),
                       dim3(threads), 0, st, row_ptr, m, nnz, P, (int)NP, tile_row);
    BHS_HIP(hipGetLastError());
    bhs::SegsumParams<T> q;
    q.m = m;
    q.nnz = nnz;
    q.NP = (int)NP;
    q.TPW = TPW;
    q.NW = (int)NW;
    q.lead_empty = -1;                                // read from tile_row[0] on the device
    q.wpb = WPB;
    q.wave_lds_bytes = (256 + 64 * S * (int)sizeof(T) + 15) / 16 * 16;
    q.alpha = alpha;
    q.x_lines = 0;
    hipLaunchKernelGGL((bhs::segsum_kernel<T, S, false, 2, true>),
                       dim3((unsigned)((NW + WPB - 1) / WPB)), dim3(64 * WPB),
                       (size_t)WPB * q.wave_lds_bytes, st, row_ptr, col, val, x, y,
                       (const int *)tile_row, sync_row, sync_val,
                       (unsigned long long *)nullptr, q);
    BHS_HIP(hipGetLastError());
    if (NW > 1) {
        hipLaunchKernelGGL(bhs::calibrate_kernel<T>, dim3((unsigned)((NW + threads - 1) / threads)),
                           dim3(threads), 0, st, (const int *)sync_row, (const T *)sync_val,
                           (int)NW, y);
        BHS_HIP(hipGetLastError());
    }
    return BHSPARSE_SUCCESS;
}

// Runs one entry point of the C ABI.  A failed host allocation
// (std::bad_alloc from the std::vector work arrays of the analysis, the
// panel layout or a batch) becomes BHSPARSE_OUT_OF_MEMORY: no C++ exception
// crosses the C boundary, so the caller's process never aborts.
template <typename F>
int guard(F &&f)
{
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return set_error(BHSPARSE_HIP_ERROR, std::string("internal error: ") + e.what());
    } catch (...) {
        return set_error(BHSPARSE_HIP_ERROR, "internal error");
    }
}

} // namespace

extern "C" {

const char *bhs_error_string(int code)
{
    switch (code) {
    case BHSPARSE_SUCCESS: return "success";
    case BHSPARSE_UNSUPPORTED_VALUE_TYPE: return "unsupported value type";
    case BHSPARSE_UNSUPPORTED_DEVICE: return "unsupported device";
    case BHSPARSE_INVALID_ARGUMENT: return "invalid argument";
    case BHSPARSE_INVALID_MATRIX: return "invalid matrix";
    case BHSPARSE_OUT_OF_MEMORY: return "out of device memory";
    case BHSPARSE_HIP_ERROR: return "HIP runtime error";
    case BHSPARSE_NOT_INITIALIZED: return "not initialized";
    default: return "unknown error";
    }
}

const char *bhs_last_error(void) { return g_last_error.c_str(); }

int bhs_device_synchronize(void)
{
    BHS_HIP(hipDeviceSynchronize());
    return BHSPARSE_SUCCESS;
}

int bhs_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int bhs_stream_copy(int device, long long bytes, int reps, double *ms)
{
    if (bytes < 16 || reps < 1 || !ms) return set_error(BHSPARSE_INVALID_ARGUMENT, "bad copy size");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return set_error(BHSPARSE_UNSUPPORTED_DEVICE, "no such device");
    BHS_HIP(hipSetDevice(device));
    const long long n16 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = BHSPARSE_SUCCESS;
    auto fail = [&](hipError_t e, const char *what) {
        rc = set_error(e == hipErrorOutOfMemory ? BHSPARSE_OUT_OF_MEMORY : BHSPARSE_HIP_ERROR,
                       std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc(&a, (size_t)n16 * 16)) != hipSuccess) fail(e, "hipMalloc");
    else if ((e = hipMalloc(&b, (size_t)n16 * 16)) != hipSuccess) fail(e, "hipMalloc");
    else if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) fail(e, "stream");
    else if ((e = create_timing_event(&e0)) != hipSuccess ||
             (e = create_timing_event(&e1)) != hipSuccess)
        fail(e, "event");
    else if ((e = hipMemsetAsync(a, 0, (size_t)n16 * 16, st)) != hipSuccess) fail(e, "memset");
    else {
        const unsigned blocks = 256 * 16;
        for (int i = 0; i < reps && rc == BHSPARSE_SUCCESS; ++i) {
            (void)hipEventRecord(e0, st);
            hipLaunchKernelGGL(bhs::copy_kernel, dim3(blocks), dim3(256), 0, st,
                               (const bhs::u32x4 *)a, (bhs::u32x4 *)b, n16);
            (void)hipEventRecord(e1, st);
            if ((e = hipEventSynchronize(e1)) != hipSuccess) { fail(e, "copy"); break; }
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            ms[i] = t;
        }
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

#define BHS_DEFINE_API(P, T)                                                                   \
    int bhs_##P##_create(bhs_handle_t *h) { return guard([&] { return create<T>(h); }); }                            \
    int bhs_##P##_init_platform(bhs_handle_t h, int d) { return guard([&] { return init_platform<T>(h, d); }); }    \
    int bhs_##P##_init_data(bhs_handle_t h, int m, int n, int nnz, const int *rp,             \
                            const int *ci, const T *v, const T *x, T *y)                      \
    {                                                                                          \
        return guard([&] { return init_data<T>(h, m, n, nnz, rp, ci, v, x, y); });                                    \
    }                                                                                          \
    int bhs_##P##_set_x(bhs_handle_t h, const T *x) { return guard([&] { return set_x<T>(h, x); }); }                \
    int bhs_##P##_set_memory_mode(bhs_handle_t h, int md) { return guard([&] { return set_memory_mode<T>(h, md); }); } \
    int bhs_##P##_get_memory_mode(bhs_handle_t h, int *md) { return guard([&] { return get_memory_mode<T>(h, md); }); } \
    int bhs_##P##_spmv(bhs_handle_t h, T a) { return guard([&] { return spmv<T>(h, a); }); }                         \
    int bhs_##P##_spmv_async(bhs_handle_t h, T a) { return guard([&] { return spmv_async<T>(h, a); }); }             \
    int bhs_##P##_get_stream(bhs_handle_t h, void **st) { return guard([&] { return get_stream<T>(h, st); }); }      \
    int bhs_##P##_get_y(bhs_handle_t h) { return guard([&] { return get_y<T>(h); }); }                               \
    int bhs_##P##_get_y_to(bhs_handle_t h, T *o) { return guard([&] { return get_y_to<T>(h, o); }); }                \
    int bhs_##P##_free_mem(bhs_handle_t h) { return guard([&] { return free_mem<T>(h); }); }                         \
    int bhs_##P##_free_platform(bhs_handle_t h) { return guard([&] { return free_platform<T>(h); }); }               \
    int bhs_##P##_destroy(bhs_handle_t h) { return guard([&] { return destroy<T>(h); }); }                           \
    int bhs_##P##_set_config(bhs_handle_t h, const bhs_config *c) { return guard([&] { return set_config<T>(h, c); }); } \
    int bhs_##P##_get_config(bhs_handle_t h, bhs_config *c) { return guard([&] { return get_config<T>(h, c); }); }   \
    int bhs_##P##_get_partition(bhs_handle_t h, int *t, int cap, int *np)                     \
    {                                                                                          \
        return guard([&] { return get_partition<T>(h, t, cap, np); });                                                \
    }                                                                                          \
    int bhs_##P##_get_stats(bhs_handle_t h, bhs_spmv_stats *o) { return guard([&] { return get_stats<T>(h, o); }); } \
    int bhs_##P##_bench(bhs_handle_t h, T a, int w, int r, long long fb, double *mt,          \
                        double *mm)                                                            \
    {                                                                                          \
        return guard([&] { return bench<T>(h, a, w, r, fb, mt, mm); });                                               \
    }                                                                                          \
    int bhs_##P##_row_features(bhs_handle_t h, double *f) { return guard([&] { return row_features<T>(h, f); }); }    \
    int bhs_##P##_bench_floor(bhs_handle_t h, int r, long long fb, double *ms)                \
    {                                                                                          \
        return guard([&] { return bench_floor<T>(h, r, fb, ms); });                                                   \
    }                                                                                          \
    int bhs_##P##_bench_event_overhead(bhs_handle_t h, int r, long long fb, double *ms)       \
    {                                                                                          \
        return guard([&] { return bench_event_overhead<T>(h, r, fb, ms); });                                          \
    }                                                                                          \
    int bhs_##P##_autotune(bhs_handle_t h, const bhs_config *c, int n, int r, long long fb,    \
                           double *us, int *best)                                              \
    {                                                                                          \
        return guard([&] { return autotune<T>(h, c, n, r, fb, us, best); });                                          \
    }                                                                                          \
    int bhs_##P##_apply_tuner(bhs_handle_t h, bhs_mlp_t net, const bhs_config *c, int k,      \
                              int *chosen)                                                     \
    {                                                                                          \
        return guard([&] { return apply_tuner<T>(h, net, c, k, chosen); });                                           \
    }                                                                                          \
    int bhs_##P##_apply_tuner_topk(bhs_handle_t h, bhs_mlp_t net, const bhs_config *c, int k, \
                                   int topk, int r, long long fb, int *chosen)                 \
    {                                                                                          \
        return guard([&] { return apply_tuner_topk<T>(h, net, c, k, topk, r, fb, chosen); });                         \
    }

BHS_DEFINE_API(s, float)
BHS_DEFINE_API(d, double)

int bhs_s_spmv_noalloc(int m, int n, int nnz, const int *row_ptr, const int *col, const float *val,
                       const float *x, float *y, float alpha, void *stream)
{
    return guard([&] { return spmv_noalloc<float>(m, n, nnz, row_ptr, col, val, x, y, alpha, stream); });
}

int bhs_d_spmv_noalloc(int m, int n, int nnz, const int *row_ptr, const int *col,
                       const double *val, const double *x, double *y, double alpha, void *stream)
{
    return guard([&] { return spmv_noalloc<double>(m, n, nnz, row_ptr, col, val, x, y, alpha, stream); });
}

#define BHS_DEFINE_BATCH_API(P, T)                                                             \
    int bhs_##P##_batch_create(bhs_batch_t *b, int device, int n_mats, const int *m,            \
                               const int *n, const int *nnz, const int *const *row_ptr,         \
                               const int *const *col, const T *const *val, const int *k)        \
    {                                                                                          \
        return guard([&] { return batch_create<T>(b, device, n_mats, m, n, nnz, row_ptr, col, val, k); });            \
    }                                                                                          \
    int bhs_##P##_batch_set_x(bhs_batch_t b, const T *const *X) { return guard([&] { return batch_set_x<T>(b, X); }); } \
    int bhs_##P##_batch_set_schedule(bhs_batch_t b, const bhs_batch_schedule *s)               \
    {                                                                                          \
        return guard([&] { return batch_set_schedule<T>(b, s); });                                                    \
    }                                                                                          \
    int bhs_##P##_batch_get_schedule(bhs_batch_t b, bhs_batch_schedule *s)                     \
    {                                                                                          \
        return guard([&] { return batch_get_schedule<T>(b, s); });                                                    \
    }                                                                                          \
    int bhs_##P##_batch_run(bhs_batch_t b, T alpha) { return guard([&] { return batch_run<T>(b, alpha, 0); }); }     \
    int bhs_##P##_batch_run_plain(bhs_batch_t b, T alpha) { return guard([&] { return batch_run<T>(b, alpha, 1); }); } \
    int bhs_##P##_batch_get_y(bhs_batch_t b, T *const *Y, int plain)                           \
    {                                                                                          \
        return guard([&] { return batch_get_y<T>(b, Y, plain); });                                                    \
    }                                                                                          \
    int bhs_##P##_batch_bench(bhs_batch_t b, T alpha, int plain, int warmup, int reps,         \
                              double *ms)                                                      \
    {                                                                                          \
        return guard([&] { return batch_bench<T>(b, alpha, plain, warmup, reps, ms); });                              \
    }                                                                                          \
    int bhs_##P##_batch_autotune(bhs_batch_t b, const bhs_batch_schedule *c, int n, int reps, \
                                 double *ms, int *best)                                        \
    {                                                                                          \
        return guard([&] { return batch_autotune<T>(b, c, n, reps, ms, best); });                                     \
    }                                                                                          \
    int bhs_##P##_batch_destroy(bhs_batch_t b) { return guard([&] { return batch_destroy<T>(b); }); }

BHS_DEFINE_BATCH_API(s, float)
BHS_DEFINE_BATCH_API(d, double)

} // extern "C"
