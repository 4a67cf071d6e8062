// Captures one bhs_d_spmv_async call into a hipGraph and replays it, for the
// segmented sum, the column-panel layout and csr-vector.  Between replays the
// program does its own eager work on the same handle (an eager spmv with
// another x, then x restored), and every replay is compared bit for bit with
// the eager result for the x it reads.  The matrix is a seeded mix of empty,
// short and long rows (rows spanning many waves exercise the calibration).
// Build: hipcc -std=c++17 -I include graph_capture.cpp -L <pkg> -lbhsparse_spmv
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bhsparse_spmv.h"

namespace {

uint64_t next(uint64_t &s)
{
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

double unit(uint64_t &s) { return (double)(next(s) >> 11) * (2.0 / 9007199254740992.0) - 1.0; }

struct Csr {
    int m, n;
    std::vector<int> rp, col;
    std::vector<double> val;
};

Csr make_matrix(int m, int n, uint64_t seed)
{
    Csr A{m, n, {0}, {}, {}};
    uint64_t s = seed;
    for (int i = 0; i < m; ++i) {
        const uint64_t r = next(s) % 100;
        int len = r < 10 ? 0 : r < 80 ? (int)(next(s) % 12) + 1 : r < 99 ? (int)(next(s) % 300)
                                                                          : (int)(next(s) % 20000);
        if (len > n) len = n;
        // sorted distinct columns: a random start and random gaps
        int c = (int)(next(s) % (uint64_t)(n - len + 1));
        for (int j = 0; j < len; ++j) {
            A.col.push_back(c);
            A.val.push_back(unit(s));
            const int room = (n - c - 1) - (len - j - 1);
            c += 1 + (room > 0 ? (int)(next(s) % (uint64_t)(room < 4 ? room : 4)) : 0);
        }
        A.rp.push_back((int)A.col.size());
    }
    return A;
}

#define CHECK(call)                                                                       \
    do {                                                                                  \
        const int rc_ = (int)(call);                                                      \
        if (rc_) {                                                                        \
            std::printf("FAIL %s -> %d (%s)\n", #call, rc_, bhs_last_error());            \
            return 1;                                                                     \
        }                                                                                 \
    } while (0)

#define CHECK_HIP(call)                                                                   \
    do {                                                                                  \
        const hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                           \
            std::printf("FAIL %s -> %s\n", #call, hipGetErrorString(e_));                 \
            return 1;                                                                     \
        }                                                                                 \
    } while (0)

int run(const Csr &A, const bhs_config &cfg, const char *name)
{
    const int nnz = A.rp.back();
    std::vector<double> x1(A.n), x2(A.n), y(A.m), ye1(A.m), ye2(A.m), yg(A.m);
    uint64_t s = 99;
    for (auto &v : x1) v = unit(s);
    for (auto &v : x2) v = unit(s);
    const double alpha = 0.75;

    bhs_handle_t h = nullptr;
    CHECK(bhs_d_create(&h));
    CHECK(bhs_d_init_platform(h, 0));
    CHECK(bhs_d_set_config(h, &cfg));
    CHECK(bhs_d_init_data(h, A.m, A.n, nnz, A.rp.data(), A.col.data(), A.val.data(), x1.data(),
                          y.data()));
    // eager results for both vectors
    CHECK(bhs_d_set_x(h, x2.data()));
    CHECK(bhs_d_spmv(h, alpha));
    CHECK(bhs_d_get_y_to(h, ye2.data()));
    CHECK(bhs_d_set_x(h, x1.data()));
    CHECK(bhs_d_spmv(h, alpha));
    CHECK(bhs_d_get_y_to(h, ye1.data()));

    void *sp = nullptr;
    CHECK(bhs_d_get_stream(h, &sp));
    hipStream_t st = reinterpret_cast<hipStream_t>(sp);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = bhs_d_spmv_async(h, alpha);
    CHECK_HIP(hipStreamEndCapture(st, &graph));
    CHECK(rc);
    CHECK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));

    int bad = 0;
    for (int rep = 0; rep < 4; ++rep) {
        const bool use_x2 = rep == 2;
        if (use_x2) CHECK(bhs_d_set_x(h, x2.data()));
        std::fill(yg.begin(), yg.end(), 12345.0);
        CHECK_HIP(hipGraphLaunch(exec, st));
        CHECK_HIP(hipStreamSynchronize(st));
        CHECK(bhs_d_get_y_to(h, yg.data()));
        const std::vector<double> &want = use_x2 ? ye2 : ye1;
        if (std::memcmp(yg.data(), want.data(), sizeof(double) * A.m) != 0) ++bad;
        // the program's own eager work between replays
        CHECK(bhs_d_set_x(h, x2.data()));
        CHECK(bhs_d_spmv(h, -1.0));
        CHECK(bhs_d_set_x(h, x1.data()));
    }
    CHECK_HIP(hipGraphExecDestroy(exec));
    CHECK_HIP(hipGraphDestroy(graph));
    CHECK(bhs_d_destroy(h));
    std::printf("%s: 4 graph replays %s\n", name, bad ? "FAIL" : "match the eager result");
    return bad ? 1 : 0;
}

} // namespace

int main()
{
    const Csr A = make_matrix(30000, 40000, 7);
    int fails = 0;
    fails += run(A, bhs_config{BHS_VARIANT_SEGSUM, 8, 4, 3, 0}, "segmented sum S=8");
    fails += run(A, bhs_config{BHS_VARIANT_SEGSUM, 4, 1, 8, 0}, "segmented sum S=4");
    fails += run(A, bhs_config{BHS_VARIANT_SEGSUM_PANELS, 8, 1, 4, 4}, "4 column panels");
    fails += run(A, bhs_config{BHS_VARIANT_CSRVEC, 8, 4, 1, 8}, "csr-vector");
    std::printf("%s\n", fails ? "FAIL" : "PASS");
    return fails ? 1 : 0;
}
