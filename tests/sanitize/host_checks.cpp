// host_checks.cpp -- the host-side C/C++ of this repository under
// AddressSanitizer + UndefinedBehaviorSanitizer, with an injectable
// allocation cap.  Built and run by tests/test_sanitize.py (g++, CPU only):
//   * the Matrix Market reader (csrc/mmio.cpp) on every fixture in
//     tests/golden/mtx, malformed ones included, in both layouts;
//   * the matrix generators (csrc/matgen.cpp) and the CPU oracle
//     (oracle/oracle.c) on small inputs;
//   * allocation failure: operator new is replaced here so that a request
//     above g_cap_bytes throws std::bad_alloc; the reader must then return
//     BHSPARSE_OUT_OF_MEMORY and the generators nullptr, never abort.
// Exit status 0 when every check holds; a sanitizer finding aborts.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <dirent.h>
#include <new>
#include <string>
#include <vector>

#include "../../include/bhs_mmio.h"
#include "../../include/bhsparse_spmv.h"
#include "../../oracle/oracle.h"

extern "C" {
bhs_csr *bhs_gen_uniform_rows(unsigned long long seed, int m, int n, int lmin, int lmax);
bhs_csr *bhs_gen_power_law(unsigned long long seed, int m, int n, double alpha, int kmax,
                           double empty_frac, long long target_nnz);
bhs_csr *bhs_gen_laplacian7(int nx, int ny, int nz);
bhs_csr *bhs_gen_block3(unsigned long long seed, int m, int bpr, int bw);
bhs_csr *bhs_gen_banded(unsigned long long seed, int m, int n, int bandwidth);
void bhs_gen_vector(unsigned long long seed, int n, double *x);
}

static size_t g_cap_bytes = 0;   // 0: no cap

void *operator new(size_t n)
{
    if (g_cap_bytes && n > g_cap_bytes) throw std::bad_alloc();
    if (void *p = std::malloc(n ? n : 1)) return p;
    throw std::bad_alloc();
}
void *operator new[](size_t n) { return operator new(n); }
void *operator new(size_t n, const std::nothrow_t &) noexcept
{
    if (g_cap_bytes && n > g_cap_bytes) return nullptr;
    return std::malloc(n ? n : 1);
}
void *operator new[](size_t n, const std::nothrow_t &t) noexcept { return operator new(n, t); }
void operator delete(void *p, const std::nothrow_t &) noexcept { std::free(p); }
void operator delete[](void *p, const std::nothrow_t &) noexcept { std::free(p); }
void operator delete(void *p) noexcept { std::free(p); }
void operator delete[](void *p) noexcept { std::free(p); }
void operator delete(void *p, size_t) noexcept { std::free(p); }
void operator delete[](void *p, size_t) noexcept { std::free(p); }

static int g_fail = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            g_fail = 1;                                                       \
        }                                                                     \
    } while (0)

static bool csr_ok(const bhs_csr *A)
{
    if (!A || A->row_ptr[0] != 0 || A->row_ptr[A->m] != A->nnz) return false;
    for (int i = 0; i < A->m; ++i) {
        if (A->row_ptr[i + 1] < A->row_ptr[i]) return false;
        for (int j = A->row_ptr[i]; j < A->row_ptr[i + 1]; ++j)
            if (A->col[j] < 0 || A->col[j] >= A->n) return false;
    }
    return true;
}

// Every oracle entry point on one matrix.
static void oracle_on(const bhs_csr *A)
{
    std::vector<double> x((size_t)A->n + 1), y((size_t)A->m + 1), y2((size_t)A->m + 1);
    bhs_gen_vector(3, A->n, x.data());
    oracle_spmv_serial_f64(A->m, A->row_ptr, A->col, A->val, x.data(), y.data());
    oracle_spmv_omp_f64(A->m, A->row_ptr, A->col, A->val, x.data(), y2.data(), 2);
    for (int i = 0; i < A->m; ++i) CHECK(y[i] == y2[i]);
    std::vector<float> vf((size_t)A->nnz + 1), xf((size_t)A->n + 1), yf((size_t)A->m + 1);
    for (int j = 0; j < A->nnz; ++j) vf[j] = (float)A->val[j];
    for (int j = 0; j < A->n; ++j) xf[j] = (float)x[j];
    oracle_spmv_serial_f32(A->m, A->row_ptr, A->col, vf.data(), xf.data(), yf.data());
    for (int P : {4, 64, 512}) {
        const int NP = (A->nnz + P - 1) / P;
        std::vector<int> tr((size_t)NP + 1);
        oracle_partition(A->m, A->nnz, A->row_ptr, P, tr.data());
        oracle_tile_stats st;
        oracle_tile_stats_compute(A->m, A->nnz, A->row_ptr, P, 3, &st);
        CHECK(st.n_tiles == NP);
    }
    std::vector<double> yr((size_t)A->m + 1);
    oracle_ref_stats rs;
    CHECK(oracle_segsum_ref(8, 4, 3, A->m, A->nnz, A->row_ptr, A->col, A->val, x.data(),
                            yr.data(), &rs) == 0);
    CHECK(oracle_rel_error_f64(A->m, A->row_ptr, A->col, A->val, x.data(), yr.data(), y.data(),
                               nullptr) <= 1e-13);
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: host_checks <mtx dir> <scratch dir>\n");
        return 2;
    }
    const std::string mtx = argv[1], tmp = argv[2];

    // 1. the reader on every fixture, both layouts
    int files = 0;
    if (DIR *d = opendir(mtx.c_str())) {
        while (dirent *e = readdir(d)) {
            const std::string f = e->d_name;
            if (f.size() < 4 || f.substr(f.size() - 4) != ".mtx") continue;
            for (int flags : {0, BHS_MM_SORT_COLUMNS | BHS_MM_MERGE_DUPLICATES}) {
                bhs_csr *A = nullptr;
                const int rc = bhs_mm_read((mtx + "/" + f).c_str(), flags, &A);
                if (rc == BHSPARSE_SUCCESS) {
                    CHECK(csr_ok(A));
                    oracle_on(A);
                    bhs_csr_free(A);
                } else {
                    CHECK(A == nullptr);
                    CHECK(rc == BHSPARSE_INVALID_ARGUMENT || rc == BHSPARSE_INVALID_MATRIX ||
                          rc == BHSPARSE_UNSUPPORTED_VALUE_TYPE);
                }
            }
            ++files;
        }
        closedir(d);
    }
    CHECK(files >= 10);

    // 2. generators and the oracle
    bhs_csr *gens[] = {bhs_gen_uniform_rows(1, 300, 400, 0, 32),
                       bhs_gen_power_law(2, 500, 500, 2.0, 400, 0.1, 3000),
                       bhs_gen_laplacian7(6, 5, 4), bhs_gen_block3(4, 301, 9, 60),
                       bhs_gen_banded(5, 200, 210, 7)};
    for (bhs_csr *A : gens) {
        CHECK(csr_ok(A));
        if (A) oracle_on(A);
        bhs_csr_free(A);
    }
    CHECK(bhs_gen_uniform_rows(1, -1, 4, 0, 2) == nullptr);
    CHECK(bhs_gen_laplacian7(2000, 2000, 2000) == nullptr);   // more rows than int holds

    // 3. allocation failure under a cap: error codes, never an abort
    const std::string big = tmp + "/big.mtx";
    if (FILE *f = std::fopen(big.c_str(), "w")) {
        std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n3000 3000 20000\n");
        for (int k = 0; k < 20000; ++k)
            std::fprintf(f, "%d %d %.17g\n", k % 3000 + 1, (k * 7919) % 3000 + 1, 0.5 + k);
        std::fclose(f);
    }
    bhs_csr *A = nullptr;
    CHECK(bhs_mm_read(big.c_str(), BHS_MM_SORT_COLUMNS, &A) == BHSPARSE_SUCCESS && csr_ok(A));
    bhs_csr_free(A);
    g_cap_bytes = 16 << 10;   // vectors of the reader outgrow 16 KiB
    A = nullptr;
    CHECK(bhs_mm_read(big.c_str(), BHS_MM_SORT_COLUMNS, &A) == BHSPARSE_OUT_OF_MEMORY);
    CHECK(A == nullptr);
    CHECK(bhs_gen_uniform_rows(1, 100000, 100, 0, 4) == nullptr);
    CHECK(bhs_gen_power_law(2, 100000, 1000, 2.0, 100, 0.1, 0) == nullptr);
    g_cap_bytes = 0;
    CHECK((A = bhs_gen_uniform_rows(1, 100000, 100, 0, 4)) != nullptr && csr_ok(A));
    bhs_csr_free(A);

    std::printf("host checks: %s (%d Matrix Market files)\n", g_fail ? "FAILED" : "ok", files);
    return g_fail;
}
