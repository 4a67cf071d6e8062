// matgen.cpp -- seeded synthetic CSR generators for the SpMV workloads.
//
// The reference builds its inputs from Matrix Market files with rand()%10
// values (spmv_cuda/main.cu:581-743).  The workloads here are synthetic and
// fully determined by a seed: every row draws from its own counter-based
// stream (splitmix64 of seed and row), so results do not depend on the number
// of host threads.  Columns inside a row are distinct and sorted; values and x
// are uniform in [-1, 1] (FP64; the FP32 build rounds them).
//
// All generators return a heap-allocated bhs_csr (free with bhs_csr_free).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <unordered_set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../include/bhs_mmio.h"   // bhs_csr

namespace {

inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** seeded from splitmix64(seed ^ stream-mix)
struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream)
    {
        uint64_t t = seed * 0xD1B54A32D192ED03ull + stream * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        for (auto &v : s) v = splitmix64(t);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform double in [0, 1) with 53 random bits
    inline double uni() { return (double)(next() >> 11) * 0x1.0p-53; }
    // uniform double in [-1, 1]
    inline double sym() { return 2.0 * uni() - 1.0; }
    // uniform integer in [0, k) (k > 0), Lemire's multiply-shift with rejection
    inline uint32_t below(uint32_t k)
    {
        uint64_t x = (uint64_t)(uint32_t)(next() >> 32) * k;
        uint32_t l = (uint32_t)x;
        if (l < k) {
            uint32_t t = (uint32_t)(-k) % k;
            while (l < t) { x = (uint64_t)(uint32_t)(next() >> 32) * k; l = (uint32_t)x; }
        }
        return (uint32_t)(x >> 32);
    }
};

// Streams: row r of a matrix uses stream (kind << 40) + r.
constexpr uint64_t kLenStream = 1ull << 40;
constexpr uint64_t kColStream = 2ull << 40;
constexpr uint64_t kValStream = 3ull << 40;
constexpr uint64_t kMiscStream = 4ull << 40;
constexpr uint64_t kXStream = 5ull << 40;

// k distinct values from [lo, lo + span), sorted (Floyd's algorithm).
void sample_sorted(Rng &g, int lo, int span, int k, int *out)
{
    if (k <= 0) return;
    if (k >= span) {
        for (int i = 0; i < span; i++) out[i] = lo + i;
        return;
    }
    if (k <= 64) {
        int cnt = 0;
        for (int j = span - k; j < span; j++) {
            int t = (int)g.below((uint32_t)j + 1);
            bool seen = false;
            for (int i = 0; i < cnt; i++) if (out[i] == t) { seen = true; break; }
            out[cnt++] = seen ? j : t;
        }
    } else if ((long long)k * 8 >= span) {
        // dense: selection over a bitmap
        std::vector<char> take(span, 0);
        int cnt = 0;
        for (int j = span - k; j < span; j++) {
            int t = (int)g.below((uint32_t)j + 1);
            if (take[t]) t = j;
            take[t] = 1;
        }
        for (int i = 0; i < span; i++) if (take[i]) out[cnt++] = i;
        for (int i = 0; i < k; i++) out[i] += lo;
        return;
    } else {
        std::unordered_set<int> seen;
        seen.reserve((size_t)k * 2);
        int cnt = 0;
        for (int j = span - k; j < span; j++) {
            int t = (int)g.below((uint32_t)j + 1);
            if (!seen.insert(t).second) { seen.insert(j); t = j; }
            out[cnt++] = t;
        }
    }
    std::sort(out, out + k);
    for (int i = 0; i < k; i++) out[i] += lo;
}

void free_csr(bhs_csr *A)
{
    if (!A) return;
    free(A->row_ptr);
    free(A->col);
    free(A->val);
    free(A);
}

// nullptr when nnz does not fit 32-bit indices or host memory runs out.
bhs_csr *alloc_csr(int m, int n, const std::vector<long long> &len)
{
    long long acc = 0;
    for (int i = 0; i < m; i++) acc += len[i];
    if (acc > 0x7fff0000LL) return nullptr;
    bhs_csr *A = (bhs_csr *)calloc(1, sizeof(bhs_csr));
    if (!A) return nullptr;
    A->m = m;
    A->n = n;
    A->nnz = (int)acc;
    A->row_ptr = (int *)malloc(sizeof(int) * ((size_t)m + 1));
    A->col = (int *)malloc(sizeof(int) * ((size_t)acc + 1));
    A->val = (double *)malloc(sizeof(double) * ((size_t)acc + 1));
    if (!A->row_ptr || !A->col || !A->val) {
        free_csr(A);
        return nullptr;
    }
    acc = 0;
    A->row_ptr[0] = 0;
    for (int i = 0; i < m; i++) { acc += len[i]; A->row_ptr[i + 1] = (int)acc; }
    return A;
}

// Runs a generator; a host allocation failure (std::bad_alloc) returns nullptr
// instead of crossing the C boundary.
template <typename F>
bhs_csr *guarded(F &&f)
{
    try {
        return f();
    } catch (...) {
        return nullptr;
    }
}

// Fill columns uniformly without replacement from [0, n) and values U[-1,1].
void fill_uniform_cols(bhs_csr *A, uint64_t seed)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < A->m; i++) {
        int b = A->row_ptr[i], e = A->row_ptr[i + 1];
        Rng gc(seed, kColStream + (uint64_t)i);
        sample_sorted(gc, 0, A->n, e - b, A->col + b);
        Rng gv(seed, kValStream + (uint64_t)i);
        for (int j = b; j < e; j++) A->val[j] = gv.sym();
    }
}

} // namespace

extern "C" {

void bhs_csr_free(bhs_csr *A) { free_csr(A); }

// x[i] = U[-1, 1], i in [0, n)
void bhs_gen_vector(unsigned long long seed, int n, double *x)
{
#pragma omp parallel for schedule(static)
    for (int blk = 0; blk < (n + 4095) / 4096; blk++) {
        Rng g(seed, kXStream + (uint64_t)blk);
        int e = std::min(n, (blk + 1) * 4096);
        for (int i = blk * 4096; i < e; i++) x[i] = g.sym();
    }
}

// Row lengths iid uniform integers in [lmin, lmax] (clipped to n).
bhs_csr *bhs_gen_uniform_rows(unsigned long long seed, int m, int n, int lmin, int lmax)
{
    return guarded([&]() -> bhs_csr * {
        if (m < 0 || n < 0 || lmin < 0 || lmax < lmin) return nullptr;
        std::vector<long long> len(m);
        for (int i = 0; i < m; i++) {
            Rng g(seed, kLenStream + (uint64_t)i);
            long long L = lmin + (long long)g.below((uint32_t)(lmax - lmin + 1));
            len[i] = std::min<long long>(L, n);
        }
        bhs_csr *A = alloc_csr(m, n, len);
        if (A) fill_uniform_cols(A, seed);
        return A;
    });
}

// Discrete power law P(k) ~ k^-alpha on [1, kmax]; floor(empty_frac*m) rows
// chosen uniformly are forced empty; the remaining lengths are scaled by one
// factor so the total is close to target_nnz (0 = no scaling), each at least 1
// and at most n.
bhs_csr *bhs_gen_power_law(unsigned long long seed, int m, int n, double alpha, int kmax,
                           double empty_frac, long long target_nnz)
{
    return guarded([&]() -> bhs_csr * {
        if (m < 0 || n <= 0 || kmax < 1 || alpha <= 0) return nullptr;
        std::vector<double> cdf(kmax);
        double acc = 0;
        for (int k = 1; k <= kmax; k++) { acc += std::pow((double)k, -alpha); cdf[k - 1] = acc; }
        for (auto &c : cdf) c /= acc;
        std::vector<long long> raw(m);
        for (int i = 0; i < m; i++) {
            Rng g(seed, kLenStream + (uint64_t)i);
            double u = g.uni();
            int k = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
            raw[i] = std::min(k, kmax);
        }
        // forced-empty rows: partial Fisher-Yates with one global stream
        std::vector<char> empty(m, 0);
        long long ne = (long long)std::floor(empty_frac * (double)m);
        if (ne > 0) {
            std::vector<int> idx(m);
            std::iota(idx.begin(), idx.end(), 0);
            Rng g(seed, kMiscStream);
            for (long long i = 0; i < ne; i++) {
                int j = (int)i + (int)g.below((uint32_t)(m - i));
                std::swap(idx[i], idx[j]);
                empty[idx[i]] = 1;
            }
        }
        double s = 1.0;
        if (target_nnz > 0) {
            double tot = 0;
            for (int i = 0; i < m; i++) if (!empty[i]) tot += (double)raw[i];
            if (tot > 0) s = (double)target_nnz / tot;
        }
        // stochastic rounding of s*k from the row's own stream keeps the expected
        // total at target_nnz
        std::vector<long long> len(m);
        for (int i = 0; i < m; i++) {
            if (empty[i]) { len[i] = 0; continue; }
            const double t = s * (double)raw[i];
            const double fl = std::floor(t);
            Rng g(seed, kMiscStream + 1 + (uint64_t)i);
            long long L = (long long)fl + (g.uni() < t - fl ? 1 : 0);
            len[i] = std::max(1LL, std::min<long long>(L, n));
        }
        bhs_csr *A = alloc_csr(m, n, len);
        if (A) fill_uniform_cols(A, seed);
        return A;
    });
}

// 7-point Laplacian on an nx*ny*nz grid, natural order, diagonal 6, off -1.
bhs_csr *bhs_gen_laplacian7(int nx, int ny, int nz)
{
    return guarded([&]() -> bhs_csr * {
        long long mm = (long long)nx * ny * nz;
        if (nx <= 0 || ny <= 0 || nz <= 0 || mm > 0x7fffffffLL) return nullptr;
        int m = (int)mm;
        std::vector<long long> len(m);
    #pragma omp parallel for schedule(static)
        for (int i = 0; i < m; i++) {
            int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
            len[i] = 1 + (x > 0) + (x < nx - 1) + (y > 0) + (y < ny - 1) + (z > 0) + (z < nz - 1);
        }
        bhs_csr *A = alloc_csr(m, m, len);
        if (!A) return nullptr;
        const long long sxy = (long long)nx * ny;
    #pragma omp parallel for schedule(static)
        for (int i = 0; i < m; i++) {
            int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
            int j = A->row_ptr[i];
            auto put = [&](long long c, double v) { A->col[j] = (int)c; A->val[j] = v; j++; };
            if (z > 0) put(i - sxy, -1.0);
            if (y > 0) put(i - nx, -1.0);
            if (x > 0) put(i - 1, -1.0);
            put(i, 6.0);
            if (x < nx - 1) put(i + 1, -1.0);
            if (y < ny - 1) put(i + nx, -1.0);
            if (z < nz - 1) put(i + sxy, -1.0);
        }
        return A;
    });
}

// Block rows of dense 3x3 blocks: m = n, nb = ceil(m/3) block rows (the last
// one truncated), `bpr` distinct block columns per block row drawn uniformly
// from |3J - 3I| <= bw, values U[-1,1].
bhs_csr *bhs_gen_block3(unsigned long long seed, int m, int bpr, int bw)
{
    return guarded([&]() -> bhs_csr * {
        if (m <= 0 || bpr <= 0 || bw < 0) return nullptr;
        const int nb = (m + 2) / 3;
        const int h = bw / 3;                       // |J - I| <= h
        std::vector<std::vector<int>> bcols(nb);
        std::vector<long long> len(m);
    #pragma omp parallel for schedule(static)
        for (int I = 0; I < nb; I++) {
            int lo = std::max(0, I - h), hi = std::min(nb - 1, I + h);
            int span = hi - lo + 1;
            int k = std::min(bpr, span);
            bcols[I].resize(k);
            Rng g(seed, kColStream + (uint64_t)I);
            sample_sorted(g, lo, span, k, bcols[I].data());
            long long rl = 0;
            for (int J : bcols[I]) rl += std::min(3, m - 3 * J);
            for (int r = 3 * I; r < std::min(m, 3 * I + 3); r++) len[r] = rl;
        }
        bhs_csr *A = alloc_csr(m, m, len);
        if (!A) return nullptr;
    #pragma omp parallel for schedule(static)
        for (int I = 0; I < nb; I++) {
            Rng g(seed, kValStream + (uint64_t)I);
            // block values row-major inside each block; row r of the block row
            // takes its three entries from each block in column order
            std::vector<double> bv(bcols[I].size() * 9);
            for (auto &v : bv) v = g.sym();
            for (int rr = 0; rr < 3; rr++) {
                int r = 3 * I + rr;
                if (r >= m) break;
                int j = A->row_ptr[r];
                for (size_t b = 0; b < bcols[I].size(); b++) {
                    int J = bcols[I][b];
                    for (int cc = 0; cc < 3; cc++) {
                        int c = 3 * J + cc;
                        if (c >= m) break;
                        A->col[j] = c;
                        A->val[j] = bv[b * 9 + rr * 3 + cc];
                        j++;
                    }
                }
            }
        }
        return A;
    });
}

// Banded rows: row i holds every column in [i - h, i + h] inside [0, n),
// bandwidth 2h+1.
bhs_csr *bhs_gen_banded(unsigned long long seed, int m, int n, int bandwidth)
{
    return guarded([&]() -> bhs_csr * {
        if (m < 0 || n <= 0 || bandwidth < 1) return nullptr;
        const int h = (bandwidth - 1) / 2;
        std::vector<long long> len(m);
        for (int i = 0; i < m; i++) {
            long long lo = std::max(0LL, (long long)i - h), hi = std::min<long long>(n - 1, (long long)i + h);
            len[i] = hi >= lo ? hi - lo + 1 : 0;
        }
        bhs_csr *A = alloc_csr(m, n, len);
        if (!A) return nullptr;
    #pragma omp parallel for schedule(static)
        for (int i = 0; i < m; i++) {
            Rng g(seed, kValStream + (uint64_t)i);
            int j = A->row_ptr[i];
            long long lo = std::max(0LL, (long long)i - h);
            for (int k = 0; k < A->row_ptr[i + 1] - A->row_ptr[i]; k++) {
                A->col[j + k] = (int)(lo + k);
                A->val[j + k] = g.sym();
            }
        }
        return A;
    });
}

// CSR from explicit arrays (copies), for tests and file readers.
bhs_csr *bhs_csr_from_arrays(int m, int n, const int *row_ptr, const int *col, const double *val)
{
    if (m < 0 || n < 0 || !row_ptr || row_ptr[m] < 0) return nullptr;
    bhs_csr *A = (bhs_csr *)calloc(1, sizeof(bhs_csr));
    if (!A) return nullptr;
    A->m = m;
    A->n = n;
    A->nnz = row_ptr[m];
    A->row_ptr = (int *)malloc(sizeof(int) * ((size_t)m + 1));
    A->col = (int *)malloc(sizeof(int) * ((size_t)A->nnz + 1));
    A->val = (double *)malloc(sizeof(double) * ((size_t)A->nnz + 1));
    if (!A->row_ptr || !A->col || !A->val) {
        free_csr(A);
        return nullptr;
    }
    memcpy(A->row_ptr, row_ptr, sizeof(int) * ((size_t)m + 1));
    memcpy(A->col, col, sizeof(int) * (size_t)A->nnz);
    memcpy(A->val, val, sizeof(double) * (size_t)A->nnz);
    return A;
}

} // extern "C"
