// mmio.cpp -- Matrix Market (coordinate) file to CSR.
//
// Follows what the reference reads and how it lays out the CSR
// (spmv_cuda/main.cu:581-733, banner and size line as in the NIST mmio.h it
// vendors, :249-362): 1-based coordinates, pattern -> 1.0, integer -> double,
// symmetric/hermitian off-diagonal entries mirrored, rows filled by a
// counting sort in file order.  Optionally sorts rows by column and merges
// duplicates, which the reference does not do.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/bhs_mmio.h"
#include "../../include/bhsparse_spmv.h"

namespace {

std::string lower(std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

// Next line that is neither a comment nor blank; false at end of file.
bool next_data_line(FILE *f, std::string &line)
{
    char buf[4096];
    while (fgets(buf, sizeof(buf), f)) {
        line.assign(buf);
        while (!line.empty() && line.back() != '\n' && !feof(f)) {   // long line
            if (!fgets(buf, sizeof(buf), f)) break;
            line.append(buf);
        }
        size_t i = line.find_first_not_of(" \t\r\n");
        if (i == std::string::npos || line[i] == '%') continue;
        return true;
    }
    return false;
}

int read_impl(const char *path, int flags, bhs_csr **out)
{
    if (!path || !out) return BHSPARSE_INVALID_ARGUMENT;
    if ((flags & BHS_MM_MERGE_DUPLICATES) && !(flags & BHS_MM_SORT_COLUMNS))
        return BHSPARSE_INVALID_ARGUMENT;
    File file;
    file.f = fopen(path, "r");
    if (!file.f) return BHSPARSE_INVALID_ARGUMENT;

    // banner: %%MatrixMarket matrix coordinate <field> <symmetry>
    char buf[1024];
    if (!fgets(buf, sizeof(buf), file.f)) return BHSPARSE_INVALID_ARGUMENT;
    char w[5][64] = {{0}};
    if (sscanf(buf, "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]) != 5)
        return BHSPARSE_INVALID_ARGUMENT;
    if (std::string(w[0]) != "%%MatrixMarket" || lower(w[1]) != "matrix" ||
        lower(w[2]) != "coordinate")
        return BHSPARSE_INVALID_ARGUMENT;
    const std::string field = lower(w[3]), sym = lower(w[4]);
    if (field == "complex") return BHSPARSE_UNSUPPORTED_VALUE_TYPE;
    const bool pattern = field == "pattern";
    if (!pattern && field != "real" && field != "double" && field != "integer")
        return BHSPARSE_INVALID_ARGUMENT;
    const bool mirror = sym == "symmetric" || sym == "hermitian";
    const bool skew = sym == "skew-symmetric";
    if (!mirror && !skew && sym != "general") return BHSPARSE_INVALID_ARGUMENT;

    // size line
    std::string line;
    if (!next_data_line(file.f, line)) return BHSPARSE_INVALID_MATRIX;
    long long m = -1, n = -1, nz = -1;
    if (sscanf(line.c_str(), "%lld %lld %lld", &m, &n, &nz) != 3 || m < 0 || n < 0 || nz < 0 ||
        m > 0x7ffffffeLL || n > 0x7fffffffLL)
        return BHSPARSE_INVALID_MATRIX;
    if ((mirror || skew) && m != n) return BHSPARSE_INVALID_MATRIX;
    // the declared count must fit the format: 32-bit indices, and no more
    // entries than the matrix (or its stored triangle) has positions
    const double positions = (mirror || skew) ? (double)m * (double)(m + 1) / 2 : (double)m * n;
    if (nz > 0x7fffffffLL || (double)nz > positions) return BHSPARSE_INVALID_MATRIX;

    // entries (storage grows with what the file holds, not with what its
    // size line claims)
    std::vector<int> ri, ci;
    std::vector<double> vi;
    const size_t hint = (size_t)std::min<long long>(nz, 1LL << 16);
    ri.reserve(hint);
    ci.reserve(hint);
    vi.reserve(hint);
    for (long long k = 0; k < nz; ++k) {
        if (!next_data_line(file.f, line)) return BHSPARSE_INVALID_MATRIX;
        long long i = 0, j = 0;
        double v = 1.0;
        const int got = pattern ? sscanf(line.c_str(), "%lld %lld", &i, &j)
                                : sscanf(line.c_str(), "%lld %lld %lf", &i, &j, &v);
        if (got != (pattern ? 2 : 3) || i < 1 || i > m || j < 1 || j > n)
            return BHSPARSE_INVALID_MATRIX;
        if (skew && i == j) return BHSPARSE_INVALID_MATRIX;   // skew: zero diagonal, not stored
        ri.push_back((int)(i - 1));
        ci.push_back((int)(j - 1));
        vi.push_back(v);
    }

    // counting sort by row; mirrored entries follow their source entry
    std::vector<long long> cnt((size_t)m + 1, 0);
    for (size_t k = 0; k < ri.size(); ++k) {
        ++cnt[(size_t)ri[k] + 1];
        if ((mirror || skew) && ri[k] != ci[k]) ++cnt[(size_t)ci[k] + 1];
    }
    for (long long r = 0; r < m; ++r) cnt[(size_t)r + 1] += cnt[(size_t)r];
    const long long total = cnt[(size_t)m];
    if (total > 0x7fffffffLL) return BHSPARSE_INVALID_MATRIX;
    std::vector<int> col((size_t)total);
    std::vector<double> val((size_t)total);
    std::vector<long long> pos(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < ri.size(); ++k) {
        long long d = pos[(size_t)ri[k]]++;
        col[(size_t)d] = ci[k];
        val[(size_t)d] = vi[k];
        if ((mirror || skew) && ri[k] != ci[k]) {
            d = pos[(size_t)ci[k]]++;
            col[(size_t)d] = ri[k];
            val[(size_t)d] = skew ? -vi[k] : vi[k];
        }
    }

    std::vector<int> rp((size_t)m + 1);
    for (long long r = 0; r <= m; ++r) rp[(size_t)r] = (int)cnt[(size_t)r];
    if (flags & BHS_MM_SORT_COLUMNS) {
        std::vector<int> perm;
        std::vector<int> c2;
        std::vector<double> v2;
        int w_at = 0;
        for (long long r = 0; r < m; ++r) {
            const int b = rp[(size_t)r], e = rp[(size_t)r + 1];
            perm.resize((size_t)(e - b));
            std::iota(perm.begin(), perm.end(), b);
            std::stable_sort(perm.begin(), perm.end(),
                             [&](int a, int c) { return col[(size_t)a] < col[(size_t)c]; });
            c2.clear();
            v2.clear();
            for (int p : perm) {
                if ((flags & BHS_MM_MERGE_DUPLICATES) && !c2.empty() && c2.back() == col[(size_t)p])
                    v2.back() += val[(size_t)p];
                else {
                    c2.push_back(col[(size_t)p]);
                    v2.push_back(val[(size_t)p]);
                }
            }
            rp[(size_t)r] = w_at;
            std::copy(c2.begin(), c2.end(), col.begin() + w_at);
            std::copy(v2.begin(), v2.end(), val.begin() + w_at);
            w_at += (int)c2.size();
        }
        rp[(size_t)m] = w_at;
    }

    bhs_csr *A = (bhs_csr *)calloc(1, sizeof(bhs_csr));
    if (!A) return BHSPARSE_OUT_OF_MEMORY;
    A->m = (int)m;
    A->n = (int)n;
    A->nnz = rp[(size_t)m];
    A->row_ptr = (int *)malloc(sizeof(int) * ((size_t)m + 1));
    A->col = (int *)malloc(sizeof(int) * ((size_t)A->nnz + 1));
    A->val = (double *)malloc(sizeof(double) * ((size_t)A->nnz + 1));
    if (!A->row_ptr || !A->col || !A->val) {
        bhs_csr_free(A);
        return BHSPARSE_OUT_OF_MEMORY;
    }
    std::copy(rp.begin(), rp.end(), A->row_ptr);
    std::copy(col.begin(), col.begin() + A->nnz, A->col);
    std::copy(val.begin(), val.begin() + A->nnz, A->val);
    *out = A;
    return BHSPARSE_SUCCESS;
}

} // namespace

// Host allocation failures (std::bad_alloc from the vectors) become an error
// code: no exception crosses the C boundary.
extern "C" int bhs_mm_read(const char *path, int flags, bhs_csr **out)
{
    try {
        return read_impl(path, flags, out);
    } catch (const std::bad_alloc &) {
        return BHSPARSE_OUT_OF_MEMORY;
    } catch (...) {
        return BHSPARSE_INVALID_MATRIX;
    }
}
