// bhs_kernels.h -- gfx950 kernels of the speculative segmented-sum SpMV.
//
// Data layout in HBM (all 0-based CSR, 32-bit indices):
//   row_ptr  (m+1) int          col  nnz_pad int (nontemporal stream)
//   val      nnz_pad T (stream)  x    n T (gathered, cached)    y  m T
//   tile_row NP+1 int  -- first row of every tile boundary, computed once
//   sync_row / sync_val  NW entries -- boundary-row partial of every wave
// nnz_pad rounds nnz up to whole tiles of the largest shape, padded with
// col = 0 / val = 0, so every lane loads its items without bounds checks.
//
// A tile is P = 64*S consecutive nonzeros (S items per lane); a wave walks
// TPW consecutive tiles carrying the open row's partial sum from tile to tile,
// as a thread bunch walks STEP tiles in the reference (spmv_cuda/common.h:39-48,
// kernels_cuda.h:64).
#pragma once
#include "bhs_device.h"

namespace bhs {

// ---------------------------------------------------------------------------
// Partition: tile_row[p] = upper_bound(row_ptr[0..m], min(p*P, nnz)) - 1.
// Same bisection as binary_search_right_boundary_kernel (utils_cuda.h:6-34)
// applied to every tile boundary (kernels_cuda.h:43-49); it depends only on
// row_ptr and P, so it runs once per matrix and tile size, not per call.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void partition_kernel(const int *__restrict__ row_ptr, int m,
                                                       int nnz, int P, int NP,
                                                       int *__restrict__ tile_row)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NP) return;
    long long k = p * (long long)P;
    const int key = k > nnz ? nnz : (int)k;
    int lo = 0, hi = m + 1;                  // search [lo, hi) of row_ptr[0..m]
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (row_ptr[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    tile_row[p] = lo - 1;
}

// Scalars of one launch (the arrays travel as separate __restrict__ kernel
// arguments, which lets the compiler use scalar loads for tile_row/row_ptr).
template <typename T>
struct SegsumParams {
    int m, nnz, NP, TPW, NW, lead_empty, wpb, wave_lds_bytes;
    T alpha;
    long long x_lines;   // > 0: x is read ahead into each XCD's L2 (x_lines 128-byte lines)
};

// One tile's column indices and values in registers.  Loads are coalesced:
// for load group j, lane l reads G = 16/sizeof(T) consecutive items starting at
// B + j*64*G + l*G (consecutive lanes read consecutive 16-byte values), so item
// k = j*G + g of lane l is nonzero B + j*64*G + l*G + g.
template <typename T, int S>
struct Tile {
    static constexpr int G = 16 / (int)sizeof(T);
    static constexpr int J = S / G;
    static_assert(S % G == 0 && S >= 4, "S must be a multiple of 16 / sizeof(T) and >= 4");
    int c[S];
    T v[S];
};

template <typename T, int S>
__device__ __forceinline__ int tile_item(int B, int lane, int k)
{
    constexpr int G = Tile<T, S>::G;
    return B + (k / G) * (kWave * G) + lane * G + (k % G);
}

template <typename T, int S>
__device__ __forceinline__ void tile_load(const int *__restrict__ col, const T *__restrict__ val,
                                          int B, int lane, Tile<T, S> &t)
{
    constexpr int G = Tile<T, S>::G, J = Tile<T, S>::J;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int e = B + j * (kWave * G) + lane * G;
        const u32x4 qv = load_nt16(val + e);
        __builtin_memcpy(&t.v[j * G], &qv, 16);
        if constexpr (G == 4) {
            const u32x4 qc = load_nt16(col + e);
            t.c[j * 4 + 0] = (int)qc.x;
            t.c[j * 4 + 1] = (int)qc.y;
            t.c[j * 4 + 2] = (int)qc.z;
            t.c[j * 4 + 3] = (int)qc.w;
        } else {
            const u32x2 qc = load_nt8(col + e);
            t.c[j * 2 + 0] = (int)qc.x;
            t.c[j * 2 + 1] = (int)qc.y;
        }
    }
}

// The same items with every index checked against nnz (items past it read
// as col 0 / val 0): the last tile of a matrix whose arrays are not padded.
template <typename T, int S>
__device__ __forceinline__ void tile_load_checked(const int *__restrict__ col,
                                                  const T *__restrict__ val, int B, int lane,
                                                  int nnz, Tile<T, S> &t)
{
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int e = tile_item<T, S>(B, lane, k);
        t.c[k] = e < nnz ? col[e] : 0;
        t.v[k] = e < nnz ? val[e] : T(0);
    }
}

// Row boundaries of the first 64*CP rows of a tile starting at row r (CP = 2
// or 4 chunk parts): lane l holds a[q] = row_ptr[r + 64q + l] for q < CP, and
// every lane holds end = row_ptr[r + 64*CP], each index clamped to m (row_ptr
// has m+1 entries, so every address is in bounds; callers ignore rows past the
// tile's last row).  A row's end, row_ptr[r + 64q + l + 1], is lane l+1's a[q]
// (lane 63: a[q+1] of lane 0, or end), so CP loads per lane cover 64*CP rows.
// Loaded one tile ahead of its use.  CP = 4 pays on matrices where many tiles
// span 128-255 rows (short rows), CP = 2 elsewhere; the host picks per matrix.
template <int CP>
struct RowChunk {
    int a[CP];
    int end;
};

template <int CP>
__device__ __forceinline__ RowChunk<CP> load_chunk(const int *__restrict__ row_ptr, int r,
                                                   int lane, int m)
{
    RowChunk<CP> c;
    static_for<CP>([&](auto qc) { c.a[qc] = row_ptr[min(r + qc * kWave + lane, m)]; });
    c.end = row_ptr[min(r + CP * kWave, m)];
    return c;
}

// row_ptr[r + 64Q + lane + 1] from a chunk loaded at row r.  Reads other
// lanes: call it with every lane of the wave active.
template <int Q, int CP>
__device__ __forceinline__ int chunk_next(const RowChunk<CP> &c, int lane)
{
    int first_after;
    if constexpr (Q + 1 < CP) first_after = read_lane(c.a[Q + 1], 0);
    else first_after = c.end;
    const int up = dpp<kDppWaveShl1>(c.a[Q], first_after);
    return lane == kWave - 1 ? first_after : up;
}

// Number of tiles p < NP whose row span tile_row[p+1] - tile_row[p] lies in
// [lo, hi), added to *count (one atomic per wave).
__global__ __launch_bounds__(256) void tile_span_count_kernel(const int *__restrict__ tile_row,
                                                              int NP, int lo, int hi,
                                                              unsigned long long *count)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (p < NP) {
        const int d = tile_row[p + 1] - tile_row[p];
        in = d >= lo && d < hi;
    }
    const uint64_t b = ballot(in);
    if (lane_id() == 0 && b) atomicAdd(count, (unsigned long long)__popcll(b));
}

// Compile-time switch between the tile code paths of segsum_wave.
template <bool B>
struct BigTag {
    static constexpr bool value = B;
};

// Given row_ptr[lo] == v (rows before lo, back to the run's start, are empty),
// the end (exclusive, at most stop+1) of the run of empty rows: the first row
// at or after lo with entries.  Wave-uniform doubling steps over row_ptr,
// closed by bisection, so a tile spanning a long run of empty rows costs
// O(log run) loads instead of one per row.
__device__ __forceinline__ int empty_run_end(const int *__restrict__ row_ptr, int lo, int v,
                                             int stop)
{
    int step = 64;
    while (lo + step <= stop + 1 && row_ptr[lo + step] == v) {
        lo += step;
        step *= 2;
    }
    int hi = min(lo + step, stop + 1);
    if (row_ptr[hi] == v) return stop + 1;   // empty up to the tile's last row
    while (hi - lo > 1) {                    // row_ptr[lo] == v < row_ptr[hi]
        const int mid = lo + (hi - lo) / 2;
        if (row_ptr[mid] == v) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Speculative segmented sum (kernels_cuda.h:7-269, wave64 form).
//
// Per tile [B, B+P) with r0 = tile_row[p], r1 = tile_row[p+1]:
//  * products val*x[col] (kernels_cuda.h:150-162): coalesced loads, the x
//    gathers all issued before any of them is waited on; the next tile's
//    col/val stream and the row_ptr values of its first 256 rows are in
//    flight while this tile is reduced, so no tile waits on a row_ptr load
//    unless it spans more than 256 rows.
//  * fast track (r0 == r1, the tile lies inside one row, :79-112): lane sums,
//    wave sum, added to the carried partial; nothing is stored until the row
//    closes.
//  * otherwise
//    - bit flags (:114-148): each lane takes rows r0+l and r0+64+l (further
//      rows, rare, in a loop) of [r0, min(r1, m-1)]; a nonempty
//      row starting at offset o in [0, P) sets bit o%S of lane o/S's word
//      (LDS atomic OR); an empty row marks the tile dirty and gets y = 0.
//    - the products go through LDS so lane l owns items [lS, lS+S) in order
//      (the reference's s_interres transpose, :150-162).
//    - carry-in (:164-172): if item 0 does not start a row it gets the
//      carried partial added; lane 0's item 0 always starts a segment.
//    - lane-sequential segmented reduction over the S items (:174-203);
//      segments that close inside a lane go straight to LDS s_seg[k].
//    - segment numbers: exclusive wave scan of per-lane head counts (:205-207).
//    - segments spanning lanes (:209-235): a segmented wave scan of the lane
//      tails replaces the reference's O(W) serial walk; the next lane with a
//      head (or lane 63 for the tile's last segment) closes the segment.
//    - speculation: segment k is row r0+k when the tile has no empty row; in
//      a dirty tile the fix-up of bhsparse_spmv_cuda.h:178-222 is done in the
//      wave: rows get their segment by rank among nonempty rows (ballot +
//      mbcnt), so no second pass over y and no host round trip.
//  * The first row of a wave, if it began in an earlier wave, is not stored:
//    its partial goes to sync_val[w] (the reference's synchronizer,
//    :98-102/:243-247) and calibrate_kernel adds it.  The wave in which a row
//    begins stores that row's partial with a plain store, so every call
//    overwrites y completely (beta = 0, repeatable), unlike the reference's
//    y[idx] += of stale values (bhsparse_spmv_cuda.h:170-173).
// ---------------------------------------------------------------------------
// One matrix as one wave sees it.  X is n x kstride and Y is m x kstride,
// row-major (a column's right-hand sides share cache lines); the wave computes
// k of them (k <= KMAX).  Rows' partials that cross waves go to
// sync_row[w] = key_base + row * kstride (the y slot of RHS 0 of the row) and
// sync_val[w * KMAX + rr].
template <typename T>
struct MatView {
    const int *row_ptr;
    const int *tile_row;
    const int *col;
    const T *val;
    const T *x;
    T *y;
    int *sync_row;
    T *sync_val;
    int m, nnz, NP, NW, lead_empty, k, kstride, TPW, key_base;
};

// x values of the tile's items for the right-hand side being reduced (v).
// One right-hand side: a gather per item.  Several (KMAX > 1): the gathers
// of the next right-hand side are issued before the current one is reduced,
// so their latency overlaps that work.
template <typename T, int S, int KMAX>
struct RhsX {
    T v[S];                                     // RHS rr
    T n0[S];                                    // KMAX > 1: RHS rr+1

    __device__ __forceinline__ void load1(const T *__restrict__ x, const Tile<T, S> &t, int ks,
                                          int rr, T (&d)[S])
    {
#pragma unroll
        for (int kk = 0; kk < S; ++kk) d[kk] = x[(long long)t.c[kk] * ks + rr];
    }
    __device__ __forceinline__ void first(const T *__restrict__ x, const Tile<T, S> &t, int ks)
    {
        load1(x, t, ks, 0, v);
    }
    // at the top of RHS rr: the loads of RHS rr+1
    __device__ __forceinline__ void issue(const T *__restrict__ x, const Tile<T, S> &t, int ks,
                                          int rr, int k)
    {
        if constexpr (KMAX > 1)
            if (rr + 1 < k) load1(x, t, ks, rr + 1, n0);
    }
    // at the end of RHS rr: v becomes RHS rr+1
    __device__ __forceinline__ void advance(int rr, int k)
    {
        if constexpr (KMAX > 1) {
            if (rr + 1 >= k) return;
#pragma unroll
            for (int kk = 0; kk < S; ++kk) v[kk] = n0[kk];
        }
    }
};

// Right-hand sides in pairs (k even, even stride, x offsets even): RHS rr and
// rr+1 of an item come from one aligned 2-wide load and are reduced together
// as 2-vectors, each component with exactly the single-RHS arithmetic; the
// next pair's loads are in flight while the current pair is reduced.
template <typename T, int S>
struct RhsPair {
    using V = typename RhsVec<T, 2>::type;
    V v[S], n[S];

    __device__ __forceinline__ void load(const T *__restrict__ x, const Tile<T, S> &t, int ks,
                                         int rr, V (&d)[S])
    {
#pragma unroll
        for (int kk = 0; kk < S; ++kk)
            d[kk] = *reinterpret_cast<const V *>(x + (long long)t.c[kk] * ks + rr);
    }
    __device__ __forceinline__ void first(const T *__restrict__ x, const Tile<T, S> &t, int ks)
    {
        load(x, t, ks, 0, v);
    }
    __device__ __forceinline__ void issue(const T *__restrict__ x, const Tile<T, S> &t, int ks,
                                          int rr, int k)
    {
        if (rr + 2 < k) load(x, t, ks, rr + 2, n);
    }
    __device__ __forceinline__ void advance(int rr, int k)
    {
        if (rr + 2 >= k) return;
#pragma unroll
        for (int kk = 0; kk < S; ++kk) v[kk] = n[kk];
    }
};

// Right-hand sides in pairs with the tile's x values staged in LDS (fp64,
// S = 4, k even): an item's k values are k*8 contiguous bytes of X (one
// 128-byte line at k = 16), cut into C = k/2 16-byte chunks, one pair each.
// Staged XCH chunks at a time (a pass), the chunks of an item come from one
// load instruction, on neighbouring lanes, so each line is one request to
// the cache per pass instead of one per pair; the pairs are then read from
// LDS.  (Loading the next pass into a second buffer while the current one is
// reduced measured slower: the extra LDS costs more resident waves than the
// overlap gains.)
//
// Layout of a pass (Cp = min(C, XCH) rounded up to a power of two, the same
// for every pass of a tile; M = 16/Cp): load instruction i in [0, 4*Cp)
// writes LDS bytes [1024 i, 1024 i + 1024), lane l at 16 l.  Lane l of
// instruction i loads chunk c of item t where, with s = l / Cp, q = l % Cp,
// r = s % M, w = s / M, e = 4 i + w, v = e / 2:
//     t = 2 (v M + r) + e % 2,    c = (q - v) mod Cp.
// A pair read has lane L take item t = 128 j + 2 L + g (fixed j, g) at one
// chunk c: lanes L = 0..15 then land in the 16 distinct 16-byte bank groups
// (r = L mod M, q = (c + L / M) mod Cp), so the reads are conflict-free.
template <int XCH>
struct XStage {
    unsigned char *buf;                            // 1024 * 4 * XCH bytes
    int *s_col;                                    // the tile's column indices
    int lc = 0, lm = 4, q = 0, w = 0, tb = 0;     // the tile's pass geometry, this lane

    // The tile's columns (lane L holds items 128j+2L+g) and the pass geometry:
    // the item this lane loads in instruction i is tb + i * 4M.
    __device__ __forceinline__ void begin_tile(const Tile<double, 4> &t, int lane, int C)
    {
        wave_sync();                               // earlier reads of s_col are done
        s_col[2 * lane] = t.c[0];
        s_col[2 * lane + 1] = t.c[1];
        s_col[128 + 2 * lane] = t.c[2];
        s_col[128 + 2 * lane + 1] = t.c[3];
        const int cw = C < XCH ? C : XCH;
        lc = cw <= 1 ? 0 : cw <= 2 ? 1 : cw <= 4 ? 2 : 3;
        lm = 4 - lc;
        const int s = lane >> lc;
        q = lane & ((1 << lc) - 1);
        w = s >> lm;
        tb = 2 * (((w >> 1) << lm) + (s & ((1 << lm) - 1))) + (w & 1);
        wave_sync();
    }
    // Issues the loads of chunks [c0, c0 + cg) of every item (pair() waits
    // for them when it first reads the pass).
    __device__ __forceinline__ void stage(const double *__restrict__ x, int ks, int c0, int cg)
    {
        const int Cm = (1 << lc) - 1;
        wave_sync();                               // earlier reads of buf are done
        const int n = 4 << lc;                     // a multiple of 4
#pragma unroll 4
        for (int i = 0; i < n; ++i) {
            const int c = (q - 2 * i - (w >> 1)) & Cm;
            if (c < cg) {
                const int cl = s_col[tb + (i << (lm + 2))];
                load_lds16(x + (long long)cl * ks + 2 * (c0 + c), buf + 1024 * i);
            }
        }
    }
    // The x values of pair rr (k even) for this lane's S = 4 items.  The
    // tile's first pass is issued at the top of the tile (so its loads are in
    // flight during the tile's flag and row work); later passes when their
    // first pair is reached.
    __device__ __forceinline__ void pair(const double *__restrict__ x, int ks, int k, int rr,
                                         int lane, double2v (&d)[4])
    {
        const int C = k >> 1, c0 = rr >> 1, cl = c0 % XCH;
        if (cl == 0) {
            if (c0 > 0) stage(x, ks, c0, min(C - c0, XCH));
            lds_dma_wait();
        }
        const int Cm = (1 << lc) - 1, Mm = (1 << lm) - 1;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int u = (kk >> 1) * 64 + lane;   // item t = 2u + (kk & 1)
            const int r = u & Mm, v = u >> lm;
            const int e = 2 * v + (kk & 1);
            const int s = r + ((e & 3) << lm);
            const int l = (s << lc) + ((cl + v) & Cm);
            d[kk] = *reinterpret_cast<const double2v *>(buf + 1024 * (e >> 2) + 16 * l);
        }
    }
};

template <int R>
struct RhsTag {
    static constexpr int value = R;
};

// BOUNDED: col/val hold exactly nnz items (no padding), so a tile reaching
// past nnz is loaded with checked indices (a wave-uniform branch).
// XCH > 0 (fp64, S = 4, KMAX > 1): right-hand-side pairs take their x values
// from the tile staged in LDS, XCH chunks per item at a time (XStage), at
// s_stage: 1 KB of column indices, then the staging buffer.
// NOEMPTY: the matrix has no empty rows (the stacked matrix of the column-panel
// layout), so no tile is dirty and the empty-row paths are left out.
template <typename T, int S, int KMAX, bool STRIDED, bool STATS, int CP, bool BOUNDED = false,
          int XCH = 0, bool NOEMPTY = false>
__device__ __forceinline__ void segsum_wave(const MatView<T> &mv, const int w, const int lane,
                                            unsigned *s_head, T *s_buf, T *s_carry, const T alpha,
                                            unsigned long long *stats,
                                            unsigned char *s_stage = nullptr, int touch = 0)
{
    constexpr int P = kWave * S;
    constexpr int G = Tile<T, S>::G;
    const int k = (KMAX == 1) ? 1 : mv.k;
    const int ks = STRIDED ? mv.kstride : 1;
    // right-hand sides in pairs: k even at an even stride, from x and y
    // offsets that keep the 2-wide loads and stores aligned; only where a
    // tile's pairs fit the registers (S * sizeof(T) <= 32 bytes per lane)
    constexpr bool kPairs = KMAX > 1 && S * sizeof(T) <= 32;
    const bool pairs = kPairs && (ks & 1) == 0 && (k & 1) == 0;
    constexpr bool kStage = XCH > 0 && kPairs && std::is_same<T, double>::value && S == 4;
    [[maybe_unused]] XStage<(XCH > 0 ? XCH : 1)> xs{kStage ? s_stage + 1024 : nullptr,
                                                   reinterpret_cast<int *>(s_stage)};
    const int *__restrict__ row_ptr = mv.row_ptr;
    const int *__restrict__ tile_row = mv.tile_row;
    const T *__restrict__ x = mv.x;
    T *__restrict__ y = mv.y;

    // Rows before the first nonzero belong to no tile: spread their zeroing.
    if (!NOEMPTY && mv.lead_empty > 0) {
        const long long lo = (long long)mv.lead_empty * w / mv.NW;
        const long long hi = (long long)mv.lead_empty * (w + 1) / mv.NW;
        for (long long r = lo + lane; r < hi; r += kWave)
            for (int rr = 0; rr < k; ++rr) y[r * ks + rr] = T(0);
    }

    const int p_begin = w * mv.TPW;
    const int p_end = min(p_begin + mv.TPW, mv.NP);
    const int E0 = p_begin * P;
    const int r_first = tile_row[p_begin];
    bool syncing = false;                       // first row began in an earlier wave
    bool did_sync = false;
    // The wave's one boundary-row partial (KMAX == 1) stays in registers and is
    // stored after the tile loop: a store inside the loop would make the
    // compiler wait for it (and, the counter being in order, for every load
    // issued since) before its data register is reused.
    int sync_key = -1;
    T sync_v1 = T(0);
    bool open = false;                          // a row continues past the last tile
    T carry1 = T(0);                            // the open row's partial (KMAX == 1)
    if (KMAX > 1) {
        if (lane < k) s_carry[lane] = T(0);
        wave_sync();
    }
    auto carry_get = [&](int rr) -> T {
        if constexpr (KMAX == 1) return carry1;
        else return s_carry[rr];
    };
    // the open row's partial of RHS rr .. rr+R-1 as a value (R = 1) or pair
    auto carry_getv = [&](auto rt, int rr) {
        constexpr int R = decltype(rt)::value;
        using V = typename RhsVec<T, R>::type;
        if constexpr (R == 1) return carry_get(rr);
        else return *reinterpret_cast<const V *>(s_carry + rr);
    };
    auto carry_setv = [&](auto rt, int rr, typename RhsVec<T, decltype(rt)::value>::type v) {
        constexpr int R = decltype(rt)::value;
        using V = typename RhsVec<T, R>::type;
        if constexpr (KMAX == 1) carry1 = v;
        else {
            if (lane == 0) *reinterpret_cast<V *>(s_carry + rr) = v;
            wave_sync();
        }
    };
    unsigned long long n_fast = 0, n_dirty = 0;

    // Row chunk of the next tile loaded one tile ahead.  The next tile's
    // stream is in flight while this one is reduced (kPrefetch) in fp32 and with
    // several right-hand sides.  In fp64 with one right-hand side a tile is
    // loaded at the top of its own iteration: there the prefetched tile holds
    // 24 more VGPRs, and the four-panel power-law tiles ran 0.4-1.9 us slower
    // with it (fp32 0.6 us faster, the plain and banded tiles alike).
    constexpr bool kPrefetch = KMAX > 1 || sizeof(T) < 8;
    Tile<T, S> cur;
    if (BOUNDED && E0 + P > mv.nnz) tile_load_checked<T, S>(mv.col, mv.val, E0, lane, mv.nnz, cur);
    else tile_load<T, S>(mv.col, mv.val, E0, lane, cur);
    constexpr int CR = CP * kWave;              // rows a chunk covers
    RowChunk<CP> rc = load_chunk<CP>(row_ptr, r_first, lane, mv.m);
    int r1 = r_first;
    // the caller's read-ahead of x (issued before the loads above) is
    // consumed here, so its register is free during the tile loop
    asm volatile("" ::"v"(touch));   // emits no instruction

    for (int p = p_begin; p < p_end; ++p) {
        const int B = p * P;
        const int r0 = r1;
        r1 = tile_row[p + 1];
        const bool general = r0 != r1;
        if constexpr (!kPrefetch) {
            if (p != p_begin) {
                if (BOUNDED && B + P > mv.nnz) tile_load_checked<T, S>(mv.col, mv.val, B, lane, mv.nnz, cur);
                else tile_load<T, S>(mv.col, mv.val, B, lane, cur);
            }
        }
        RhsX<T, S, KMAX> xr;
        [[maybe_unused]] RhsPair<T, S> xp;
        if constexpr (kStage) {
            if (pairs) {
                xs.begin_tile(cur, lane, k >> 1);
                xs.stage(x, ks, 0, min(k >> 1, XCH));   // waited for by the first pair
            } else {
                xr.first(x, cur, ks);
            }
        } else if constexpr (kPairs) {
            if (pairs) xp.first(x, cur, ks);
            else xr.first(x, cur, ks);
        } else {
            xr.first(x, cur, ks);
        }
        const RowChunk<CP> cc = rc;
        if (p + 1 < p_end) rc = load_chunk<CP>(row_ptr, r1, lane, mv.m);
        __builtin_amdgcn_sched_barrier(0);
        [[maybe_unused]] Tile<T, S> nxt;
        if (kPrefetch && p + 1 < p_end) {         // wave-uniform; the last tile loads nothing more
            const int Bn = B + P;
            if (BOUNDED && Bn + P > mv.nnz) tile_load_checked<T, S>(mv.col, mv.val, Bn, lane, mv.nnz, nxt);
            else tile_load<T, S>(mv.col, mv.val, Bn, lane, nxt);
        }

        const bool cont = read_lane(cc.a[0], 0) < B;   // item 0 continues row r0 (lane 0 holds row r0)
        if (p == p_begin) syncing = cont;
        const bool tail = B + P > mv.nnz;

        // products of the tile for RHS rr .. rr+R-1 (items past nnz masked)
        auto products = [&](auto rt, const auto &xv, auto (&prod)[S]) {
#pragma unroll
            for (int kk = 0; kk < S; ++kk) {
                const auto pk = cur.v[kk] * xv[kk];
                prod[kk] = tail ? keep_or_zero(pk, tile_item<T, S>(B, lane, kk) < mv.nnz) : pk;
            }
        };
        if (!general) {                          // fast track
            auto fast_rhs = [&](auto rt, int rr, const auto &xv) __attribute__((always_inline)) {
                using V = typename RhsVec<T, decltype(rt)::value>::type;
                V prod[S];
                products(rt, xv, prod);
                V s = prod[0];
#pragma unroll
                for (int kk = 1; kk < S; ++kk) s += prod[kk];
                s = wave_sum(s);
                carry_setv(rt, rr, cont ? carry_getv(rt, rr) + s : s);
            };
            bool done = false;
            if constexpr (kStage) {
                if (pairs) {
                    for (int rr = 0; rr < k; rr += 2) {
                        double2v xv[4];
                        xs.pair(x, ks, k, rr, lane, xv);
                        fast_rhs(RhsTag<2>{}, rr, xv);
                    }
                    done = true;
                }
            } else if constexpr (kPairs) {
                if (pairs) {
                    for (int rr = 0; rr < k; rr += 2) {
                        xp.issue(x, cur, ks, rr, k);
                        fast_rhs(RhsTag<2>{}, rr, xp.v);
                        xp.advance(rr, k);
                    }
                    done = true;
                }
            }
            if (!done)
                for (int rr = 0; rr < k; ++rr) {
                    xr.issue(x, cur, ks, rr, k);
                    fast_rhs(RhsTag<1>{}, rr, xr.v);
                    xr.advance(rr, k);
                }
            open = true;
            if (STATS) ++n_fast;
        } else {
            // A tile of fewer than CR rows takes every row boundary from
            // its chunk; only longer tiles load more row_ptr values, in
            // loops.  The two cases are separate instantiations, so the loops
            // (and the waits they need) are not on the common path at all.
            auto general_tile = [&](auto big) __attribute__((always_inline)) {
                constexpr bool BIG = decltype(big)::value;
                // ---- bit flags (once for all right-hand sides) ------------------
                const int stop = (r1 == mv.m) ? mv.m - 1 : r1;
                // nonempty rows of the chunk, one ballot per 64 rows (kept as
                // wave-wide masks, not per-lane registers, until the output)
                uint64_t mne[CP];
                bool lane_dirty = false;
                wave_sync();                         // previous tile's LDS reads are done
                s_head[lane] = 0u;
                wave_sync();
                // (empty rows of the chunk get their zero in the output step,
                // after this tile's loads have been consumed)
                static_for<CP>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    mne[q] = 0ull;
                    if (q > 0 && r0 + q * kWave > stop) return;   // wave-uniform: no rows here
                    // the shift reads the next lane: it runs on the whole
                    // wave, before any lane-dependent condition masks lanes off
                    const bool have = r0 + q * kWave + lane <= stop;
                    bool ne = have;
                    if constexpr (!NOEMPTY) ne = have && chunk_next<q>(cc, lane) > cc.a[q];
                    lane_dirty = lane_dirty || (have && !ne);
                    mne[q] = ballot(ne);
                    if (ne) {
                        const int o = cc.a[q] - B;
                        if (o >= 0 && o < P) atomicOr(&s_head[o / S], 1u << (o % S));
                    }
                });
                if constexpr (BIG)
                for (int rb = r0 + CR; rb <= stop;) {   // >= CR rows
                    const int rq = min(rb + lane, stop);
                    const int rp = row_ptr[rq];
                    if constexpr (NOEMPTY) {
                        if (rb + lane <= stop) {
                            const int o = rp - B;
                            if (o >= 0 && o < P) atomicOr(&s_head[o / S], 1u << (o % S));
                        }
                        rb += kWave;
                        continue;
                    }
                    const int rpn = row_ptr[rq + 1];
                    const int vb = row_ptr[min(rb + kWave, stop + 1)];
                    if (rb + kWave <= stop && read_lane(rp, 0) == vb) {
                        // 64 empty rows with more rows after them: skip the whole run
                        const int re = empty_run_end(row_ptr, rb + kWave, vb, stop);
                        lane_dirty = true;
                        for (int r = rb + lane; r < re; r += kWave)
                            for (int rr = 0; rr < k; ++rr) y[(long long)r * ks + rr] = T(0);
                        rb = re;
                        continue;
                    }
                    if (rb + lane <= stop) {
                        if (rpn > rp) {
                            const int o = rp - B;
                            if (o >= 0 && o < P) atomicOr(&s_head[o / S], 1u << (o % S));
                        } else {
                            lane_dirty = true;
                            for (int rr = 0; rr < k; ++rr) y[(long long)(rb + lane) * ks + rr] = T(0);
                        }
                    }
                    rb += kWave;
                }
                const bool dirty = !NOEMPTY && ballot(lane_dirty) != 0ull;
                wave_sync();
                unsigned hb = s_head[lane];
                if (lane == 0) hb |= 1u;

                // ---- segment numbers -----------------------------------------
                const int nh = __popc(hb);
                const int incl = wave_incl_scan(nh, lane);
                const int seg_base = incl - nh;
                const int K = read_lane(incl, kWave - 1);
                // row_ptr[r1] sits in the chunk when r1 < r0+CR
                int rp_r1;
                if (!BIG || r1 - r0 < CR) {
                    // read the lane out of every part, then pick one: a
                    // select among registers by a run-time index would be
                    // compiled as an indexed load from scratch memory
                    const int d = r1 - r0, q = d / kWave, l = d % kWave;
                    rp_r1 = read_lane(cc.a[0], l);
                    static_for<CP>([&](auto qc) {
                        if constexpr (decltype(qc)::value > 0) {
                            const int v = read_lane(cc.a[qc], l);
                            if (q == qc) rp_r1 = v;
                        }
                    });
                } else {
                    rp_r1 = row_ptr[min(r1, mv.m)];
                }
                const bool partial_tail = (r1 < mv.m) && (rp_r1 < B + P);
                const int kb = syncing ? 1 : 0;
                const int ke = partial_tail ? K - 1 : K;

                if (STATS && dirty) ++n_dirty;

                auto general_rhs = [&](auto rt, int rr, const auto &xv) __attribute__((always_inline)) {
                    constexpr int R = decltype(rt)::value;
                    using V = typename RhsVec<T, R>::type;
                    V *sb = reinterpret_cast<V *>(s_buf);
                    const V zero = static_cast<V>(T(0));
                    auto store_y = [&](long long row, V val) {
                        *reinterpret_cast<V *>(y + row * ks + rr) = val;
                    };
                    V prod[S];
                    products(rt, xv, prod);
                    // products -> LDS in tile order, back lane-contiguous
                    wave_sync();
    #pragma unroll
                    for (int j = 0; j < S / G; ++j) {
                        V *dst = sb + j * (kWave * G) + lane * G;
    #pragma unroll
                        for (int g = 0; g < G; ++g) dst[g] = prod[j * G + g];
                    }
                    wave_sync();
                    V pc[S];
    #pragma unroll
                    for (int kk = 0; kk < S; ++kk) pc[kk] = sb[lane * S + kk];
                    wave_sync();
                    if (lane == 0 && cont) pc[0] = pc[0] + carry_getv(rt, rr);

                    // ---- lane-sequential segmented reduction -----------------
                    V acc = zero, lead = zero;
                    bool seen = false;
                    int sidx = seg_base;
    #pragma unroll
                    for (int kk = 0; kk < S; ++kk) {
                        if ((hb >> kk) & 1u) {
                            if (seen) { sb[sidx] = acc; ++sidx; }
                            else { lead = acc; seen = true; }
                            acc = zero;
                        }
                        acc += pc[kk];
                    }

                    // ---- segments spanning lanes -----------------------------
                    const uint64_t present = ballot(seen);  // bit 0 always set
                    const V I = wave_segscan(acc, present, lane);
                    const V Iprev = dpp<kDppWaveShr1>(I);     // lane i <- lane i-1 (lane 0: unused)
                    if (seen && lane > 0) sb[seg_base - 1] = Iprev + lead;
                    if (lane == kWave - 1) sb[K - 1] = I;
                    wave_sync();

                    // ---- output ------------------------------------------------
                    if (!dirty) {
                        for (int kk = kb + lane; kk < ke; kk += kWave)
                            store_y(r0 + kk, alpha * sb[kk]);
                    } else {
                        // a nonempty row's segment is its rank among the
                        // tile's nonempty rows; an empty row gets 0
                        int running = 0;
                        static_for<CP>([&](auto qc) {
                            constexpr int q = decltype(qc)::value;
                            if (q > 0 && r0 + q * kWave > stop) return;   // wave-uniform
                            const int row = r0 + q * kWave + lane;
                            const bool ne = (mne[q] >> lane) & 1ull;
                            const int rank = running + popc_below(mne[q]);
                            if (ne && rank >= kb && rank < ke)
                                store_y(row, alpha * sb[rank]);
                            else if (row <= stop && !ne)
                                store_y(row, zero);
                            running += __popcll(mne[q]);
                        });
                        if constexpr (BIG)
                        for (int rb = r0 + CR; rb <= stop;) {   // >= CR rows
                            const int rq = min(rb + lane, stop);
                            const int rp = row_ptr[rq], rpn = row_ptr[rq + 1];
                            const int vb = row_ptr[min(rb + kWave, stop + 1)];
                            if (rb + kWave <= stop && read_lane(rp, 0) == vb) {
                                rb = empty_run_end(row_ptr, rb + kWave, vb, stop);   // no segments
                                continue;
                            }
                            const int r = rb + lane;
                            const bool ne = r <= stop && rpn > rp;
                            const uint64_t msk = ballot(ne);
                            const int rank = running + popc_below(msk);
                            if (ne && rank >= kb && rank < ke)
                                store_y(r, alpha * sb[rank]);
                            running += __popcll(msk);
                            rb += kWave;
                        }
                    }
                    if (syncing) {
                        if constexpr (KMAX == 1) sync_v1 = alpha * s_buf[0];
                        else if (lane == 0)
                            *reinterpret_cast<V *>(mv.sync_val + (long long)w * KMAX + rr) =
                                alpha * sb[0];
                    }
                    carry_setv(rt, rr, sb[K - 1]);
                };
                bool done = false;
                if constexpr (kStage) {
                    if (pairs) {
                        for (int rr = 0; rr < k; rr += 2) {
                            double2v xv[4];
                            xs.pair(x, ks, k, rr, lane, xv);
                            general_rhs(RhsTag<2>{}, rr, xv);
                        }
                        done = true;
                    }
                } else if constexpr (kPairs) {
                    if (pairs) {
                        for (int rr = 0; rr < k; rr += 2) {
                            xp.issue(x, cur, ks, rr, k);
                            general_rhs(RhsTag<2>{}, rr, xp.v);
                            xp.advance(rr, k);
                        }
                        done = true;
                    }
                }
                if (!done)
                    for (int rr = 0; rr < k; ++rr) {
                        xr.issue(x, cur, ks, rr, k);
                        general_rhs(RhsTag<1>{}, rr, xr.v);
                        xr.advance(rr, k);
                    }
                if (syncing) {
                    sync_key = mv.key_base + r0 * ks;
                    syncing = false;
                    did_sync = true;
                }
                open = partial_tail;
                // the long-tile path ends with nothing in flight, so registers
                // its row_ptr loads used carry no pending write into the next
                // tile (which would otherwise wait for every load in flight)
                if constexpr (BIG) __builtin_amdgcn_s_waitcnt(0);
            };
            if (r1 - r0 < CR) general_tile(BigTag<false>{});
            else general_tile(BigTag<true>{});
        }
        if constexpr (kPrefetch) cur = nxt;
    }

    if (open) {
        const int r_open = tile_row[p_end];
        if (syncing) {
            sync_key = mv.key_base + r_open * ks;
            if constexpr (KMAX == 1) sync_v1 = alpha * carry_get(0);
            else if (lane == 0)
                for (int rr = 0; rr < k; ++rr)
                    mv.sync_val[(long long)w * KMAX + rr] = alpha * carry_get(rr);
            did_sync = true;
        } else if (lane == 0) {
            for (int rr = 0; rr < k; ++rr) y[(long long)r_open * ks + rr] = alpha * carry_get(rr);
        }
    }
    if (lane == 0) {
        mv.sync_row[w] = did_sync ? sync_key : -1;
        if (KMAX == 1 && did_sync) mv.sync_val[w] = sync_v1;
    }
    if (STATS && lane == 0) {
        atomicAdd(&stats[0], n_fast);
        atomicAdd(&stats[1], n_dirty);
        atomicAdd(&stats[2], did_sync ? 1ull : 0ull);
    }
}

// q.lead_empty < 0: read the number of leading empty rows from tile_row[0]
// (the first tile's row), for launches that have not copied it to the host.
template <typename T, int S, bool STATS, int CP, bool BOUNDED = false, bool NOEMPTY = false>
__global__ __launch_bounds__(512) void segsum_kernel(
    const int *__restrict__ row_ptr, const int *__restrict__ col, const T *__restrict__ val,
    const T *__restrict__ x, T *__restrict__ y, const int *__restrict__ tile_row,
    int *__restrict__ sync_row, T *__restrict__ sync_val, unsigned long long *__restrict__ stats,
    const SegsumParams<T> q)
{
    extern __shared__ double smem_raw[];
    const int lane = lane_id();
    // q.x_lines > 0 (x small enough for every L2): the blocks of each XCD
    // (dealt round-robin, block b on XCD b % 8) read one word of every
    // 128-byte line of x between them, so the first tile's gathers hit L2.
    int touch = 0;
    {
        const long long t = (long long)(blockIdx.x / 8) * blockDim.x + threadIdx.x;
        if (t < q.x_lines) touch = reinterpret_cast<const int *>(x)[t * (128 / sizeof(int))];
    }
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int blk = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int w = blk * q.wpb + wib;
    if (w >= q.NW) return;
    unsigned char *wave_lds = reinterpret_cast<unsigned char *>(smem_raw) + wib * q.wave_lds_bytes;
    const int lead_empty = q.lead_empty >= 0 ? q.lead_empty : tile_row[0];
    MatView<T> mv{row_ptr, tile_row, col, val, x, y, sync_row, sync_val,
                  q.m, q.nnz, q.NP, q.NW, lead_empty, 1, 1, q.TPW, 0};
    segsum_wave<T, S, 1, false, STATS, CP, BOUNDED, 0, NOEMPTY>(mv, w, lane, reinterpret_cast<unsigned *>(wave_lds),
                                       reinterpret_cast<T *>(wave_lds + 256), nullptr, q.alpha,
                                       stats, nullptr, touch);
}

// ---------------------------------------------------------------------------
// Data-movement floor of a matrix: the main kernel's loads and gathers (same
// coalesced tile loads, same x gathers) with no segmentation and no y -- one
// wave per tile, one partial per lane.  Its time bounds what any CSR SpMV of
// the matrix can reach on the device.
// ---------------------------------------------------------------------------
template <typename T, int S>
__global__ __launch_bounds__(256) void floor_kernel(const int *__restrict__ col,
                                                   const T *__restrict__ val,
                                                   const T *__restrict__ x, int NP,
                                                   T *__restrict__ out)
{
    constexpr int P = kWave * S;
    const int lane = lane_id();
    const int p = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (p >= NP) return;
    Tile<T, S> t;
    tile_load<T, S>(col, val, p * P, lane, t);
    T xv[S];
#pragma unroll
    for (int k = 0; k < S; ++k) xv[k] = x[t.c[k]];
    T s = t.v[0] * xv[0];
#pragma unroll
    for (int k = 1; k < S; ++k) s += t.v[k] * xv[k];
    out[(long long)p * kWave + lane] = s;
}

// ---------------------------------------------------------------------------
// Calibration (bhsparse_spmv_cuda.h:166-173 moved to the device).  Every wave
// w wrote sync_row[w] (-1: none) and sync_val[w]; consecutive waves with the
// same row form a chain whose partials must be added, in wave order, to the
// value the row's starting wave stored.  One wave handles 64 consecutive
// entries: a segmented scan over runs of equal row gives each run's sum, the
// lane ending a run adds it to y; a run reaching past the window is extended
// window by window (wave sums).  Runs continuing from the previous window
// belong to the wave of that window.
// ---------------------------------------------------------------------------
// One window of 64 sync entries [base, base+64) for right-hand side rr.
// key(e) = sync_row[e] if entry e holds a partial of RHS rr, else -1.
template <typename T, int KMAX>
__device__ __forceinline__ void calibrate_window(const int *__restrict__ sync_row,
                                                 const T *__restrict__ sync_val,
                                                 const int *__restrict__ sync_k, int NW, int base,
                                                 int rr, int lane, T *__restrict__ y)
{
    auto key = [&](int e) -> int {
        if (e < 0 || e >= NW) return -1;
        const int r = sync_row[e];
        if (KMAX > 1 && r >= 0 && rr >= sync_k[e]) return -1;
        return r;
    };
    const int e = base + lane;
    const int r = key(e);
    const T v = r >= 0 ? sync_val[(long long)e * KMAX + rr] : T(0);
    const int before = key(base - 1);
    int rprev = dpp<kDppWaveShr1>(r);                 // lane i <- lane i-1
    if (lane == 0) rprev = before;
    const uint64_t starts = ballot(r != rprev) | 1ull;
    const T s = wave_segscan(v, starts, lane);
    const int rnext = dpp<kDppWaveShl1>(r);           // lane i <- lane i+1 (lane 63: unused)
    const bool run_end = (lane == kWave - 1) || (rnext != r);
    // the run holding lane 0 continues an earlier window's run?
    const int r0w = read_lane(r, 0);
    const bool lead_cont = r0w >= 0 && before == r0w;
    // run of lane 63 continuing into later windows
    const int r63 = read_lane(r, kWave - 1);
    T ext = T(0);
    if (r63 >= 0) {
        for (int nb = base + kWave; nb < NW; nb += kWave) {
            const int e2 = nb + lane;
            const int r2 = key(e2);
            const uint64_t same = ballot(r2 == r63);
            const int len = (same == ~0ull) ? kWave : __builtin_ctzll(~same);
            if (len == 0) break;
            const T v2 = lane < len ? sync_val[(long long)e2 * KMAX + rr] : T(0);
            ext = ext + wave_sum(v2);
            if (len < kWave) break;
        }
    }
    // lane's run starts at the last set bit of `starts` at or below it
    const uint64_t upto = starts & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    const int run_start = 63 - __builtin_clzll(upto);
    const bool skip = (run_start == 0) && lead_cont;
    if (run_end && r >= 0 && !skip) {
        const T tot = (lane == kWave - 1) ? s + ext : s;
        y[(long long)r + rr] = y[(long long)r + rr] + tot;
    }
}

// ---------------------------------------------------------------------------
// Calibration (bhsparse_spmv_cuda.h:166-173 moved to the device).  Every wave
// w wrote sync_row[w] (-1: none) and sync_val[w]; consecutive waves with the
// same row form a chain whose partials must be added, in wave order, to the
// value the row's starting wave stored.  One wave handles 64 consecutive
// entries: a segmented scan over runs of equal row gives each run's sum, the
// lane ending a run adds it to y; a run reaching past the window is extended
// window by window (wave sums).  Runs continuing from the previous window
// belong to the wave of that window.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void calibrate_kernel(const int *__restrict__ sync_row,
                                                       const T *__restrict__ sync_val, int NW,
                                                       T *__restrict__ y)
{
    const int lane = lane_id();
    const int base = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * kWave;
    if (base >= NW) return;
    calibrate_window<T, 1>(sync_row, sync_val, nullptr, NW, base, 0, lane, y);
}

// ---------------------------------------------------------------------------
// Batched solves: many independent CSR matrices with up to KMAX right-hand
// sides each, in one persistent launch.  Every wave takes (matrix, wave-range)
// items from a global queue (one atomic per item) until the queue is empty --
// each wave exits once the counter passes n_items, so the grid always drains.
// The per-tile arithmetic is segsum_wave's, so a matrix/RHS pair gets exactly
// the bits a plain launch with the same tile shape gives.
// ---------------------------------------------------------------------------
struct BatchMeta {
    long long rp_off, tr_off, el_off, x_off, y_off;
    int sync_off, m, nnz, NP, NW, lead_empty, k, pad;
};

template <typename T, int S, int KMAX, int XCH>
__global__ __launch_bounds__(512) void batched_kernel(
    const int *__restrict__ row_ptr, const int *__restrict__ col, const T *__restrict__ val,
    const T *__restrict__ x, T *__restrict__ y, const int *__restrict__ tile_row,
    int *__restrict__ sync_row, T *__restrict__ sync_val, int *__restrict__ sync_k,
    const BatchMeta *__restrict__ meta, const int2 *__restrict__ items, int n_items,
    int *__restrict__ queue, int TPW, T alpha, int wave_lds_bytes)
{
    extern __shared__ double smem_raw[];
    const int lane = lane_id();
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    unsigned char *wave_lds = reinterpret_cast<unsigned char *>(smem_raw) + wib * wave_lds_bytes;
    unsigned *s_head = reinterpret_cast<unsigned *>(wave_lds);
    T *s_buf = reinterpret_cast<T *>(wave_lds + 256);   // room for 64*S pairs
    T *s_carry = s_buf + 2 * kWave * S;
    // XCH > 0: column indices and the x staging buffer after the carry
    unsigned char *s_stage =
        wave_lds + (256 + 2 * kWave * S * sizeof(T) + KMAX * sizeof(T) + 15) / 16 * 16;
    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(queue, 1);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= n_items) break;
        const int2 iw = items[it];
        const BatchMeta md = meta[iw.x];
        MatView<T> mv{row_ptr + md.rp_off, tile_row + md.tr_off, col + md.el_off,
                      val + md.el_off, x + md.x_off, y + md.y_off,
                      sync_row + md.sync_off, sync_val + (long long)md.sync_off * KMAX,
                      md.m, md.nnz, md.NP, md.NW, md.lead_empty, md.k, md.k, TPW,
                      (int)md.y_off};
        segsum_wave<T, S, KMAX, true, false, 2, false, XCH>(mv, iw.y, lane, s_head, s_buf,
                                                           s_carry, alpha, nullptr, s_stage);
        if (lane == 0) sync_k[md.sync_off + iw.y] = md.k;
        wave_sync();                              // LDS is reused by the next item
    }
}

// Calibration of a batch: sync entries of each matrix start on a multiple of
// 64, so every window lies inside one matrix and matches the windows of that
// matrix's own plain calibration; keys are y slots (unique across the batch).
template <typename T, int KMAX>
__global__ __launch_bounds__(256) void calibrate_batched_kernel(const int *__restrict__ sync_row,
                                                               const T *__restrict__ sync_val,
                                                               const int *__restrict__ sync_k,
                                                               int NW, T *__restrict__ y)
{
    const int lane = lane_id();
    const int base = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * kWave;
    if (base >= NW) return;
    const int e = base + lane;
    int kw = (e < NW && sync_row[e] >= 0) ? sync_k[e] : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(kw, off, kWave);
        kw = o > kw ? o : kw;
    }
    kw = __builtin_amdgcn_readfirstlane(kw);
    for (int rr = 0; rr < kw; ++rr)
        calibrate_window<T, KMAX>(sync_row, sync_val, sync_k, NW, base, rr, lane, y);
}

// y[i] = sum over panels p (in order) of row i's panel partial, +0 where row i
// has no entry in panel p.  The last step of the column-panel variant; every
// partial already carries alpha.  A wave covers GROUPS groups of 64 rows; for
// group g, bit p of mask[i] says row i has entries in panel p, and that
// partial is stacked row base[g*NPAN+p] + (rows of the group below i with bit
// p set), a ballot and a popcount.  Reads m bytes of mask instead of an
// NPAN x m index map.  All partials of the wave's rows are loaded before any
// is added, so each wave has GROUPS*NPAN loads in flight.
//
// The calibration of the stacked matrix is done here too (no calibrate_kernel
// launch on this layout).  A stacked row whose nonzeros cross wave boundaries
// has a chain of consecutive waves that each left a partial in sync_val; the
// chains depend only on the stacked matrix and the tile shape, so they are
// listed once per analysis as ChainRec records: combine wave cw has
// kChainSlots slots chains[cw * kChainSlots + j] (count 0: empty), filled in
// order, and the last slot's `next` links any further records.  The slots are
// loaded with the masks, and the first 64 partials of every slot's chain with
// the panel partials, so that a combine wave's chains add no round trip to
// memory unless it has more than kChainSlots of them or one of more than 64
// waves.  A chain's total is summed 64 entries at a time (wave_sum of each
// chunk, the chunk sums in order) and added to that partial before the panel
// sum: y_stack[R] + total, the value calibrate_kernel would have stored.  A
// stacked row has at most one chain, so chains never meet in one partial.
struct ChainRec {
    int first;   // first wave of the chain (its sync_val index)
    int count;   // waves in the chain (0: no record)
    int where;   // (row - first row of the combine wave) << 4 | panel
    int next;    // index of the next record of the same combine wave, -1: none
};
constexpr int kCombineGroups = 2;   // 64-row groups per combine wave
constexpr int kChainSlots = 4;      // chain records per combine wave loaded up front

template <typename T, int NPAN, int GROUPS>
__global__ __launch_bounds__(256) void combine_panels_kernel(const T *__restrict__ ystack,
                                                            const unsigned char *__restrict__ mask,
                                                            const int *__restrict__ base, int m,
                                                            T *__restrict__ y,
                                                            const ChainRec *__restrict__ chains,
                                                            const T *__restrict__ sync_val)
{
    const int lane = lane_id();
    const int cw = __builtin_amdgcn_readfirstlane(
        (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const long long g0 = (long long)cw * GROUPS;
    if (g0 * kWave >= m) return;                      // whole wave past the last row
    // chain slots, masks and bases first, so that no wait for them also waits
    // for partials already in flight
    ChainRec slot{0, 0, 0, -1};
    if (lane < kChainSlots) slot = chains[(long long)cw * kChainSlots + lane];
    unsigned mk[GROUPS];
    int gb[GROUPS][NPAN];
#pragma unroll
    for (int h = 0; h < GROUPS; ++h) {
        const long long g = g0 + h;
        const long long i = g * kWave + lane;
        mk[h] = i < m ? (unsigned)mask[i] : 0u;          // no bits past the last row
        const int *b = base + (g * kWave < m ? g : g0) * NPAN;
#pragma unroll
        for (int p = 0; p < NPAN; ++p) gb[h][p] = b[p];
    }
    T v[GROUPS][NPAN];
#pragma unroll
    for (int h = 0; h < GROUPS; ++h) {
#pragma unroll
        for (int p = 0; p < NPAN; ++p) {
            const bool has = (mk[h] >> p) & 1u;
            const uint64_t bits = ballot(has);         // every lane of the wave is active
            const int pos = has ? gb[h][p] + popc_below(bits) : 0;
            v[h][p] = has ? ystack[pos] : T(0);
        }
    }
    // the first chunk of every slot's chain, in flight with the partials
    int c_first[kChainSlots], c_count[kChainSlots], c_where[kChainSlots];
    T c0[kChainSlots];
#pragma unroll
    for (int j = 0; j < kChainSlots; ++j) {
        c_first[j] = read_lane(slot.first, j);
        c_count[j] = read_lane(slot.count, j);
        c_where[j] = read_lane(slot.where, j);
        c0[j] = lane < c_count[j] ? sync_val[(long long)c_first[j] + lane] : T(0);
    }
    auto add_chain = [&](int first, int count, int where, T s) {
        T tot = wave_sum(s);
        for (int j = kWave; j < count; j += kWave)
            tot = tot + wave_sum(j + lane < count ? sync_val[(long long)first + j + lane] : T(0));
        const int row = where >> 4, pan = where & 15;
#pragma unroll
        for (int h = 0; h < GROUPS; ++h)
#pragma unroll
            for (int p = 0; p < NPAN; ++p)
                if (pan == p && row == h * kWave + lane) v[h][p] = v[h][p] + tot;
    };
#pragma unroll
    for (int j = 0; j < kChainSlots; ++j)
        if (c_count[j] > 0) add_chain(c_first[j], c_count[j], c_where[j], c0[j]);
    for (int nx = read_lane(slot.next, kChainSlots - 1); nx >= 0;) {   // wave-uniform
        const ChainRec c = chains[nx];
        add_chain(c.first, c.count, c.where,
                  lane < c.count ? sync_val[(long long)c.first + lane] : T(0));
        nx = c.next;
    }
#pragma unroll
    for (int h = 0; h < GROUPS; ++h) {
        const long long i = (g0 + h) * kWave + lane;
        T s = v[h][0];
#pragma unroll
        for (int p = 1; p < NPAN; ++p) s += v[h][p];    // panel order
        if (i < m) y[i] = s;
    }
}

// ---------------------------------------------------------------------------
// Device build of the column-panel copy (count -> scan -> scatter).  Panel of
// column c: c / pw.  The stacked matrix has one row per nonempty (panel, row)
// pair, panel-major; within a stacked row the entries keep their order in A.
// ---------------------------------------------------------------------------
// cnt[p*m + i] = entries of row i in panel p; mask[i] bit p = (cnt > 0).  One
// wave per row, 64 entries at a time (a row may hold 10^5 entries).
template <int NPAN>
__global__ __launch_bounds__(256) void panel_count_kernel(const int *__restrict__ row_ptr,
                                                         const int *__restrict__ col, int m,
                                                         int pw, int *__restrict__ cnt,
                                                         unsigned char *__restrict__ mask)
{
    const int lane = lane_id();
    const long long i = __builtin_amdgcn_readfirstlane(
        (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    if (i >= m) return;
    const int b = row_ptr[i], e = row_ptr[i + 1];
    int c[NPAN];
#pragma unroll
    for (int p = 0; p < NPAN; ++p) c[p] = 0;
    for (int j = b; j < e; j += kWave) {
        const int q = j + lane < e ? col[j + lane] / pw : -1;
#pragma unroll
        for (int p = 0; p < NPAN; ++p) c[p] += __popcll(ballot(q == p));
    }
    if (lane == 0) {
        unsigned mk = 0;
#pragma unroll
        for (int p = 0; p < NPAN; ++p) {
            cnt[(long long)p * m + i] = c[p];
            mk |= (c[p] > 0 ? 1u : 0u) << p;
        }
        mask[i] = (unsigned char)mk;
    }
}

// With flag[q] = (cnt[q] > 0) and rows[q] / nzs[q] the exclusive sums of the
// flags / counts over q = p*m + i: the stacked row of a nonempty pair is
// rows[q]; its start is nzs[q]; base[g*NPAN + p] = rows[p*m + 64g] is the first
// stacked row of panel p at or after row 64g (read only for groups that have
// rows in panel p); owner[R] = the row of A that stacked row R belongs to.
template <int NPAN>
__global__ __launch_bounds__(256) void panel_index_kernel(const int *__restrict__ cnt,
                                                         const int *__restrict__ rows,
                                                         const int *__restrict__ nzs, int m,
                                                         int *__restrict__ prow_ptr,
                                                         int *__restrict__ owner,
                                                         int *__restrict__ base)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)NPAN * m) return;
    const int p = (int)(q / m), i = (int)(q % m);
    if (cnt[q] > 0) {
        prow_ptr[rows[q]] = nzs[q];
        owner[rows[q]] = i;
    }
    if ((i & 63) == 0) base[(long long)(i >> 6) * NPAN + p] = rows[q];
}

// Entries of row i to their stacked rows, in order: one wave per row, 64
// entries at a time; an entry's place is its panel's running offset plus the
// number of lower lanes in the same panel (ballot + mbcnt).
template <typename T, int NPAN>
__global__ __launch_bounds__(256) void panel_scatter_kernel(
    const int *__restrict__ row_ptr, const int *__restrict__ col, const T *__restrict__ val,
    int m, int pw, const int *__restrict__ cnt, const int *__restrict__ nzs,
    int *__restrict__ pcol, T *__restrict__ pval)
{
    const int lane = lane_id();
    const long long i = __builtin_amdgcn_readfirstlane(
        (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    if (i >= m) return;
    const int b = row_ptr[i], e = row_ptr[i + 1];
    int off[NPAN];
#pragma unroll
    for (int p = 0; p < NPAN; ++p) off[p] = nzs[(long long)p * m + i];
    for (int j = b; j < e; j += kWave) {
        const bool in = j + lane < e;
        const int c = in ? col[j + lane] : 0;
        const T v = in ? val[j + lane] : T(0);
        const int q = in ? c / pw : -1;
        int d = 0;
#pragma unroll
        for (int p = 0; p < NPAN; ++p) {
            const uint64_t bits = ballot(q == p);
            if (q == p) d = off[p] + popc_below(bits);
            off[p] += __popcll(bits);
        }
        if (in) {
            pcol[d] = c;
            pval[d] = v;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T *__restrict__ p, long long n, T v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Writes `n16` 16-byte words; used to push earlier data out of L2 and MALL.
__global__ __launch_bounds__(256) void flush_kernel(u32x4 *__restrict__ p, long long n16, unsigned v)
{
    const u32x4 z = {v, v, v, v};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = z;
}

// Plain copy of n16 16-byte words (bandwidth probe and counter calibration).
__global__ __launch_bounds__(256) void copy_kernel(const u32x4 *__restrict__ src,
                                                  u32x4 *__restrict__ dst, long long n16)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = __builtin_nontemporal_load(src + i);
}

// ---------------------------------------------------------------------------
// csr-vector (kernels_cuda.h:272-342, CUSP v0.4.0; wave64 form of
// csr_vector_spmv_kernels.cl:10-78): TPV lanes per row, lane-strided partial
// sums, then a shuffle tree replacing the LDS tree; rows longer than 64 with
// TPV = 64 start at an aligned index (:307-320).
// ---------------------------------------------------------------------------
template <typename T, int TPV>
__global__ __launch_bounds__(256) void csrvec_kernel(int m, const int *__restrict__ row_ptr,
                                                    const int *__restrict__ col,
                                                    const T *__restrict__ val,
                                                    const T *__restrict__ x, T *__restrict__ y,
                                                    T alpha, long long x_lines)
{
    const int tl = threadIdx.x & (TPV - 1);
    const long long vid = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / TPV;
    const long long nv = (long long)gridDim.x * blockDim.x / TPV;
    // x_lines > 0 (x small enough for every L2): the blocks of each XCD (dealt
    // round-robin, block b on XCD b % 8) read one word of every 128-byte line
    // of x between them, so x is on its way into that XCD's L2 while row_ptr
    // and the entries are loaded, and the gathers after them hit it.
    int touch = 0;
    {
        const long long t = (long long)(blockIdx.x / 8) * blockDim.x + threadIdx.x;
        if (t < x_lines) touch = reinterpret_cast<const int *>(x)[t * (128 / sizeof(int))];
    }
    for (long long row = vid; row < m; row += nv) {
        const int rs = row_ptr[row], re = row_ptr[row + 1];
        T sum = T(0);
        if (TPV == kWave && re - rs > kWave) {
            int jj = rs - (rs & (TPV - 1)) + tl;
            if (jj >= rs && jj < re) sum += val[jj] * x[col[jj]];
            for (jj += TPV; jj < re; jj += TPV) sum += val[jj] * x[col[jj]];
        } else {
            for (int jj = rs + tl; jj < re; jj += TPV) sum += val[jj] * x[col[jj]];
        }
#pragma unroll
        for (int off = TPV / 2; off > 0; off >>= 1) sum += __shfl_down(sum, off, TPV);
        if (tl == 0) y[row] = alpha * sum;
    }
    asm volatile("" ::"v"(touch));   // keeps the touch load; emits no instruction
}

// ---------------------------------------------------------------------------
// Row-length statistics for the tuner: sums of len and len^2, max len, empty
// rows and the summed column span (last - first + 1) of nonempty rows, as
// exact 64-bit integers (one atomic per block).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void row_stats_kernel(int m, const int *__restrict__ row_ptr,
                                                       const int *__restrict__ col,
                                                       unsigned long long *__restrict__ out)
{
    unsigned long long s1 = 0, s2 = 0, se = 0, ssp = 0;
    unsigned mx = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (long long)gridDim.x * blockDim.x) {
        const int b = row_ptr[i], e = row_ptr[i + 1];
        const unsigned L = (unsigned)(e - b);
        s1 += L;
        s2 += (unsigned long long)L * L;
        mx = L > mx ? L : mx;
        if (L == 0) se++;
        else ssp += (unsigned long long)(col[e - 1] - col[b] + 1);
    }
    __shared__ unsigned long long red[4][4];
    __shared__ unsigned redm[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off, kWave);
        s2 += __shfl_xor(s2, off, kWave);
        se += __shfl_xor(se, off, kWave);
        ssp += __shfl_xor(ssp, off, kWave);
        unsigned o = __shfl_xor(mx, off, kWave);
        mx = o > mx ? o : mx;
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = s1; red[wv][1] = s2; red[wv][2] = se; red[wv][3] = ssp;
        redm[wv] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int i = 1; i < nw; ++i) {
            s1 += red[i][0]; s2 += red[i][1]; se += red[i][2]; ssp += red[i][3];
            mx = redm[i] > mx ? redm[i] : mx;
        }
        atomicAdd(&out[0], s1);
        atomicAdd(&out[1], s2);
        atomicAdd(&out[2], se);
        atomicAdd(&out[3], ssp);
        atomicMax(&out[4], (unsigned long long)mx);
    }
}

} // namespace bhs
