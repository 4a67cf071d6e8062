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

template <typename T>
struct SegsumParams {
    const int *row_ptr;
    const int *col;
    const T *val;
    const T *x;
    T *y;
    const int *tile_row;
    int *sync_row;
    T *sync_val;
    unsigned long long *stats;   // [fast, dirty, sync] when counting
    int m, nnz, NP, TPW, NW, lead_empty, wpb, wave_lds_bytes;
    T alpha;
};

// Loads this lane's S items (e0 .. e0+S-1) and forms val*x[col]; items at or
// past nnz contribute 0 (kernels_cuda.h:88, :161).
template <typename T, int S>
__device__ __forceinline__ void load_products(const int *__restrict__ col,
                                              const T *__restrict__ val,
                                              const T *__restrict__ x, int e0, int nnz,
                                              T (&prod)[S])
{
    static_assert(S % 4 == 0, "S must be a multiple of 4");
    constexpr int VPQ = 16 / (int)sizeof(T);   // values per 16-byte load
    int c[S];
    T v[S];
#pragma unroll
    for (int i = 0; i < S / 4; ++i) {
        u32x4 q = load_nt16(col + e0 + 4 * i);
        c[4 * i + 0] = (int)q.x;
        c[4 * i + 1] = (int)q.y;
        c[4 * i + 2] = (int)q.z;
        c[4 * i + 3] = (int)q.w;
    }
#pragma unroll
    for (int i = 0; i < S / VPQ; ++i) {
        u32x4 q = load_nt16(val + e0 + VPQ * i);
        __builtin_memcpy(&v[VPQ * i], &q, 16);
    }
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const T xv = x[c[k]];
        prod[k] = (e0 + k < nnz) ? v[k] * xv : T(0);
    }
}

// ---------------------------------------------------------------------------
// Speculative segmented sum (kernels_cuda.h:7-269, wave64 form).
//
// Per tile [B, B+P) with r0 = tile_row[p], r1 = tile_row[p+1]:
//  * fast track (r0 == r1, the tile lies inside one row, :79-112): lane sums,
//    wave sum, added to the carried partial; nothing is stored until the row
//    closes.
//  * otherwise
//    - bit flags (:114-148): lanes walk rows [r0, min(r1, m-1)]; a nonempty
//      row starting at offset o in [0, P) sets bit o%S of lane o/S's word
//      (LDS atomic OR); an empty row marks the tile dirty and gets y = 0.
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
template <typename T, int S, bool STATS>
__global__ __launch_bounds__(512) void segsum_kernel(const SegsumParams<T> q)
{
    constexpr int P = kWave * S;
    extern __shared__ double smem_raw[];
    const int lane = lane_id();
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int blk = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int w = blk * q.wpb + wib;
    if (w >= q.NW) return;

    unsigned char *wave_lds = reinterpret_cast<unsigned char *>(smem_raw) + wib * q.wave_lds_bytes;
    unsigned *s_head = reinterpret_cast<unsigned *>(wave_lds);
    T *s_seg = reinterpret_cast<T *>(wave_lds + 256);

    // Rows before the first nonzero belong to no tile: spread their zeroing.
    if (q.lead_empty > 0) {
        const long long lo = (long long)q.lead_empty * w / q.NW;
        const long long hi = (long long)q.lead_empty * (w + 1) / q.NW;
        for (long long r = lo + lane; r < hi; r += kWave) q.y[r] = T(0);
    }

    const int p_begin = w * q.TPW;
    const int p_end = min(p_begin + q.TPW, q.NP);
    const int E0 = p_begin * P;
    const int r_first = q.tile_row[p_begin];
    bool syncing = q.row_ptr[r_first] < E0;    // first row began in an earlier wave
    bool did_sync = false;
    bool open = false;                          // a row continues past the last tile
    T carry = T(0);
    unsigned long long n_fast = 0, n_dirty = 0;

    for (int p = p_begin; p < p_end; ++p) {
        const int B = p * P;
        const int r0 = q.tile_row[p];
        const int r1 = q.tile_row[p + 1];
        const bool cont = q.row_ptr[r0] < B;    // item 0 continues row r0

        T prod[S];
        load_products<T, S>(q.col, q.val, q.x, B + lane * S, q.nnz, prod);

        if (r0 == r1) {                          // fast track
            T s = prod[0];
#pragma unroll
            for (int k = 1; k < S; ++k) s += prod[k];
            s = wave_sum(s);
            carry = cont ? carry + s : s;
            open = true;
            if (STATS) ++n_fast;
            continue;
        }

        // ---- bit flags ----------------------------------------------------
        const int stop = (r1 == q.m) ? q.m - 1 : r1;
        wave_sync();                             // previous tile's LDS reads are done
        s_head[lane] = 0u;
        wave_sync();
        bool lane_dirty = false;
        for (int r = r0 + lane; r <= stop; r += kWave) {
            const int rp = q.row_ptr[r], rpn = q.row_ptr[r + 1];
            if (rpn > rp) {
                const int o = rp - B;
                if (o >= 0 && o < P) atomicOr(&s_head[o / S], 1u << (o % S));
            } else {
                lane_dirty = true;
                q.y[r] = T(0);
            }
        }
        const bool dirty = ballot(lane_dirty) != 0ull;
        wave_sync();
        unsigned hb = s_head[lane];
        if (lane == 0) {
            hb |= 1u;
            if (cont) prod[0] = prod[0] + carry;
        }

        // ---- segment numbers ---------------------------------------------
        const int nh = __popc(hb);
        const int incl = wave_incl_scan(nh, lane);
        const int seg_base = incl - nh;
        const int K = read_lane(incl, kWave - 1);

        // ---- lane-sequential segmented reduction -------------------------
        T acc = T(0), lead = T(0);
        bool seen = false;
        int sidx = seg_base;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            if ((hb >> k) & 1u) {
                if (seen) { s_seg[sidx] = acc; ++sidx; }
                else { lead = acc; seen = true; }
                acc = T(0);
            }
            acc += prod[k];
        }

        // ---- segments spanning lanes -------------------------------------
        const uint64_t present = ballot(seen);  // bit 0 always set
        const T I = wave_segscan(acc, present, lane);
        const T Iprev = __shfl_up(I, 1, kWave);
        if (seen && lane > 0) s_seg[seg_base - 1] = Iprev + lead;
        if (lane == kWave - 1) s_seg[K - 1] = I;
        wave_sync();

        // ---- output ------------------------------------------------------
        const bool partial_tail = (r1 < q.m) && (q.row_ptr[r1] < B + P);
        const int kb = syncing ? 1 : 0;
        const int ke = partial_tail ? K - 1 : K;
        if (!dirty) {
            for (int k = kb + lane; k < ke; k += kWave) q.y[r0 + k] = q.alpha * s_seg[k];
        } else {
            if (STATS) ++n_dirty;
            int running = 0;
            for (int rb = r0; rb <= stop; rb += kWave) {
                const int r = rb + lane;
                bool ne = false;
                if (r <= stop) ne = q.row_ptr[r + 1] > q.row_ptr[r];
                const uint64_t msk = ballot(ne);
                const int rank = running + popc_below(msk);
                if (ne && rank >= kb && rank < ke) q.y[r] = q.alpha * s_seg[rank];
                running += __popcll(msk);
            }
        }
        if (syncing) {
            if (lane == 0) {
                q.sync_row[w] = r0;
                q.sync_val[w] = q.alpha * s_seg[0];
            }
            syncing = false;
            did_sync = true;
        }
        carry = s_seg[K - 1];
        open = partial_tail;
    }

    if (open) {
        const int r_open = q.tile_row[p_end];
        if (syncing) {
            if (lane == 0) {
                q.sync_row[w] = r_open;
                q.sync_val[w] = q.alpha * carry;
            }
            did_sync = true;
        } else if (lane == 0) {
            q.y[r_open] = q.alpha * carry;
        }
    }
    if (!did_sync && lane == 0) q.sync_row[w] = -1;
    if (STATS && lane == 0) {
        atomicAdd(&q.stats[0], n_fast);
        atomicAdd(&q.stats[1], n_dirty);
        atomicAdd(&q.stats[2], did_sync ? 1ull : 0ull);
    }
}

// ---------------------------------------------------------------------------
// Calibration (bhsparse_spmv_cuda.h:166-173 moved to the device): each row
// whose partials are spread over several waves is finished by the thread of
// the first wave in its chain, adding the partials in wave order to the value
// the row's starting wave stored -- the reference's y[idx] += val loop order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void calibrate_kernel(const int *__restrict__ sync_row,
                                                       const T *__restrict__ sync_val, int NW,
                                                       T *__restrict__ y)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= NW) return;
    const int r = sync_row[w];
    if (r < 0) return;
    if (w > 0 && sync_row[w - 1] == r) return;
    T acc = y[r];
    for (int i = w; i < NW && sync_row[i] == r; ++i) acc += sync_val[i];
    y[r] = acc;
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
                                                    T alpha)
{
    const int tl = threadIdx.x & (TPV - 1);
    const long long vid = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / TPV;
    const long long nv = (long long)gridDim.x * blockDim.x / TPV;
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
