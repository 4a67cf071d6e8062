// bhs_device.h -- wave64 building blocks for gfx950 (CDNA4).
//
// The reference exchanges values between lanes through shared memory with no
// synchronisation and relies on 32-lane lockstep (spmv_cuda/kernels_cuda.h:161,
// :181-196, :214-227).  Here every cross-lane step is an explicit wave-level
// primitive (ballot, mbcnt, ds_bpermute shuffles, readlane) or an LDS access
// ordered by wave_sync().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace bhs {

constexpr int kWave = 64;

// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N-1>{}):
// a loop whose index is a constant expression in the body, so small register
// arrays indexed by it are never placed in scratch memory.
template <typename F, int... I>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}

template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ int lane_id()
{
    return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Orders this wave's LDS traffic: every LDS write issued before the call is
// visible to every lane of the wave after it.  (LDS operations of one wave
// execute in order; the fences, restricted to the LDS address space, stop the
// compiler from moving LDS operations across the point while leaving global
// loads free to be scheduled early.)  Never a workgroup barrier: waves of a
// block run independent tile ranges.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ int popc_below(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }

__device__ __forceinline__ int read_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ double read_lane(double v, int l)
{
    union { double d; int i[2]; } u;
    u.d = v;
    u.i[0] = __builtin_amdgcn_readlane(u.i[0], l);
    u.i[1] = __builtin_amdgcn_readlane(u.i[1], l);
    return u.d;
}

__device__ __forceinline__ float read_lane(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// DPP lane moves (GFX9 encodings).  64-bit values move as two 32-bit halves.
enum : int {
    kDppXor1 = 0xB1,          // quad_perm [1,0,3,2]
    kDppXor2 = 0x4E,          // quad_perm [2,3,0,1]
    kDppRowShr1 = 0x111,
    kDppRowShr2 = 0x112,
    kDppRowShr4 = 0x114,
    kDppRowShr8 = 0x118,
    kDppRowMirror = 0x140,    // lane i <- lane 15-i inside each row of 16
    kDppRowHalfMirror = 0x141,// lane i <- lane 7-i inside each half row of 8
    kDppRowBcast15 = 0x142,   // row r+1 <- lane 15 of row r
    kDppRowBcast31 = 0x143,   // rows 2, 3 <- lane 31
    kDppWaveShl1 = 0x130,     // lane i <- lane i+1 across the wave (lane 63 keeps `old`)
    kDppWaveShr1 = 0x138,     // lane i <- lane i-1 across the wave (lane 0 keeps `old`)
};

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp(int v, int old = 0)
{
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    union { double d; int i[2]; } u, r;
    u.d = v;
    r.i[0] = __builtin_amdgcn_update_dpp(0, u.i[0], CTRL, 0xF, 0xF, false);
    r.i[1] = __builtin_amdgcn_update_dpp(0, u.i[1], CTRL, 0xF, 0xF, false);
    return r.d;
}

template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Two right-hand sides carried together (the batched path): every operation
// on a pair acts on each component exactly as on a single value, so
// component c holds the bits a single-RHS pass would.
typedef double double2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

template <typename T, int R>
struct RhsVec;
template <typename T>
struct RhsVec<T, 1> {
    using type = T;
};
template <>
struct RhsVec<double, 2> {
    using type = double2v;
};
template <>
struct RhsVec<float, 2> {
    using type = float2v;
};

template <int CTRL>
__device__ __forceinline__ double2v dpp(double2v v)
{
    return double2v{dpp<CTRL>(v.x), dpp<CTRL>(v.y)};
}

template <int CTRL>
__device__ __forceinline__ float2v dpp(float2v v)
{
    return float2v{dpp<CTRL>(v.x), dpp<CTRL>(v.y)};
}

__device__ __forceinline__ double2v read_lane(double2v v, int l)
{
    return double2v{read_lane(v.x, l), read_lane(v.y, l)};
}

__device__ __forceinline__ float2v read_lane(float2v v, int l)
{
    return float2v{read_lane(v.x, l), read_lane(v.y, l)};
}

// lane i <- lane i-d (lanes below d keep their own value)
template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d)
{
    return __shfl_up(v, d, kWave);
}

template <>
__device__ __forceinline__ double2v shfl_up(double2v v, int d)
{
    return double2v{__shfl_up(v.x, d, kWave), __shfl_up(v.y, d, kWave)};
}

template <>
__device__ __forceinline__ float2v shfl_up(float2v v, int d)
{
    return float2v{__shfl_up(v.x, d, kWave), __shfl_up(v.y, d, kWave)};
}

// Inclusive prefix sum of an int over the 64 lanes (DPP row shifts, then the
// row broadcasts carry the running total across rows of 16).
__device__ __forceinline__ int wave_incl_scan(int v, int /*lane*/)
{
    v += dpp<kDppRowShr1>(v);
    v += dpp<kDppRowShr2>(v);
    v += dpp<kDppRowShr4>(v);
    v += dpp<kDppRowShr8>(v);
    v += dpp<kDppRowBcast15, 0xA>(v);
    v += dpp<kDppRowBcast31, 0xC>(v);
    return v;
}

// Sum over the 64 lanes, the same value in every lane.  Pairwise tree:
// quads ((v0+v1)+(v2+v3)), then half rows (Q0+Q1), rows (H0+H1) -- each step
// pairs mirror-image partners, so both partners compute the same sum -- and
// finally (R0+R1)+(R2+R3) over the four rows.
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    v = v + dpp<kDppXor1>(v);
    v = v + dpp<kDppXor2>(v);
    v = v + dpp<kDppRowHalfMirror>(v);
    v = v + dpp<kDppRowMirror>(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Segmented inclusive scan across lanes, left to right.  Lanes whose bit is
// set in `head` start a new segment; lane 0 must be a head.  Returns, for lane
// i, the sum of v over [s, i] where s is the last head at or before i.
// Hillis-Steele doubling, stopping as soon as every lane's window reaches a
// head (no step at all when every lane is a head).
template <typename T>
__device__ __forceinline__ T wave_segscan(T v, uint64_t head, int lane)
{
    uint64_t cov = head;   // bit i: lane i's window already contains a head
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        if (cov == ~0ull) break;
        // lane i <- lane i-d (lane 0 is a head and never uses its value)
        T up = d == 1 ? dpp<kDppWaveShr1>(v) : shfl_up(v, d);
        if (!((cov >> lane) & 1ull)) v = up + v;
        cov |= cov << d;
    }
    return v;
}

// 16-byte nontemporal vector load (global_load_dwordx4 ... nt): the values
// and column indices are streamed once per call and should not displace x
// from the caches.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u32x4 load_nt16(const void *p)
{
    return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
}

__device__ __forceinline__ u32x2 load_nt8(const void *p)
{
    return __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(p));
}

// 16 bytes from global memory straight into LDS (global_load_lds_dwordx4, no
// registers): the lanes of one call write one contiguous 1 KB block, lane l at
// lds_block + 16*l.  lds_block is wave-uniform.  Completion is counted by
// vmcnt; lds_dma_wait() makes the block readable by every lane of the wave.
__device__ __forceinline__ void load_lds16(const void *src, void *lds_block)
{
    __builtin_amdgcn_global_load_lds(
        (__attribute__((address_space(1))) const void *)src,
        (__attribute__((address_space(3))) void *)lds_block, 16, 0, 0);
}

__device__ __forceinline__ void lds_dma_wait()
{
    __builtin_amdgcn_s_waitcnt(0);
    wave_sync();
}

// v if keep, else +0, without a branch (bitwise select keeps loads feeding v
// unconditional).
__device__ __forceinline__ double keep_or_zero(double v, bool keep)
{
    long long b = __builtin_bit_cast(long long, v);
    return __builtin_bit_cast(double, keep ? b : 0LL);
}

__device__ __forceinline__ float keep_or_zero(float v, bool keep)
{
    int b = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, keep ? b : 0);
}

__device__ __forceinline__ double2v keep_or_zero(double2v v, bool keep)
{
    return double2v{keep_or_zero(v.x, keep), keep_or_zero(v.y, keep)};
}

__device__ __forceinline__ float2v keep_or_zero(float2v v, bool keep)
{
    return float2v{keep_or_zero(v.x, keep), keep_or_zero(v.y, keep)};
}

// Block index remap so that each XCD (blocks are dealt round-robin over the
// 8 XCDs) works on one contiguous range of blocks: neighbouring tiles share
// row_ptr, y and (for banded matrices) x lines in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    constexpr int X = 8;
    const int xcd = b % X, local = b / X;
    const int full = nb / X, rem = nb % X;
    return xcd < rem ? xcd * (full + 1) + local : rem * (full + 1) + (xcd - rem) * full + local;
}

} // namespace bhs
