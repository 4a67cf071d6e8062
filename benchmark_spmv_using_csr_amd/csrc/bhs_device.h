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

namespace bhs {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id()
{
    return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Orders this wave's LDS traffic: every LDS write issued before the call is
// visible to every lane of the wave after it.  (LDS operations of one wave
// execute in order; the fences stop the compiler from moving memory operations
// across the point.)  Never a workgroup barrier: waves of a block run
// independent tile ranges.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

// Inclusive prefix sum of an int over the 64 lanes.
__device__ __forceinline__ int wave_incl_scan(int v, int lane)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        int t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

// Sum over the 64 lanes (xor butterfly; every lane gets the same value
// because each step adds the same two operands in both partner lanes).
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
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
        T up = __shfl_up(v, d, kWave);
        if (!((cov >> lane) & 1ull)) v = up + v;
        cov |= cov << d;
    }
    return v;
}

// 16-byte nontemporal vector load (global_load_dwordx4 ... nt): the values
// and column indices are streamed once per call and should not displace x
// from the caches.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ u32x4 load_nt16(const void *p)
{
    return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
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
