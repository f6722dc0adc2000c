// Wave-level block-row helpers shared by the block kernels (BSR SpMV, the
// block LU-SGS apply, the FlowSys block sweeps).
//
// One wave (64 lanes, one wave per workgroup) serves G = 64 / B block rows;
// lane (g, r) = (lane / B, lane % B) owns scalar row r of block row g.  The
// wave first copies the blocks its rows use from HBM into LDS with loads in
// which consecutive lanes read consecutive words (a block row's blocks are
// contiguous in memory), then every lane reads its own block row's entries
// from LDS.  A block is row-major, so one lane's scalar row is B contiguous
// values: a thread-per-block-row kernel makes each lane stream its own 8*B*B
// bytes, and each wave-wide load touches up to 64 different cache lines;
// here each load touches 4-5.  Each lane keeps the summation order of its
// scalar row, so results equal those of a thread-per-row kernel bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace ucfd {

// LDS ordering inside one wave: LDS operations of a wave execute in order,
// the fences keep the compiler from moving LDS accesses across this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// exclusive prefix sum of v over the 64 lanes of the wave
__device__ __forceinline__ int wave_excl_scan(int v)
{
    const int lane = threadIdx.x & 63;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x - v;
}

// one-wave workgroups remapped so that each XCD owns runs of 32 consecutive
// waves, the 8 XCDs neighbouring runs (hardware deals workgroup ids
// round-robin over the XCDs): adjacent rows share one L2 while the XCDs
// work on nearby rows (3-4 % faster block sweeps than one contiguous eighth
// per XCD).  The last incomplete group of runs keeps the identity.
__device__ __forceinline__ int wave_xcd_index()
{
    constexpr int RUN = 32;
    const int b = blockIdx.x, nblk = gridDim.x;
    const int full = nblk/(8*RUN)*(8*RUN);
    if (b >= full) return b;
    const int x = b % 8, k = b / 8;
    return (k/RUN)*(8*RUN) + x*RUN + k%RUN;
}

}  // namespace ucfd
