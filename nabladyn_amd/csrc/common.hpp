// Shared definitions of the gfx950 kernels: index type, status bits, the rebuild gate, wave64
// helpers, a block-wide scan, and the C-ABI error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/nabladyn_amd.h"

namespace ndn {

using index_t = std::int32_t;

constexpr int kWave = 64;        // CDNA wavefront
constexpr int kBlock = 256;      // 4 waves
constexpr int kStatusInvalidCell = NDN_STATUS_INVALID_CELL;
constexpr int kStatusNonFinite = NDN_STATUS_NON_FINITE;
constexpr int kStatusOverflow = NDN_STATUS_OVERFLOW;
constexpr int kStatusFullOverflow = NDN_STATUS_FULL_OVERFLOW;
constexpr int kStatusHaloOverflow = NDN_STATUS_HALO_OVERFLOW;

inline unsigned int blocks_for(std::int64_t items, int threads = kBlock)
{
    return static_cast<unsigned int>((items + threads - 1) / threads);
}

// Whether a rebuild kernel works: a first-frame build passes no flag and always does; an update
// works only when the reuse check asked for a rebuild and no status bit is set.
__device__ __forceinline__ bool rebuilding(const index_t* rebuild, const index_t* status)
{
    return rebuild == nullptr || (rebuild[0] != 0 && status[0] == 0);
}

__device__ __forceinline__ int lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ int bits_below(std::uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

__device__ __forceinline__ std::uint64_t ballot(bool predicate)
{
    return __ballot(predicate);
}

// The value `lane` (wave-uniform) holds: a scalar read, no LDS round trip.
__device__ __forceinline__ int read_lane(int value, int lane)
{
    return __builtin_amdgcn_readlane(value, lane);
}

__device__ __forceinline__ float read_lane(float value, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(value), lane));
}

__device__ __forceinline__ double read_lane(double value, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(value), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(value), lane);
    return __hiloint2double(hi, lo);
}

template <typename V>
__device__ __forceinline__ V wave_inclusive_sum(V value)
{
    const int lane = lane_id();
#pragma unroll
    for (int offset = 1; offset < kWave; offset <<= 1)
    {
        const V other = __shfl_up(value, offset, kWave);
        if (lane >= offset) value += other;
    }
    return value;
}

template <typename V>
__device__ __forceinline__ V wave_sum(V value)
{
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1)
    {
        value += __shfl_xor(value, offset, kWave);
    }
    return value;
}

// Runs of equal keys in neighbouring lanes (the active lanes a prefix of the wave, as in a
// grid-stride loop over consecutive items): whether this lane starts its run, the run's length
// (for a starting lane) and the run's first lane (for every lane). Binning counts with one
// atomic per run of atoms in one bin: a frame sorted at the last build, or atoms given in a
// spatial order, come in such runs.
struct LaneRun
{
    bool head;
    int length;
    int first;
};

__device__ __forceinline__ LaneRun equal_runs(index_t key)
{
    const int lane = lane_id();
    const std::uint64_t active = ballot(true);
    const index_t before = __shfl_up(key, 1, kWave);
    LaneRun run;
    run.head = lane == 0 || before != key;
    // inactive lanes end runs too
    const std::uint64_t heads = ballot(run.head) | ~active;
    const std::uint64_t higher = lane == kWave - 1 ? 0ull : heads >> (lane + 1);
    run.length = (higher == 0 ? kWave : lane + 1 + __builtin_ctzll(higher)) - lane;
    const std::uint64_t lower = lane == kWave - 1 ? heads : heads & ((2ull << lane) - 1ull);
    run.first = 63 - __builtin_clzll(lower);
    return run;
}

template <typename V>
__device__ __forceinline__ V wave_max(V value)
{
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1)
    {
        const V other = __shfl_xor(value, offset, kWave);
        value = other > value ? other : value;
    }
    return value;
}

// Exclusive scan over the block (blockDim.x a multiple of 64, at most 1024). Returns this
// thread's exclusive prefix; *total receives the block sum. `shared` holds 16 values.
template <typename V>
__device__ __forceinline__ V block_exclusive_sum(V value, V* shared, V* total)
{
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const int waves = blockDim.x / kWave;
    const V inclusive = wave_inclusive_sum(value);
    if (lane == kWave - 1) shared[wave] = inclusive;
    __syncthreads();
    V before = 0;
    V sum = 0;
    for (int w = 0; w < waves; ++w)
    {
        const V s = shared[w];
        if (w < wave) before += s;
        sum += s;
    }
    __syncthreads();
    *total = sum;
    return before + inclusive - value;
}

template <typename V>
__device__ __forceinline__ V block_sum(V value, V* shared)
{
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const int waves = blockDim.x / kWave;
    value = wave_sum(value);
    if (lane == 0) shared[wave] = value;
    __syncthreads();
    V sum = 0;
    for (int w = 0; w < waves; ++w) sum += shared[w];
    __syncthreads();
    return sum;
}

// ---- C-ABI error plumbing -------------------------------------------------------------------

void set_error(const std::string& message);
const char* last_error();

#define NDN_REQUIRE(condition, message)                                                          \
    do                                                                                           \
    {                                                                                            \
        if (!(condition))                                                                        \
        {                                                                                        \
            ::ndn::set_error(message);                                                           \
            return NDN_ERR_ARGUMENT;                                                             \
        }                                                                                        \
    } while (0)

#define NDN_LAUNCHED(name)                                                                       \
    do                                                                                           \
    {                                                                                            \
        const hipError_t launch_error = hipGetLastError();                                       \
        if (launch_error != hipSuccess)                                                          \
        {                                                                                        \
            ::ndn::set_error(std::string(name) + ": " + hipGetErrorString(launch_error));        \
            return NDN_ERR_HIP;                                                                  \
        }                                                                                        \
    } while (0)

}  // namespace ndn
