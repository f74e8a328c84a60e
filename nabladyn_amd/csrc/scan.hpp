// A gated exclusive scan of int32 values for the list builds: three launches (tile sums, one
// block scanning the tile sums, tiles applied), each of which returns at once when the build
// is not running, so a captured update pays three empty launches on steps without a rebuild.
// The tile sums live after the list header in the caller's scan_temp buffer.
#pragma once

#include "common.hpp"

namespace ndn {

constexpr int kScanPerThread = 4;
constexpr int kScanTile = kBlock * kScanPerThread;  // 1024 values per block

// Bytes of the list header at the start of scan_temp; the tile sums follow it. Bytes 64-575 hold
// 64 slots of per-block maxima of the squared displacement at the last reuse check (the bits of
// a nonnegative T, widened to 64 bits), so no slot takes many atomics.
constexpr std::int64_t kHeaderBytes = 1024;
constexpr int kMovedSlots = 64;

struct ListHeader
{
    index_t old_total;  // pair_total before this build: columns past it hold padding
};

__device__ __forceinline__ unsigned long long* moved_slot(ListHeader* header, unsigned block)
{
    return reinterpret_cast<unsigned long long*>(reinterpret_cast<std::uint8_t*>(header) + 64) +
           (block % kMovedSlots);
}

inline std::int64_t scan_tiles(std::int64_t items)
{
    return (items + kScanTile - 1) / kScanTile;
}

inline std::int64_t scan_bytes(std::int64_t items)
{
    return kHeaderBytes + 4 * (scan_tiles(items) + 1);
}

// The bodies take the block (a launch over a batch of systems passes each system's own).
static __device__ __forceinline__ void scan_tile_sums_body(const index_t* __restrict__ input,
                                                           std::int64_t count,
                                                           index_t* __restrict__ tile_sums,
                                                           index_t block)
{
    __shared__ index_t shared[16];
    const std::int64_t base = static_cast<std::int64_t>(block) * kScanTile;
    index_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k)
    {
        const std::int64_t i = base + threadIdx.x + static_cast<std::int64_t>(k) * kBlock;
        if (i < count) local += input[i];
    }
    const index_t sum = block_sum(local, shared);
    if (threadIdx.x == 0) tile_sums[block] = sum;
}

static __global__ void __launch_bounds__(kBlock) scan_tile_sums_kernel(
    const index_t* __restrict__ input, std::int64_t count, index_t* __restrict__ tile_sums,
    const index_t* rebuild, const index_t* status)
{
    if (!rebuilding(rebuild, status)) return;
    scan_tile_sums_body(input, count, tile_sums, static_cast<index_t>(blockIdx.x));
}

static __device__ __forceinline__ void scan_tile_prefix_body(index_t* __restrict__ tile_sums,
                                                             std::int64_t tiles)
{
    __shared__ index_t shared[16];
    index_t carry = 0;
    for (std::int64_t base = 0; base < tiles; base += kScanTile)
    {
        index_t v[kScanPerThread];
        index_t local = 0;
#pragma unroll
        for (int k = 0; k < kScanPerThread; ++k)
        {
            const std::int64_t i = base + threadIdx.x * kScanPerThread + k;
            v[k] = i < tiles ? tile_sums[i] : 0;
            local += v[k];
        }
        index_t total;
        index_t running = carry + block_exclusive_sum(local, shared, &total);
#pragma unroll
        for (int k = 0; k < kScanPerThread; ++k)
        {
            const std::int64_t i = base + threadIdx.x * kScanPerThread + k;
            if (i < tiles) tile_sums[i] = running;
            running += v[k];
        }
        carry += total;
        __syncthreads();
    }
}

static __global__ void __launch_bounds__(kBlock) scan_tile_prefix_kernel(
    index_t* __restrict__ tile_sums, std::int64_t tiles, const index_t* rebuild,
    const index_t* status)
{
    if (!rebuilding(rebuild, status)) return;
    scan_tile_prefix_body(tile_sums, tiles);
}

// out may alias in: each block reads its whole tile before writing it.
static __device__ __forceinline__ void scan_apply_body(const index_t* input, std::int64_t count,
                                                       const index_t* __restrict__ tile_prefix,
                                                       index_t* output, index_t block)
{
    __shared__ index_t shared[16];
    const std::int64_t base =
        static_cast<std::int64_t>(block) * kScanTile + threadIdx.x * kScanPerThread;
    index_t v[kScanPerThread];
    index_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k)
    {
        v[k] = base + k < count ? input[base + k] : 0;
        local += v[k];
    }
    index_t total;
    index_t running = tile_prefix[block] + block_exclusive_sum(local, shared, &total);
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k)
    {
        if (base + k < count) output[base + k] = running;
        running += v[k];
    }
}

static __global__ void __launch_bounds__(kBlock) scan_apply_kernel(
    const index_t* input, std::int64_t count, const index_t* __restrict__ tile_prefix,
    index_t* output, const index_t* rebuild, const index_t* status)
{
    if (!rebuilding(rebuild, status)) return;
    scan_apply_body(input, count, tile_prefix, output, static_cast<index_t>(blockIdx.x));
}

// Exclusive scan of `count` values; returns NDN_OK or NDN_ERR_HIP.
inline int exclusive_scan(const index_t* input, index_t* output, std::int64_t count,
                          std::uint8_t* scan_temp, const index_t* rebuild, const index_t* status,
                          hipStream_t stream)
{
    const std::int64_t tiles = scan_tiles(count);
    index_t* tile_sums = reinterpret_cast<index_t*>(scan_temp + kHeaderBytes);
    scan_tile_sums_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(
        input, count, tile_sums, rebuild, status);
    NDN_LAUNCHED("scan_tile_sums");
    scan_tile_prefix_kernel<<<1, kBlock, 0, stream>>>(tile_sums, tiles, rebuild, status);
    NDN_LAUNCHED("scan_tile_prefix");
    scan_apply_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(
        input, count, tile_sums, output, rebuild, status);
    NDN_LAUNCHED("scan_apply");
    return NDN_OK;
}

}  // namespace ndn
