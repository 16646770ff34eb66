// Queue binning: the iterate kernel stores one 64-bit record per walker and
// pass (see flame_kernels.h for the layout); these kernels fold a chunk of
// records into the histogram without contended global atomics:
//
//   1. count    per (tile, partition) record counts; a partition is a fixed
//               contiguous range of records handled by one workgroup
//   2. scan     exclusive scan over [tile][partition] -> write offsets, and
//               the table of (tile, segment) work items for step 4
//   3. scatter  counting sort by tile: each partition sorts 4096-record
//               pieces in LDS and writes tile runs to their offsets
//   4. bin      one workgroup per (tile, segment of <= 2^18 records) sums
//               the records in a 64x64 LDS tile, then adds the non-zero
//               pixels into the histogram with one atomic per channel
//
// All integer sums: the histogram equals the one direct atomics give, bit
// for bit, whatever the order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "static_kernels.hpp"

namespace rfk::sk {

namespace {

typedef unsigned long long u64;

constexpr uint32_t kTileLog2 = 6;                 // 64x64-pixel tiles
constexpr uint32_t kTilePix = 1u << (2 * kTileLog2);
constexpr uint32_t kNoTile = 0xFFFFFu;
constexpr uint32_t kPartThreads = 1024;
constexpr uint32_t kPiece = 4096;                 // records sorted in LDS at a time
constexpr uint32_t kSegment = 1u << 18;           // records per bin work item (u32 LDS sums cannot wrap)

__device__ __forceinline__ uint32_t rec_tile(u64 r) { return uint32_t(r >> 44); }

// Lanes of a wave that share the wave leader's key add once, by the first of
// them; the others (usually a minority for real flames, the whole wave for a
// concentrated one) add on their own. Returns the rank of the lane's record
// among those counted for its key in this call.
__device__ __forceinline__ uint32_t wave_count(uint32_t* counters, uint32_t key, bool valid) {
    const u64 vm = __ballot(valid);
    if (vm == 0ull) return 0u;
    const uint32_t lead = uint32_t(__shfl(int(key), int(__builtin_ctzll(vm))));
    const u64 same = __ballot(valid && key == lead);
    const uint32_t lane = __lane_id();
    uint32_t rank = 0;
    if (valid && key == lead) {
        const uint32_t first = uint32_t(__builtin_ctzll(same));
        const uint32_t below = uint32_t(__popcll(same & ((1ull << lane) - 1ull)));
        uint32_t base = 0;
        if (lane == first) base = atomicAdd(&counters[key], uint32_t(__popcll(same)));
        base = __shfl(base, int(first));
        rank = base + below;
    } else if (valid) {
        rank = atomicAdd(&counters[key], 1u);
    }
    return rank;
}

__global__ void __launch_bounds__(kPartThreads)
    qcount_kernel(const u64* __restrict__ rec, size_t n, size_t per_part, uint32_t ntiles, uint32_t nparts,
                  uint32_t* __restrict__ counts) {
    extern __shared__ uint32_t s_cnt[];
    for (uint32_t t = threadIdx.x; t < ntiles; t += blockDim.x) s_cnt[t] = 0;
    __syncthreads();
    const size_t b0 = size_t(blockIdx.x) * per_part, b1 = std::min(n, b0 + per_part);
    for (size_t base = b0; base < b1; base += blockDim.x) {
        const size_t i = base + threadIdx.x;
        const u64 r = i < b1 ? rec[i] : ~0ull;
        const uint32_t tile = rec_tile(r);
        (void)wave_count(s_cnt, tile, tile < ntiles);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < ntiles; t += blockDim.x) counts[size_t(t) * nparts + blockIdx.x] = s_cnt[t];
}

// One workgroup. counts [ntiles][nparts] -> exclusive offsets in place;
// segments: (tile, begin, end) per work item of step 4, nseg[0] = count.
__global__ void __launch_bounds__(1024)
    qscan_kernel(uint32_t* __restrict__ counts, uint32_t ntiles, uint32_t nparts, uint4* __restrict__ segments,
                 uint32_t max_segments, uint32_t* __restrict__ nseg) {
    __shared__ uint32_t s_sum[1024];
    __shared__ uint32_t s_segs[1024];
    const uint32_t n = ntiles * nparts;
    const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = threadIdx.x * per, hi = std::min(n, lo + per);
    uint32_t sum = 0;
    for (uint32_t i = lo; i < hi; i++) sum += counts[i];
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    // inclusive Hillis-Steele scan of the 1024 partial sums
    for (uint32_t d = 1; d < blockDim.x; d <<= 1) {
        const uint32_t v = threadIdx.x >= d ? s_sum[threadIdx.x - d] : 0u;
        __syncthreads();
        s_sum[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? s_sum[threadIdx.x - 1] : 0u;
    for (uint32_t i = lo; i < hi; i++) {
        const uint32_t c = counts[i];
        counts[i] = run;
        run += c;
    }
    const uint32_t total = s_sum[blockDim.x - 1];
    __syncthreads();
    // segments: tile t covers [off(t,0), off(t+1,0)) split into kSegment pieces;
    // tiles are handed out in contiguous ranges, one range per thread
    const uint32_t tper = (ntiles + blockDim.x - 1) / blockDim.x;
    const uint32_t t0 = threadIdx.x * tper, t1 = std::min(ntiles, t0 + tper);
    uint32_t nmine = 0;
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t b = counts[size_t(t) * nparts];
        const uint32_t e = t + 1 < ntiles ? counts[size_t(t + 1) * nparts] : total;
        nmine += (e - b + kSegment - 1) / kSegment;
    }
    s_segs[threadIdx.x] = nmine;
    __syncthreads();
    for (uint32_t d = 1; d < blockDim.x; d <<= 1) {
        const uint32_t v = threadIdx.x >= d ? s_segs[threadIdx.x - d] : 0u;
        __syncthreads();
        s_segs[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t k = threadIdx.x ? s_segs[threadIdx.x - 1] : 0u;
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t b = counts[size_t(t) * nparts];
        const uint32_t e = t + 1 < ntiles ? counts[size_t(t + 1) * nparts] : total;
        for (uint32_t s = b; s < e; s += kSegment) {
            if (k < max_segments) segments[k] = make_uint4(t, s, std::min(e, s + kSegment), 0u);
            k++;
        }
    }
    if (threadIdx.x == blockDim.x - 1) nseg[0] = std::min(s_segs[blockDim.x - 1], max_segments);
}

// LDS: cursor[ntiles] (global write position of this partition per tile),
// lcount[ntiles], lstart[ntiles], piece[kPiece] records.
__global__ void __launch_bounds__(kPartThreads)
    qscatter_kernel(const u64* __restrict__ rec, size_t n, size_t per_part, uint32_t ntiles, uint32_t nparts,
                    const uint32_t* __restrict__ offsets, u64* __restrict__ sorted) {
    extern __shared__ uint32_t smem[];
    uint32_t* cursor = smem;
    uint32_t* lcount = cursor + ntiles;
    uint32_t* lstart = lcount + ntiles;
    u64* piece = reinterpret_cast<u64*>(lstart + ntiles + (ntiles & 1u));
    __shared__ uint32_t s_wave[kPartThreads / 64];
    for (uint32_t t = threadIdx.x; t < ntiles; t += blockDim.x) cursor[t] = offsets[size_t(t) * nparts + blockIdx.x];
    const size_t b0 = size_t(blockIdx.x) * per_part, b1 = std::min(n, b0 + per_part);
    constexpr uint32_t kPer = kPiece / kPartThreads;
    for (size_t base = b0; base < b1; base += kPiece) {
        for (uint32_t t = threadIdx.x; t < ntiles; t += blockDim.x) lcount[t] = 0;
        __syncthreads();
        u64 r[kPer];
        uint32_t rank[kPer];
#pragma unroll
        for (uint32_t j = 0; j < kPer; j++) {
            const size_t i = base + size_t(j) * kPartThreads + threadIdx.x;
            r[j] = i < b1 ? rec[i] : ~0ull;
            const uint32_t tile = rec_tile(r[j]);
            rank[j] = wave_count(lcount, tile, tile < ntiles);
        }
        __syncthreads();
        // exclusive scan of lcount -> lstart (ntiles entries, blockDim.x threads)
        const uint32_t per = (ntiles + blockDim.x - 1) / blockDim.x;
        const uint32_t lo = threadIdx.x * per, hi = std::min(ntiles, lo + per);
        uint32_t sum = 0;
        for (uint32_t t = lo; t < hi; t++) sum += lcount[t];
        // wave-level inclusive scan of the per-thread sums, then across waves
        uint32_t incl = sum;
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t v = __shfl_up(incl, d);
            if (__lane_id() >= d) incl += v;
        }
        if (__lane_id() == 63) s_wave[threadIdx.x >> 6] = incl;
        __syncthreads();
        uint32_t wave_base = 0;
        for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) wave_base += s_wave[w];
        uint32_t run = wave_base + incl - sum;
        for (uint32_t t = lo; t < hi; t++) {
            lstart[t] = run;
            run += lcount[t];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kPer; j++) {
            const uint32_t tile = rec_tile(r[j]);
            if (tile < ntiles) piece[lstart[tile] + rank[j]] = r[j];
        }
        __syncthreads();
        uint32_t used = 0;
        for (uint32_t w = 0; w < kPartThreads / 64; w++) used += s_wave[w];
        // sorted order: consecutive threads write consecutive addresses of a run
        for (uint32_t k = threadIdx.x; k < used; k += kPartThreads) {
            const u64 v = piece[k];
            const uint32_t tile = rec_tile(v);
            sorted[size_t(cursor[tile]) + (k - lstart[tile])] = v;
        }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < ntiles; t += blockDim.x) cursor[t] += lcount[t];
        __syncthreads();
    }
}

template <class HistT>
__device__ __forceinline__ void add_bin(HistT* __restrict__ hist, size_t p, uint32_t r, uint32_t g, uint32_t b,
                                        uint32_t w, bool packed64) {
    HistT* h = hist + 4 * p;
    if constexpr (sizeof(HistT) == 4) {
        if (packed64) {
            u64* h64 = reinterpret_cast<u64*>(h);
            if (r | g) atomicAdd(h64, u64(r) | (u64(g) << 32));
            if (b | w) atomicAdd(h64 + 1, u64(b) | (u64(w) << 32));
            return;
        }
    }
    if (r) atomicAdd(h + 0, HistT(r));
    if (g) atomicAdd(h + 1, HistT(g));
    if (b) atomicAdd(h + 2, HistT(b));
    if (w) atomicAdd(h + 3, HistT(w));
}

template <class HistT>
__global__ void __launch_bounds__(1024)
    qbin_kernel(const u64* __restrict__ sorted, const uint4* __restrict__ segments, const uint32_t* __restrict__ nseg,
                uint32_t width, uint32_t height, uint32_t tiles_x, HistT* __restrict__ hist, bool packed64) {
    __shared__ uint32_t acc[4][kTilePix];
    if (blockIdx.x >= nseg[0]) return;
    const uint4 sg = segments[blockIdx.x];
    for (uint32_t p = threadIdx.x; p < kTilePix; p += blockDim.x) {
        acc[0][p] = 0u;
        acc[1][p] = 0u;
        acc[2][p] = 0u;
        acc[3][p] = 0u;
    }
    __syncthreads();
    for (uint32_t i = sg.y + threadIdx.x; i < sg.z; i += blockDim.x) {
        const u64 v = sorted[i];
        const uint32_t local = uint32_t(v >> 32) & (kTilePix - 1u);
        const uint32_t c = uint32_t(v);
        atomicAdd(&acc[0][local], c & 0xFFu);
        atomicAdd(&acc[1][local], (c >> 8) & 0xFFu);
        atomicAdd(&acc[2][local], (c >> 16) & 0xFFu);
        atomicAdd(&acc[3][local], c >> 24);
    }
    __syncthreads();
    const uint32_t tx = sg.x % tiles_x, ty = sg.x / tiles_x;
    for (uint32_t p = threadIdx.x; p < kTilePix; p += blockDim.x) {
        const uint32_t r = acc[0][p], g = acc[1][p], b = acc[2][p], w = acc[3][p];
        if ((r | g | b | w) == 0u) continue;
        const uint32_t x = (tx << kTileLog2) + (p & 63u), y = (ty << kTileLog2) + (p >> kTileLog2);
        if (x < width && y < height) add_bin(hist, size_t(y) * width + x, r, g, b, w, packed64);
    }
}

}  // namespace

uint32_t queue_tiles_x(uint32_t width) { return (width + 63u) >> kTileLog2; }
uint32_t queue_tiles(uint32_t width, uint32_t height) { return queue_tiles_x(width) * ((height + 63u) >> kTileLog2); }
uint32_t queue_max_segments(size_t records, uint32_t ntiles) {
    return uint32_t((records + kSegment - 1) / kSegment) + ntiles;
}

hipError_t queue_bin(const unsigned long long* records, size_t n, unsigned long long* sorted, uint32_t* counts,
                     uint32_t nparts, void* segments, uint32_t max_segments, uint32_t* nseg, uint32_t width,
                     uint32_t height, uint32_t* hist, bool hist64, bool packed64, hipStream_t s) {
    if (!n) return hipSuccess;
    const uint32_t tiles_x = queue_tiles_x(width), ntiles = queue_tiles(width, height);
    if (ntiles >= kNoTile) return hipErrorInvalidValue;
    // `parts` contiguous partitions of whole 4096-record pieces; the count
    // table is [tile][parts] (at most ntiles * nparts entries)
    const size_t per_part = ((n + nparts - 1) / nparts + kPiece - 1) / kPiece * kPiece;
    const uint32_t parts = uint32_t((n + per_part - 1) / per_part);
    qcount_kernel<<<parts, kPartThreads, ntiles * sizeof(uint32_t), s>>>(records, n, per_part, ntiles, parts,
                                                                          counts);
    qscan_kernel<<<1, 1024, 0, s>>>(counts, ntiles, parts, reinterpret_cast<uint4*>(segments), max_segments, nseg);
    const size_t lds = size_t(3 * ntiles + 1) * sizeof(uint32_t) + kPiece * sizeof(u64) + 8;
    qscatter_kernel<<<parts, kPartThreads, lds, s>>>(records, n, per_part, ntiles, parts, counts, sorted);
    const uint32_t grid = max_segments;
    if (hist64)
        qbin_kernel<<<grid, 1024, 0, s>>>(sorted, reinterpret_cast<const uint4*>(segments), nseg, width, height,
                                          tiles_x, reinterpret_cast<u64*>(hist), false);
    else
        qbin_kernel<<<grid, 1024, 0, s>>>(sorted, reinterpret_cast<const uint4*>(segments), nseg, width, height,
                                          tiles_x, hist, packed64);
    return hipGetLastError();
}

}  // namespace rfk::sk
