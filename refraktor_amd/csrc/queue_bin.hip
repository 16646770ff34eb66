// Queue binning: the iterate kernel stores one 64-bit record per walker and
// pass (layout in flame_kernels.h), block-major: block b's records of a chunk
// are contiguous, and the block counts its records per 64x64-pixel screen
// tile in LDS (counts [tile][block]). These kernels fold a chunk of records
// into the histogram without contended global atomics:
//
//   1. qscan_col  per tile: exclusive prefix of its counts over the blocks
//                 (in place) and the tile's total
//   2. qseg       exclusive scan of the tile totals -> tile starts, and the
//                 (tile, segment of <= 2^17 or 2^18 records) work items of step 4
//   3. qscatter   counting sort by tile: each block's records, 4096 at a time,
//                 ranked in LDS and written as tile runs at their offsets,
//                 colour (u32) and pixel in tile (u16) in separate arrays
//   4. qbin       one workgroup per work item sums its records in a 64x64
//                 LDS tile, then adds the non-zero pixels into the histogram:
//                 plain read-add-write when the work item is the tile's only
//                 one, one atomic per channel otherwise
//
// Several frames in one chunk (animation batches): the parts come in groups
// of `ncols`, group g holding only records of one frame (group_frame[g]),
// whose tiles are counted as chunk tiles [g * tpf, (g + 1) * tpf); the
// counts are [chunk tile][ncols]. A single frame is one group of all parts.
//
// All integer sums: the histogram equals the one direct atomics give, bit
// for bit, whatever the order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "static_kernels.hpp"

namespace rfk::sk {

namespace {

typedef unsigned long long u64;

constexpr uint32_t kTileLog2 = 6;  // 64x64-pixel tiles
constexpr uint32_t kTilePix = 1u << (2 * kTileLog2);
constexpr uint32_t kNoTile = 0xFFFFFu;
constexpr uint32_t kThreads = 1024;
// the counting sort: 512 threads rank 4096 records at a time, so that
// several sorting workgroups share a CU and one's barriers overlap another's
// memory traffic (2.5 % off a 2^30-iteration bin call against 1024 / 8192)
constexpr uint32_t kScatterThreads = 512, kScatterPiece = 4096;
// records per bin work item (u32 LDS sums cannot wrap): 2^18 for chunks
// of more than 2^26 records, 2^17 below, where the hottest tiles' work
// items would otherwise leave most CUs idle at the end (a 2^25-record
// denoised frame's tile sums 107 -> 90 us; 2^17 at 2^27-record chunks
// costs more in shared-tile atomics than it balances: -0.8 % plain frames)
constexpr uint32_t kSegmentLog2Max = 18, kSegmentLog2Min = 17;
constexpr uint64_t kSmallChunk = 1ull << 26;
__host__ __device__ constexpr uint32_t segment_log2(uint64_t records) {
    return records <= kSmallChunk ? kSegmentLog2Min : kSegmentLog2Max;
}
// copies of a work item's LDS tile sums: lane l adds into copy l % kCopies,
// so the lanes of one wave that draw into the same pixel (a flame's hottest
// pixels take 10-13 % of their tile's records) split over kCopies addresses
// and their LDS atomics serialise half as deep (the hottest tile's work items
// of a 2^25-iteration 1080p frame 125 -> 80 us)
constexpr uint32_t kCopies = 2;
constexpr size_t kBinLds = size_t(kCopies) * 2 * kTilePix * sizeof(unsigned long long);

__device__ __forceinline__ uint32_t rec_tile(u64 r) { return uint32_t(r >> 44); }

// Block-wide exclusive scan of one value per thread (blockDim.x = T).
template <uint32_t T = kThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t* total) {
    uint32_t incl = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(incl, d);
        if (__lane_id() >= d) incl += u;
    }
    if (__lane_id() == 63) s_wave[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < T / 64; w++) {
        const uint32_t s = s_wave[w];
        if (w < (threadIdx.x >> 6)) base += s;
        all += s;
    }
    __syncthreads();
    if (total) *total = all;
    return base + incl - v;
}

// grid = ntiles: row t of counts [tile][ncols] -> exclusive prefix in
// place, tile_total[t] = its sum.
__global__ void __launch_bounds__(kThreads)
    qscan_col_kernel(uint32_t* __restrict__ counts, uint32_t ncols, uint32_t* __restrict__ tile_total) {
    __shared__ uint32_t s_wave[kThreads / 64];
    uint32_t* col = counts + size_t(blockIdx.x) * ncols;
    const uint32_t per = (ncols + kThreads - 1) / kThreads;
    const uint32_t lo = threadIdx.x * per, hi = std::min(ncols, lo + per);
    uint32_t sum = 0;
    for (uint32_t i = lo; i < hi; i++) sum += col[i];
    uint32_t total = 0;
    uint32_t run = block_exclusive_scan(sum, s_wave, &total);
    for (uint32_t i = lo; i < hi; i++) {
        const uint32_t c = col[i];
        col[i] = run;
        run += c;
    }
    if (threadIdx.x == 0) tile_total[blockIdx.x] = total;
}

// One workgroup: tile_start = exclusive scan of tile_total; work items
// (tile, begin, end, exclusive) of at most 2^seg_log2 records, nseg[0] =
// their count. The full items come first and the tiles' shorter remainders
// after them, so the workgroups that start last are the short ones (longest
// first).
__global__ void __launch_bounds__(kThreads)
    qseg_kernel(const uint32_t* __restrict__ tile_total, uint32_t ntiles, uint32_t* __restrict__ tile_start,
                uint4* __restrict__ segments, uint32_t max_segments, uint32_t* __restrict__ nseg,
                uint32_t seg_log2) {
    __shared__ uint32_t s_wave[kThreads / 64];
    const uint32_t seg = 1u << seg_log2;
    const uint32_t per = (ntiles + kThreads - 1) / kThreads;
    const uint32_t lo = threadIdx.x * per, hi = std::min(ntiles, lo + per);
    uint32_t sum = 0, nfull = 0, nrem = 0;
    for (uint32_t t = lo; t < hi; t++) {
        sum += tile_total[t];
        nfull += tile_total[t] / seg;
        nrem += (tile_total[t] % seg) != 0u;
    }
    uint32_t run = block_exclusive_scan(sum, s_wave, nullptr);
    uint32_t full_all = 0, rem_all = 0;
    uint32_t kf = block_exclusive_scan(nfull, s_wave, &full_all);
    uint32_t kr = full_all + block_exclusive_scan(nrem, s_wave, &rem_all);
    for (uint32_t t = lo; t < hi; t++) {
        const uint32_t n = tile_total[t];
        tile_start[t] = run;
        const uint32_t nf = n / seg, exclusive = (n <= seg) ? 1u : 0u;
        for (uint32_t s = 0; s < nf; s++, kf++)
            if (kf < max_segments) segments[kf] = make_uint4(t, run + s * seg, run + (s + 1) * seg, exclusive);
        if (n % seg) {
            if (kr < max_segments) segments[kr] = make_uint4(t, run + nf * seg, run + n, exclusive);
            kr++;
        }
        run += n;
    }
    if (threadIdx.x == 0) nseg[0] = std::min(full_all + rem_all, max_segments);
}

// grid = nparts: partition p = records [p * per_part, (p + 1) * per_part),
// column p % ncols of the counts of group p / ncols, whose ntiles tiles (the
// records' tile fields) are chunk tiles from (p / ncols) * ntiles on.
// LDS: cursor[ntiles], lcount[ntiles], lstart[ntiles], piece[PIECE]. T
// threads sort PIECE records at a time.
template <uint32_t T, uint32_t PIECE>
__global__ void __launch_bounds__(T)
    qscatter_kernel(const u64* __restrict__ rec, size_t per_part, uint32_t ntiles, uint32_t ncols,
                    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ tile_start,
                    uint32_t* __restrict__ sorted_col, uint16_t* __restrict__ sorted_loc) {
    extern __shared__ uint32_t smem[];
    uint32_t* cursor = smem;
    uint32_t* lcount = cursor + ntiles;
    uint32_t* lstart = lcount + ntiles;
    u64* piece = reinterpret_cast<u64*>(lstart + ntiles + (ntiles & 1u));
    __shared__ uint32_t s_wave[T / 64];
    const uint32_t group = blockIdx.x / ncols, col = blockIdx.x - group * ncols;
    const uint32_t tile0 = group * ntiles;
    for (uint32_t t = threadIdx.x; t < ntiles; t += T)
        cursor[t] = tile_start[tile0 + t] + offsets[size_t(tile0 + t) * ncols + col];
    const size_t b0 = size_t(blockIdx.x) * per_part, b1 = b0 + per_part;
    constexpr uint32_t kPer = PIECE / T;
    const uint32_t per = (ntiles + T - 1) / T;
    const uint32_t lo = threadIdx.x * per, hi = std::min(ntiles, lo + per);
    // all loads of a piece are issued before any of them is used, and the
    // next piece's loads are in flight while this one is ranked and written
    u64 nxt[kPer];
    auto load_piece = [&](size_t base) {
#pragma unroll
        for (uint32_t j = 0; j < kPer; j++) {
            const size_t i = base + size_t(j) * T + threadIdx.x;
            nxt[j] = i < b1 ? rec[i] : ~0ull;
        }
    };
    load_piece(b0);
    for (size_t base = b0; base < b1; base += PIECE) {
        for (uint32_t t = threadIdx.x; t < ntiles; t += T) lcount[t] = 0;
        u64 r[kPer];
#pragma unroll
        for (uint32_t j = 0; j < kPer; j++) r[j] = nxt[j];
        if (base + PIECE < b1) load_piece(base + PIECE);
        __syncthreads();
        // a record's rank among its tile's records in this piece, one LDS
        // add with return each (which of a tile's records comes first does
        // not matter: the tile's sums are integer)
        uint32_t rank[kPer];
#pragma unroll
        for (uint32_t j = 0; j < kPer; j++) {
            const uint32_t tile = rec_tile(r[j]);
            rank[j] = tile < ntiles ? atomicAdd(&lcount[tile], 1u) : 0u;
        }
        __syncthreads();
        uint32_t sum = 0;
        for (uint32_t t = lo; t < hi; t++) sum += lcount[t];
        uint32_t used = 0;
        uint32_t run = block_exclusive_scan<T>(sum, s_wave, &used);
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
        // sorted order: consecutive threads write consecutive addresses of a run
        for (uint32_t k = threadIdx.x; k < used; k += T) {
            const u64 v = piece[k];
            const uint32_t tile = rec_tile(v);
            const size_t pos = size_t(cursor[tile]) + (k - lstart[tile]);
            sorted_col[pos] = uint32_t(v);
            sorted_loc[pos] = uint16_t((v >> 32) & (kTilePix - 1u));
        }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < ntiles; t += T) cursor[t] += lcount[t];
    }
}

template <class HistT>
__device__ __forceinline__ void add_bin(HistT* __restrict__ hist, size_t p, uint32_t r, uint32_t g, uint32_t b,
                                        uint32_t w, bool packed64, bool exclusive) {
    HistT* h = hist + 4 * p;
    if (exclusive) {
        // the only work item of its tile: nobody else touches these bins now
        h[0] += HistT(r);
        h[1] += HistT(g);
        h[2] += HistT(b);
        h[3] += HistT(w);
        return;
    }
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

// LDS sums of a work item: (r | g << 32) and (b | w << 32) per pixel, in
// kCopies copies (dynamic LDS). A work item holds at most 2^18 records of
// 8-bit channels, so each 32-bit half stays below 2^26 and never carries into
// its neighbour. Chunk tile t is
// tile t % tpf of frame group_frame[t / tpf] (frame 0 without a list), whose
// bins start frame_elems elements after the previous frame's.
template <class HistT>
__global__ void __launch_bounds__(kThreads)
    qbin_kernel(const uint32_t* __restrict__ sorted_col, const uint16_t* __restrict__ sorted_loc,
                const uint4* __restrict__ segments, const uint32_t* __restrict__ nseg,
                uint32_t width, uint32_t height, uint32_t tiles_x, uint32_t tpf,
                const uint32_t* __restrict__ group_frame, size_t frame_elems, HistT* __restrict__ hist,
                bool packed64) {
    static_assert(kSegmentLog2Max <= 24, "32-bit halves of the LDS sums must not carry");
    extern __shared__ u64 qbin_lds[];
    u64(*copies)[2][kTilePix] = reinterpret_cast<u64(*)[2][kTilePix]>(qbin_lds);
    u64(*acc)[kTilePix] = copies[__lane_id() % kCopies];
    if (blockIdx.x >= nseg[0]) return;
    const uint4 sg = segments[blockIdx.x];
    for (uint32_t p = threadIdx.x; p < kCopies * 2 * kTilePix; p += kThreads) qbin_lds[p] = 0ull;
    __syncthreads();
    // U records per thread in flight before the first LDS add
    constexpr uint32_t U = 8;
    for (uint32_t i0 = sg.y + threadIdx.x; i0 < sg.z; i0 += kThreads * U) {
        uint32_t col[U], loc[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = i0 + u * kThreads;
            col[u] = i < sg.z ? sorted_col[i] : 0u;
            loc[u] = i < sg.z ? sorted_loc[i] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            if (i0 + u * kThreads >= sg.z) break;
            const uint32_t local = loc[u];
            const uint32_t c = col[u];
            atomicAdd(&acc[0][local], u64(c & 0xFFu) | (u64((c >> 8) & 0xFFu) << 32));
            atomicAdd(&acc[1][local], u64((c >> 16) & 0xFFu) | (u64(c >> 24) << 32));
        }
    }
    __syncthreads();
    const uint32_t group = sg.x / tpf, tile = sg.x - group * tpf;
    if (group_frame) hist += size_t(group_frame[group]) * frame_elems;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const bool exclusive = sg.w != 0u;
    for (uint32_t p = threadIdx.x; p < kTilePix; p += kThreads) {
        u64 rg = 0ull, bw = 0ull;
#pragma unroll
        for (uint32_t c = 0; c < kCopies; c++) {
            rg += copies[c][0][p];
            bw += copies[c][1][p];
        }
        if ((rg | bw) == 0ull) continue;
        const uint32_t r = uint32_t(rg), g = uint32_t(rg >> 32), b = uint32_t(bw), w = uint32_t(bw >> 32);
        const uint32_t x = (tx << kTileLog2) + (p & 63u), y = (ty << kTileLog2) + (p >> kTileLog2);
        if (x < width && y < height) add_bin(hist, size_t(y) * width + x, r, g, b, w, packed64, exclusive);
    }
}

}  // namespace

uint32_t queue_tiles_x(uint32_t width) { return (width + 63u) >> kTileLog2; }
uint32_t queue_tiles(uint32_t width, uint32_t height) { return queue_tiles_x(width) * ((height + 63u) >> kTileLog2); }
// work items of the smallest size, whatever a call of up to `records`
// records uses
uint32_t queue_max_segments(size_t records, uint32_t ntiles) {
    constexpr size_t seg = size_t(1) << kSegmentLog2Min;
    return uint32_t((records + seg - 1) / seg) + ntiles;
}

hipError_t queue_bin_groups(const unsigned long long* records, size_t per_part, uint32_t nparts, uint32_t ncols,
                            unsigned long long* sorted, uint32_t* counts, uint32_t* tile_scratch, void* segments,
                            uint32_t max_segments, uint32_t* nseg, uint32_t width, uint32_t height,
                            const uint32_t* group_frame, size_t frame_elems, uint32_t* hist, bool hist64,
                            bool packed64, hipStream_t s) {
    if (!per_part || !nparts) return hipSuccess;
    if (!ncols || nparts % ncols) return hipErrorInvalidValue;
    const uint32_t tiles_x = queue_tiles_x(width), tpf = queue_tiles(width, height);
    const uint64_t ntiles = uint64_t(tpf) * (nparts / ncols);  // chunk tiles
    if (tpf >= kNoTile || ntiles >= (1ull << 31)) return hipErrorInvalidValue;
    uint32_t* tile_total = tile_scratch;
    uint32_t* tile_start = tile_scratch + ntiles;
    qscan_col_kernel<<<uint32_t(ntiles), kThreads, 0, s>>>(counts, ncols, tile_total);
    qseg_kernel<<<1, kThreads, 0, s>>>(tile_total, uint32_t(ntiles), tile_start, reinterpret_cast<uint4*>(segments),
                                       max_segments, nseg, segment_log2(uint64_t(per_part) * nparts));
    const size_t lds = size_t(3 * tpf + 1) * sizeof(uint32_t) + kScatterPiece * sizeof(u64) + 8;
    // more than 64 KiB of dynamic LDS (gfx950 has 160 KiB per workgroup)
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&qscatter_kernel<kScatterThreads, kScatterPiece>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    // sorted records, colour and pixel-in-tile apart (the tile is the work
    // item's): 6 of the 8 bytes per record, in the same allocation
    const size_t n = per_part * nparts;
    uint32_t* sorted_col = reinterpret_cast<uint32_t*>(sorted);
    uint16_t* sorted_loc = reinterpret_cast<uint16_t*>(sorted_col + n);
    qscatter_kernel<kScatterThreads, kScatterPiece><<<nparts, kScatterThreads, lds, s>>>(
        records, per_part, tpf, ncols, counts, tile_start, sorted_col, sorted_loc);
    const void* bin_fn = hist64 ? reinterpret_cast<const void*>(&qbin_kernel<u64>)
                                : reinterpret_cast<const void*>(&qbin_kernel<uint32_t>);
    e = hipFuncSetAttribute(bin_fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(kBinLds));
    if (e != hipSuccess) return e;
    if (hist64)
        qbin_kernel<<<max_segments, kThreads, kBinLds, s>>>(sorted_col, sorted_loc, reinterpret_cast<const uint4*>(segments),
                                                      nseg, width, height, tiles_x, tpf, group_frame, frame_elems,
                                                      reinterpret_cast<u64*>(hist), false);
    else
        qbin_kernel<<<max_segments, kThreads, kBinLds, s>>>(sorted_col, sorted_loc, reinterpret_cast<const uint4*>(segments),
                                                      nseg, width, height, tiles_x, tpf, group_frame, frame_elems,
                                                      hist, packed64);
    return hipGetLastError();
}

hipError_t queue_bin(const unsigned long long* records, size_t per_part, uint32_t nparts, unsigned long long* sorted,
                     uint32_t* counts, uint32_t* tile_scratch, void* segments, uint32_t max_segments, uint32_t* nseg,
                     uint32_t width, uint32_t height, uint32_t* hist, bool hist64, bool packed64, hipStream_t s) {
    return queue_bin_groups(records, per_part, nparts, nparts, sorted, counts, tile_scratch, segments, max_segments,
                            nseg, width, height, nullptr, 0, hist, hist64, packed64, s);
}

}  // namespace rfk::sk
