// Iterate kernels appended after the generated per-flame code. The generated
// part defines RFK_TPB, RFK_TPB_BITS, RFK_NUM_XFORMS, RFK_HAS_FINAL,
// RFK_USE_CHAOS and the functions rfk_select_xform, rfk_dispatch and
// rfk_precalc_sample.
//
// Scheme (replaces refactor.cu:176-210, 247-343, 388-475):
//  * one workgroup = one temporal sample = RFK_TPB walkers, parameters read
//    through scalar loads from the sample's row of `params`;
//  * every wave applies one xform per pass, drawn from the block's Philox
//    stream (xaos flames pick per walker instead);
//  * after each pass the block's points are permuted through LDS with a
//    counter-keyed bijection so a walker's next xform is decorrelated from its
//    wave neighbours (the reference uses host-built shuffle tables);
//  * every in-bounds draw is four integer atomics into the RGBA u32 bin.
// rfk_iterate runs one launch of `npasses` for every sample; rfk_iterate_batch
// is a persistent kernel that pulls (frame, block) work items from a queue so
// frames with uneven budgets share one launch.
#pragma once

#ifndef RFK_ATOMIC64
#define RFK_ATOMIC64 0
#endif
#ifndef RFK_BIN_TILE_LOG2
#define RFK_BIN_TILE_LOG2 0
#endif
#ifndef RFK_HIST64
#define RFK_HIST64 0
#endif
#ifndef RFK_BIN_QUEUE
#define RFK_BIN_QUEUE 0
#endif

// Queue binning: instead of atomics, every pass of every walker stores one
// 64-bit draw record at queue[block][pass][lane] (coalesced, no atomics) and
// the block counts its records per 64x64-pixel screen tile in LDS; the
// library's queue kernels then counting-sort the records by tile and sum
// them per tile in LDS. Record layout:
//   bits  0..31  R | G << 8 | B << 16 | W << 24   (fixed-point increments)
//   bits 32..43  pixel inside the tile (y * 64 + x)
//   bits 44..63  tile index (ty * tiles_x + tx); all ones = no draw
#define RFK_QTILE_LOG2 6
#define RFK_QEMPTY 0xFFFFFFFFFFFFFFFFull

RFK_INLINE u64 rfk_queue_record(const rfk_draw_t& d, u32 tiles_x) {
    const u32 tile = (d.py >> RFK_QTILE_LOG2) * tiles_x + (d.px >> RFK_QTILE_LOG2);
    const u32 local = ((d.py & 63u) << RFK_QTILE_LOG2) | (d.px & 63u);
    const u32 col = d.r | (d.g << 8) | (d.b << 16) | (d.w << 24);
    return u64(col) | (u64(local) << 32) | (u64(tile) << 44);
}

#if RFK_BIN_QUEUE
// per-block record counts per tile (dynamic LDS, a.ntiles entries)
extern __shared__ u32 rfk_qcount[];

// count a record: one LDS add per lane. Lanes of one tile serialise in the
// LDS, which this VALU-bound kernel has to spare; adding them up in the wave
// first (ballots, a lane count, two branches) cost 2.4 % of a 2^30-iteration
// bin call in VALU and SALU issue.
RFK_INLINE void rfk_queue_count(u32 tile, bool valid) {
    if (valid) atomicAdd(&rfk_qcount[tile], 1u);
}
#endif

// Where a block's queue records and tile counts go (indices into a.queue and
// a.qcounts): rec(j) = first record of its j-th binned pass (RFK_TPB lanes),
// passes() = record passes it owns (those it does not run are left empty),
// cnt(k) = its count of tile k. One type per launch layout, so each kernel
// computes its indices from values it keeps anyway.
// rfk_iterate: part = block, npasses passes each, counts [tile][grid].
struct rfk_qdst_grid {
    u32 block, npasses;
    RFK_INLINE size_t rec(u32 j) const { return ((size_t)block * npasses + j) * RFK_TPB; }
    RFK_INLINE size_t cnt(u32 k) const { return (size_t)k * gridDim.x + block; }
    RFK_INLINE u32 passes() const { return npasses; }
};
// rfk_iterate_qbatch: part = workgroup, q_passes passes each, counts of
// slot g at [g * ntiles + tile][cols], column = the walker block.
struct rfk_qdst_chunk {
    u32 part, q_passes, cnt0, cols;
    RFK_INLINE size_t rec(u32 j) const { return ((size_t)part * q_passes + j) * RFK_TPB; }
    RFK_INLINE size_t cnt(u32 k) const { return cnt0 + (size_t)k * cols; }
    RFK_INLINE u32 passes() const { return q_passes; }
};

// bin element: u32 fixed point (wraps past 2^32 / 255 full-opacity draws in
// one pixel) or u64 for budgets that can put more than that into a pixel
#if RFK_HIST64
typedef u64 hist_t;
#else
typedef u32 hist_t;
#endif

// LDS bin cache: 2^RFK_BIN_TILE_LOG2 direct-mapped slots per block. The first
// bin to claim a free slot owns it until the next flush; its draws accumulate
// with LDS atomics, draws of other bins mapping there go straight to global
// memory. Flushes every RFK_TILE_FLUSH passes and at the end of the block
// keep a slot's u32 sums below 2^32 (RFK_TPB * RFK_TILE_FLUSH * 255).
// Integer sums make the histogram identical with or without the cache.
#define RFK_TILE (1u << RFK_BIN_TILE_LOG2)
#define RFK_TILE_EMPTY 0xFFFFFFFFu
#define RFK_TILE_FLUSH 64u

// Per-wave tables of the wave-uniform Philox draws: rows [0, 16) hold the
// xform picks of 16 groups of 4 passes, rows [16, 48) the permutation keys of
// 32 groups of 2 passes. A wave fills its own rows on the vector unit, one
// group per lane, when its pass leaves the table, and then reads one row per
// pass instead of running Philox on the scalar unit.
#define RFK_PICK_ROWS 16u
#define RFK_PERM_ROWS 32u
struct rfk_shared {
    rfk_pal_lut pal[256];
    Real pts[2][3][RFK_TPB];
    u32x4 rng[RFK_TPB / 64][RFK_PICK_ROWS + RFK_PERM_ROWS];
    u64 hits, draws, bad, start;
    u64 published;  // hits already added to the call's counter (queue binning)
    u32 stop, item;
#if RFK_BIN_TILE_LOG2
    u32 tkey[RFK_TILE];
    u32 tacc[4][RFK_TILE];
#endif
};

#if RFK_BIN_TILE_LOG2
RFK_INLINE u32 rfk_tile_slot(u32 pixel) { return (pixel * 0x9E3779B1u) >> (32 - RFK_BIN_TILE_LOG2); }

RFK_INLINE void rfk_tile_clear(rfk_shared& sm) {
    for (u32 s = threadIdx.x; s < RFK_TILE; s += RFK_TPB) {
        sm.tkey[s] = RFK_TILE_EMPTY;
        sm.tacc[0][s] = 0u;
        sm.tacc[1][s] = 0u;
        sm.tacc[2][s] = 0u;
        sm.tacc[3][s] = 0u;
    }
}

// every thread of the block; LDS must be quiescent (callers sync before)
RFK_INLINE void rfk_tile_flush(rfk_shared& sm, hist_t* __restrict__ hist) {
    for (u32 s = threadIdx.x; s < RFK_TILE; s += RFK_TPB) {
        const u32 key = sm.tkey[s];
        if (key != RFK_TILE_EMPTY) {
            rfk_draw_t d;
            d.pixel = key;
            d.r = sm.tacc[0][s];
            d.g = sm.tacc[1][s];
            d.b = sm.tacc[2][s];
            d.w = sm.tacc[3][s];
            d.hit = true;
            rfk_commit_draw(hist, d);
            sm.tkey[s] = RFK_TILE_EMPTY;
            sm.tacc[0][s] = 0u;
            sm.tacc[1][s] = 0u;
            sm.tacc[2][s] = 0u;
            sm.tacc[3][s] = 0u;
        }
    }
}

RFK_INLINE void rfk_tile_draw(rfk_shared& sm, hist_t* __restrict__ hist, const rfk_draw_t& d) {
    const u32 s = rfk_tile_slot(d.pixel);
    const u32 prev = atomicCAS(&sm.tkey[s], RFK_TILE_EMPTY, d.pixel);
    if (prev == RFK_TILE_EMPTY || prev == d.pixel) {
        atomicAdd(&sm.tacc[0][s], d.r);
        atomicAdd(&sm.tacc[1][s], d.g);
        atomicAdd(&sm.tacc[2][s], d.b);
        atomicAdd(&sm.tacc[3][s], d.w);
    } else {
        rfk_commit_draw(hist, d);
    }
}
#endif


// One block of walkers for `nfuse` un-binned passes followed by `npasses`
// passes (binned when flags has RFK_FLAG_DRAW), starting at pass `pass0`.
//   P, pal:   the sample's parameters and palette (global memory)
//   block:    block id inside its stream (RNG counters), key1: stream id
//   hist:     the frame's histogram
//   init:     seed walkers (Hammersley over `total` walkers) instead of loading
//   wx..wl:   walker state arrays (null: the block keeps its walkers only
//             for this call)
//   q:        queue binning: the block's records and tile counts
// Returns the passes done (bailouts may stop early); counters land in sm.
template <class QDst>
RFK_INLINE u32 rfk_block(const rfk_iter_args& a, rfk_shared& sm, param_ptr P, const u32* pal,
                         u32 block, u32 key1, hist_t* hist, bool init, u32 total, u32 pass0, u32 nfuse, u32 npasses,
                         u32 flags, float* wx, float* wy, float* wc, u32* wl, const QDst& q) {
    const u32 t = threadIdx.x;
    const u32 gid = block * RFK_TPB + t;
    const u32 wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const bool draw = (flags & RFK_FLAG_DRAW) != 0u;
    const bool trace = (flags & RFK_FLAG_TRACE) != 0u;
    const bool bailout = (flags & RFK_FLAG_BAILOUT) != 0u;

    for (u32 i = t; i < 256u; i += RFK_TPB) sm.pal[i] = rfk_pal_lut_entry(pal, i);
    if (t == 0) {
        sm.hits = 0;
        sm.published = 0;
        sm.draws = 0;
        sm.bad = 0;
        sm.stop = 0;
        sm.start = wall_clock64();
#if RFK_BIN_QUEUE
        // a bin call spans several launches: its clock starts with the first
        // block of the first launch, and a stop seen by any block ends it
        if (bailout) {
            atomicCAS(&a.counters[RFK_CNT_ORIGIN], 0ull, sm.start);
            if (__hip_atomic_load(&a.counters[RFK_CNT_STOPPED], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull)
                sm.stop = 1u;
        }
#endif
    }
#if RFK_BIN_TILE_LOG2
    if (draw) rfk_tile_clear(sm);
#endif
    float S[6];
#pragma unroll
    for (int k = 0; k < 6; k++) S[k] = float(P[k]);

    Real x, y, c;
    u32 last = 255u;
    if (init) {
        // Hammersley start positions in pixel space mapped through plane_space
        // (refactor.cu:301-308), colour from the walker's init stream.
        v2 h = rfk_hammersley(gid, total);
        v2 pix = {(h.x + Real(1)) / Real(2) * Real(a.width), (h.y + Real(1)) / Real(2) * Real(a.height)};
        v2 pl = fl::affine(P + 6, pix);
        x = pl.x;
        y = pl.y;
        c = rfk_u01(philox4x32_10({0u, gid, 0u, RFK_TAG_INIT}, a.seed, key1).x);
    } else {
        x = Real(wx[gid]);
        y = Real(wy[gid]);
        c = Real(wc[gid]);
#if RFK_USE_CHAOS
        if (wl) last = wl[gid];
#endif
    }
    __syncthreads();

    rng_t rs;
    rs.init(a.seed, key1, gid, RFK_TAG_LANE);
    u64 hits = 0, draws = 0, bad = 0;
    u32 done = 0, buf = 0;
    u32 pick_base = 0u, perm_base = 0u;  // first group held by the wave's tables
    const u32 lane = __lane_id();

    const u32 ntotal = nfuse + npasses;
#if RFK_BIN_QUEUE
    if (draw)
        for (u32 k = t; k < a.ntiles; k += RFK_TPB) rfk_qcount[k] = 0u;
    __syncthreads();
    const u32 nrun = sm.stop ? 0u : ntotal;
#else
    const u32 nrun = ntotal;
#endif
    for (u32 i = 0; i < nrun; i++) {
        const u32 pass = pass0 + i;
        rs.begin(pass);
        v3 pt = {x, y, c};
        Real op;
        u32 xid;
#if RFK_USE_CHAOS
        // per-walker Markov choice (flame.tpl:125-144): waves run every xform
        // present among their lanes, one uniform xform at a time
        xid = rfk_select_xform_chaos(P, last, rs.rand01());
        op = Real(0);
        v3 res = pt;
        for (u32 xf = 0; xf < RFK_NUM_XFORMS; xf++) {
            if (__builtin_amdgcn_ballot_w64(xid == xf) == 0ull) continue;
            if (xid == xf) {
                v3 q = pt;
                op = rfk_dispatch(xf, P, q, rs);
                res = q;
            }
        }
        pt = res;
        last = xid;
#else
        if (i == 0 || (pass >> 2) - pick_base >= RFK_PICK_ROWS) {
            pick_base = pass >> 2;
            // the wave's earlier reads of these rows are issued before the
            // writes (LDS executes one wave's instructions in order)
            __asm__ volatile("" ::: "memory");
            if (lane < RFK_PICK_ROWS)
                sm.rng[wave][lane] = philox4x32_10({pick_base + lane, block, wave, RFK_TAG_PICK}, a.seed, key1);
            __asm__ volatile("" ::: "memory");
        }
        const u32x4 prow = sm.rng[wave][(pass >> 2) - pick_base];
        const u32 pj = pass & 3u;
        const u32 pick = __builtin_amdgcn_readfirstlane(pj == 0 ? prow.x : pj == 1 ? prow.y : pj == 2 ? prow.z : prow.w);
        xid = rfk_select_xform(P, rfk_u01(pick));
        xid = __builtin_amdgcn_readfirstlane(xid);
        op = rfk_dispatch(xid, P, pt, rs);
#endif
        if (rfk_badvalue(pt.x) || rfk_badvalue(pt.y)) {
            // reset to a uniform point in [-1,1]^2 with zero opacity (refactor.cu:195-199)
            pt.x = rs.rand01() * Real(2) - Real(1);
            pt.y = rs.rand01() * Real(2) - Real(1);
            op = Real(0);
            bad++;
        }
        if (draw && i >= nfuse) {
            v3 d = pt;
#if RFK_HAS_FINAL
            // the final xform acts on a copy; opacity stays the chosen xform's (refactor.cu:351-354)
            rfk_dispatch(RFK_NUM_XFORMS, P, d, rs);
#endif
            rfk_draw_t dr = rfk_make_draw(S, sm.pal, a.width, a.height, float(d.x), float(d.y), float(d.z), float(op));
#if RFK_BIN_QUEUE
            u64 rec = RFK_QEMPTY;
#endif
            if (dr.hit) {
#if RFK_BIN_QUEUE
                // opacities above 1 overflow the 8-bit fields: bin those directly
                if ((dr.r | dr.g | dr.b | dr.w) <= 255u) rec = rfk_queue_record(dr, a.tiles_x);
                else rfk_commit_draw(hist, dr);
#elif RFK_BIN_TILE_LOG2
                rfk_tile_draw(sm, hist, dr);
#else
                rfk_commit_draw(hist, dr);
#endif
                hits += dr.w;
                draws++;
            }
#if RFK_BIN_QUEUE
            a.queue[q.rec(i - nfuse) + t] = rec;
            rfk_queue_count(u32(rec >> 44), rec != RFK_QEMPTY);
#endif
        }
#if RFK_BIN_TILE_LOG2
        if (draw && ((i + 1u) % RFK_TILE_FLUSH) == 0u) {
            __syncthreads();
            rfk_tile_flush(sm, hist);
            __syncthreads();
        }
#endif
        if (trace) {
            float* tr = a.trace + ((size_t)i * a.total_walkers + gid) * 4u;
            tr[0] = float(pt.x);
            tr[1] = float(pt.y);
            tr[2] = float(pt.z);
            tr[3] = float(op);
            if (a.trace_xid) {
#if RFK_USE_CHAOS
                a.trace_xid[(size_t)i * a.total_walkers + gid] = xid;
#else
                if ((t & 63u) == 0u) a.trace_xid[(size_t)i * (a.total_walkers / 64u) + gid / 64u] = xid;
#endif
            }
        }
#if RFK_USE_CHAOS
        x = pt.x;
        y = pt.y;
        c = pt.z;
#else
        if (i == 0 || (pass >> 1) - perm_base >= RFK_PERM_ROWS) {
            perm_base = pass >> 1;
            __asm__ volatile("" ::: "memory");
            if (lane < RFK_PERM_ROWS)
                sm.rng[wave][RFK_PICK_ROWS + lane] =
                    philox4x32_10({perm_base + lane, block, 0u, RFK_TAG_PERM}, a.seed, key1);
            __asm__ volatile("" ::: "memory");
        }
        const u32x4 krow = sm.rng[wave][RFK_PICK_ROWS + (pass >> 1) - perm_base];
        const u32 wa = __builtin_amdgcn_readfirstlane((pass & 1u) ? krow.z : krow.x);
        const u32 wb = __builtin_amdgcn_readfirstlane((pass & 1u) ? krow.w : krow.y);
        const u32 slot = rfk_permute(t, wa, wb, RFK_TPB_BITS);
        sm.pts[buf][0][slot] = pt.x;
        sm.pts[buf][1][slot] = pt.y;
        sm.pts[buf][2][slot] = pt.z;
        __syncthreads();
        x = sm.pts[buf][0][t];
        y = sm.pts[buf][1][t];
        c = sm.pts[buf][2][t];
        buf ^= 1u;
#endif
        done++;

        if (bailout && ((i + 1u) & (a.check_every - 1u)) == 0u) {
            // bailouts (refactor.cu:443-458): quality target per block, time
            // limit on the device clock, host stop flag in coherent host memory
            atomicAdd(&sm.hits, hits);
            hits = 0;
            __syncthreads();
            if (t == 0) {
                bool stop = a.stop_flag &&
                            __hip_atomic_load((u32*)a.stop_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
#if RFK_BIN_QUEUE
                // whole-call criteria: clock since the call's first block,
                // hits of finished blocks plus this block's, any block stopped
                const u64 origin = __hip_atomic_load(&a.counters[RFK_CNT_ORIGIN], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                if (a.time_limit_ticks && wall_clock64() - origin >= a.time_limit_ticks) stop = true;
                // publish this block's new hits and compare the call's running total
                const u64 delta = sm.hits - sm.published;
                sm.published = sm.hits;
                const u64 total = atomicAdd(&a.counters[RFK_CNT_HITS], delta) + delta;
                if (a.quality_target && total >= a.quality_target) stop = true;
                if (__hip_atomic_load(&a.counters[RFK_CNT_STOPPED], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull)
                    stop = true;
#else
                if (a.time_limit_ticks && wall_clock64() - sm.start >= a.time_limit_ticks) stop = true;
                if (a.quality_target && sm.hits >= a.quality_target) stop = true;
#endif
                sm.stop = stop ? 1u : 0u;
            }
            __syncthreads();
            if (sm.stop) break;
        }
    }
#if RFK_BIN_TILE_LOG2
    if (draw) {
        __syncthreads();
        rfk_tile_flush(sm, hist);
    }
#endif
#if RFK_BIN_QUEUE
    // passes a bailout skipped leave no draws behind
    if (draw) {
        for (u32 j = (done > nfuse ? done - nfuse : 0u); j < q.passes(); j++) a.queue[q.rec(j) + t] = RFK_QEMPTY;
        __syncthreads();
        for (u32 k = t; k < a.ntiles; k += RFK_TPB) a.qcounts[q.cnt(k)] = rfk_qcount[k];
    }
#endif

    if (wx) {
        wx[gid] = float(x);
        wy[gid] = float(y);
        wc[gid] = float(c);
#if RFK_USE_CHAOS
        if (wl) wl[gid] = last;
#endif
    }
    atomicAdd(&sm.hits, hits);
    if (draws) atomicAdd(&sm.draws, draws);
    if (bad) atomicAdd(&sm.bad, bad);
    __syncthreads();
    return done;
}

extern "C" __global__ void rfk_precalc(rfk_iter_args a) {
    const u32 s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nsamples) return;
    Real* P = (Real*)a.params + (size_t)s * a.param_stride;
    rng_t rs;
    rs.init(a.seed, a.row_stream ? a.row_stream[s] : a.stream, 0u, RFK_TAG_JITTER);
    rs.begin(a.row_local ? a.row_local[s] : s);
    rfk_precalc_sample(P, rs);
}

// f32 iterate kernels are held to 64 VGPRs: eight waves per SIMD hide the
// per-pass barrier and LDS latency better than seven (measured: the 1080p
// 2^30-iteration bin call 15.6 -> 15.2 ms)
#ifdef RFK_DOUBLE
#define RFK_WAVES_PER_SIMD 1
#else
#define RFK_WAVES_PER_SIMD 8
#endif
extern "C" __global__ __launch_bounds__(RFK_TPB, RFK_WAVES_PER_SIMD) void rfk_iterate(rfk_iter_args a) {
    __shared__ rfk_shared sm;
    const u32 block = blockIdx.x;
    const param_ptr P = (param_ptr)((const Real*)a.params + (size_t)block * a.param_stride);
    const bool init = (a.flags & RFK_FLAG_INIT) != 0u;
    const u32 pass0 = init ? 0u : a.block_pass[block];
    const rfk_qdst_grid q = {block, a.npasses};
    const u32 done = rfk_block(a, sm, P, a.palettes + (size_t)block * 256u, block, a.stream, (hist_t*)a.hist, init,
                               a.total_walkers, pass0, 0u, a.npasses, a.flags, a.wx, a.wy, a.wc, a.wlast, q);
    if (threadIdx.x == 0) {
        a.block_pass[block] = pass0 + done;
        atomicAdd(&a.counters[RFK_CNT_HITS], sm.hits - sm.published);
        atomicAdd(&a.counters[RFK_CNT_PASSES], (u64)done * RFK_TPB);
        atomicMin(&a.counters[RFK_CNT_START], sm.start);
        atomicMax(&a.counters[RFK_CNT_STOP], wall_clock64());
        if (sm.stop) atomicAdd(&a.counters[RFK_CNT_STOPPED], 1ull);
        if (sm.draws) atomicAdd(&a.counters[RFK_CNT_DRAWS], sm.draws);
        if (sm.bad) atomicAdd(&a.counters[RFK_CNT_BAD], sm.bad);
    }
}

#if !RFK_BIN_QUEUE
// Persistent animation kernel: work item i (popped from a.queue in the
// host-chosen order, longest first) is block b of frame f. Each item seeds its
// walkers, runs the frame's fuse passes without drawing and then the frame's
// binned passes into the frame's histogram; streams are keyed by (seed, f).
extern "C" __global__ __launch_bounds__(RFK_TPB) void rfk_iterate_batch(rfk_iter_args a, rfk_batch_args b) {
    __shared__ rfk_shared sm;
    const rfk_qdst_grid q = {0u, 0u};  // unused: atomic binning
    for (;;) {
        if (threadIdx.x == 0) sm.item = atomicAdd(b.queue, 1u);
        __syncthreads();
        const u32 item = sm.item;
        __syncthreads();
        if (item >= b.nitems) break;  // uniform: every wave of the group leaves together
        const u32 f = b.item_frame[item], blk = b.item_block[item];
        const u32 row = b.frame_row0[f] + blk;
        const param_ptr P = (param_ptr)((const Real*)a.params + (size_t)row * a.param_stride);
        const u32* pal = a.palettes + (size_t)row * 256u;
        hist_t* hist = (hist_t*)a.hist + (size_t)f * b.frame_pixels * 4u;
        const u32 total = b.frame_blocks[f] * RFK_TPB;
        const u32 done = rfk_block(a, sm, P, pal, blk, f, hist, true, total, 0u, b.fuse, b.frame_passes[f],
                                   RFK_FLAG_DRAW, nullptr, nullptr, nullptr, nullptr, q) - b.fuse;
        if (threadIdx.x == 0) {
            u64* fc = b.frame_counters + (size_t)f * 4u;
            atomicAdd(&fc[0], sm.hits);
            atomicAdd(&fc[1], sm.draws);
            atomicAdd(&fc[2], (u64)done * RFK_TPB);
            atomicAdd(&fc[3], sm.bad);
        }
        __syncthreads();
    }
}
#else
// Animation batch with queue binning, one launch per chunk. Block b runs
// slot q.slot0 + b / q.cols (a frame and its next slice of passes: slots
// hold frame, first pass, binned passes, fuse passes) on walker block
// b % q.cols of that frame, whose walkers persist between slices in
// a.wx/wy/wc at [frame][block][lane]. Its records go to part b of the chunk
// (q.passes passes each) and its tile counts to the rows of slot b / q.cols:
// chunk tiles [(b / q.cols) * ntiles, ...). Streams are keyed by (seed,
// frame) and passes are numbered as in one uninterrupted render, so every
// frame equals the atomic batch's.
extern "C" __global__ __launch_bounds__(RFK_TPB, RFK_WAVES_PER_SIMD) void rfk_iterate_qbatch(rfk_iter_args a,
                                                                                              rfk_qbatch_args q) {
    __shared__ rfk_shared sm;
    const u32 g = blockIdx.x / q.cols, blk = blockIdx.x - g * q.cols;
    const u32* sl = q.slots + (size_t)(q.slot0 + g) * 4u;
    const u32 f = sl[0], pass0 = sl[1], npasses = sl[2], nfuse = sl[3];
    const u32 row = f * q.cols + blk;
    const param_ptr P = (param_ptr)((const Real*)a.params + (size_t)row * a.param_stride);
    const u32* pal = a.palettes + (size_t)row * 256u;
    hist_t* hist = (hist_t*)a.hist + (size_t)f * q.frame_pixels * 4u;
    const size_t w0 = (size_t)f * q.cols * RFK_TPB;
    const rfk_qdst_chunk qd = {blockIdx.x, q.passes, g * a.ntiles * q.cols + blk, q.cols};
    const u32 done = rfk_block(a, sm, P, pal, blk, f, hist, pass0 == 0u, q.cols * RFK_TPB, pass0, nfuse, npasses,
                               RFK_FLAG_DRAW, a.wx + w0, a.wy + w0, a.wc + w0, nullptr, qd) - nfuse;
    if (threadIdx.x == 0) {
        u64* fc = q.frame_counters + (size_t)f * 4u;
        atomicAdd(&fc[0], sm.hits);
        atomicAdd(&fc[1], sm.draws);
        atomicAdd(&fc[2], (u64)done * RFK_TPB);
        atomicAdd(&fc[3], sm.bad);
    }
}
#endif
