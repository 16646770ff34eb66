// Iterate kernels appended after the generated per-flame code. The generated
// part defines RFK_TPB, RFK_TPB_BITS, RFK_NUM_XFORMS, RFK_HAS_FINAL,
// RFK_USE_CHAOS and the functions rfk_select_xform, rfk_dispatch and
// rfk_precalc_sample.
//
// Scheme (replaces refactor.cu:176-210, 247-343, 388-475):
//  * one block = one temporal sample = RFK_TPB walkers, parameters read
//    through scalar loads from the block's row of `params`;
//  * every wave applies one xform per pass, drawn from the block's Philox
//    stream (xaos flames pick per walker instead);
//  * after each pass the block's points are permuted through LDS with a
//    counter-keyed bijection so a walker's next xform is decorrelated from its
//    wave neighbours (the reference uses host-built shuffle tables);
//  * every in-bounds draw is four integer atomics into the RGBA u32 bin.
#pragma once

#ifndef RFK_ATOMIC64
#define RFK_ATOMIC64 0
#endif

RFK_INLINE u32 rfk_wave_pick(u32 pass, u32 block, u32 wave, u32 k0, u32 k1, u32x4& cache, bool refresh) {
    if (refresh) cache = philox4x32_10({pass >> 2, block, wave, RFK_TAG_PICK}, k0, k1);
    const u32 j = pass & 3u;
    return j == 0 ? cache.x : j == 1 ? cache.y : j == 2 ? cache.z : cache.w;
}

extern "C" __global__ void rfk_precalc(rfk_iter_args a) {
    const u32 s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nsamples) return;
    Real* P = (Real*)a.params + (size_t)s * a.param_stride;
    rng_t rs;
    rs.init(a.seed, a.stream, 0u, RFK_TAG_JITTER);
    rs.begin(s);
    rfk_precalc_sample(P, rs);
}

extern "C" __global__ __launch_bounds__(RFK_TPB) void rfk_iterate(rfk_iter_args a) {
    __shared__ u32 s_pal[256];
    __shared__ Real s_pts[2][3][RFK_TPB];
    __shared__ u64 s_hits, s_draws, s_bad;
    __shared__ u32 s_stop;
    __shared__ u64 s_start;

    const u32 t = threadIdx.x;
    const u32 block = blockIdx.x;
    const u32 gid = block * RFK_TPB + t;
    const u32 wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const Real* __restrict__ P = (const Real*)a.params + (size_t)block * a.param_stride;
    const bool draw = (a.flags & RFK_FLAG_DRAW) != 0u;
    const bool trace = (a.flags & RFK_FLAG_TRACE) != 0u;
    const bool bailout = (a.flags & RFK_FLAG_BAILOUT) != 0u;

    for (u32 i = t; i < 256u; i += RFK_TPB) s_pal[i] = a.palettes[(size_t)block * 256u + i];
    if (t == 0) {
        s_hits = 0;
        s_draws = 0;
        s_bad = 0;
        s_stop = 0;
        s_start = wall_clock64();
    }
    float S[6];
#pragma unroll
    for (int k = 0; k < 6; k++) S[k] = float(P[k]);

    u32 pass0;
    Real x, y, c;
    u32 last = 255u;
    if (a.flags & RFK_FLAG_INIT) {
        // Hammersley start positions in pixel space mapped through plane_space
        // (refactor.cu:301-308), colour from the walker's init stream.
        pass0 = 0u;
        v2 h = rfk_hammersley(gid, a.total_walkers);
        v2 pix = {(h.x + Real(1)) / Real(2) * Real(a.width), (h.y + Real(1)) / Real(2) * Real(a.height)};
        v2 pl = fl::affine(P + 6, pix);
        x = pl.x;
        y = pl.y;
        c = rfk_u01(philox4x32_10({0u, gid, 0u, RFK_TAG_INIT}, a.seed, a.stream).x);
    } else {
        pass0 = a.block_pass[block];
        x = Real(a.wx[gid]);
        y = Real(a.wy[gid]);
        c = Real(a.wc[gid]);
    }
#if RFK_USE_CHAOS
    if (!(a.flags & RFK_FLAG_INIT) && a.wlast) last = a.wlast[gid];
#endif
    __syncthreads();

    rng_t rs;
    rs.init(a.seed, a.stream, gid, RFK_TAG_LANE);
    u64 hits = 0, draws = 0, bad = 0;
    u32 done = 0, buf = 0;
    u32x4 pick_cache = {0u, 0u, 0u, 0u}, perm_cache = {0u, 0u, 0u, 0u};

    for (u32 i = 0; i < a.npasses; i++) {
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
        xid = rfk_select_xform(P, rfk_wave_pick(pass, block, wave, a.seed, a.stream, pick_cache, i == 0 || (pass & 3u) == 0u));
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
        if (draw) {
            v3 d = pt;
#if RFK_HAS_FINAL
            // the final xform acts on a copy; opacity stays the chosen xform's (refactor.cu:351-354)
            rfk_dispatch(RFK_NUM_XFORMS, P, d, rs);
#endif
            rfk_draw_t dr = rfk_make_draw(S, s_pal, a.width, a.height, float(d.x), float(d.y), float(d.z), float(op));
            if (dr.hit) {
                rfk_commit_draw(a.hist, dr);
                hits += dr.w;
                draws++;
            }
        }
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
        if (i == 0 || (pass & 1u) == 0u) perm_cache = philox4x32_10({pass >> 1, block, 0u, RFK_TAG_PERM}, a.seed, a.stream);
        const u32 wa = (pass & 1u) ? perm_cache.z : perm_cache.x;
        const u32 wb = (pass & 1u) ? perm_cache.w : perm_cache.y;
        const u32 slot = rfk_permute(t, wa, wb, RFK_TPB_BITS);
        s_pts[buf][0][slot] = pt.x;
        s_pts[buf][1][slot] = pt.y;
        s_pts[buf][2][slot] = pt.z;
        __syncthreads();
        x = s_pts[buf][0][t];
        y = s_pts[buf][1][t];
        c = s_pts[buf][2][t];
        buf ^= 1u;
#endif
        done++;

        if (bailout && ((i + 1u) & (a.check_every - 1u)) == 0u) {
            // bailouts (refactor.cu:443-458): quality target per block, time
            // limit on the device clock, host stop flag in coherent host memory
            atomicAdd(&s_hits, hits);
            hits = 0;
            __syncthreads();
            if (t == 0) {
                bool stop = a.stop_flag &&
                            __hip_atomic_load((u32*)a.stop_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
                if (a.time_limit_ticks && wall_clock64() - s_start >= a.time_limit_ticks) stop = true;
                if (a.quality_target && s_hits >= a.quality_target) stop = true;
                s_stop = stop ? 1u : 0u;
            }
            __syncthreads();
            if (s_stop) break;
        }
    }

    a.wx[gid] = float(x);
    a.wy[gid] = float(y);
    a.wc[gid] = float(c);
#if RFK_USE_CHAOS
    if (a.wlast) a.wlast[gid] = last;
#endif
    atomicAdd(&s_hits, hits);
    if (draws) atomicAdd(&s_draws, draws);
    if (bad) atomicAdd(&s_bad, bad);
    __syncthreads();
    if (t == 0) {
        a.block_pass[block] = pass0 + done;
        atomicAdd(&a.counters[RFK_CNT_HITS], s_hits);
        atomicAdd(&a.counters[RFK_CNT_PASSES], (u64)done * RFK_TPB);
        atomicMin(&a.counters[RFK_CNT_START], s_start);
        atomicMax(&a.counters[RFK_CNT_STOP], wall_clock64());
        if (s_stop) atomicAdd(&a.counters[RFK_CNT_STOPPED], 1ull);
        if (s_draws) atomicAdd(&a.counters[RFK_CNT_DRAWS], s_draws);
        if (s_bad) atomicAdd(&a.counters[RFK_CNT_BAD], s_bad);
    }
}
