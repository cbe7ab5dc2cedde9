// MI355X (gfx950) per-tick spiking step: kernels, host orchestration and the C ABI declared in
// include/snncuda_hip.h.
//
// One tick is a fixed launch sequence on the handle's stream; all sizes live in device memory, so
// a run of many ticks never waits on the host (the reference's host loop does three blocking reads
// and syncs per tick, src/backends/cuda/CudaBackend.cu:3672-3799).
//
//   plan      1 block      delivery-order groups of the tick (stage-major, see snn_step.h)
//   segments  grid-stride  one (fire, lag group) or inject run per segment
//   scan      3 kernels    segment event counts -> global event positions g
//   count     wave/segment rank of each event within its target's bin (one atomic per event)
//   scan      3 kernels    bin offsets per target
//   scatter   wave/segment events into bins; edge-parallel code patterns (one code per synapse)
//   fold      thread/neuron sort a bin by g and apply it sequentially
//   scan      3 kernels    fire bitmap -> fire ranks
//   firewrite grid-stride  rank-ordered fire list into the fire ring
//   post      wave/fire    post-synaptic STDP scan over the neuron's incoming plastic records
//   advance   1 thread     tick counters
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "snn_net.h"
#include "snn_step.h"
#include "../../include/snncuda_hip.h"

#ifndef SNNCUDA_VERSION
#define SNNCUDA_VERSION "0.1.0"
#endif

using namespace snn;

namespace {

thread_local std::string g_last_error;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kScanBlocks = 1024;
static_assert(kScanBlocks == (int)kFireBlocks, "k_fire_write's blocks are the blocks the fold counts fires in");
constexpr int kPBlock = 1024;  // partition count / scatter workgroups: 16 waves per CU
// scatter workgroups at most: k_hscan scans four counts per thread of a 256-thread block, and the
// fold's LDS table holds every workgroup's run start
constexpr uint32_t kMaxScatterWgs = 1024;

// ------------------------------------------------------------------------------------ scans

struct PlainSrc {
    const uint32_t* p;
    const uint32_t* n_dev;
    uint32_t n_static;
    __device__ uint32_t count() const { return n_dev ? *n_dev : n_static; }
    __device__ uint32_t load(uint32_t i) const { return p[i]; }
};


__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// exclusive block scan of one value per thread; returns the block total through *total
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == kWave - 1) wsum[wid] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
        if (w < wid) base += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

// exclusive scan across a block of NT threads (one value each)
template <int NT>
__device__ uint32_t block_excl_scan_n(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[NT / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == kWave - 1) wsum[wid] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int w = 0; w < NT / kWave; ++w) {
        if (w < wid) base += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

template <class Src>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(Src src, uint32_t* partials) {
    const uint32_t n = src.count();
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = min(n, blockIdx.x * chunk);
    const uint32_t hi = min(n, lo + chunk);
    uint32_t s = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s += src.load(i);
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// sum over the block of two values per thread (kBlock threads)
__device__ __forceinline__ void block_sum2(uint32_t& a, uint32_t& b) {
    __shared__ uint32_t wa[kBlock / kWave], wb[kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        a += __shfl_down(a, o, kWave);
        b += __shfl_down(b, o, kWave);
    }
    if (lane == 0) {
        wa[wid] = a;
        wb[wid] = b;
    }
    __syncthreads();
    a = b = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
        a += wa[w];
        b += wb[w];
    }
    __syncthreads();
}

// This block's base (the sum of the block partials before it) and their total, for up to 4 * kBlock
// partials: each thread's loads are issued together.
__device__ __forceinline__ void partials_prefix(const uint32_t* partials, uint32_t nparts, uint32_t& before,
                                                uint32_t& all) {
    static_assert(kScanBlocks <= 4 * kBlock, "four partials per thread");
    uint32_t v[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t j = threadIdx.x + u * kBlock;
        v[u] = j < nparts ? partials[j] : 0u;
    }
    before = 0;
    all = 0;
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        all += v[u];
        before += threadIdx.x + u * kBlock < blockIdx.x ? v[u] : 0u;
    }
    block_sum2(before, all);
}

// Down-sweep of the scan. Each block forms its base from the block partials itself (the sum of
// those before it) and block 0 the grand total: *total, and when cap is exceeded err_flag raised
// and *zero_on_overflow and *total zeroed (so the tick does no further work).
template <class Src>
__global__ void __launch_bounds__(kBlock) k_scan_down(Src src, const uint32_t* partials, uint32_t* out, uint32_t* total,
                                                      uint64_t cap, uint32_t err_flag, uint32_t* err,
                                                      uint32_t* zero_on_overflow) {
    const uint32_t n = src.count();
    uint32_t before, all;
    partials_prefix(partials, gridDim.x, before, all);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (total) *total = all;
        if (cap && all > cap) {
            atomicOr(err, err_flag);
            if (zero_on_overflow) *zero_on_overflow = 0;
            if (total) *total = 0;
        }
    }
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = min(n, blockIdx.x * chunk);
    const uint32_t hi = min(n, lo + chunk);
    uint32_t base = before;
    for (uint32_t off = lo; off < hi; off += blockDim.x) {
        const uint32_t i = off + threadIdx.x;
        const uint32_t v = i < hi ? src.load(i) : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, &tot);
        if (i < hi) out[i] = base + ex;
        base += tot;
    }
}

// ------------------------------------------------------------------------------------ step kernels

__device__ __forceinline__ uint32_t tick_of(const Step& S) { return S.tick != kNone ? S.tick : S.ctl[0]; }

// Counters of finished tick T (its delivered / fired counts, the fire-log head, the raster fill)
// and the device tick. One thread.
__device__ void advance_tick(const Step& S, uint32_t T) {
    const uint32_t F = S.ctl[2];
    const uint32_t D = S.seg_count[1];
    S.tick_delivered[T - S.inj_t0] = D;
    S.tick_fired[T - S.inj_t0] = F;
    S.totals[0] += D;
    S.totals[1] += F;
    S.totals[2] += S.tick_fcount[T % (S.lmax + 1)];  // fires actually written to the ring
    if (S.raster && !(S.ctl[3] & 1u)) S.totals[3] += F;
    S.ctl[0] = T + 1;
}

// k_segments' split of the tick's nseg segments: block b holds [b * chunk, (b + 1) * chunk)
__device__ __forceinline__ uint32_t seg_chunk(uint32_t nseg) { return (nseg + kScanBlocks - 1) / kScanBlocks; }

constexpr uint32_t kMaxGroups = 2 * 127 + 1;  // group table held in LDS: 2 * lmax + 1 with lmax <= 127

// The tick's delivery-order groups and their first segments, by one wave, into grp[0..ng) (LDS or
// global memory); returns the segment count (0 after an error or past the segment capacity).
template <typename G>
__device__ uint32_t plan_groups(const Step& S, uint32_t T, G* grp, bool report) {
    const uint32_t ng = 2 * S.lmax + 1;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint32_t base = 0;
    for (uint32_t off = 0; off < ng; off += kWave) {
        const uint32_t gi = off + lane;
        Group g;
        g.nseg = 0;
        if (gi < ng) group_info(S, T, gi, g);
        const uint32_t incl = wave_incl_scan(g.nseg);
        if (gi < ng) {
            g.first = base + incl - g.nseg;
            grp[gi] = g;
        }
        base += __shfl(incl, kWave - 1, kWave);
    }
    if (base > S.seg_cap) {
        if (report && lane == 0) atomicOr(&S.ctl[1], (uint32_t)kErrSegments);
        base = 0;
    }
    if (S.ctl[1]) base = 0;  // after any error the run is invalid: stop doing work
    return base;
}

// The plan of a tick whose group table does not fit LDS (lmax > 127): one wave writes the groups
// to global memory (k_segments<true> reads them there). With fused_advance it first finishes the
// previous tick's counters.
__global__ void __launch_bounds__(kWave) k_plan(Step S) {
    const uint32_t T = tick_of(S);
    if (S.fused_advance && threadIdx.x == 0 && T > S.inj_t0) advance_tick(S, T - 1);
    const uint32_t base = plan_groups(S, T, S.groups, true);
    if (threadIdx.x == 0) {
        S.seg_count[0] = base;
        S.seg_count[1] = 0;
    }
}

// First kernel of a tick (kScanBlocks blocks). Every block builds the tick's group table in LDS
// (one wave; GLOBAL: k_plan built it in global memory), then takes a contiguous range of the
// segments (one (fire, lag group) or inject run each), writes their descriptors and the range's
// event count for the scan. With fused_advance, block 0 first finishes the previous tick's counters.
template <bool GLOBAL>
__global__ void __launch_bounds__(kBlock) k_segments(Step S) {
    const uint32_t T = tick_of(S);
    const uint32_t ng = 2 * S.lmax + 1;
    __shared__ Group lgrp[GLOBAL ? 1 : kMaxGroups];
    __shared__ uint32_t nseg_s;
    const Group* grp = GLOBAL ? S.groups : lgrp;
    // the fire counts of this tick's fire-list blocks start at zero (the previous tick's fire-list
    // kernel has read them)
    if (threadIdx.x == 0) S.fire_part[blockIdx.x] = 0;
    if constexpr (!GLOBAL) {
        if (threadIdx.x < kWave) {
            if (S.fused_advance && blockIdx.x == 0 && threadIdx.x == 0 && T > S.inj_t0) advance_tick(S, T - 1);
            const uint32_t base = plan_groups(S, T, lgrp, blockIdx.x == 0);
            if (threadIdx.x == 0) {
                nseg_s = base;
                if (blockIdx.x == 0) {
                    S.seg_count[0] = base;
                    S.seg_count[1] = 0;
                }
            }
        }
    } else {
        if (threadIdx.x == 0) nseg_s = S.seg_count[0];
    }
    __syncthreads();
    const uint32_t nseg = nseg_s;
    // the block's contiguous range of segments (seg_first gives k_pcount the same split), in rounds
    // of blockDim.x: their descriptors, and each one's first event relative to the block's first
    const uint32_t chunk = seg_chunk(nseg);
    const uint32_t lo = min(nseg, blockIdx.x * chunk);
    const uint32_t hi = min(nseg, lo + chunk);
    uint32_t run = 0;
    for (uint32_t q0 = lo; q0 < hi; q0 += blockDim.x) {
        const uint32_t q = q0 + threadIdx.x;
        uint32_t cnt = 0;
        if (q < hi) {
            // last group whose first segment is <= q (groups are in order, empty ones share `first`)
            uint32_t a = 0, b = ng;
            while (a < b) {
                const uint32_t mid = (a + b) >> 1;
                if (grp[mid].first <= q) a = mid + 1; else b = mid;
            }
            const Group g = grp[a - 1];
            uint64_t v0;
            uint32_t mult;
            segment_info(S, g, q - g.first, v0, cnt, mult);
            S.seg_v0[q] = v0;
            S.seg_mult[q] = mult;
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan(cnt, &tot);
        if (q < hi) S.seg_loc[q] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0) S.partials[blockIdx.x] = run;
}

// Event ranges: workgroup w owns events [D*w/nwg, D*(w+1)/nwg) (a contiguous range of g), caches the
// descriptors of the segments covering it in LDS and lets its lanes walk consecutive events, so the
// CSR reads are coalesced and the load is balanced whatever the fan-out distribution.
constexpr uint32_t kSegChunk = 512;

constexpr uint32_t kSegTab = 1024;  // segment lookup table entries per chunk
struct SegCache {
    uint32_t off[kSegChunk + 1];
    uint64_t v0[kSegChunk];
    uint32_t mult[kSegChunk];
    uint32_t tab[kSegTab];  // the segment of every 2^shift-th position of the chunk's range
};

// last segment whose first event is <= g (segments of zero events share their first position)
__device__ __forceinline__ uint32_t seg_find(const uint32_t* off, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

// Positions [ga, gb) of a cached chunk of nq segments (off[0] <= ga, gb <= off[nq]) take their
// segment from a table of every 2^shift-th position's segment and a short forward walk, instead of
// a binary search each: a block builds the table once per chunk (kSegTab searches at most).
__device__ __forceinline__ uint32_t seg_tab_shift(uint32_t ga, uint32_t gb) {
    const uint32_t span = ga < gb ? gb - ga : 0u;
    uint32_t sh = 0;
    while ((span >> sh) >= kSegTab) ++sh;  // ceil(span / 2^sh) <= kSegTab
    return sh;
}
__device__ __forceinline__ void seg_tab_build(SegCache& c, uint32_t nq, uint32_t ga, uint32_t gb, uint32_t sh) {
    const uint32_t nent = ga < gb ? ((gb - ga + (1u << sh) - 1) >> sh) : 0u;
    for (uint32_t j = threadIdx.x; j < nent; j += blockDim.x) c.tab[j] = seg_find(c.off, nq, ga + (j << sh));
    __syncthreads();
}
__device__ __forceinline__ uint32_t seg_lookup(const SegCache& c, uint32_t ga, uint32_t sh, uint32_t g) {
    uint32_t k = c.tab[(g - ga) >> sh];
    while (c.off[k + 1] <= g) ++k;  // stops at nq - 1 at the latest: g < off[nq]
    return k;
}

// The last q in [0, n) with a[q] <= g, for a non-decreasing a with a[0] <= g. Up to kSerialFind
// entries one thread's binary search (at most 12 dependent loads) is cheaper; above, the whole block
// finds it in two parallel passes: a sample of up to blockDim.x evenly spaced entries, then the
// entries between the two neighbouring samples that enclose the answer (those <= g are a prefix of
// each). Every thread of the block calls it, once per kernel, and gets the result.
constexpr uint32_t kSerialFind = 4096;
__device__ __forceinline__ uint32_t block_last_le(const uint32_t* a, uint32_t n, uint32_t g) {
    if (n <= kSerialFind) {  // block-uniform
        __shared__ uint32_t found;
        if (threadIdx.x == 0) {
            uint32_t lo = 1, hi = n;  // a[0] <= g
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a[mid] <= g) lo = mid + 1; else hi = mid;
            }
            found = lo - 1;
        }
        __syncthreads();
        return found;
    }
    const uint32_t B = blockDim.x, t = threadIdx.x;
    const uint32_t m = min(B, n);
    auto at = [&](uint32_t k) { return (uint32_t)((uint64_t)k * n / m); };  // sample k's index; at(m) = n
    const uint32_t c1 = (uint32_t)__syncthreads_count(t < m && a[at(t)] <= g);  // >= 1: a[0] <= g
    const uint32_t lo = at(c1 - 1), hi = c1 < m ? at(c1) : n;
    uint32_t c2 = 0;
    for (uint32_t q0 = lo + 1; q0 < hi; q0 += B) {
        const uint32_t q = q0 + t;
        c2 += (uint32_t)__syncthreads_count(q < hi && a[q] <= g);
    }
    return lo + c2;
}

// Per-workgroup histogram of the tick's events over target partitions (LDS atomics only).
//
// The tick's segment offsets have no scan kernel of their own: k_segments left each segment's first
// event relative to its block (seg_loc) and each block's event count (partials). Every workgroup
// forms the exclusive prefix of the block counts in LDS (pre), so segment q's first event is
// pre[q / chunk] + seg_loc[q]; together the workgroups write those offsets to seg_off for the
// scatter, and block 0 writes the tick's event count (or, past the event capacity, raises the
// error and zeroes the tick's segments and events, so nothing further is done).
__global__ void __launch_bounds__(kPBlock) k_pcount(Step S) {
    static_assert(kPBlock == kScanBlocks, "one block count per thread");
    extern __shared__ uint32_t h[];
    __shared__ SegCache cache;
    __shared__ uint32_t pre[kScanBlocks + 1];
    const uint32_t nseg_all = S.seg_count[0];
    uint32_t total;
    {
        const uint32_t ex = block_excl_scan_n<kPBlock>(S.partials[threadIdx.x], &total);
        pre[threadIdx.x] = ex;
        if (threadIdx.x == 0) pre[kScanBlocks] = total;
    }
    const uint64_t cap = S.ev_cap < kMaxEventsPerTick ? S.ev_cap : (uint64_t)kMaxEventsPerTick;
    const bool over = total > cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (over) {
            atomicOr(&S.ctl[1], (uint32_t)kErrEvents);
            S.seg_count[0] = 0;
            S.seg_count[1] = 0;
        } else {
            S.seg_count[1] = total;
        }
    }
    const uint32_t nseg = over ? 0u : nseg_all;
    const uint64_t D = over ? 0u : total;
    const uint32_t chunk = seg_chunk(nseg_all);
    for (uint32_t p = threadIdx.x; p < S.nparts; p += blockDim.x) h[p] = 0;
    __syncthreads();
    // this workgroup's share of the scatter's segment offsets
    {
        const uint32_t qe = (uint32_t)((uint64_t)nseg * (blockIdx.x + 1) / gridDim.x);
        for (uint32_t q = (uint32_t)((uint64_t)nseg * blockIdx.x / gridDim.x) + threadIdx.x; q < qe; q += blockDim.x)
            S.seg_off[q] = pre[q / chunk] + S.seg_loc[q];
    }
    const uint32_t shift = S.pshift;
    // kCountUnroll events per thread at a time: their target loads are all in flight before the
    // histogram updates
    constexpr uint32_t kCountUnroll = 4;
    const uint32_t g0 = (uint32_t)(D * blockIdx.x / gridDim.x);
    const uint32_t g1 = (uint32_t)(D * (blockIdx.x + 1) / gridDim.x);
    // the last segment whose first event is <= g0: in the last block (of those holding segments)
    // whose first event is <= g0, the last segment whose offset is <= g0
    __shared__ uint32_t qblk;
    if (threadIdx.x == 0 && g0 < g1) {
        const uint32_t nblk = (nseg + chunk - 1) / chunk;
        uint32_t a = 1, b = nblk;  // pre[0] = 0 <= g0
        while (a < b) {
            const uint32_t mid = (a + b) >> 1;
            if (pre[mid] <= g0) a = mid + 1; else b = mid;
        }
        qblk = a - 1;
    }
    __syncthreads();
    uint32_t qa = nseg;
    if (g0 < g1) {  // block-uniform
        const uint32_t first = qblk * chunk;  // its seg_loc is 0
        qa = first + block_last_le(S.seg_loc + first, min(nseg, first + chunk) - first, g0 - pre[qblk]);
    }
    while (g0 < g1) {
        const uint32_t qb = min(nseg, qa + kSegChunk);
        const uint32_t nq = qb - qa;
        for (uint32_t k = threadIdx.x; k < nq; k += blockDim.x) {
            cache.off[k] = pre[(qa + k) / chunk] + S.seg_loc[qa + k];
            cache.v0[k] = S.seg_v0[qa + k];
        }
        if (threadIdx.x == 0) cache.off[nq] = qb < nseg ? pre[qb / chunk] + S.seg_loc[qb] : (uint32_t)D;
        __syncthreads();
        const uint32_t ga = max(g0, cache.off[0]);
        const uint32_t gb = min(g1, cache.off[nq]);
        const uint32_t sh = seg_tab_shift(ga, gb);
        seg_tab_build(cache, nq, ga, gb, sh);
        for (uint32_t g = ga + threadIdx.x; g < gb; g += kCountUnroll * blockDim.x) {
            uint32_t x[kCountUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kCountUnroll; ++u) {
                const uint32_t gg = g + u * blockDim.x;
                x[u] = 0xFFFFFFFFu;
                if (gg < gb) {
                    const uint32_t k = seg_lookup(cache, ga, sh, gg);
                    x[u] = event_target(S, cache.v0[k], gg - cache.off[k]);
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kCountUnroll; ++u)
                if (x[u] != 0xFFFFFFFFu) atomicAdd(&h[x[u] >> shift], 1u);
        }
        __syncthreads();
        if (gb >= g1 || qb >= nseg) break;
        qa = qb;
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < S.nparts; p += blockDim.x) S.hist[(uint64_t)p * S.nwg + blockIdx.x] = h[p];
}

// Per partition (one block each): the scatter workgroups' event counts -> their offsets within the
// partition's bucket, in place, and the partition's total.
__global__ void __launch_bounds__(kBlock) k_hscan(Step S) {
    const uint32_t p = blockIdx.x;
    uint32_t* h = S.hist + (uint64_t)p * S.nwg;
    const uint32_t per = (S.nwg + kBlock - 1) / kBlock;  // <= 4 (nwg <= 1024)
    uint32_t v[4], sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t w = threadIdx.x * per + q;
        v[q] = (q < per && w < S.nwg) ? h[w] : 0u;
        sum += v[q];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(sum, &tot);
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t w = threadIdx.x * per + q;
        if (q < per && w < S.nwg) h[w] = ex;
        ex += v[q];
    }
    if (threadIdx.x == 0) S.ptotal[p] = tot;
}

// The block's events in rounds of EPT * blockDim.x consecutive positions, EPT per thread: every
// thread runs the same number of rounds (f gets ok = false past the end), so f may contain block
// barriers.
template <bool WITH_MULT, uint32_t EPT, typename F>
__device__ void for_block_rounds(const Step& S, SegCache& c, F&& f) {
    const uint32_t nseg = S.seg_count[0];
    const uint64_t D = S.seg_count[1];
    const uint32_t g0 = (uint32_t)(D * blockIdx.x / gridDim.x);
    const uint32_t g1 = (uint32_t)(D * (blockIdx.x + 1) / gridDim.x);
    if (g0 >= g1) return;  // block-uniform
    uint32_t qa = block_last_le(S.seg_off, nseg, g0);  // seg_off[0] = 0 <= g0
    while (true) {
        const uint32_t qb = min(nseg, qa + kSegChunk);
        const uint32_t nq = qb - qa;
        for (uint32_t k = threadIdx.x; k < nq; k += blockDim.x) {
            c.off[k] = S.seg_off[qa + k];
            c.v0[k] = S.seg_v0[qa + k];
            if (WITH_MULT) c.mult[k] = S.seg_mult[qa + k];
        }
        if (threadIdx.x == 0) c.off[nq] = qb < nseg ? S.seg_off[qb] : (uint32_t)D;
        __syncthreads();
        const uint32_t ga = max(g0, c.off[0]);
        const uint32_t gb = min(g1, c.off[nq]);
        const uint32_t sh = seg_tab_shift(ga, gb);
        seg_tab_build(c, nq, ga, gb, sh);
        for (uint32_t gr = ga; gr < gb; gr += EPT * blockDim.x) {
            bool ok[EPT];
            uint64_t v0[EPT];
            uint32_t o[EPT], m[EPT], ii[EPT];
#pragma unroll
            for (uint32_t u = 0; u < EPT; ++u) {
                const uint32_t g = gr + u * blockDim.x + threadIdx.x;
                ok[u] = g < gb;
                const uint32_t k = ok[u] ? seg_lookup(c, ga, sh, g) : 0u;
                v0[u] = ok[u] ? c.v0[k] : 0ull;
                o[u] = ok[u] ? c.off[k] : 0u;
                m[u] = (ok[u] && WITH_MULT) ? c.mult[k] : 0u;
                ii[u] = ok[u] ? g - c.off[k] : 0u;
            }
            f(ok, v0, o, m, ii);
        }
        __syncthreads();
        if (gb >= g1 || qb >= nseg) break;
        qa = qb;
    }
}

// Entries move between registers, LDS buffers and buckets as 32-bit words in plain arrays, which stay in
// registers (an array of 12-byte entries is kept in scratch memory instead).
template <typename E>
__device__ __forceinline__ void load_words(const E* p, uint32_t* w) {
#pragma unroll
    for (uint32_t j = 0; j < sizeof(E) / 4; ++j) __builtin_memcpy(&w[j], (const char*)p + 4 * j, 4);
}
template <typename E>
__device__ __forceinline__ void store_words(E* p, const uint32_t* w) {
#pragma unroll
    for (uint32_t j = 0; j < sizeof(E) / 4; ++j) __builtin_memcpy((char*)p + 4 * j, &w[j], 4);
}

// Place each event in its partition's bucket: the workgroup's run in partition p starts at the
// scanned base of (p, workgroup). Records (the event and its 16-bit target index within the
// partition, in two arrays) are gathered per partition in LDS, `chunk` at a time, and written out
// as one coalesced chunk (a record that finds its partition's buffer full is written straight to
// its place), so the bucket writes leave the workgroup as whole segments instead of one scattered
// store per event. Positions within a run are unordered; the fold orders by g.
constexpr uint32_t kScatterEpt = 2;  // events per thread per round
// at most 64 VGPRs: two 1024-thread workgroups per CU (8 waves per SIMD)
template <bool WIDE, bool CODE_HERE, bool CODE2>
__global__ void __launch_bounds__(kPBlock) __attribute__((amdgpu_waves_per_eu(8))) k_pscatter(Step S) {
    typedef typename EntryOf<WIDE, CODE_HERE>::type E;
    extern __shared__ uint32_t dyn[];
    __shared__ SegCache cache;
    const uint32_t np = S.nparts, ch = S.chunk;
    uint32_t* base = dyn;              // next bucket position of this workgroup's run in partition p
    uint32_t* cnt = dyn + np;          // records taken since the last flush of partition p
    E* buf = (E*)(dyn + 2 * np);       // [np][ch] events
    uint16_t* buft = (uint16_t*)(buf + (size_t)np * ch);  // [np][ch] target indices
    {
        // bucket starts: exclusive scan of the partition totals, one per thread per block scan
        uint32_t run = 0;
        for (uint32_t p0 = 0; p0 < np; p0 += blockDim.x) {
            const uint32_t p = p0 + threadIdx.x;
            // the workgroup's offset within the partition is requested with the total, before the scan
            const uint32_t v = p < np ? S.ptotal[p] : 0u;
            const uint32_t own = p < np ? S.hist[(uint64_t)p * S.nwg + blockIdx.x] : 0u;
            uint32_t tot;
            const uint32_t ex = run + block_excl_scan_n<kPBlock>(v, &tot);
            if (p < np) {
                base[p] = ex + own;
                cnt[p] = 0;
                if (blockIdx.x == 0) S.pstart[p] = ex;  // for the fold
            }
            run += tot;
        }
    }
    __syncthreads();
    const uint32_t T = tick_of(S);
    const uint32_t shift = S.pshift;
    const uint32_t mask = (1u << shift) - 1;
    E* out = (E*)S.prec;
    uint16_t* outt = S.ptl;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
    uint32_t err = 0;
    for_block_rounds<CODE_HERE, kScatterEpt>(S, cache, [&](const bool* ok, const uint64_t* v0, const uint32_t* g0,
                                                          const uint32_t* mult, const uint32_t* i) {
        uint32_t x[kScatterEpt];
        // wide entries as words, not entries: an array of 12- or 16-byte entries would not stay in
        // registers (8-byte entries do, and keep their faster code)
        typedef typename std::conditional<sizeof(E) == 8, E, uint32_t[sizeof(E) / 4]>::type Held;
        Held ew[kScatterEpt];
        auto put = [](E* dst, const Held& h) {
            if constexpr (sizeof(E) == 8) *dst = h;
            else store_words(dst, h);
        };
#pragma unroll
        for (uint32_t u = 0; u < kScatterEpt; ++u) {
            if (!ok[u]) continue;
            if constexpr (sizeof(E) == 8) {
                err |= make_event<WIDE, CODE_HERE, CODE2>(S, v0[u], g0[u], mult[u], i[u], T, x[u], ew[u]);
            } else {
                E e;
                err |= make_event<WIDE, CODE_HERE, CODE2>(S, v0[u], g0[u], mult[u], i[u], T, x[u], e);
                load_words(&e, ew[u]);
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kScatterEpt; ++u) {
            if (!ok[u]) continue;
            const uint16_t tl = (uint16_t)(x[u] & mask);
            const uint32_t p = x[u] >> shift;
            const uint32_t slot = atomicAdd(&cnt[p], 1u);
            if (slot < ch) {
                put(buf + (p * ch + slot), ew[u]);
                buft[p * ch + slot] = tl;
            } else {
                put(out + (base[p] + slot), ew[u]);
                outt[base[p] + slot] = tl;
            }
        }
        if (ch == 0) return;
        __syncthreads();
        // write out every full buffer: wave w owns partitions w, w + nwaves, ...; its lanes check 64
        // of them at once and the wave writes each full one, one lane per record
        for (uint32_t k0 = 0; k0 * nwaves * kWave < np; ++k0) {
            const uint32_t pl = wave + nwaves * (lane + kWave * k0);
            uint64_t m = __ballot(pl < np && cnt[pl] >= ch);
            while (m) {
                const uint32_t p = wave + nwaves * ((uint32_t)__builtin_ctzll(m) + kWave * k0);
                m &= m - 1;
                const uint32_t c = cnt[p];
                if (lane < ch) {
                    out[base[p] + lane] = buf[p * ch + lane];
                    outt[base[p] + lane] = buft[p * ch + lane];
                }
                if (lane == 0) {
                    base[p] += c;
                    cnt[p] = 0;
                }
            }
        }
        __syncthreads();
    });
    // the partial buffers (the same ownership; only the partitions holding records are visited)
    for (uint32_t k0 = 0; k0 * nwaves * kWave < np; ++k0) {
        const uint32_t pl = wave + nwaves * (lane + kWave * k0);
        uint64_t m = __ballot(pl < np && cnt[pl] > 0);
        while (m) {
            const uint32_t p = wave + nwaves * ((uint32_t)__builtin_ctzll(m) + kWave * k0);
            m &= m - 1;
            const uint32_t c = min(cnt[p], ch);
            if (lane < c) {
                out[base[p] + lane] = buf[p * ch + lane];
                outt[base[p] + lane] = buft[p * ch + lane];
            }
        }
    }
    if (err) atomicOr(&S.ctl[1], err);
}

// LDS of the 1024-thread fold: a 96 KB stage, two 16 KB count arrays (one per chunk in flight) and a
// 16 KB table (scatter-workgroup run starts, or the sparse path's touched-target ranks)
constexpr uint32_t kStageBytes = 98304;
constexpr uint32_t kMaxPartition = 4096;  // neurons per partition at most (partition shift <= 12)
constexpr int kSmallFold = 256;              // the small fold workgroup (sparse workloads)
constexpr uint32_t kSmallStageBytes = 16384;  // its LDS stage: four workgroups per CU

// Dense partition in rounds of BLOCK targets, each round staging its own targets' events (in LDS,
// or global scratch when they do not fit); off holds the partition's per-target counts. The
// 256-thread fold's dense path, and the 1024-thread fold's when a scatter run exceeds its stage.
template <int BLOCK, bool WIDE, bool STDP, bool CODE_HERE, bool CODE2, typename E>
__device__ __forceinline__ void fold_rounds(const Step& S, E* lds_stage, uint32_t* aux, uint32_t* off, uint32_t c,
                                            uint32_t start, uint32_t T) {
    constexpr uint32_t kRound = BLOCK;
    constexpr uint32_t kCap = (BLOCK == kPBlock ? kStageBytes : kSmallStageBytes) / sizeof(E);
    constexpr uint32_t kScanPer = 1024 / BLOCK;
    constexpr uint32_t kUnroll = WIDE ? 4 : 8;
    constexpr uint32_t kNoTarget = 0xFFFFu;
    const uint32_t PS = 1u << S.pshift;
    const uint32_t p = blockIdx.x;
    const E* src = (const E*)S.prec + start;
    const uint16_t* tls = S.ptl + start;
    const bool narrow_g = S.seg_count[1] <= S.narrow_max;
    uint32_t err = 0;
    uint32_t run = 0;
    for (uint32_t t0 = 0; t0 < PS; t0 += kScanPer * BLOCK) {
        const uint32_t tb = t0 + threadIdx.x * kScanPer;
        uint32_t v[kScanPer], sum = 0;
#pragma unroll
        for (uint32_t q = 0; q < kScanPer; ++q) {
            v[q] = off[tb + q];
            sum += v[q];
        }
        uint32_t tot;
        uint32_t ex = run + block_excl_scan_n<BLOCK>(sum, &tot);
#pragma unroll
        for (uint32_t q = 0; q < kScanPer; ++q) {
            off[tb + q] = ex;
            ex += v[q];
        }
        run += tot;
    }
    if (threadIdx.x == 0) off[PS] = run;
    __syncthreads();
    uint32_t* fill = aux;
    const uint32_t rounds = PS / kRound;
    // one round staged at st: instantiated once on the LDS stage and once on global scratch, so the
    // common case addresses the stage with LDS instructions rather than flat ones
    auto round = [&](uint32_t r, uint32_t lo, E* st) {
        const uint32_t t = r * kRound + threadIdx.x;
        for (uint32_t i0 = threadIdx.x; i0 < c; i0 += kUnroll * BLOCK) {
            uint32_t tv[kUnroll];
            uint32_t ev[kUnroll][sizeof(E) / 4];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) tv[u] = i0 + u * BLOCK < c ? tls[i0 + u * BLOCK] : kNoTarget;
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u)
                if (tv[u] / kRound == r) load_words(src + i0 + u * BLOCK, ev[u]);
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u)
                if (tv[u] / kRound == r) store_words(st + atomicAdd(&fill[tv[u] % kRound], 1u), ev[u]);
        }
        __syncthreads();
        const uint32_t x = (p << S.pshift) + t;
        const uint32_t k = off[t + 1] - off[t];
        if (x < S.n && k) err |= fold_events<WIDE, STDP, CODE_HERE, CODE2, BLOCK == kPBlock>(S, st + (off[t] - lo), k, x, T, narrow_g);
        __syncthreads();
    };
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t lo = off[r * kRound], hi = off[(r + 1) * kRound];
        fill[threadIdx.x] = off[r * kRound + threadIdx.x] - lo;
        __syncthreads();
        if (hi - lo <= kCap) round(r, lo, lds_stage);
        else round(r, lo, (E*)S.stage + start + lo);
    }
    if (err) atomicOr(&S.ctl[1], err);
}

// Dense partition of the 1024-thread fold: the bucket in chunks of whole scatter-workgroup runs,
// each counting-sorted by target into the LDS stage and folded into its targets' registers.
//   runs: LDS table of the nwg run starts (relative to the partition's first event), then c; no
//         run is longer than the stage
//   off:  two arrays of PS + 1 counters (the chunk being placed and the next one)
template <bool WIDE, bool STDP, bool CODE_HERE, bool CODE2, typename E>
__device__ __forceinline__ void fold_chunks(const Step& S, E* lds_stage, uint32_t* runs, uint32_t* off, uint32_t c,
                                            uint32_t start, uint32_t T) {
    constexpr uint32_t BLOCK = kPBlock;
    // a chunk of at most 8 records per thread (held in registers while the previous chunk is folded)
    constexpr uint32_t R = (kStageBytes / sizeof(E) / BLOCK) < 8 ? (kStageBytes / sizeof(E) / BLOCK) : 8;
    constexpr uint32_t kCap = R * BLOCK;
    constexpr uint32_t TPT = kMaxPartition / BLOCK;  // targets per thread at most
    constexpr uint32_t W = sizeof(E) / 4;
    constexpr uint32_t kNoTarget = 0xFFFFu;
    static_assert(kCap % BLOCK == 0, "the stage holds whole rounds of records");
    const uint32_t PS = 1u << S.pshift;
    const uint32_t tpt = PS / BLOCK;  // 1, 2 or 4 (partition shift 10 to 12)
    const bool narrow_g = S.seg_count[1] <= S.narrow_max;
    const uint32_t p = blockIdx.x;
    const uint32_t nwg = S.nwg;
    const E* src = (const E*)S.prec + start;
    const uint16_t* tls = S.ptl + start;
    for (uint32_t t = threadIdx.x; t < 2 * (PS + 1); t += BLOCK) off[t] = 0;
    // End of the chunk that starts at the run start lo: the last run start within lo + kCap (no run
    // exceeds kCap, so it is past lo). runs[0] = 0 <= lo and runs[nwg] = c keep the search in range.
    auto chunk_end = [&](uint32_t lo) -> uint32_t {
        uint32_t a = 0, b = nwg + 1;
        while (a < b) {
            const uint32_t m = (a + b) >> 1;
            if (runs[m] <= lo + kCap) a = m + 1; else b = m;
        }
        return runs[a - 1];
    };
    uint32_t tv[R];
    uint32_t ev[R][W];
    // positions a + threadIdx.x + u * BLOCK (u < R) below b into registers
    auto load = [&](uint32_t a, uint32_t b) {
#pragma unroll
        for (uint32_t u = 0; u < R; ++u) {
            const uint32_t i = a + threadIdx.x + u * BLOCK;
            tv[u] = i < b ? tls[i] : kNoTarget;
            if (i < b) load_words(src + i, ev[u]);
        }
    };
    // the first chunk's records are requested before the targets' state, so both round trips overlap
    uint32_t lo = 0;
    uint32_t hi = chunk_end(0);
    load(lo, hi);
    // thread's targets: threadIdx.x + j * BLOCK, so the state loads and stores are coalesced
    FoldTarget f[TPT];
#pragma unroll
    for (uint32_t j = 0; j < TPT; ++j) {
        const uint32_t x = (p << S.pshift) + threadIdx.x + j * BLOCK;
        f[j].k = 0;
        if (j < tpt && x < S.n) fold_begin(S, f[j], x, T);
    }
    __syncthreads();
    uint32_t err = 0;
    uint32_t* oc = off;           // counts, then places, of the current chunk
    uint32_t* on = off + PS + 1;  // the next chunk's counts
    while (lo < c) {
#pragma unroll
        for (uint32_t u = 0; u < R; ++u)
            if (tv[u] != kNoTarget) atomicAdd(&oc[tv[u]], 1u);
        __syncthreads();
        {
            // exclusive scan of the chunk's counts, tpt consecutive counters per thread
            uint32_t v[TPT], sum = 0;
            const uint32_t tb = threadIdx.x * tpt;
#pragma unroll
            for (uint32_t q = 0; q < TPT; ++q) {
                v[q] = q < tpt ? oc[tb + q] : 0u;
                sum += v[q];
            }
            uint32_t tot;
            uint32_t ex = block_excl_scan_n<BLOCK>(sum, &tot);
#pragma unroll
            for (uint32_t q = 0; q < TPT; ++q) {
                if (q < tpt) oc[tb + q] = ex;
                ex += v[q];
            }
        }
        __syncthreads();
        // place the chunk, zero the next chunk's counters and start loading it, then fold this one
#pragma unroll
        for (uint32_t u = 0; u < R; ++u)
            if (tv[u] != kNoTarget) store_words(lds_stage + atomicAdd(&oc[tv[u]], 1u), ev[u]);
        for (uint32_t t = threadIdx.x; t < PS + 1; t += BLOCK) on[t] = 0;
        const uint32_t nlo = hi;
        const uint32_t nhi = nlo < c ? chunk_end(nlo) : nlo;
        load(nlo, nhi);
        __syncthreads();
        // after the placement oc[t] is the end of target t's events, oc[t - 1] their start
#pragma unroll
        for (uint32_t j = 0; j < TPT; ++j) {
            const uint32_t t = threadIdx.x + j * BLOCK;
            const uint32_t x = (p << S.pshift) + t;
            if (j < tpt && x < S.n) {
                const uint32_t e = oc[t], s = t ? oc[t - 1] : 0u;
                if (e > s) err |= fold_piece<WIDE, STDP, CODE_HERE, CODE2>(S, f[j], lds_stage + s, e - s, x, T, narrow_g);
            }
        }
        uint32_t* const t = oc;
        oc = on;
        on = t;
        lo = nlo;
        hi = nhi;
    }
    // the temporal headers of all the thread's targets are loaded before any is finished
    TemporalHdr h[TPT];
#pragma unroll
    for (uint32_t j = 0; j < TPT; ++j) {
        const uint32_t x = (p << S.pshift) + threadIdx.x + j * BLOCK;
        if (j < tpt && x < S.n) fold_end_load(S, f[j], x, h[j]);
    }
#pragma unroll
    for (uint32_t j = 0; j < TPT; ++j) {
        const uint32_t x = (p << S.pshift) + threadIdx.x + j * BLOCK;
        if (j < tpt && x < S.n) err |= fold_end_apply(S, f[j], x, T, h[j]);
    }
    if (err) atomicOr(&S.ctl[1], err);
}
// One workgroup per target partition.
//
// Dense partitions (1024-thread fold): the bucket is read once, in chunks of whole scatter-workgroup
// runs. The scatter workgroups own contiguous ranges of g in workgroup order, so a chunk holds, for
// every target, events of larger g than the chunks before it. Each chunk is counting-sorted by
// target into the LDS stage (a run larger than the stage goes through global scratch) and every
// thread folds the chunk's events of its targets (PS / 1024 of them, state in registers across the
// chunks). The next chunk's records are loaded while the current one is folded.
//
// Partitions whose events all fit the stage: one counting sort over the partition and one round
// over the touched targets only. The 256-thread fold folds dense partitions in rounds of 256
// targets, each round staging its own targets' events.
template <int BLOCK, bool WIDE, bool STDP, bool CODE_HERE, bool CODE2>
__global__ void __launch_bounds__(BLOCK) k_pfold(Step S) {
    typedef typename EntryOf<WIDE, CODE_HERE>::type E;
    constexpr uint32_t kRound = BLOCK;  // targets folded per round (one per thread)
    constexpr uint32_t kCap = (BLOCK == kPBlock ? kStageBytes : kSmallStageBytes) / sizeof(E);
    __shared__ E lds_stage[kCap];
    // 1024-thread fold: run starts of the scatter workgroups (dense) or per touched target (by rank,
    // at most kCap) its index | next stage slot << 16 (sparse); 256-thread fold: also each round
    // target's next stage slot
    constexpr uint32_t kAux = BLOCK == kPBlock ? kMaxPartition : (kCap > kRound ? kCap : kRound);
    __shared__ uint32_t aux[kAux];
    // PS + 1 event counts per target, then their places (the 1024-thread fold: two such arrays)
    extern __shared__ uint32_t off[];
    const uint32_t PS = 1u << S.pshift;
    // counters per thread in the scans over the partition's targets: 1,024 per block scan (the
    // partition holds 1,024 to 4,096 targets)
    constexpr uint32_t kScanPer = 1024 / BLOCK;
    const uint32_t p = blockIdx.x;
    const uint32_t T = tick_of(S);
    const uint32_t start = S.pstart[p];
    const uint32_t c = S.ptotal[p];
    if (S.page_hits && threadIdx.x == 0 && c) {
        // the events reaching each learned-pattern page of the partition (paged pool residency)
        const uint32_t x0 = p << S.pshift, x1 = min(S.n, (p + 1) << S.pshift) - 1;
        for (uint32_t q = x0 >> kRefPageShift; q <= (x1 >> kRefPageShift); ++q) atomicAdd(&S.page_hits[q], c);
    }
    const E* src = (const E*)S.prec + start;
    const uint16_t* tls = S.ptl + start;
    constexpr uint32_t kNoTarget = 0xFFFFu;
    if constexpr (BLOCK == kPBlock) {
        if (c > kCap) {
            // chunks of whole scatter-workgroup runs, unless one run alone exceeds the stage (a burst
            // concentrated on few partitions): then rounds of 1024 targets as below
            const uint32_t nwg = S.nwg;
            bool big = false;
            for (uint32_t w = threadIdx.x; w < nwg; w += BLOCK) {
                aux[w] = S.hist[(uint64_t)p * nwg + w];
                const uint32_t nxt = w + 1 < nwg ? S.hist[(uint64_t)p * nwg + w + 1] : c;
                big |= nxt - aux[w] > kCap;
            }
            if (threadIdx.x == 0) aux[nwg] = c;
            if (!__syncthreads_or(big)) {
                fold_chunks<WIDE, STDP, CODE_HERE, CODE2>(S, lds_stage, aux, off, c, start, T);
                return;
            }
        }
    }
    for (uint32_t t = threadIdx.x; t < PS; t += blockDim.x) off[t] = 0;
    __syncthreads();
    // the passes over the bucket load kUnroll records per thread before using them, so that many
    // loads are in flight per wave
    constexpr uint32_t kUnroll = WIDE ? 4 : 8;
    for (uint32_t i0 = threadIdx.x; i0 < c; i0 += kUnroll * BLOCK) {
        uint32_t tv[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) tv[u] = i0 + u * BLOCK < c ? tls[i0 + u * BLOCK] : kNoTarget;
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u)
            if (tv[u] != kNoTarget) atomicAdd(&off[tv[u]], 1u);
    }
    __syncthreads();
    uint32_t err = 0;
    const bool narrow_g = S.seg_count[1] <= S.narrow_max;
    if (c <= kCap) {
        // Sparse partition: every event fits the stage at once, so one round folds the touched
        // targets only, one per thread. One scan gives both the targets' places (low 16 bits) and
        // the touched targets' ranks (high 16 bits); c <= kCap < 2^16 keeps the halves apart.
        uint32_t run = 0;
        for (uint32_t t0 = 0; t0 < PS; t0 += kScanPer * BLOCK) {
            const uint32_t tb = t0 + threadIdx.x * kScanPer;
            uint32_t v[kScanPer], sum = 0;
#pragma unroll
            for (uint32_t q = 0; q < kScanPer; ++q) {
                v[q] = off[tb + q];
                sum += v[q] | (v[q] ? 1u << 16 : 0u);
            }
            uint32_t tot;
            uint32_t ex = run + block_excl_scan_n<BLOCK>(sum, &tot);
#pragma unroll
            for (uint32_t q = 0; q < kScanPer; ++q) {
                off[tb + q] = ex;
                if (v[q]) aux[ex >> 16] = (tb + q) | (ex << 16);
                ex += v[q] | (v[q] ? 1u << 16 : 0u);
            }
            run += tot;
        }
        if (threadIdx.x == 0) off[PS] = run;
        const uint32_t ntouched = run >> 16;
        __syncthreads();
        for (uint32_t i0 = threadIdx.x; i0 < c; i0 += kUnroll * BLOCK) {
            uint32_t tv[kUnroll];
            uint32_t ev[kUnroll][sizeof(E) / 4];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                const uint32_t i = i0 + u * BLOCK;
                tv[u] = i < c ? tls[i] : kNoTarget;
                if (i < c) load_words(src + i, ev[u]);
            }
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u)
                if (tv[u] != kNoTarget) store_words(lds_stage + (atomicAdd(&aux[off[tv[u]] >> 16], 1u << 16) >> 16), ev[u]);
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < ntouched; j += blockDim.x) {
            const uint32_t t = aux[j] & 0xFFFFu;
            const uint32_t x = (p << S.pshift) + t;
            const uint32_t at = off[t] & 0xFFFFu;
            if (x < S.n) err |= fold_events<WIDE, STDP, CODE_HERE, CODE2>(S, lds_stage + at, (off[t + 1] & 0xFFFFu) - at, x, T, narrow_g);
        }
        if (err) atomicOr(&S.ctl[1], err);
        return;
    }
    fold_rounds<BLOCK, WIDE, STDP, CODE_HERE, CODE2>(S, lds_stage, aux, off, c, start, T);
}

constexpr uint32_t kFireCodeBlocks = 32;  // k_fire_write's blocks for the per-source code patterns
// Rank-ordered fire list of tick T into the fire ring (and the raster). kScanBlocks blocks, each over
// a contiguous range of the fire bitmap's words (fire_block): the fold counted each block's fires
// as it set their bits, so a block's first rank is the sum of the counts before it and F their
// total.
__global__ void __launch_bounds__(kBlock) k_fire_write(Step S) {
    const uint32_t T = tick_of(S);
    if (blockIdx.x >= kScanBlocks) {
        // the last kFireCodeBlocks blocks: the per-source code patterns of the fires of T - lmax,
        // which have now reached every synapse (their ring entries are kept until this tick is
        // written, so this runs beside the fire list)
        if (!S.code_lazy || T < S.lmax) return;
        const uint32_t te = T - S.lmax;
        const uint32_t R = S.lmax + 1;
        const uint32_t so = te % R;
        const uint32_t Fo = S.tick_fcount[so];
        const uint64_t st = S.tick_fstart[so];
        uint32_t err = 0;
        for (uint32_t f = (blockIdx.x - kScanBlocks) * blockDim.x + threadIdx.x; f < Fo; f += kFireCodeBlocks * blockDim.x) {
            const uint64_t idx = (st + f) % S.fire_cap;
            const uint32_t m = S.firemult[idx];
            if (m) err |= code_observe2(S.src_code[S.firelog[idx]], te, m, S.tp_window, S.tp_sim, S.tp_max_refs);
        }
        if (err) atomicOr(&S.ctl[1], err);
        return;
    }
    // the ring and raster state is loaded first, so its round trip overlaps the others
    const uint64_t head = S.totals[2];
    const uint64_t rfill = S.totals[3];
    const uint32_t rflags = S.ctl[3];
    const uint32_t R = S.lmax + 1;
    // entries of ticks older than T+1-lmax may be overwritten (older than T-lmax with lazy code
    // patterns, whose fires the last blocks apply)
    const uint32_t keep = S.code_lazy ? S.lmax : S.lmax - 1;  // ticks kept besides T
    const uint64_t oldest = keep == 0 ? head : (T >= keep ? S.tick_fstart[(T - keep) % R] : 0);
    const uint32_t nwords = (S.seg_count[1] + 31u) >> 5;
    const uint32_t chunk = (nwords + kScanBlocks - 1) / kScanBlocks;
    const uint32_t lo = min(nwords, blockIdx.x * chunk);
    const uint32_t hi = min(nwords, lo + chunk);
    // the block's first bitmap word is loaded while the partials are summed
    uint32_t next = lo + threadIdx.x < hi ? S.fire_bits[lo + threadIdx.x] : 0u;
    uint32_t before, all;
    partials_prefix(S.fire_part, kScanBlocks, before, all);
    const uint32_t F = all;
    const bool fits = head + F - oldest <= S.fire_cap;
    // the raster is a recording aid: when it is full, recording stops and the run goes on
    const bool rfits = S.raster && !(rflags & 1u) && rfill + F <= S.raster_cap;
    // Each round's fire positions are listed in rank order in LDS first, then written by all
    // threads: a word holding many fires (the inject events' words often do) would otherwise keep
    // one thread walking its bits with a memory round trip per fire.
    __shared__ uint32_t list[kBlock * 32];
    uint32_t base = before;
    for (uint32_t off = lo; off < hi; off += blockDim.x) {
        const uint32_t wi = off + threadIdx.x;
        uint32_t bits = next;
        next = wi + blockDim.x < hi ? S.fire_bits[wi + blockDim.x] : 0u;
        uint32_t tot;
        uint32_t r = block_excl_scan(__popc(bits), &tot);
        if (!tot) continue;  // block-uniform
        if (bits) S.fire_bits[wi] = 0;
        while (bits) {
            const uint32_t b = __ffs(bits) - 1;
            bits &= bits - 1;
            list[r++] = wi * 32 + b;
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < tot; j += blockDim.x) {
            const uint32_t g = list[j];
            const uint32_t x = S.fire_neuron[g];
            const uint64_t pos = base + j;
            if (fits) {
                const uint64_t idx = (head + pos) % S.fire_cap;
                S.firelog[idx] = x;
                S.firemult[idx] = S.fire_mult[g];
            }
            if (rfits) {
                S.raster[rfill + pos] = x;
                if (S.raster_evt) S.raster_evt[rfill + pos] = S.fire_evt[g];
            }
        }
        base += tot;
        __syncthreads();  // the list is refilled by the next round
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S.ctl[2] = F;
        if (!fits) atomicOr(&S.ctl[1], (uint32_t)kErrFireLog);
        if (S.raster && !rfits) atomicOr(&S.ctl[3], 1u);
        S.tick_fstart[T % R] = head;
        S.tick_fcount[T % R] = fits ? F : 0;
    }
}

// post-synaptic scan: one wave per (first) fire of tick T, kPostBatch records per lane loaded before
// any is updated, so each lane has that many 32-byte records in flight
constexpr uint32_t kPostBatch = 4;
__global__ void __launch_bounds__(kBlock) k_post(Step S) {
    const uint32_t T = tick_of(S);
    const uint32_t R = S.lmax + 1;
    const uint32_t F = S.tick_fcount[T % R];
    const uint64_t start = S.tick_fstart[T % R];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
    for (uint32_t f = wave; f < F; f += nwaves) {
        const uint64_t idx = (start + f) % S.fire_cap;
        const uint32_t m = S.firemult[idx];
        if (!m) continue;
        const uint32_t x = S.firelog[idx];
        const uint64_t lo = S.in_ptr[x], hi = S.in_ptr[x + 1];
        for (uint64_t t0 = lo + lane; t0 < hi; t0 += kPostBatch * kWave) {
            StdpRec r[kPostBatch];
#pragma unroll
            for (uint32_t u = 0; u < kPostBatch; ++u) {
                const uint64_t t = t0 + u * kWave;
                if (t < hi) r[u] = S.stdp[t];
            }
#pragma unroll
            for (uint32_t u = 0; u < kPostBatch; ++u) {
                const uint64_t t = t0 + u * kWave;
                if (t < hi && post_scan_apply(S, r[u], T, m)) S.stdp[t] = r[u];
            }
        }
    }
}

// the counters of tick T (every tick of a HIP graph replay; the last tick of a run of stream
// launches, whose earlier ticks are finished by the next tick's k_segments)
__global__ void k_advance(Step S) { advance_tick(S, tick_of(S)); }

// Absolute start of every (source, lag) group, row[s] + grp[s * W + j] (k_segments reads one of
// these pairs per segment instead of the row start and the group offsets).
__global__ void k_group_starts(const uint64_t* row, const uint32_t* grp, uint64_t n, uint32_t W, uint64_t* out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n * W; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = row[i / W] + grp[i];
}

// the STDP records' update counters, record order
__global__ void k_stdp_counts(const StdpRec* stdp, uint64_t nt, uint32_t* upd, uint32_t* ltp) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < nt; t += (uint64_t)gridDim.x * blockDim.x) {
        upd[t] = stdp[t].upd;
        ltp[t] = stdp[t].ltp;
    }
}

// readback gathers (connectome order)
__global__ void k_gather_weights(const uint32_t* slot_of, const float* w, const uint32_t* tpos, const StdpRec* stdp,
                                 uint64_t e, float* out) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < e; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = slot_of[j];
        const uint32_t t = tpos ? tpos[s] : kNone;
        out[j] = t != kNone ? stdp[t].weight : w[s];
    }
}

template <bool CODE2>
__global__ void k_gather_codes(const uint32_t* slot_of, const CodeRec2* c2, const CodeRecN* cn, uint64_t e,
                               uint32_t* match, uint32_t* learned) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < e; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = slot_of[j];
        if (CODE2) {
            const uint32_t w1 = c2[s].w1;
            match[j] = w1 & 0x0FFFFFFFu;
            learned[j] = (w1 >> 30) & 1u;
        } else {
            match[j] = cn[s].match;
            learned[j] = (cn[s].flags >> 4) & 0xFu;
        }
    }
}

// last_pre / last_post of every synapse (SynapseRuntimeState, reference
// include/snncuda/runtime/SynapseDendriteProcessor.h:25-26): plastic synapses keep them in their
// STDP record; for the others last_post is the target's last fire tick (every fire stamps all
// incoming synapses, NetworkPropagator.cpp:95-102) and last_pre is the last arrival tick minus the
// delay (SynapseDendriteProcessor.cpp:95,125).
__global__ void k_gather_ticks(const uint32_t* slot_of, const uint32_t* tpos, const StdpRec* stdp, const CodeRec2* c2,
                               const CodeRecN* cn, const uint16_t* delay, const uint32_t* tgt,
                               const uint32_t* last_fired, uint64_t e, int64_t* last_pre, int64_t* last_post) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < e; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = slot_of[j];
        const uint32_t t = tpos ? tpos[s] : kNone;
        uint32_t pre, post;
        if (t != kNone) {
            pre = stdp[t].last_pre;
            post = stdp[t].last_post;
        } else {
            const uint32_t a = c2 ? code_last_arrival2(c2[s]) : code_last_arrival(cn[s]);
            const uint32_t dl = delay[s];
            pre = a == kNone ? kNone : (a >= dl ? a - dl : a);
            post = last_fired[tgt[s]];
        }
        last_pre[j] = pre == kNone ? -1 : (int64_t)pre;
        last_post[j] = post == kNone ? -1 : (int64_t)post;
    }
}

// Lazy code patterns at readback. The fires of ticks [t_first, t_first + nk) (the last lmax ticks)
// are not in the source records yet: win[s * lmax + k] = multiplicity of s's fires at t_first + k,
// snap[s * lmax + k] = s's record with those fires applied up to tick t_first + k.
__global__ void k_code_window(Step S, uint32_t t_first, uint32_t* win) {
    const uint32_t k = blockIdx.y, te = t_first + k, slot = te % (S.lmax + 1);
    const uint32_t F = S.tick_fcount[slot];
    const uint64_t st = S.tick_fstart[slot];
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += gridDim.x * blockDim.x) {
        const uint64_t idx = (st + f) % S.fire_cap;
        const uint32_t m = S.firemult[idx];
        if (m) win[(uint64_t)S.firelog[idx] * S.lmax + k] = m;
    }
}

__global__ void k_code_snap(Step S, uint32_t t_first, uint32_t nk, const uint32_t* win, CodeRec2* snap) {
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < S.n; s += (uint64_t)gridDim.x * blockDim.x) {
        CodeRec2 r = S.src_code[s];
        for (uint32_t k = 0; k < nk; ++k) {
            const uint32_t m = win[s * S.lmax + k];
            if (m) (void)code_observe2(r, t_first + k, m, S.tp_window, S.tp_sim, S.tp_max_refs);
            snap[s * S.lmax + k] = r;
        }
    }
}

// Every synapse slot's record: its source's record after the fires of ticks <= T_end - lag, moved
// from fire time to arrival time (window start + lag).
__global__ void k_code_materialize(Step S, uint32_t T_end, uint32_t t_first, const CodeRec2* snap, uint64_t e,
                                   CodeRec2* out) {
    const uint32_t W = S.lmax + 1;
    for (uint64_t slot = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; slot < e; slot += (uint64_t)gridDim.x * blockDim.x) {
        // source: last s with row[s] <= slot (one code per synapse: rows index slots)
        uint32_t lo = 0, hi = S.n;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (S.row[mid] <= slot) lo = mid; else hi = mid;
        }
        const uint32_t s = lo;
        const uint32_t off = (uint32_t)(slot - S.row[s]);
        // lag: last L with grp[s * W + L - 1] <= off
        uint32_t a = 1, b = S.lmax;
        while (a < b) {
            const uint32_t mid = (a + b + 1) >> 1;
            if (S.grp[(uint64_t)s * W + mid - 1] <= off) a = mid; else b = mid - 1;
        }
        const uint32_t L = a;
        CodeRec2 r = S.src_code[s];
        if (T_end >= L && T_end - L >= t_first) r = snap[(uint64_t)s * S.lmax + (T_end - L - t_first)];
        if ((r.w1 >> 28) & 3u) r.start += L;
        out[slot] = r;
    }
}

__global__ void k_gather_u32(const uint32_t* slot_of, const uint32_t* src, uint64_t e, uint32_t* out) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < e; j += (uint64_t)gridDim.x * blockDim.x)
        out[j] = src[slot_of[j]];
}

// ------------------------------------------------------------------------------------ synthetic E/I networks
// Device generator for the seeded E/I network family of snncuda_amd/workloads.py (ei_network), for
// one network or a batch of them laid out as their union: every draw is the same counter hash, so a
// generated network is identical to the host-built one. Connectome order is source-major (edge j of
// union source s has index srow[s] + j); slots are sorted by (source, lag of code 0, j) and virtual
// edges by (source, lag, j, code index), exactly as build_host_net sorts them.

constexpr uint32_t kGenBins = 128;  // lags 1..127 (delay_max <= 63, code_max <= 63)

struct GenNet {
    uint64_t n, nall, n_exc, noff;  // network neurons, plus sources, excitatory, first union neuron
    uint64_t key[10];               // stream keys of the network's seed (index = stream)
    float w_exc, w_inh;
    uint32_t delay_max, n_comps;
    uint32_t comps[4];
    uint32_t ncode, code_max, plastic, pad;
};

struct GenParams {
    const GenNet* nets;
    uint32_t n_nets;
    uint32_t W;  // lmax + 1
    uint64_t nall;
    const uint64_t* srow;  // slot (connectome) row starts, nall + 1
    const uint64_t* vrow;  // virtual-edge row starts, nall + 1
    uint32_t* grp;
    uint32_t* tgt;
    float* w;
    uint8_t* meta;
    uint16_t* delay;
    uint32_t* slot_of;
    uint32_t* ve;        // null when every synapse has one code
    uint8_t* expected;   // null when uniform
    uint32_t* incnt;
    const uint64_t* in_ptr;
    uint32_t* tpos;
    StdpRec* stdp;
    uint32_t* tpos_conn;
};

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
constexpr uint64_t kGold = 0x9E3779B97F4A7C15ull;
__device__ __forceinline__ uint64_t ghash(uint64_t key, uint64_t idx) { return mix64(key + idx * kGold); }
__device__ __forceinline__ uint32_t gbelow(uint64_t h, uint64_t n) { return (uint32_t)(((h >> 32) * n) >> 32); }
__device__ __forceinline__ float gunif24(uint64_t h) { return (float)(uint32_t)(h >> 40) * 0x1p-24f; }

struct GenEdge {
    uint32_t tgt, delay;
    float w;
    uint8_t meta;
};

// the network of union neuron s (nets sorted by noff)
__device__ __forceinline__ uint32_t gen_net_of(const GenParams& P, uint64_t s) {
    uint32_t lo = 0, hi = P.n_nets - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (P.nets[mid].noff <= s) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t gen_delay(const GenNet& N, uint64_t s, uint64_t ctr) {
    return 1 + gbelow(ghash(N.key[s < N.n ? 5 : 7], ctr), N.delay_max);
}

__device__ __forceinline__ uint32_t gen_code(const GenNet& N, uint64_t ctr, uint32_t c) {
    return N.code_max ? gbelow(ghash(N.key[8], (ctr << 4) + c), N.code_max + 1) : 0u;
}

// network-local target
__device__ __forceinline__ uint32_t gen_target(const GenNet& N, uint64_t s, uint64_t ctr) {
    if (s < N.n) {
        uint32_t t = gbelow(ghash(N.key[1], ctr), N.n - 1);
        return t + (t >= s ? 1u : 0u);
    }
    return gbelow(ghash(N.key[2], ctr), N.n);
}

// plastic synapses of a generated network (snn_ei_desc::plastic): none, all, or all but those of
// inhibitory neurons; s is the network-local source
__device__ __forceinline__ bool gen_plastic(const GenNet& N, uint64_t s) {
    return N.plastic == 1 || (N.plastic == 2 && !(s < N.n && s >= N.n_exc));
}

// edge j of network-local source s
__device__ GenEdge gen_edge(const GenNet& N, uint64_t s, uint64_t j) {
    const uint64_t ctr = (s << 24) + j;
    GenEdge g;
    g.tgt = (uint32_t)(N.noff + gen_target(N, s, ctr));
    g.delay = gen_delay(N, s, ctr);
    uint32_t comp, rec;
    if (s < N.n) {
        const bool exc = s < N.n_exc;
        const float u = gunif24(ghash(N.key[3], ctr));
        g.w = exc ? u * N.w_exc : u * N.w_inh;
        rec = exc ? kAmpa : kGabaA;
        comp = N.comps[gbelow(ghash(N.key[4], ctr), N.n_comps)];
    } else {
        g.w = gunif24(ghash(N.key[6], ctr)) * 0.5f;
        rec = kAmpa;
        comp = kBasal;
    }
    g.meta = (uint8_t)(rec | (comp << 2) | ((gen_plastic(N, s) ? 1u : 0u) << 4));
    return g;
}

__device__ __forceinline__ uint32_t gen_lag(uint32_t delay, uint32_t code) {
    const uint32_t l = delay + code;
    return l == 0 ? 1u : l;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// exclusive scan of a wave's kGenBins-entry LDS histogram, in place
__device__ __forceinline__ void wave_bins_excl_scan(uint32_t* h, uint32_t lane) {
    const uint32_t a = h[2 * lane], b = h[2 * lane + 1];
    const uint32_t incl = wave_incl_scan(a + b);
    wave_lds_sync();
    h[2 * lane] = incl - a - b;
    h[2 * lane + 1] = incl - b;
    wave_lds_sync();
}

// Stable placement of one chunk of up to 64 items (lanes in item order) by key: item gets
// run[key] + its rank among this chunk's lanes with the same key; run[key] advances.
__device__ __forceinline__ uint32_t wave_place(bool ok, uint32_t key, uint32_t* run, uint32_t lane) {
    uint64_t rest = __ballot(ok);
    uint32_t pos = 0;
    const uint64_t lt = (1ull << lane) - 1;
    while (rest) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(rest);
        const uint32_t k = __shfl(key, (int)leader, kWave);
        const uint64_t m = __ballot(ok && key == k);
        const uint32_t base = run[k];
        if (ok && key == k) pos = base + (uint32_t)__popcll(m & lt);
        wave_lds_sync();
        if (lane == leader) run[k] = base + (uint32_t)__popcll(m);
        wave_lds_sync();
        rest &= ~m;
    }
    return pos;
}

// pass 1: per-source lag histogram of the virtual edges -> group table; in-degree counts of plastic
// synapses
__global__ void __launch_bounds__(kBlock) k_gen_count(GenParams P) {
    __shared__ uint32_t hist[kBlock / kWave][kGenBins];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wv = threadIdx.x / kWave;
    const uint64_t s = (uint64_t)blockIdx.x * (kBlock / kWave) + wv;
    if (s >= P.nall) return;  // whole wave; no block barriers below
    uint32_t* h = hist[wv];
    h[2 * lane] = 0;
    h[2 * lane + 1] = 0;
    wave_lds_sync();
    const GenNet& N = P.nets[gen_net_of(P, s)];
    const uint64_t sl = s - N.noff;
    const uint64_t len = P.srow[s + 1] - P.srow[s];
    for (uint64_t j = lane; j < len; j += kWave) {
        const uint64_t ctr = (sl << 24) + j;
        const uint32_t d = gen_delay(N, sl, ctr);
        for (uint32_t c = 0; c < N.ncode; ++c) atomicAdd(&h[gen_lag(d, gen_code(N, ctr, c))], 1u);
        if (gen_plastic(N, sl)) atomicAdd(&P.incnt[N.noff + gen_target(N, sl, ctr)], 1u);
    }
    wave_lds_sync();
    wave_bins_excl_scan(h, lane);
    // h[L] = virtual edges of lag < L; group L starts at h[L]
    for (uint32_t L = 1 + lane; L < P.W; L += kWave) P.grp[s * P.W + L - 1] = h[L];
    if (lane == 0) P.grp[s * P.W + P.W - 1] = (uint32_t)(P.vrow[s + 1] - P.vrow[s]);
}

// pass 2: stable placement of each source's synapses by lag of code 0 and of its virtual edges by
// lag, plus target-major STDP records
__global__ void __launch_bounds__(kBlock) k_gen_place(GenParams P) {
    __shared__ uint32_t runs[kBlock / kWave][kGenBins];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wv = threadIdx.x / kWave;
    const uint64_t s = (uint64_t)blockIdx.x * (kBlock / kWave) + wv;
    if (s >= P.nall) return;  // whole wave; no block barriers below
    uint32_t* run = runs[wv];
    const GenNet& N = P.nets[gen_net_of(P, s)];
    const uint64_t sl = s - N.noff;
    const uint64_t r0 = P.srow[s], len = P.srow[s + 1] - r0;
    // slots: histogram of lag of code 0, scanned, then stable placement in connectome order
    run[2 * lane] = 0;
    run[2 * lane + 1] = 0;
    wave_lds_sync();
    for (uint64_t j = lane; j < len; j += kWave) {
        const uint64_t ctr = (sl << 24) + j;
        atomicAdd(&run[gen_lag(gen_delay(N, sl, ctr), gen_code(N, ctr, 0))], 1u);
    }
    wave_lds_sync();
    wave_bins_excl_scan(run, lane);
    for (uint64_t c = 0; c < len; c += kWave) {
        const uint64_t j = c + lane;
        const bool ok = j < len;
        GenEdge g{};
        uint32_t l0 = 0;
        if (ok) {
            g = gen_edge(N, sl, j);
            l0 = gen_lag(g.delay, gen_code(N, (sl << 24) + j, 0));
        }
        const uint32_t pos = wave_place(ok, l0, run, lane);
        if (ok) {
            const uint64_t slot = r0 + pos;
            P.tgt[slot] = g.tgt;
            P.w[slot] = g.w;
            P.meta[slot] = g.meta;
            P.delay[slot] = (uint16_t)g.delay;
            if (P.expected) P.expected[slot] = (uint8_t)(N.ncode > 2 ? N.ncode : 2);
            P.slot_of[r0 + j] = (uint32_t)slot;
            if (g.meta & 16u) {
                const uint64_t tp = P.in_ptr[g.tgt] + atomicAdd(&P.incnt[g.tgt], 1u);
                P.tpos[slot] = (uint32_t)tp;
                StdpRec r{};
                r.weight = g.w;
                r.last_pre = kNone;
                r.last_post = kNone;
                r.aux = 0;
                r.hi = 2.0f;  // max_weight 2.0 on every synapse: min(2, 2)
                P.stdp[tp] = r;
                P.tpos_conn[tp] = (uint32_t)(r0 + j);
            } else if (P.tpos) {
                P.tpos[slot] = kNone;
            }
        }
    }
    if (!P.ve) return;
    // virtual edges in (j, code) order, placed stably by lag; each refers to its synapse's slot
    __threadfence();  // this wave's slot_of stores are read back below
    wave_lds_sync();
    for (uint32_t L = lane; L < kGenBins; L += kWave) run[L] = (L >= 1 && L < P.W) ? P.grp[s * P.W + L - 1] : 0u;
    wave_lds_sync();
    const uint64_t v0 = P.vrow[s], nv = P.vrow[s + 1] - v0;
    for (uint64_t c = 0; c < nv; c += kWave) {
        const uint64_t v = c + lane;
        const bool ok = v < nv;
        uint32_t L = 0;
        uint64_t j = 0;
        if (ok) {
            j = v / N.ncode;
            const uint32_t k = (uint32_t)(v - j * N.ncode);
            const uint64_t ctr = (sl << 24) + j;
            L = gen_lag(gen_delay(N, sl, ctr), gen_code(N, ctr, k));
        }
        const uint32_t pos = wave_place(ok, L, run, lane);
        if (ok) P.ve[v0 + pos] = P.slot_of[r0 + j];
    }
}

__global__ void k_widen(const uint32_t* in, uint64_t n, uint64_t* out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}


// ------------------------------------------------------------------------------------ host side

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

template <typename T>
T* dev_alloc(std::vector<DevBuf>& owned, size_t count, bool zero = true) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    check(hipMalloc(&p, bytes), "hipMalloc");
    if (zero) check(hipMemset(p, 0, bytes), "hipMemset");
    owned.push_back({p, bytes});
    return (T*)p;
}

template <typename T, typename A>
T* dev_upload(std::vector<DevBuf>& owned, const std::vector<T, A>& v) {
    T* p = dev_alloc<T>(owned, v.size(), false);
    if (!v.empty()) check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return p;
}

// Upload of the packed arrays while the host pack goes on. Each array is queued as soon as the pack
// has finished it (PackStage); a worker thread moves it chunk by chunk through a ring of pinned
// staging buffers: a parallel memcpy into a free buffer, then an async copy from it on the
// uploader's own stream, so the DMA of the slot arrays runs during the group-table and STDP passes.
// A network below kStreamMinSynapses copies directly with hipMemcpy (the pinned ring would cost more
// than it saves). The queued host arrays must stay alive until finish() (or the destructor) returns.
constexpr uint64_t kStreamMinSynapses = 1ull << 22;
class Uploader {
public:
    Uploader(int device, bool streamed) : device_(device), streamed_(streamed) {
        if (!streamed_) return;
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
        for (int k = 0; k < kRing; ++k) {
            check(hipHostMalloc(&ring_[k], kChunk, hipHostMallocDefault), "hipHostMalloc");
            check(hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming), "hipEventCreate");
        }
        worker_ = std::thread([this] { run(); });
    }
    ~Uploader() {
        try {
            finish();
        } catch (...) {
        }
    }
    Uploader(const Uploader&) = delete;
    Uploader& operator=(const Uploader&) = delete;

    template <typename T, typename A>
    T* upload(std::vector<DevBuf>& owned, const std::vector<T, A>& v) {
        if (!streamed_) return dev_upload(owned, v);
        T* p = dev_alloc<T>(owned, v.size(), false);
        if (!v.empty()) {
            std::lock_guard<std::mutex> lk(m_);
            q_.push_back(Job{(char*)p, (const char*)v.data(), v.size() * sizeof(T)});
            cv_.notify_one();
        }
        return p;
    }

    // every queued copy done; rethrows the worker's error
    void finish() {
        if (!streamed_) return;
        {
            std::lock_guard<std::mutex> lk(m_);
            closing_ = true;
            cv_.notify_one();
        }
        if (worker_.joinable()) worker_.join();
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (int k = 0; k < kRing; ++k) {
            if (ev_[k]) (void)hipEventDestroy(ev_[k]);
            if (ring_[k]) (void)hipHostFree(ring_[k]);
            ev_[k] = nullptr;
            ring_[k] = nullptr;
        }
        if (stream_) (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
        streamed_ = false;
        if (!error_.empty()) throw HipError(error_);
    }

private:
    static constexpr size_t kChunk = 64ull << 20;
    static constexpr int kRing = 3;
    static constexpr unsigned kCopyThreads = 4;
    struct Job {
        char* dst;
        const char* src;
        size_t bytes;
    };

    void run() {
        try {
            check(hipSetDevice(device_), "hipSetDevice");
            uint64_t k = 0;
            while (true) {
                Job j;
                {
                    std::unique_lock<std::mutex> lk(m_);
                    cv_.wait(lk, [&] { return closing_ || !q_.empty(); });
                    if (q_.empty()) return;
                    j = q_.front();
                    q_.pop_front();
                }
                for (size_t off = 0; off < j.bytes; off += kChunk, ++k) {
                    const size_t len = std::min(kChunk, j.bytes - off);
                    const int slot = (int)(k % kRing);
                    check(hipEventSynchronize(ev_[slot]), "hipEventSynchronize");  // the buffer's last copy
                    char* buf = (char*)ring_[slot];
                    const char* src = j.src + off;
                    std::thread part[kCopyThreads - 1];
                    const size_t per = (len + kCopyThreads - 1) / kCopyThreads;
                    for (unsigned t = 1; t < kCopyThreads; ++t) {
                        const size_t a = std::min(len, t * per), b = std::min(len, (t + 1) * per);
                        part[t - 1] = std::thread([=] { if (b > a) std::memcpy(buf + a, src + a, b - a); });
                    }
                    std::memcpy(buf, src, std::min(len, per));
                    for (auto& th : part) th.join();
                    check(hipMemcpyAsync(j.dst + off, buf, len, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync H2D");
                    check(hipEventRecord(ev_[slot], stream_), "hipEventRecord");
                }
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(m_);
            error_ = e.what();
            q_.clear();
        }
    }

    int device_;
    bool streamed_;
    hipStream_t stream_ = nullptr;
    void* ring_[kRing] = {};
    hipEvent_t ev_[kRing] = {};
    std::thread worker_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<Job> q_;
    bool closing_ = false;
    std::string error_;
};

struct PendingInject {
    uint32_t neuron;
    uint64_t tick;
    float weight;
    uint64_t stage;
    uint64_t seq;
};

constexpr int kProfileSelect = 0x40000000;  // snn_net_set_profiling: time only the kernels in the mask
enum KernelId { KPlan, KSegments, KScanSeg, KCount, KScanTgt, KScatter, KFold, KScanFire, KFireWrite, KPost, KAdvance, KCount_ };
// "scan_segments" and "scan_fires" have no launches of their own: k_pcount forms the segment
// offsets, and the fold counts the fire-list blocks' fires
const char* kKernelNames[KCount_] = {"plan", "segments", "scan_segments", "count", "scan_partitions", "scatter",
                                     "fold", "scan_fires", "fire_write", "post", "advance"};

}  // namespace

struct snn_net {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<DevBuf> owned;
    Step S{};
    uint32_t n = 0;
    uint64_t e = 0;
    bool wide = false, stdp = false, code_here = true, code2 = true;
    uint32_t grid = 1024;
    uint32_t fold_threads = 1024;  // the fold variant (Schedule::fold)
    uint32_t fold_grid = 1;
    uint32_t* partials = nullptr;
    // host-side run state
    uint64_t next_tick = 0;
    uint64_t inject_seq = 0;
    std::vector<PendingInject> pending;
    std::vector<uint64_t> tick_delivered, tick_fired;
    // per-run device buffers
    std::vector<DevBuf> run_owned;
    uint32_t* stat_buf = nullptr;
    uint64_t stat_cap = 0;
    // readback helpers
    uint32_t* slot_of = nullptr;
    uint32_t* tpos_conn_dev = nullptr;
    uint64_t n_tpos = 0;
    // the next tick to launch, and whether the last launched tick still needs its counters
    // finished (stream launches leave them to the next tick's first kernel)
    uint64_t tick_next = 0;
    bool advance_pending = false;
    // profiling
    bool profiling = false;
    uint32_t prof_mask = ~0u;           // kernels timed (bit = KernelId)
    uint32_t prof_every = 1;            // time them on every prof_every-th tick
    std::vector<hipEvent_t> event_pool;
    double kms[KCount_] = {0};
    uint64_t klaunch[KCount_] = {0};
    // launch schedule (results are identical under every schedule)
    uint32_t ncu = 256;
    uint32_t graph_ticks = 0;     // ticks per HIP graph replay; 0: stream launches
    uint32_t hist_cap_parts = 0;  // hist holds hist_cap_parts * max_nwg counters
    uint32_t max_nwg = 0;
    uint32_t tune_trial = 0;      // > 0: the next long enough run_until autotunes with chunks of this many ticks
    std::string tune_log;
    // fire-event recording (fixed-spike adjoint input): every inject delivered, in delivery order
    bool record = false;
    std::vector<PendingInject> inj_log;
    // paged learned-pattern pool (Step::ref_page): every page's rings in pinned host memory, a device
    // cache of cache_pages pages; slot_of_page[q] is page q's cache slot (-1: on the host), table
    // mirrors the device page table
    bool paged = false;
    uint32_t npages = 0, cache_pages = 0;
    Run* host_runs = nullptr;
    RefMeta* host_meta = nullptr;
    Run* cache_runs = nullptr;
    RefMeta* cache_meta = nullptr;
    std::vector<int32_t> slot_of_page, page_of_slot;
    std::vector<RefPage> table;
    uint64_t page_moves = 0;

    ~snn_net() {
        if (device >= 0) (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (host_runs) (void)hipHostFree(host_runs);
        if (host_meta) (void)hipHostFree(host_meta);
        for (auto e : event_pool) (void)hipEventDestroy(e);
        for (auto& b : run_owned) (void)hipFree(b.p);
        for (auto& b : owned) (void)hipFree(b.p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Independent networks run as one network: their disjoint union. Every state write of a tick is
// local to one target and its incoming synapses, and the order of the events that reach a target is
// the order they have in that target's own network, so each network's results equal a run of that
// network alone (tests check this against the oracle per network).
struct snn_batch {
    std::unique_ptr<snn_net> net;
    std::vector<uint64_t> noff, eoff;  // n_nets + 1 neuron / synapse offsets into the union
    std::vector<uint32_t> inj_net;     // injects through snn_batch_inject: network and tick (clamped
    std::vector<uint64_t> inj_tick;    // to the next poll tick, as the scheduler does)
};

namespace {

template <class Src>
void launch_scan(snn_net* net, Src src, uint32_t* out, uint32_t* total, uint64_t cap, uint32_t err_flag,
                 uint32_t* zero_on_overflow) {
    hipLaunchKernelGGL(k_scan_reduce<Src>, dim3(kScanBlocks), dim3(kBlock), 0, net->stream, src, net->partials);
    hipLaunchKernelGGL(k_scan_down<Src>, dim3(kScanBlocks), dim3(kBlock), 0, net->stream, src, (const uint32_t*)net->partials,
                       out, total, cap, err_flag, &net->S.ctl[1], zero_on_overflow);
}

// LDS of the scatter besides its 12 KB segment cache: base and count per partition plus the
// write-combining buffers; two scatter workgroups fit in a CU's 160 KB
constexpr uint32_t kScatterDynLds = 64 * 1024;

template <bool WIDE, bool CODE_HERE, bool CODE2>
void launch_scatter(snn_net* net, Step& S) {
    typedef typename EntryOf<WIDE, CODE_HERE>::type E;
    // up to 256 bytes of events per flush, as many as the LDS holds (26 per partition at 244
    // partitions): longer flushed runs leave fewer partly written cache lines
    const uint32_t want = WIDE ? 16u : 32u;
    const uint64_t room = kScatterDynLds > 8ull * S.nparts ? kScatterDynLds - 8ull * S.nparts : 0;
    constexpr uint64_t kRecBytes = sizeof(E) + sizeof(uint16_t);
    uint32_t ch = (uint32_t)std::min<uint64_t>(want, room / ((uint64_t)S.nparts * kRecBytes));
    if (ch < 4) ch = 0;
    S.chunk = ch;
    net->S.chunk = ch;
    const size_t lds = 8ull * S.nparts + (size_t)S.nparts * ch * kRecBytes;
    hipLaunchKernelGGL((k_pscatter<WIDE, CODE_HERE, CODE2>), dim3(S.nwg), dim3(kPBlock), lds, net->stream, S);
}

template <bool WIDE, bool STDP, bool CODE_HERE, bool CODE2>
void launch_fold(snn_net* net, const Step& S) {
    // per-target counters: one array for the 256-thread fold, two for the 1024-thread fold's chunks
    const size_t lds = ((1u << S.pshift) + 1) * 4;
    if (net->fold_threads == (uint32_t)kSmallFold)
        hipLaunchKernelGGL((k_pfold<kSmallFold, WIDE, STDP, CODE_HERE, CODE2>), dim3(S.nparts), dim3(kSmallFold),
                           lds, net->stream, S);
    else
        hipLaunchKernelGGL((k_pfold<kPBlock, WIDE, STDP, CODE_HERE, CODE2>), dim3(S.nparts), dim3(kPBlock), 2 * lds,
                           net->stream, S);
}

// The per-slot compact code records: the network's own, or with per-source code patterns a
// temporary copy materialized from the source records and the fire ring's last lmax ticks.
const CodeRec2* code2_records(snn_net* net, std::vector<DevBuf>& tmp) {
    Step& S = net->S;
    if (!S.code_lazy) return S.code2;
    const uint64_t nt = net->next_tick;
    const uint32_t T_end = nt ? (uint32_t)(nt - 1) : 0u;
    const uint32_t t_first = nt == 0 ? 1u : (T_end >= S.lmax ? T_end - S.lmax + 1 : 0u);
    const uint32_t nk = nt == 0 ? 0u : T_end - t_first + 1;
    uint32_t* win = dev_alloc<uint32_t>(tmp, (uint64_t)S.n * S.lmax);
    CodeRec2* snap = dev_alloc<CodeRec2>(tmp, (uint64_t)S.n * S.lmax, false);
    if (nk) hipLaunchKernelGGL(k_code_window, dim3(64, nk), dim3(kBlock), 0, net->stream, S, t_first, win);
    hipLaunchKernelGGL(k_code_snap, dim3(1024), dim3(kBlock), 0, net->stream, S, t_first, nk, (const uint32_t*)win, snap);
    CodeRec2* out = dev_alloc<CodeRec2>(tmp, net->e, false);
    hipLaunchKernelGGL(k_code_materialize, dim3(4096), dim3(kBlock), 0, net->stream, S, T_end, t_first,
                       (const CodeRec2*)snap, net->e, out);
    check(hipGetLastError(), "code records");
    return out;
}

// once per process: let the scatter kernels take more than the default dynamic LDS
void allow_scatter_lds() {
    static std::once_flag once;
    std::call_once(once, [] {
        const int bytes = (int)kScatterDynLds;
        (void)hipFuncSetAttribute((const void*)k_pscatter<false, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_pscatter<false, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_pscatter<true, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_pscatter<true, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_pscatter<true, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_pscatter<true, false, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    });
}

void dispatch_scatter(snn_net* net, Step& S) {
    // (wide, code_here, code2); code_here needs one code per synapse
    if (!net->wide) {
        if (net->code2) launch_scatter<false, true, true>(net, S); else launch_scatter<false, true, false>(net, S);
    } else if (net->code_here) {
        if (net->code2) launch_scatter<true, true, true>(net, S); else launch_scatter<true, true, false>(net, S);
    } else {
        if (net->code2) launch_scatter<true, false, true>(net, S); else launch_scatter<true, false, false>(net, S);
    }
}

void dispatch_fold(snn_net* net, const Step& S) {
    if (!net->wide) {
        if (net->code2) launch_fold<false, false, true, true>(net, S); else launch_fold<false, false, true, false>(net, S);
        return;
    }
#define SNN_FOLD(ST, CH)                                                         \
    if (net->stdp == ST && net->code_here == CH) {                             \
        if (net->code2) launch_fold<true, ST, CH, true>(net, S);                  \
        else launch_fold<true, ST, CH, false>(net, S);                            \
        return;                                                                \
    }
    SNN_FOLD(true, true)
    SNN_FOLD(true, false)
    SNN_FOLD(false, true)
    SNN_FOLD(false, false)
#undef SNN_FOLD
}

// HIP events around the profiled kernels (all, or those selected by the profiling mask), taken
// from a pool owned by the handle so that a timed run creates none.
struct Timer {
    snn_net* net;
    bool enabled = true;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    int cur = -1;
    hipEvent_t a = nullptr;
    hipEvent_t take() {
        if (net->event_pool.empty()) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            return e;
        }
        hipEvent_t e = net->event_pool.back();
        net->event_pool.pop_back();
        return e;
    }
    void begin(int id) {
        cur = -1;
        if (!enabled || !net->profiling || !((net->prof_mask >> id) & 1u)) return;
        if (net->prof_every > 1 && net->tick_next % net->prof_every != 0) return;
        cur = id;
        a = take();
        check(hipEventRecord(a, net->stream), "hipEventRecord");
    }
    void end() {
        if (cur < 0) return;
        hipEvent_t b = take();
        check(hipEventRecord(b, net->stream), "hipEventRecord");
        marks.push_back({cur, {a, b}});
        cur = -1;
    }
    void collect() {
        for (auto& m : marks) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, m.second.first, m.second.second), "hipEventElapsedTime");
            net->kms[m.first] += ms;
            net->klaunch[m.first] += 1;
            net->event_pool.push_back(m.second.first);
            net->event_pool.push_back(m.second.second);
        }
        marks.clear();
    }
};

// One tick's launches. Stream launches carry the tick as an argument and leave the tick's counters
// to the next tick's k_segments (or to flush_advance); a HIP graph capture (graph) reads the tick
// from device memory and finishes each tick with k_advance, so every replay steps on.
void launch_tick(snn_net* net, Timer& tm, bool graph = false) {
    Step S = net->S;
    hipStream_t st = net->stream;
    if (graph) {
        S.tick = kNone;
        S.fused_advance = 0;
    } else {
        S.tick = (uint32_t)net->tick_next;
        S.fused_advance = net->advance_pending ? 1u : 0u;
    }
    if (2 * S.lmax + 1 > kMaxGroups) {
        // the group table does not fit LDS: planned by one wave in global memory
        tm.begin(KPlan);
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(kWave), 0, st, S);
        tm.end();
        S.fused_advance = 0;
        tm.begin(KSegments);
        hipLaunchKernelGGL(k_segments<true>, dim3(kScanBlocks), dim3(kBlock), 0, st, S);
        tm.end();
    } else {
        tm.begin(KSegments);
        hipLaunchKernelGGL(k_segments<false>, dim3(kScanBlocks), dim3(kBlock), 0, st, S);
        tm.end();
    }
    tm.begin(KCount);
    hipLaunchKernelGGL(k_pcount, dim3(S.nwg), dim3(kPBlock), S.nparts * 4, st, S);
    tm.end();
    tm.begin(KScanTgt);
    hipLaunchKernelGGL(k_hscan, dim3(S.nparts), dim3(kBlock), 0, st, S);
    tm.end();
    tm.begin(KScatter);
    dispatch_scatter(net, S);
    tm.end();
    tm.begin(KFold);
    dispatch_fold(net, S);
    tm.end();
    tm.begin(KFireWrite);
    hipLaunchKernelGGL(k_fire_write, dim3(kScanBlocks + kFireCodeBlocks), dim3(kBlock), 0, st, S);
    tm.end();
    if (net->stdp) {
        tm.begin(KPost);
        hipLaunchKernelGGL(k_post, dim3(net->grid), dim3(kBlock), 0, st, S);
        tm.end();
    }
    if (graph) {
        tm.begin(KAdvance);
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, st, S);
        tm.end();
    } else {
        net->advance_pending = true;
        net->tick_next += 1;
    }
}

// The counters of the last tick launched by stream launches, when the next tick's k_segments will
// not finish them (end of a run, before graph replays, before a trial is read back).
void flush_advance(snn_net* net, Timer& tm) {
    if (!net->advance_pending) return;
    Step S = net->S;
    S.tick = (uint32_t)(net->tick_next - 1);
    tm.begin(KAdvance);
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, net->stream, S);
    tm.end();
    net->advance_pending = false;
}

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return fail(SNN_EDEVICE, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(SNN_EINVAL, e.what());
    } catch (const std::bad_alloc&) {
        return fail(SNN_EINVAL, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(SNN_EDEVICE, e.what());
    }
}

// ------------------------------------------------------------------------------------ launch schedules
// A schedule changes only how a tick is launched (partition size, scatter workgroups, ticks per HIP
// graph replay); every schedule gives the same results, so it may change between any two ticks.

struct Schedule {
    uint32_t pshift, nwg, graph;
    uint32_t fold;  // fold workgroup threads: 1024, or 256 (more workgroups per CU, small LDS stage)
};

Schedule current_schedule(const snn_net* net) { return {net->S.pshift, net->S.nwg, net->graph_ticks, net->fold_threads}; }

void apply_schedule(snn_net* net, const Schedule& sc) {
    Step& S = net->S;
    if (sc.pshift < 10 || sc.pshift > 12) throw std::invalid_argument("partition_shift must be in [10, 12]");
    const uint64_t parts = ((uint64_t)S.n + (1ull << sc.pshift) - 1) >> sc.pshift;
    if (parts > net->hist_cap_parts || parts > kMaxPartitions) throw std::invalid_argument("partition_shift too small");
    if (sc.nwg < 1 || sc.nwg > net->max_nwg) throw std::invalid_argument("scatter_workgroups out of range");
    if (sc.graph > 1024) throw std::invalid_argument("graph_ticks above 1024");
    if (sc.fold != (uint32_t)kPBlock && sc.fold != (uint32_t)kSmallFold)
        throw std::invalid_argument("fold_threads must be 1024 or 256");
    S.pshift = sc.pshift;
    S.nparts = (uint32_t)parts;
    S.nwg = sc.nwg;
    net->graph_ticks = sc.graph;
    net->fold_threads = sc.fold;
}

// The launches of graph_ticks consecutive ticks captured as one HIP graph. Kernel arguments are
// fixed for the length of one run_until call, so a graph lives for one call.
struct TickGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Schedule sc{};
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~TickGraph() { reset(); }
};

void ensure_graph(snn_net* net, uint32_t t0, TickGraph& g) {
    const Schedule sc = current_schedule(net);
    if (g.exec && g.sc.pshift == sc.pshift && g.sc.nwg == sc.nwg && g.sc.graph == sc.graph && g.sc.fold == sc.fold) return;
    g.reset();
    Timer off{net};
    off.enabled = false;
    check(hipStreamBeginCapture(net->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    for (uint32_t i = 0; i < sc.graph; ++i) launch_tick(net, off, true);
    check(hipStreamEndCapture(net->stream, &g.graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    g.sc = sc;
}

// k ticks of the current run: graph replays where graph_ticks fit, stream launches for the rest.
void run_ticks(snn_net* net, uint32_t t0, uint64_t k, Timer& tm, TickGraph& g) {
    uint64_t i = 0;
    const uint32_t G = net->profiling ? 0 : net->graph_ticks;
    if (G && k >= G) {
        flush_advance(net, tm);
        ensure_graph(net, t0, g);
        for (; i + G <= k; i += G) {
            check(hipGraphLaunch(g.exec, net->stream), "hipGraphLaunch");
            net->tick_next += G;
        }
    }
    for (; i < k; ++i) launch_tick(net, tm);
    flush_advance(net, tm);
}

// Online autotuning on the run's own ticks. Coordinate search, one schedule field at a time (fold
// workgroup size, then scatter workgroups, then partition size, then graph length) around the best
// so far; every candidate runs `trial` ticks in each of two interleaved rounds, timed with HIP
// events on the stream, except that a candidate more than 25% behind the best after the first
// round is not run again (its score is the minimum over its rounds, so it could not win). The
// score is the time per delivered event when the trial delivered at least 1e5 events per tick
// (throughput-bound: activity drifts between trials), else the time per tick (launch-bound); the
// lowest score wins. Returns the ticks consumed (0 when the run is too short, in which case tuning
// stays pending for a later call).
uint64_t autotune(snn_net* net, uint32_t t0, uint64_t nt, Timer& tm, TickGraph& g) {
    const uint32_t trial = net->tune_trial;
    auto graph_cands = [&](Schedule b) {
        std::vector<Schedule> c{b, b};
        c[0].graph = 0;
        c[1].graph = std::min<uint32_t>(16, trial);
        return c;
    };
    auto nwg_cands = [&](Schedule b) {
        std::vector<Schedule> c;
        for (uint32_t m : {net->ncu / 2, net->ncu, 2 * net->ncu, 4 * net->ncu}) {
            if (m < 1 || m > net->max_nwg) continue;
            b.nwg = m;
            c.push_back(b);
        }
        return c;
    };
    auto fold_cands = [&](Schedule b) {
        std::vector<Schedule> c{b, b};
        c[0].fold = (uint32_t)kPBlock;
        c[1].fold = (uint32_t)kSmallFold;
        return c;
    };
    auto part_cands = [&](Schedule b) {
        std::vector<Schedule> c;
        for (uint32_t ps = 10; ps <= 12; ++ps) {
            const uint64_t parts = ((uint64_t)net->S.n + (1ull << ps) - 1) >> ps;
            if (parts > net->hist_cap_parts || parts > kMaxPartitions) continue;
            if (ps > 10 && net->S.n <= (1u << (ps - 1))) continue;  // one partition already
            b.pshift = ps;
            c.push_back(b);
        }
        return c;
    };
    const Schedule start = current_schedule(net);
    const uint64_t need = 2ull * trial *
                          (graph_cands(start).size() + nwg_cands(start).size() + part_cands(start).size() + 2);
    if (trial == 0 || nt < need + trial) return 0;
    hipEvent_t a = nullptr, b = nullptr;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    uint64_t used = 0;
    int per_event = -1;
    Schedule best = start;
    std::string log;
    try {
        for (int dim = 0; dim < 4; ++dim) {
            const std::vector<Schedule> cands = dim == 0   ? fold_cands(best)
                                                : dim == 1 ? nwg_cands(best)
                                                : dim == 2 ? part_cands(best)
                                                           : graph_cands(best);
            std::vector<float> t(cands.size(), 1e30f);
            for (int round = 0; round < 2; ++round) {
                const float lead = *std::min_element(t.begin(), t.end());
                for (size_t c = 0; c < cands.size(); ++c) {
                    if (round > 0 && t[c] > 1.25f * lead) continue;
                    apply_schedule(net, cands[c]);
                    if (cands[c].graph && trial >= cands[c].graph) ensure_graph(net, t0, g);
                    uint64_t d0 = 0, d1 = 0;
                    check(hipMemcpy(&d0, net->S.totals, 8, hipMemcpyDeviceToHost), "read totals");
                    check(hipEventRecord(a, net->stream), "hipEventRecord");
                    run_ticks(net, t0, trial, tm, g);
                    check(hipEventRecord(b, net->stream), "hipEventRecord");
                    check(hipEventSynchronize(b), "hipEventSynchronize");
                    float ms = 0;
                    check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
                    check(hipMemcpy(&d1, net->S.totals, 8, hipMemcpyDeviceToHost), "read totals");
                    const double ev = (double)(d1 - d0);
                    if (per_event < 0) per_event = ev >= 1e5 * trial ? 1 : 0;  // decided by the first trial
                    const float score = per_event ? (float)(ms * 1e6 / std::max(ev, 1.0)) : ms / trial;
                    t[c] = std::min(t[c], score);
                    used += trial;
                }
            }
            size_t k = 0;
            for (size_t c = 1; c < cands.size(); ++c)
                if (t[c] < t[k]) k = c;
            best = cands[k];
            for (size_t c = 0; c < cands.size(); ++c)
                log += "shift=" + std::to_string(cands[c].pshift) + " wgs=" + std::to_string(cands[c].nwg) +
                       " graph=" + std::to_string(cands[c].graph) + " fold=" + std::to_string(cands[c].fold) +
                       " score=" + std::to_string(t[c]) +
                       (c == k ? " *" : "") + "\n";
        }
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    apply_schedule(net, best);
    net->tune_log = log;
    net->tune_trial = 0;
    return used;
}

std::vector<float> host_decay_table() {
    // expf(-(float)dt / tau) as glibc computes it on this host, for tau = 5, 100, 10, 150 (receptor
    // codes 0..3); the reference's decay is current * std::exp(-elapsed / tau).
    static const float taus[4] = {5.0f, 100.0f, 10.0f, 150.0f};
    std::vector<float> t(4 * (size_t)kDecayTable);
    for (int c = 0; c < 4; ++c)
        for (uint32_t dt = 0; dt < kDecayTable; ++dt) {
            volatile float el = (float)dt;
            t[(size_t)c * kDecayTable + dt] = std::exp(-el / taus[c]);
        }
    return t;
}

void host_stdp_tables(std::vector<float>& ltp, std::vector<float>& ltd) {
    // StdpRule::weight_delta (src/learning/STDP.cpp:20-35): 0.01f*exp(-dt/20), -0.012f*exp(dt/20)
    ltp.assign(kStdpTable, 0.0f);
    ltd.assign(kStdpTable, 0.0f);
    for (uint32_t a = 1; a < kStdpTable; ++a) {
        volatile float dt = (float)a;
        ltp[a] = 0.01f * std::exp(-dt / 20.0f);
        volatile float ndt = -(float)a;
        ltd[a] = -0.012f * std::exp(ndt / 20.0f);
    }
}


int check_device(const snn_options& o) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SNN_ENODEVICE, "no HIP device");
    if (o.device < 0 || o.device >= ndev) return fail(SNN_EINVAL, "device ordinal out of range");
    return 0;
}

// device, stream, modes and per-neuron state
void open_net(snn_net* net, const snn_options& o, uint32_t n, uint64_t e, uint32_t lmax, bool plastic, bool ve_identity,
              bool code_compact) {
    net->device = o.device;
    check(hipSetDevice(o.device), "hipSetDevice");
    check(hipStreamCreate(&net->stream), "hipStreamCreate");
    net->n = n;
    net->e = e;
    net->stdp = plastic;
    net->code_here = ve_identity;
    net->wide = plastic || !ve_identity;
    net->code2 = code_compact;
    auto& own = net->owned;
    Step& S = net->S;
    S.n = n;
    S.lmax = lmax;
    S.pre_count = o.track_pre_count ? dev_alloc<uint32_t>(own, e) : nullptr;
    S.dyn = dev_alloc<NeuronDyn>(own, n);
    S.th = dev_alloc<TemporalHdr>(own, n);
    if (o.pattern_runs > kRunsMax) throw std::invalid_argument("pattern_runs above " + std::to_string(kRunsMax));
    S.run_cap = o.pattern_runs ? (uint32_t)o.pattern_runs : kRuns;
    S.cand = dev_alloc<Run>(own, (uint64_t)n * S.run_cap);
    S.last_fired = dev_alloc<uint32_t>(own, n, false);
    check(hipMemset(S.last_fired, 0xFF, (size_t)n * 4), "hipMemset");
    S.decay_tab = dev_upload(own, host_decay_table());
    std::vector<float> ltp, ltd;
    host_stdp_tables(ltp, ltd);
    S.ltp = dev_upload(own, ltp);
    S.ltd = dev_upload(own, ltd);
}

// capacities and per-tick work buffers
void finish_net(snn_net* net, const snn_options& o, uint64_t nve, const std::vector<NeuronParams>& params) {
    auto& own = net->owned;
    Step& S = net->S;
    const uint32_t n = S.n;
    {
        // the groups' absolute starts, when they take at most 1/16 of the free device memory
        // (SNNCUDA_NO_GROUP_STARTS=1 leaves them out, so the path of larger networks can be checked)
        const uint64_t words = (uint64_t)n * (S.lmax + 1);
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const char* no_gs = std::getenv("SNNCUDA_NO_GROUP_STARTS");
        if (words * 8 <= free_b / 16 && !(no_gs && std::atoi(no_gs))) {
            uint64_t* gs = dev_alloc<uint64_t>(own, words, false);
            hipLaunchKernelGGL(k_group_starts, dim3(2048), dim3(kBlock), 0, net->stream, S.row, S.grp, (uint64_t)n,
                               S.lmax + 1, gs);
            check(hipGetLastError(), "k_group_starts");
            S.gstart = gs;
        }
    }
    S.tp_uniform = 1;
    for (const auto& p : params)
        if (p.window != params[0].window || p.max_refs != params[0].max_refs ||
            std::memcmp(&p.sim_thr, &params[0].sim_thr, 4) != 0) {
            S.tp_uniform = 0;
            break;
        }
    S.tp_window = params[0].window;
    S.tp_max_refs = params[0].max_refs;
    S.tp_sim = params[0].sim_thr;
    S.thr_uniform = 1;
    for (const auto& p : params)
        if (std::memcmp(&p.threshold, &params[0].threshold, 4) != 0) {
            S.thr_uniform = 0;
            break;
        }
    S.thr_all = params[0].threshold;
    // learned-pattern pool: as many slots per neuron as the largest max_reference_patterns (or the
    // option), within a quarter of the device's memory; a neuron that learns beyond its slots
    // reports SNN_ECAPACITY
    {
        uint64_t want = 1;
        for (const auto& p : params) want = std::max<uint64_t>(want, p.max_refs);
        if (o.learned_pattern_slots) want = std::min<uint64_t>(want, o.learned_pattern_slots);
        want = std::min<uint64_t>(want, kHeadMask);
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const uint64_t R = S.run_cap;
        const uint64_t per = R * sizeof(Run) + sizeof(RefMeta);
        const uint64_t budget = total_b / 4;
        const uint32_t npages = (uint32_t)((n + (1u << kRefPageShift) - 1) >> kRefPageShift);
        const uint64_t page_bytes = ((uint64_t)1 << kRefPageShift) * want * per;
        uint64_t cache = o.pattern_device_pages;
        if (cache == 0 && (uint64_t)n * want * per > budget) cache = std::max<uint64_t>(1, budget / page_bytes);
        if (cache == 0 || cache >= npages) {
            S.ref_cap = (uint32_t)want;
            S.ref_runs = dev_alloc<Run>(own, (uint64_t)n * S.ref_cap * R);
            S.ref_meta = dev_alloc<RefMeta>(own, (uint64_t)n * S.ref_cap);
        } else {
            // every page in pinned host memory (the kernels reach it directly), `cache` pages also in a
            // device cache; the run moves the pages most events reach into the cache between ticks
            S.ref_cap = (uint32_t)want;
            const uint64_t slots = (uint64_t)npages << kRefPageShift;
            const uint64_t page_slots = (uint64_t)1 << kRefPageShift;
            net->paged = true;
            net->npages = npages;
            net->cache_pages = (uint32_t)cache;
            void* hr = nullptr;
            void* hm = nullptr;
            check(hipHostMalloc(&hr, slots * want * R * sizeof(Run), hipHostMallocMapped), "hipHostMalloc (pattern pool)");
            net->host_runs = (Run*)hr;
            check(hipHostMalloc(&hm, slots * want * sizeof(RefMeta), hipHostMallocMapped), "hipHostMalloc (pattern pool)");
            net->host_meta = (RefMeta*)hm;
            std::memset(hr, 0, slots * want * R * sizeof(Run));
            std::memset(hm, 0, slots * want * sizeof(RefMeta));
            Run* dr = nullptr;
            RefMeta* dm = nullptr;
            check(hipHostGetDevicePointer((void**)&dr, hr, 0), "hipHostGetDevicePointer");
            check(hipHostGetDevicePointer((void**)&dm, hm, 0), "hipHostGetDevicePointer");
            net->cache_runs = dev_alloc<Run>(own, cache * page_slots * want * R);
            net->cache_meta = dev_alloc<RefMeta>(own, cache * page_slots * want);
            net->table.resize(npages);
            for (uint32_t q = 0; q < npages; ++q)
                net->table[q] = RefPage{dr + (uint64_t)q * page_slots * want * R, dm + (uint64_t)q * page_slots * want};
            net->slot_of_page.assign(npages, -1);
            net->page_of_slot.assign(cache, -1);
            S.ref_page = dev_upload(own, net->table);
            S.page_hits = dev_alloc<uint32_t>(own, npages);
        }
    }
    // code pattern records: per source when every synapse has one code offset and every target the
    // same temporal parameters (SNNCUDA_CODE_PER_SYNAPSE=1 keeps them per synapse), else per synapse
    // the per-target sorts take one 32-bit key per entry when every g of the tick is below 2^27;
    // SNNCUDA_WIDE_G=1 sends every tick through the sorts of full entries (the path of ticks with
    // more events), so that path can be checked on small networks
    const char* wide = std::getenv("SNNCUDA_WIDE_G");
    S.narrow_max = (wide && std::atoi(wide)) ? 0u : (1u << 27);
    const char* eager = std::getenv("SNNCUDA_CODE_PER_SYNAPSE");
    S.code_lazy = (net->code_here && net->code2 && S.tp_uniform && !(eager && std::atoi(eager))) ? 1u : 0u;
    if (S.code_lazy) S.src_code = dev_alloc<CodeRec2>(own, n);
    else if (net->code2) S.code2 = dev_alloc<CodeRec2>(own, net->e);
    else S.codeN = dev_alloc<CodeRecN>(own, net->e);
    const uint64_t fire_cap = o.fire_log_capacity ? o.fire_log_capacity : std::max<uint64_t>(1u << 16, 8ull * n);
    uint64_t ev_cap = o.max_events_per_tick ? o.max_events_per_tick
                                            : std::max<uint64_t>(1u << 20, std::min<uint64_t>(2 * nve, 1ull << 26));
    ev_cap = std::min<uint64_t>(ev_cap, kMaxEventsPerTick);
    ev_cap = (ev_cap + 31) & ~31ull;
    const uint64_t seg_cap = std::min<uint64_t>(fire_cap + S.lmax + 1, 0xFFFFFFF0ull);
    S.fire_cap = fire_cap;
    S.firelog = dev_alloc<uint32_t>(own, fire_cap);
    S.firemult = dev_alloc<uint32_t>(own, fire_cap);
    S.tick_fstart = dev_alloc<uint64_t>(own, S.lmax + 1);
    S.tick_fcount = dev_alloc<uint32_t>(own, S.lmax + 1);
    S.groups = dev_alloc<Group>(own, 2 * (size_t)S.lmax + 1);
    S.seg_count = dev_alloc<uint32_t>(own, 4);
    S.seg_cap = seg_cap;
    S.ev_cap = ev_cap;
    S.seg_v0 = dev_alloc<uint64_t>(own, seg_cap, false);
    S.seg_loc = dev_alloc<uint32_t>(own, seg_cap, false);  // (seg_n is the CPU model's)
    S.seg_off = dev_alloc<uint32_t>(own, seg_cap, false);
    S.seg_mult = dev_alloc<uint32_t>(own, seg_cap, false);
    // Target partitions of 2048 neurons (4096 from 512k neurons): few enough buckets that the
    // scatter's write-combining buffers hold a full chunk for each; folded in rounds of 1024
    // neurons. Networks of at most 1024 neurons use a single 1024-neuron partition. Measured on the
    // 1M-neuron workload: 4096-neuron partitions and two scatter workgroups per CU are fastest.
    uint32_t pshift = default_partition_shift(n);
    if (const char* e = std::getenv("SNNCUDA_PART_SHIFT")) pshift = std::max(10, std::min(12, std::atoi(e)));
    if (partitions_of(n, pshift) > kMaxPartitions) pshift = 12;
    S.pshift = pshift;
    S.nparts = (uint32_t)(((uint64_t)n + (1ull << pshift) - 1) >> pshift);
    int ncu = 256;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, o.device) == hipSuccess) ncu = pr.multiProcessorCount;
    net->ncu = (uint32_t)std::max(1, ncu);
    S.nwg = 2 * net->ncu;
    if (const char* e = std::getenv("SNNCUDA_SCATTER_WGS")) S.nwg = (uint32_t)std::max(1, std::atoi(e));
    S.nwg = std::min<uint32_t>(S.nwg, kMaxScatterWgs);
    // the histogram holds the largest schedule the autotuner may pick: 1024-neuron partitions and
    // four workgroups per CU
    net->max_nwg = std::min<uint32_t>(kMaxScatterWgs, std::max<uint32_t>(S.nwg, 4 * net->ncu));
    net->hist_cap_parts = (uint32_t)(((uint64_t)n + 1023) >> 10);
    S.hist = dev_alloc<uint32_t>(own, (uint64_t)net->hist_cap_parts * net->max_nwg);
    S.prec = net->wide ? (void*)dev_alloc<Entry16>(own, ev_cap, false) : (void*)dev_alloc<Entry8>(own, ev_cap, false);
    S.ptl = dev_alloc<uint16_t>(own, ev_cap, false);
    S.stage = net->wide ? (void*)dev_alloc<Entry16>(own, ev_cap, false) : (void*)dev_alloc<Entry8>(own, ev_cap, false);
    S.fire_bits = dev_alloc<uint32_t>(own, ev_cap / 32);
    S.fire_part = dev_alloc<uint32_t>(own, kFireBlocks);
    S.fire_neuron = dev_alloc<uint32_t>(own, ev_cap, false);
    S.fire_mult = dev_alloc<uint32_t>(own, ev_cap, false);
    S.ctl = dev_alloc<uint32_t>(own, 4);
    S.totals = dev_alloc<uint64_t>(own, 4);
    S.raster_cap = o.raster_capacity ? o.raster_capacity : (1ull << 22);
    S.raster = dev_alloc<uint32_t>(own, S.raster_cap, false);
    if (!net->partials) net->partials = dev_alloc<uint32_t>(own, kScanBlocks);
    S.partials = net->partials;
    S.ptotal = dev_alloc<uint32_t>(own, net->hist_cap_parts);
    S.pstart = dev_alloc<uint32_t>(own, net->hist_cap_parts);
    S.tick = kNone;
    S.fused_advance = 0;
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, o.device) == hipSuccess) cus = prop.multiProcessorCount;
    net->grid = (uint32_t)std::max(64, cus * 4);
    net->fold_grid = (uint32_t)std::min<uint64_t>((n + kBlock - 1) / kBlock, 65535ull * 4);
    allow_scatter_lds();
    check(hipDeviceSynchronize(), "create");
}

NetDesc to_net_desc(const snn_net_desc& x) {
    NetDesc d;
    d.n_neurons = x.n_neurons;
    d.threshold = x.threshold;
    d.window = x.window;
    d.sim_thr = x.sim_thr;
    d.max_refs = x.max_refs;
    d.n_synapses = x.n_synapses;
    d.source = x.source;
    d.target = x.target;
    d.weight = x.weight;
    d.max_weight = x.max_weight;
    d.delay = x.delay;
    d.compartment = x.compartment;
    d.receptor = x.receptor;
    d.plastic = x.plastic;
    d.code_ptr = x.code_ptr;
    d.codes = x.codes;
    return d;
}

std::unique_ptr<snn_net> create_net(const NetDesc& d, const snn_options& o) {
    if (d.n_neurons && (!d.threshold || !d.window || !d.sim_thr || !d.max_refs))
        throw std::invalid_argument("null neuron array");
    if (d.n_synapses && (!d.source || !d.target || !d.weight || !d.max_weight || !d.delay || !d.compartment ||
                         !d.receptor || !d.plastic))
        throw std::invalid_argument("null synapse array");
    (void)default_partition_shift(d.n_neurons);  // a network too large for one device fails here

    // net before h before up: on any error the uploader drains first, then the host arrays it reads
    // go, then the device buffers
    std::unique_ptr<snn_net> net(new snn_net());
    net->device = o.device;
    check(hipSetDevice(o.device), "hipSetDevice");
    auto& own = net->owned;
    Step& S = net->S;
    HostNet h;
    // SNNCUDA_CREATE_TIMING=1 prints the phases of the call to stderr
    const bool timing = std::getenv("SNNCUDA_CREATE_TIMING") != nullptr;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto t_pack = t0, t_up = t0;
    {
        Uploader up(o.device, d.n_synapses >= kStreamMinSynapses);
        build_host_net_into(h, d, 0, [&](PackStage stage, const HostNet& hh) {
            switch (stage) {
            case PackStage::Slots:
                S.tgt = up.upload(own, hh.tgt);
                S.w = up.upload(own, hh.w);
                S.meta = up.upload(own, hh.meta);
                S.delay = up.upload(own, hh.delay);
                S.expected = hh.expected.empty() ? nullptr : up.upload(own, hh.expected);
                S.np = up.upload(own, hh.params);
                net->slot_of = up.upload(own, hh.slot_of);
                break;
            case PackStage::Groups:
                S.row = up.upload(own, hh.row);
                S.grp = up.upload(own, hh.grp);
                S.ve = hh.ve_identity ? nullptr : up.upload(own, hh.ve);
                break;
            case PackStage::Stdp:
                if (hh.any_plastic) {
                    S.tpos = up.upload(own, hh.tpos);
                    S.in_ptr = up.upload(own, hh.in_ptr);
                    S.stdp = up.upload(own, hh.stdp);
                    net->tpos_conn_dev = up.upload(own, hh.tpos_conn);
                    net->n_tpos = hh.stdp.size();
                }
                break;
            }
        });
        t_pack = clk::now();
        up.finish();
        t_up = clk::now();
    }
    S.expected_all = h.expected_all ? h.expected_all : 2;
    open_net(net.get(), o, h.n, h.e, h.lmax, h.any_plastic, h.ve_identity, h.code_compact);
    finish_net(net.get(), o, h.nve, h.params);
    if (timing) {
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        auto sec = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        std::fprintf(stderr, "snn_net_create: %llu synapses, pack %.3f s, upload after the pack %.3f s, device state %.3f s (%s upload)\n",
                     (unsigned long long)d.n_synapses, sec(t0, t_pack), sec(t_pack, t_up), sec(t_up, clk::now()),
                     d.n_synapses >= kStreamMinSynapses ? "streamed" : "direct");
    }
    return net;
}

uint64_t stream_key_of(uint64_t seed, uint64_t stream) { return mix64(seed * kGold + stream); }

// Generates networks g[0..m) on the device as one union network (m = 1: a single network). noff /
// eoff receive the m + 1 neuron and synapse offsets.
std::unique_ptr<snn_net> generate_nets(const snn_ei_desc* g, uint64_t m, const snn_options& o,
                                       std::vector<uint64_t>& noff, std::vector<uint64_t>& eoff) {
    noff.assign(m + 1, 0);
    eoff.assign(m + 1, 0);
    std::vector<GenNet> nets(m);
    uint32_t lmax = 1, exp_all = 0;
    bool identity = true, exp_uniform = true, all_one = true, any_plastic = false;
    for (uint64_t k = 0; k < m; ++k) {
        const snn_ei_desc& d = g[k];
        const std::string at = " (network " + std::to_string(k) + ")";
        if (d.n < 2 || !d.fanout) throw std::invalid_argument("need at least two network neurons and a fan-out array" + at);
        if (d.delay_max < 1 || d.delay_max > 63) throw std::invalid_argument("delay_max must be in [1, 63]" + at);
        if (d.n_comps < 1 || d.n_comps > 4) throw std::invalid_argument("n_comps must be in [1, 4]" + at);
        if (d.window > 65535 || d.max_refs < 1) throw std::invalid_argument("bad temporal parameters" + at);
        if (d.n_codes > kCodeMax || d.code_max > 63 || (d.n_codes == 0 && d.code_max != 0))
            throw std::invalid_argument("n_codes must be in [1, 8] and code_max in [0, 63]" + at);
        for (uint32_t c = 0; c < d.n_comps; ++c)
            if (d.comps[c] > 3) throw std::invalid_argument("bad compartment" + at);
        if (d.source_fanout >= (1u << 24)) throw std::invalid_argument("fan-out must be below 2^24" + at);
        GenNet& N = nets[k];
        N.n = d.n;
        N.nall = d.n + d.n_sources;
        N.n_exc = d.n_exc;
        N.noff = noff[k];
        for (uint32_t st = 0; st < 10; ++st) N.key[st] = stream_key_of(d.seed, st);
        N.w_exc = d.w_exc;
        N.w_inh = d.w_inh;
        N.delay_max = d.delay_max;
        N.n_comps = d.n_comps;
        for (int c = 0; c < 4; ++c) N.comps[c] = d.comps[c];
        N.ncode = d.n_codes ? d.n_codes : 1;
        N.code_max = d.code_max;
        if (d.plastic < 0 || d.plastic > 2) throw std::invalid_argument("plastic must be 0, 1 or 2");
        N.plastic = (uint32_t)d.plastic;
        noff[k + 1] = noff[k] + N.nall;
        if (noff[k + 1] >= 0xFFFFFFFFull) throw std::invalid_argument("too many neurons");
        lmax = std::max<uint32_t>(lmax, d.delay_max + d.code_max);
        if (N.ncode != 1 || N.code_max != 0) identity = false;
        const uint32_t ex = N.ncode > 2 ? N.ncode : 2;
        if (k == 0) exp_all = ex;
        else if (ex != exp_all) exp_uniform = false;
        if (d.max_refs != 1) all_one = false;
        any_plastic |= N.plastic != 0;
    }
    const uint64_t n = noff[m];
    (void)default_partition_shift(n);  // a network too large for one device fails here
    const uint32_t W = lmax + 1;
    if (W > kGenBins) throw std::invalid_argument("delay_max + code_max above 127");
    if (n * W > (1ull << 31)) throw std::invalid_argument("neurons x (max lag + 1) exceeds the group-table limit");
    std::vector<uint64_t> srow(n + 1, 0), vrow(n + 1, 0);
    for (uint64_t k = 0; k < m; ++k) {
        const snn_ei_desc& d = g[k];
        const uint64_t b = noff[k];
        for (uint64_t i = 0; i < nets[k].nall; ++i) {
            const uint64_t f = i < d.n ? d.fanout[i] : d.source_fanout;
            if (f >= (1u << 24)) throw std::invalid_argument("fan-out must be below 2^24");
            srow[b + i + 1] = srow[b + i] + f;
            vrow[b + i + 1] = vrow[b + i] + f * nets[k].ncode;
        }
        eoff[k + 1] = srow[noff[k + 1]];
    }
    const uint64_t e = srow[n], nve = vrow[n];
    if (e >= 0xFFFFFFFFull) throw std::invalid_argument("more than 2^32-2 synapses on one device");

    std::unique_ptr<snn_net> net(new snn_net());
    const bool compact = exp_uniform && exp_all == 2 && all_one;
    open_net(net.get(), o, (uint32_t)n, e, lmax, any_plastic, identity, compact);
    auto& own = net->owned;
    Step& S = net->S;
    hipStream_t st = net->stream;
    GenParams gp{};
    gp.nets = dev_upload(own, nets);
    gp.n_nets = (uint32_t)m;
    gp.W = W;
    gp.nall = n;
    gp.srow = dev_upload(own, srow);
    S.row = dev_upload(own, vrow);
    gp.vrow = S.row;
    gp.grp = dev_alloc<uint32_t>(own, n * W);
    S.grp = gp.grp;
    gp.ve = identity ? nullptr : dev_alloc<uint32_t>(own, nve, false);
    S.ve = gp.ve;
    gp.tgt = dev_alloc<uint32_t>(own, e, false);
    gp.w = dev_alloc<float>(own, e, false);
    gp.meta = dev_alloc<uint8_t>(own, e, false);
    gp.delay = dev_alloc<uint16_t>(own, e, false);
    S.tgt = gp.tgt;
    S.w = gp.w;
    S.meta = gp.meta;
    S.delay = gp.delay;
    gp.expected = exp_uniform ? nullptr : dev_alloc<uint8_t>(own, e, false);
    S.expected = gp.expected;
    S.expected_all = exp_uniform ? exp_all : 2;
    net->slot_of = dev_alloc<uint32_t>(own, e, false);
    gp.slot_of = net->slot_of;
    const uint32_t rows_per_block = kBlock / kWave;
    const dim3 ggrid((uint32_t)((n + rows_per_block - 1) / rows_per_block));
    if (any_plastic) gp.incnt = dev_alloc<uint32_t>(own, n);
    hipLaunchKernelGGL(k_gen_count, ggrid, dim3(kBlock), 0, st, gp);
    if (any_plastic) {
        // target-major STDP records
        uint32_t* in32 = dev_alloc<uint32_t>(own, n + 1);
        net->partials = dev_alloc<uint32_t>(own, kScanBlocks);
        launch_scan(net.get(), PlainSrc{gp.incnt, nullptr, (uint32_t)n}, in32, &in32[n], 0, 0, nullptr);
        uint64_t* in_ptr = dev_alloc<uint64_t>(own, n + 1);
        hipLaunchKernelGGL(k_widen, dim3(1024), dim3(kBlock), 0, st, in32, n + 1, in_ptr);
        check(hipMemsetAsync(gp.incnt, 0, n * 4, st), "memset");
        uint32_t n_tpos = 0;
        check(hipMemcpyAsync(&n_tpos, &in32[n], 4, hipMemcpyDeviceToHost, st), "read plastic count");
        check(hipStreamSynchronize(st), "generate");
        S.in_ptr = in_ptr;
        gp.in_ptr = in_ptr;
        gp.tpos = dev_alloc<uint32_t>(own, e, false);
        gp.stdp = dev_alloc<StdpRec>(own, n_tpos, false);
        gp.tpos_conn = dev_alloc<uint32_t>(own, n_tpos, false);
        S.tpos = gp.tpos;
        S.stdp = gp.stdp;
        net->tpos_conn_dev = gp.tpos_conn;
        net->n_tpos = n_tpos;
    }
    hipLaunchKernelGGL(k_gen_place, ggrid, dim3(kBlock), 0, st, gp);
    std::vector<NeuronParams> params(n);
    for (uint64_t k = 0; k < m; ++k) {
        const snn_ei_desc& d = g[k];
        for (uint64_t i = 0; i < nets[k].nall; ++i) {
            NeuronParams& q = params[noff[k] + i];
            q.threshold = i < d.n ? d.threshold : d.source_threshold;
            q.window = d.window;
            q.sim_thr = d.sim_thr;
            q.max_refs = d.max_refs;
        }
    }
    S.np = dev_upload(own, params);
    check(hipStreamSynchronize(st), "generate");
    finish_net(net.get(), o, nve, params);
    return net;
}

}  // namespace

// the thread's last-error message, shared with the library's other translation units
void snn_internal_set_error(const std::string& msg) { g_last_error = msg; }

// ------------------------------------------------------------------------------------ C ABI

namespace {

// per STDP record (record order): changed updates and those with delta > 0
void read_stdp_counts(snn_net* net, uint32_t* upd, uint32_t* ltp) {
    const uint64_t nt = net->n_tpos;
    std::vector<DevBuf> tmp;
    uint32_t* du = dev_alloc<uint32_t>(tmp, nt, false);
    uint32_t* dl = dev_alloc<uint32_t>(tmp, nt, false);
    const uint64_t nb = std::min<uint64_t>((nt + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_stdp_counts, dim3((uint32_t)nb), dim3(kBlock), 0, net->stream, (const StdpRec*)net->S.stdp, nt, du, dl);
    check(hipGetLastError(), "k_stdp_counts");
    check(hipMemcpyAsync(upd, du, nt * 4, hipMemcpyDeviceToHost, net->stream), "read stdp");
    check(hipMemcpyAsync(ltp, dl, nt * 4, hipMemcpyDeviceToHost, net->stream), "read stdp");
    check(hipStreamSynchronize(net->stream), "read stdp");
}

}  // namespace

extern "C" {

const char* snncuda_version(void) { return SNNCUDA_VERSION; }
int snncuda_cuda_available(void) { return 1; }

int snn_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* snn_last_error(void) { return g_last_error.c_str(); }

int snn_net_create(const snn_net_desc* desc, const snn_options* opts, snn_net** out) {
    if (!desc || !out) return fail(SNN_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&]() -> int {
        snn_options o{};
        if (opts) o = *opts;
        const int rc = check_device(o);
        if (rc) return rc;
        *out = create_net(to_net_desc(*desc), o).release();
        return SNN_OK;
    });
}

int snn_net_create_ei(const snn_ei_desc* g, const snn_options* opts, snn_net** out) {
    if (!g || !out) return fail(SNN_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&]() -> int {
        snn_options o{};
        if (opts) o = *opts;
        const int rc = check_device(o);
        if (rc) return rc;
        std::vector<uint64_t> noff, eoff;
        *out = generate_nets(g, 1, o, noff, eoff).release();
        return SNN_OK;
    });
}

int snn_batch_create_ei(const snn_ei_desc* nets, uint64_t n_nets, const snn_options* opts, snn_batch** out) {
    if (!nets || !out) return fail(SNN_EINVAL, "null argument");
    *out = nullptr;
    if (n_nets == 0) return fail(SNN_EINVAL, "empty batch");
    return guarded([&]() -> int {
        snn_options o{};
        if (opts) o = *opts;
        const int rc = check_device(o);
        if (rc) return rc;
        std::unique_ptr<snn_batch> b(new snn_batch());
        b->net = generate_nets(nets, n_nets, o, b->noff, b->eoff);
        *out = b.release();
        return SNN_OK;
    });
}

int snn_net_destroy(snn_net* net) {
    delete net;
    return SNN_OK;
}

int snn_net_record_fire_events(snn_net* net, int enabled) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    if (net->next_tick != 0) return fail(SNN_EINVAL, "fire events can only be recorded from tick 0");
    return guarded([&]() -> int {
        check(hipSetDevice(net->device), "hipSetDevice");
        net->record = enabled != 0;
        if (net->record && !net->S.fire_evt) {
            net->S.fire_evt = dev_alloc<uint32_t>(net->owned, net->S.ev_cap, false);
            net->S.raster_evt = dev_alloc<uint32_t>(net->owned, net->S.raster_cap, false);
        }
        if (!net->record) {
            net->S.fire_evt = nullptr;
            net->S.raster_evt = nullptr;
        }
        return SNN_OK;
    });
}

int snn_net_read_fire_events(snn_net* net, uint64_t* n_out, uint32_t* event_index, uint64_t cap) {
    if (!net || !n_out) return fail(SNN_EINVAL, "null argument");
    if (!net->S.raster_evt) return fail(SNN_EINVAL, "fire events were not recorded (snn_net_record_fire_events)");
    return guarded([&]() -> int {
        check(hipSetDevice(net->device), "hipSetDevice");
        uint64_t tot[4];
        check(hipMemcpy(tot, net->S.totals, sizeof(tot), hipMemcpyDeviceToHost), "read totals");
        uint32_t ctl[4];
        check(hipMemcpy(ctl, net->S.ctl, sizeof(ctl), hipMemcpyDeviceToHost), "read ctl");
        *n_out = tot[3];
        if (event_index && tot[3]) {
            const uint64_t m = std::min(cap, tot[3]);
            check(hipMemcpy(event_index, net->S.raster_evt, m * 4, hipMemcpyDeviceToHost), "read fire events");
        }
        if (ctl[3] & 1u) return fail(SNN_ECAPACITY, "raster capacity reached: only the first fires were recorded");
        return SNN_OK;
    });
}

int snn_net_read_inject_log(snn_net* net, uint64_t* n_out, uint32_t* neuron, uint64_t* tick, uint64_t* stage,
                            float* weight, uint64_t cap) {
    if (!net || !n_out) return fail(SNN_EINVAL, "null argument");
    if (!net->record) return fail(SNN_EINVAL, "injects were not recorded (snn_net_record_fire_events)");
    *n_out = net->inj_log.size();
    const uint64_t m = std::min<uint64_t>(cap, net->inj_log.size());
    for (uint64_t i = 0; i < m; ++i) {
        const PendingInject& q = net->inj_log[i];
        if (neuron) neuron[i] = q.neuron;
        if (tick) tick[i] = q.tick;
        if (stage) stage[i] = q.stage;
        if (weight) weight[i] = q.weight;
    }
    return SNN_OK;
}

int snn_net_get_schedule(snn_net* net, snn_schedule* out) {
    if (!net || !out) return fail(SNN_EINVAL, "null argument");
    out->partition_shift = net->S.pshift;
    out->scatter_workgroups = net->S.nwg;
    out->graph_ticks = net->graph_ticks;
    out->fold_threads = net->fold_threads;
    return SNN_OK;
}

int snn_net_set_schedule(snn_net* net, const snn_schedule* sc) {
    if (!net || !sc) return fail(SNN_EINVAL, "null argument");
    return guarded([&]() -> int {
        apply_schedule(net, Schedule{sc->partition_shift, sc->scatter_workgroups, sc->graph_ticks,
                                     sc->fold_threads ? sc->fold_threads : (uint32_t)kPBlock});
        return SNN_OK;
    });
}

int snn_net_autotune(snn_net* net, uint32_t trial_ticks) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    if (trial_ticks > 4096) return fail(SNN_EINVAL, "trial_ticks above 4096");
    net->tune_trial = trial_ticks;
    return SNN_OK;
}

int snn_net_tune_log(snn_net* net, char* buf, uint64_t cap) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    if (buf && cap) {
        const size_t m = std::min<size_t>(cap - 1, net->tune_log.size());
        std::memcpy(buf, net->tune_log.data(), m);
        buf[m] = 0;
    }
    return (int)std::min<size_t>(net->tune_log.size(), 0x7FFFFFFF);
}

int snn_batch_create(const snn_net_desc* nets, uint64_t n_nets, const snn_options* opts, snn_batch** out) {
    if (!out || (n_nets && !nets)) return fail(SNN_EINVAL, "null argument");
    *out = nullptr;
    if (n_nets == 0) return fail(SNN_EINVAL, "empty batch");
    return guarded([&]() -> int {
        snn_options o{};
        if (opts) o = *opts;
        const int rc = check_device(o);
        if (rc) return rc;
        std::unique_ptr<snn_batch> b(new snn_batch());
        b->noff.assign(n_nets + 1, 0);
        b->eoff.assign(n_nets + 1, 0);
        uint64_t ncodes = 0;
        bool any_codes = false;
        for (uint64_t k = 0; k < n_nets; ++k) {
            const snn_net_desc& x = nets[k];
            if (x.n_neurons == 0) return fail(SNN_EINVAL, "network " + std::to_string(k) + " has no neurons");
            if (!x.threshold || !x.window || !x.sim_thr || !x.max_refs)
                return fail(SNN_EINVAL, "null neuron array in network " + std::to_string(k));
            if (x.n_synapses && (!x.source || !x.target || !x.weight || !x.max_weight || !x.delay || !x.compartment ||
                                 !x.receptor || !x.plastic))
                return fail(SNN_EINVAL, "null synapse array in network " + std::to_string(k));
            for (uint64_t j = 0; j < x.n_synapses; ++j)
                if (x.source[j] >= x.n_neurons || x.target[j] >= x.n_neurons)
                    return fail(SNN_EINVAL, "synapse endpoint out of range in network " + std::to_string(k));
            b->noff[k + 1] = b->noff[k] + x.n_neurons;
            b->eoff[k + 1] = b->eoff[k] + x.n_synapses;
            if (x.code_ptr) {
                any_codes = true;
                ncodes += x.code_ptr[x.n_synapses] - x.code_ptr[0];
            }
        }
        const uint64_t n = b->noff[n_nets], e = b->eoff[n_nets];
        if (n >= 0xFFFFFFFFull) return fail(SNN_EINVAL, "more than 2^32-2 neurons in the batch");
        std::vector<float> thr(n), sim(n), w(e), mw(e);
        std::vector<uint64_t> win(n), mref(n), cptr(any_codes ? e + 1 : 0);
        std::vector<uint32_t> src(e), tgt(e), dl(e), codes;
        std::vector<uint8_t> comp(e), rec(e), pl(e);
        codes.reserve(ncodes);
        if (any_codes) cptr[0] = 0;
        for (uint64_t k = 0; k < n_nets; ++k) {
            const snn_net_desc& x = nets[k];
            const uint64_t n0 = b->noff[k], e0 = b->eoff[k];
            std::copy(x.threshold, x.threshold + x.n_neurons, thr.begin() + n0);
            std::copy(x.window, x.window + x.n_neurons, win.begin() + n0);
            std::copy(x.sim_thr, x.sim_thr + x.n_neurons, sim.begin() + n0);
            std::copy(x.max_refs, x.max_refs + x.n_neurons, mref.begin() + n0);
            for (uint64_t j = 0; j < x.n_synapses; ++j) {
                src[e0 + j] = (uint32_t)(x.source[j] + n0);
                tgt[e0 + j] = (uint32_t)(x.target[j] + n0);
            }
            std::copy(x.weight, x.weight + x.n_synapses, w.begin() + e0);
            std::copy(x.max_weight, x.max_weight + x.n_synapses, mw.begin() + e0);
            std::copy(x.delay, x.delay + x.n_synapses, dl.begin() + e0);
            std::copy(x.compartment, x.compartment + x.n_synapses, comp.begin() + e0);
            std::copy(x.receptor, x.receptor + x.n_synapses, rec.begin() + e0);
            std::copy(x.plastic, x.plastic + x.n_synapses, pl.begin() + e0);
            if (any_codes)
                for (uint64_t j = 0; j < x.n_synapses; ++j) {
                    if (x.code_ptr)  // an empty list stands for {0}, as a missing one does
                        codes.insert(codes.end(), x.codes + x.code_ptr[j], x.codes + x.code_ptr[j + 1]);
                    cptr[e0 + j + 1] = codes.size();
                }
        }
        NetDesc d;
        d.n_neurons = n;
        d.threshold = thr.data();
        d.window = win.data();
        d.sim_thr = sim.data();
        d.max_refs = mref.data();
        d.n_synapses = e;
        d.source = src.data();
        d.target = tgt.data();
        d.weight = w.data();
        d.max_weight = mw.data();
        d.delay = dl.data();
        d.compartment = comp.data();
        d.receptor = rec.data();
        d.plastic = pl.data();
        d.code_ptr = any_codes ? cptr.data() : nullptr;
        d.codes = any_codes ? codes.data() : nullptr;
        b->net = create_net(d, o);
        *out = b.release();
        return SNN_OK;
    });
}

int snn_batch_destroy(snn_batch* b) {
    delete b;
    return SNN_OK;
}

int snn_batch_network(snn_batch* b, snn_net** net) {
    if (!b || !net) return fail(SNN_EINVAL, "null argument");
    *net = b->net.get();
    return SNN_OK;
}

int snn_batch_offsets(snn_batch* b, uint64_t* n_nets, uint64_t* neuron_off, uint64_t* synapse_off) {
    if (!b) return fail(SNN_EINVAL, "null handle");
    if (n_nets) *n_nets = b->noff.size() - 1;
    if (neuron_off) std::copy(b->noff.begin(), b->noff.end(), neuron_off);
    if (synapse_off) std::copy(b->eoff.begin(), b->eoff.end(), synapse_off);
    return SNN_OK;
}

int snn_batch_inject(snn_batch* b, const uint32_t* net, const uint32_t* neuron, const uint64_t* tick,
                     const float* weight, uint64_t n) {
    if (!b || (n && (!net || !neuron || !tick || !weight))) return fail(SNN_EINVAL, "null argument");
    const uint64_t nn = b->noff.size() - 1;
    std::vector<uint32_t> g(n);
    for (uint64_t i = 0; i < n; ++i) {
        if (net[i] >= nn) return fail(SNN_EINVAL, "inject network out of range");
        if (neuron[i] >= b->noff[net[i] + 1] - b->noff[net[i]]) return fail(SNN_EINVAL, "inject neuron out of range");
        g[i] = (uint32_t)(b->noff[net[i]] + neuron[i]);
    }
    const int rc = snn_net_inject(b->net.get(), g.data(), tick, weight, n);
    if (rc) return rc;
    for (uint64_t i = 0; i < n; ++i) {
        b->inj_net.push_back(net[i]);
        b->inj_tick.push_back(std::max<uint64_t>(tick[i], b->net->next_tick));
    }
    return SNN_OK;
}

int snn_batch_run_until(snn_batch* b, uint64_t inclusive_tick) {
    if (!b) return fail(SNN_EINVAL, "null handle");
    return snn_net_run_until(b->net.get(), inclusive_tick);
}

int snn_batch_net_counts(snn_batch* b, uint64_t* delivered, uint64_t* fired) {
    if (!b) return fail(SNN_EINVAL, "null handle");
    snn_net* net = b->net.get();
    if (delivered && !net->S.pre_count) return fail(SNN_EINVAL, "per-network delivered counts need track_pre_count");
    return guarded([&]() -> int {
        const uint64_t nn = b->noff.size() - 1;
        if (fired) {
            std::vector<uint64_t> sc(net->n);
            const int rc = snn_net_read_neurons(net, nullptr, nullptr, nullptr, sc.data(), nullptr, nullptr, nullptr);
            if (rc) return rc;
            for (uint64_t k = 0; k < nn; ++k) {
                uint64_t f = 0;
                for (uint64_t i = b->noff[k]; i < b->noff[k + 1]; ++i) f += sc[i];
                fired[k] = f;
            }
        }
        if (delivered) {
            std::vector<uint64_t> pc(net->e);
            const int rc = snn_net_read_synapses(net, nullptr, pc.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
            if (rc) return rc;
            for (uint64_t k = 0; k < nn; ++k) {
                uint64_t d = 0;
                for (uint64_t j = b->eoff[k]; j < b->eoff[k + 1]; ++j) d += pc[j];
                delivered[k] = d;
            }
            for (size_t i = 0; i < b->inj_net.size(); ++i)
                if (b->inj_tick[i] < net->next_tick) delivered[b->inj_net[i]] += 1;
        }
        return SNN_OK;
    });
}

int snn_net_inject(snn_net* net, const uint32_t* neuron, const uint64_t* tick, const float* weight, uint64_t n) {
    if (!net || (n && (!neuron || !tick || !weight))) return fail(SNN_EINVAL, "null argument");
    for (uint64_t i = 0; i < n; ++i)
        if (neuron[i] >= net->n) return fail(SNN_EINVAL, "inject neuron out of range");
    if (net->pending.capacity() < net->pending.size() + n)  // geometric growth for many small calls
        net->pending.reserve(std::max<size_t>(net->pending.size() + n, 2 * net->pending.capacity()));
    for (uint64_t i = 0; i < n; ++i) {
        // SpikeScheduler::schedule clamps a late event to the next poll tick
        const uint64_t t = std::max<uint64_t>(tick[i], net->next_tick);
        net->pending.push_back({neuron[i], t, weight[i], net->next_tick, net->inject_seq++});
    }
    return SNN_OK;
}

}  // extern "C"

namespace {

// Staging of snn_net_run_until_static: the run segment's inject tables and per-tick counters in a
// fixed device array instead of buffers allocated per call. Every call on a device shares that
// device's copy of the array (calls that overlap on one device race); it holds the last call's
// contents, all of which a call overwrites before its kernels read them.
constexpr size_t kStaticScratchBytes = 4u << 20;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

// Device address of `scratch` on each device, looked up once per device with that device current
// (the host-side name is only a stand-in for it).
uint8_t* scratch_on(int device) {
    constexpr int kMaxDevices = 64;
    static std::mutex m;
    static uint8_t* ptr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device ordinal out of range");
    std::lock_guard<std::mutex> lock(m);
    if (!ptr[device]) {
        check(hipSetDevice(device), "hipSetDevice");
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        ptr[device] = (uint8_t*)p;
    }
    return ptr[device];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// run_until with the segment's buffers allocated per call (static_staging false), or staged in
// `scratch` with nothing allocated or freed on the device (true: stream launches only, no HIP
// graph, no autotuning, no profiling events)
// Paged learned-pattern pool: the cache holds the pages most events reached since the last update
// (ties: pages already held, then lower pages). Evicted pages are copied back to their host place and
// loaded ones into the freed slots, on the step's stream before the next ticks; the page table
// follows. Results do not depend on where a page is.
void update_residency(snn_net* net) {
    if (!net->paged) return;
    Step& S = net->S;
    const uint32_t np = net->npages, nc = net->cache_pages;
    std::vector<uint32_t> hits(np);
    check(hipMemcpyAsync(hits.data(), S.page_hits, (size_t)np * 4, hipMemcpyDeviceToHost, net->stream), "read page hits");
    check(hipStreamSynchronize(net->stream), "hipStreamSynchronize");
    std::vector<uint32_t> order(np);
    for (uint32_t q = 0; q < np; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        if (hits[a] != hits[b]) return hits[a] > hits[b];
        return (net->slot_of_page[a] >= 0) > (net->slot_of_page[b] >= 0);
    });
    std::vector<uint8_t> want(np, 0);
    for (uint32_t k = 0; k < nc && k < np; ++k) want[order[k]] = 1;
    const uint64_t page_slots = ((uint64_t)1 << kRefPageShift) * S.ref_cap;
    const uint64_t page_runs = page_slots * S.run_cap;  // a page's learned-pattern runs
    auto host_runs = [&](uint32_t q) { return net->host_runs + (uint64_t)q * page_runs; };
    auto host_meta = [&](uint32_t q) { return net->host_meta + (uint64_t)q * page_slots; };
    Run* dr = nullptr;
    RefMeta* dm = nullptr;
    check(hipHostGetDevicePointer((void**)&dr, net->host_runs, 0), "hipHostGetDevicePointer");
    check(hipHostGetDevicePointer((void**)&dm, net->host_meta, 0), "hipHostGetDevicePointer");
    bool moved = false;
    for (uint32_t slot = 0; slot < nc; ++slot) {  // evictions
        const int32_t q = net->page_of_slot[slot];
        if (q < 0 || want[q]) continue;
        check(hipMemcpyAsync(host_runs(q), net->cache_runs + slot * page_runs, page_runs * sizeof(Run),
                             hipMemcpyDeviceToHost, net->stream), "page out");
        check(hipMemcpyAsync(host_meta(q), net->cache_meta + slot * page_slots, page_slots * sizeof(RefMeta),
                             hipMemcpyDeviceToHost, net->stream), "page out");
        net->table[q] = RefPage{dr + (uint64_t)q * page_runs, dm + (uint64_t)q * page_slots};
        net->slot_of_page[q] = -1;
        net->page_of_slot[slot] = -1;
        moved = true;
        ++net->page_moves;
    }
    uint32_t free_slot = 0;
    for (uint32_t k = 0; k < nc && k < np; ++k) {  // loads into free slots
        const uint32_t q = order[k];
        if (net->slot_of_page[q] >= 0) continue;
        while (net->page_of_slot[free_slot] >= 0) ++free_slot;
        check(hipMemcpyAsync(net->cache_runs + free_slot * page_runs, host_runs(q), page_runs * sizeof(Run),
                             hipMemcpyHostToDevice, net->stream), "page in");
        check(hipMemcpyAsync(net->cache_meta + free_slot * page_slots, host_meta(q), page_slots * sizeof(RefMeta),
                             hipMemcpyHostToDevice, net->stream), "page in");
        net->table[q] = RefPage{net->cache_runs + free_slot * page_runs, net->cache_meta + free_slot * page_slots};
        net->slot_of_page[q] = (int32_t)free_slot;
        net->page_of_slot[free_slot] = (int32_t)q;
        moved = true;
        ++net->page_moves;
    }
    if (moved)
        check(hipMemcpyAsync(S.ref_page, net->table.data(), (size_t)np * sizeof(RefPage), hipMemcpyHostToDevice,
                             net->stream), "page table");
    check(hipMemsetAsync(S.page_hits, 0, (size_t)np * 4, net->stream), "clear page hits");
    // the table and the host copies stay valid until the stream has used them
    check(hipStreamSynchronize(net->stream), "hipStreamSynchronize");
}

int run_until_impl(snn_net* net, uint64_t inclusive_tick, bool static_staging) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    if (inclusive_tick < net->next_tick) return SNN_OK;
    return guarded([&]() -> int {
        if (inclusive_tick + net->S.lmax + 2 >= 0xFFFFFFFFull) return fail(SNN_EINVAL, "tick beyond 2^32");
        check(hipSetDevice(net->device), "hipSetDevice");
        const uint64_t t0 = net->next_tick;
        const uint64_t nt = inclusive_tick - t0 + 1;
        // injects delivered inside this segment, ordered by (tick, stage, call order). `pending` is
        // in call order and a call's stage (the clock when it was made) never decreases along it,
        // so a stable counting sort by tick gives that order: iptr holds each tick's first place.
        // Every pending tick is >= t0 (snn_net_inject clamps to the clock; earlier runs took the rest).
        const std::vector<PendingInject>& pend = net->pending;
        std::vector<uint64_t> iptr(nt + 1, 0);
        for (const auto& p : pend)
            if (p.tick <= inclusive_tick) iptr[p.tick - t0 + 1]++;
        for (uint64_t t = 0; t < nt; ++t) iptr[t + 1] += iptr[t];
        // staging layout of snn_net_run_until_static, checked before anything changes
        const size_t ni = iptr[nt];
        const size_t o_w = align256(ni * 4), o_stage = o_w + align256(ni * 4), o_ptr = o_stage + align256(ni * 4);
        const size_t o_del = o_ptr + align256((nt + 1) * 8), o_fir = o_del + align256(nt * 4);
        const size_t need = o_fir + align256(nt * 4);
        if (static_staging && need > kStaticScratchBytes)
            return fail(SNN_EINVAL, "snn_net_run_until_static: the run segment (" + std::to_string(nt) + " ticks, " +
                                        std::to_string(ni) + " injects) needs " + std::to_string(need) +
                                        " bytes of staging; the static scratch holds " +
                                        std::to_string(kStaticScratchBytes) + " (use shorter segments)");
        std::vector<uint32_t> itgt(ni), istage(ni);
        std::vector<float> iw(ni);
        std::vector<uint64_t> cursor(iptr.begin(), iptr.end() - 1);
        std::vector<size_t> order(net->record ? ni : 0);  // place -> index in pend, for the inject log
        std::vector<PendingInject> later;
        later.reserve(pend.size() - ni);
        for (size_t i = 0; i < pend.size(); ++i) {
            const PendingInject& p = pend[i];
            if (p.tick > inclusive_tick) {
                later.push_back(p);
                continue;
            }
            const uint64_t k = cursor[p.tick - t0]++;
            itgt[k] = p.neuron;
            iw[k] = p.weight;
            istage[k] = (uint32_t)p.stage;
            if (net->record) order[k] = i;
        }
        if (net->record)
            for (size_t k = 0; k < ni; ++k) net->inj_log.push_back(pend[order[k]]);
        net->pending.swap(later);
        Step& S = net->S;
        if (static_staging) {
            uint8_t* base = scratch_on(net->device);
            check(hipSetDevice(net->device), "hipSetDevice");
            if (ni) {
                check(hipMemcpy(base, itgt.data(), ni * 4, hipMemcpyHostToDevice), "stage injects");
                check(hipMemcpy(base + o_w, iw.data(), ni * 4, hipMemcpyHostToDevice), "stage injects");
                check(hipMemcpy(base + o_stage, istage.data(), ni * 4, hipMemcpyHostToDevice), "stage injects");
            }
            check(hipMemcpy(base + o_ptr, iptr.data(), (nt + 1) * 8, hipMemcpyHostToDevice), "stage injects");
            S.inj_tgt = (const uint32_t*)base;
            S.inj_w = (const float*)(base + o_w);
            S.inj_stage = (const uint32_t*)(base + o_stage);
            S.inj_ptr = (const uint64_t*)(base + o_ptr);
            // every tick of the segment writes its two counters (k_advance), so they need no clearing
            S.tick_delivered = (uint32_t*)(base + o_del);
            S.tick_fired = (uint32_t*)(base + o_fir);
        } else {
            for (auto& b : net->run_owned) (void)hipFree(b.p);
            net->run_owned.clear();
            S.inj_tgt = dev_upload(net->run_owned, itgt);
            S.inj_w = dev_upload(net->run_owned, iw);
            S.inj_stage = dev_upload(net->run_owned, istage);
            S.inj_ptr = dev_upload(net->run_owned, iptr);
            S.tick_delivered = dev_alloc<uint32_t>(net->run_owned, nt);
            S.tick_fired = dev_alloc<uint32_t>(net->run_owned, nt);
        }
        S.inj_t0 = (uint32_t)t0;
        S.inj_nticks = (uint32_t)nt;
        {
            const uint32_t tick32 = (uint32_t)t0;
            check(hipMemcpy(S.ctl, &tick32, 4, hipMemcpyHostToDevice), "set tick");
        }
        Timer tm{net};
        net->tick_next = t0;
        net->advance_pending = false;
        if (static_staging) {
            tm.enabled = false;
            for (uint64_t t = 0; t < nt; ++t) launch_tick(net, tm);
            flush_advance(net, tm);
        } else if (net->profiling) {
            for (uint64_t t = 0; t < nt; ++t) {
                launch_tick(net, tm);
                if (tm.marks.size() > 4096) {
                    check(hipStreamSynchronize(net->stream), "sync");
                    tm.collect();
                }
            }
            flush_advance(net, tm);
        } else {
            TickGraph g;
            const uint64_t done = net->tune_trial ? autotune(net, (uint32_t)t0, nt, tm, g) : 0;
            run_ticks(net, (uint32_t)t0, nt - done, tm, g);
        }
        check(hipGetLastError(), "launch");
        check(hipStreamSynchronize(net->stream), "run");
        tm.collect();
        uint32_t ctl[4];
        check(hipMemcpy(ctl, S.ctl, sizeof(ctl), hipMemcpyDeviceToHost), "read ctl");
        std::vector<uint32_t> td(nt), tf(nt);
        check(hipMemcpy(td.data(), S.tick_delivered, nt * 4, hipMemcpyDeviceToHost), "read stats");
        check(hipMemcpy(tf.data(), S.tick_fired, nt * 4, hipMemcpyDeviceToHost), "read stats");
        net->tick_delivered.resize(t0);
        net->tick_fired.resize(t0);
        for (uint64_t t = 0; t < nt; ++t) {
            net->tick_delivered.push_back(td[t]);
            net->tick_fired.push_back(tf[t]);
        }
        net->next_tick = inclusive_tick + 1;
        if (ctl[1]) {
            std::string m = "device capacity exceeded during run (flags 0x" + std::to_string(ctl[1]) + "):";
            if (ctl[1] & kErrEvents) m += " events-per-tick";
            if (ctl[1] & kErrSegments) m += " segments-per-tick";
            if (ctl[1] & kErrFireLog) m += " fire-log";
            if (ctl[1] & kErrTemporal) m += " neuron-temporal-pattern";
            if (ctl[1] & kErrCode) m += " synapse-code-pattern";
            if (ctl[1] & kErrRaster) m += " raster";
            if (ctl[1] & kErrCounter) m += " match-counter";
            return fail(SNN_ECAPACITY, m);
        }
        return SNN_OK;
    });
}

}  // namespace

extern "C" {

// A paged learned-pattern pool updates its device cache every kPageEpoch ticks: the run is cut into
// segments of that many ticks (cutting a run changes no result).
constexpr uint64_t kPageEpoch = 64;
int run_paged(snn_net* net, uint64_t inclusive_tick, bool static_staging) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    if (!net->paged) return run_until_impl(net, inclusive_tick, static_staging);
    while (net->next_tick <= inclusive_tick) {
        const int rc = guarded([&]() -> int {
            check(hipSetDevice(net->device), "hipSetDevice");
            update_residency(net);
            return SNN_OK;
        });
        if (rc != SNN_OK) return rc;
        const uint64_t end = std::min(inclusive_tick, net->next_tick + kPageEpoch - 1);
        const int r = run_until_impl(net, end, static_staging);
        if (r != SNN_OK) return r;
    }
    return SNN_OK;
}

int snn_net_run_until(snn_net* net, uint64_t inclusive_tick) { return run_paged(net, inclusive_tick, false); }

int snn_net_run_until_static(snn_net* net, uint64_t inclusive_tick) { return run_paged(net, inclusive_tick, true); }

int snn_net_pattern_paging(snn_net* net, uint64_t* pages, uint64_t* device_pages, uint64_t* moves) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    const uint64_t all = ((uint64_t)net->n + (1u << kRefPageShift) - 1) >> kRefPageShift;
    if (pages) *pages = net->paged ? net->npages : all;
    if (device_pages) *device_pages = net->paged ? net->cache_pages : 0;
    if (moves) *moves = net->page_moves;
    return SNN_OK;
}

int snn_net_counts(snn_net* net, uint64_t* delivered, uint64_t* fired, uint64_t* next_tick) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    return guarded([&]() -> int {
        uint64_t tot[4];
        check(hipSetDevice(net->device), "hipSetDevice");
        check(hipMemcpy(tot, net->S.totals, sizeof(tot), hipMemcpyDeviceToHost), "read totals");
        if (delivered) *delivered = tot[0];
        if (fired) *fired = tot[1];
        if (next_tick) *next_tick = net->next_tick;
        return SNN_OK;
    });
}

int snn_net_read_tick_counts(snn_net* net, uint64_t first_tick, uint64_t n, uint64_t* delivered, uint64_t* fired) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    if (first_tick + n > net->tick_delivered.size()) return fail(SNN_EINVAL, "tick range not processed yet");
    for (uint64_t i = 0; i < n; ++i) {
        if (delivered) delivered[i] = net->tick_delivered[first_tick + i];
        if (fired) fired[i] = net->tick_fired[first_tick + i];
    }
    return SNN_OK;
}

int snn_net_read_raster(snn_net* net, uint64_t* n_out, uint32_t* neurons, uint64_t cap) {
    if (!net || !n_out) return fail(SNN_EINVAL, "null argument");
    return guarded([&]() -> int {
        check(hipSetDevice(net->device), "hipSetDevice");
        uint64_t tot[4];
        check(hipMemcpy(tot, net->S.totals, sizeof(tot), hipMemcpyDeviceToHost), "read totals");
        uint32_t ctl[4];
        check(hipMemcpy(ctl, net->S.ctl, sizeof(ctl), hipMemcpyDeviceToHost), "read ctl");
        *n_out = tot[3];
        if (neurons && tot[3]) {
            const uint64_t m = std::min(cap, tot[3]);
            check(hipMemcpy(neurons, net->S.raster, m * 4, hipMemcpyDeviceToHost), "read raster");
        }
        if (ctl[3] & 1u) return fail(SNN_ECAPACITY, "raster capacity reached: only the first fires were recorded");
        return SNN_OK;
    });
}

int snn_net_read_neurons(snn_net* net, float* membrane, float* currents, float* taus, uint64_t* spike_count,
                         int64_t* last_fired, uint64_t* temporal_match, uint64_t* temporal_learned) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    return guarded([&]() -> int {
        check(hipSetDevice(net->device), "hipSetDevice");
        const uint32_t n = net->n;
        std::vector<NeuronDyn> dyn(n);
        std::vector<TemporalHdr> th(n);
        std::vector<uint32_t> lf(n);
        check(hipMemcpy(dyn.data(), net->S.dyn, n * sizeof(NeuronDyn), hipMemcpyDeviceToHost), "read dyn");
        check(hipMemcpy(th.data(), net->S.th, n * sizeof(TemporalHdr), hipMemcpyDeviceToHost), "read temporal");
        check(hipMemcpy(lf.data(), net->S.last_fired, n * 4, hipMemcpyDeviceToHost), "read last fired");
        static const float tau_of[4] = {5.0f, 100.0f, 10.0f, 150.0f};
        for (uint32_t i = 0; i < n; ++i) {
            if (membrane) membrane[i] = dyn[i].membrane;
            for (int k = 0; k < 4; ++k) {
                if (currents) currents[(size_t)k * n + i] = dyn[i].cur[k];
                if (taus) taus[(size_t)k * n + i] = tau_of[(dyn[i].taus >> (2 * k)) & 3u];
            }
            if (spike_count) spike_count[i] = dyn[i].spike_count;
            if (last_fired) last_fired[i] = lf[i] == kNone ? -1 : (int64_t)lf[i];
            if (temporal_match) temporal_match[i] = th[i].match;
            if (temporal_learned) temporal_learned[i] = th[i].nlearned;
        }
        return SNN_OK;
    });
}

int snn_net_read_synapses(snn_net* net, float* weight, uint64_t* pre_count, uint64_t* update_count,
                          int64_t* last_pre, int64_t* last_post, uint64_t* code_match, uint64_t* code_learned) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    if (pre_count && !net->S.pre_count) return fail(SNN_EINVAL, "pre_count needs track_pre_count");
    return guarded([&]() -> int {
        check(hipSetDevice(net->device), "hipSetDevice");
        const uint64_t e = net->e;
        if (e == 0) return SNN_OK;
        const dim3 g(1024), b(kBlock);
        std::vector<DevBuf> tmp;
        struct Free {
            std::vector<DevBuf>& v;
            ~Free() { for (auto& x : v) (void)hipFree(x.p); }
        } fr{tmp};
        if (weight) {
            float* d = dev_alloc<float>(tmp, e, false);
            hipLaunchKernelGGL(k_gather_weights, g, b, 0, net->stream, net->slot_of, net->S.w, net->S.tpos, net->S.stdp, e, d);
            check(hipMemcpyAsync(weight, d, e * 4, hipMemcpyDeviceToHost, net->stream), "read weights");
        }
        const CodeRec2* c2 =
            net->code2 && (code_match || code_learned || last_pre || last_post) ? code2_records(net, tmp) : nullptr;
        if (code_match || code_learned) {
            uint32_t* m = dev_alloc<uint32_t>(tmp, e, false);
            uint32_t* l = dev_alloc<uint32_t>(tmp, e, false);
            if (net->code2)
                hipLaunchKernelGGL(k_gather_codes<true>, g, b, 0, net->stream, net->slot_of, c2, net->S.codeN, e, m, l);
            else
                hipLaunchKernelGGL(k_gather_codes<false>, g, b, 0, net->stream, net->slot_of, net->S.code2, net->S.codeN, e, m, l);
            std::vector<uint32_t> hm(e), hl(e);
            check(hipMemcpyAsync(hm.data(), m, e * 4, hipMemcpyDeviceToHost, net->stream), "read codes");
            check(hipMemcpyAsync(hl.data(), l, e * 4, hipMemcpyDeviceToHost, net->stream), "read codes");
            check(hipStreamSynchronize(net->stream), "sync");
            for (uint64_t j = 0; j < e; ++j) {
                if (code_match) code_match[j] = hm[j];
                if (code_learned) code_learned[j] = hl[j];
            }
        }
        if (pre_count) {
            uint32_t* d = dev_alloc<uint32_t>(tmp, e, false);
            hipLaunchKernelGGL(k_gather_u32, g, b, 0, net->stream, net->slot_of, net->S.pre_count, e, d);
            std::vector<uint32_t> h(e);
            check(hipMemcpyAsync(h.data(), d, e * 4, hipMemcpyDeviceToHost, net->stream), "read pre counts");
            check(hipStreamSynchronize(net->stream), "sync");
            for (uint64_t j = 0; j < e; ++j) pre_count[j] = h[j];
        }
        if (last_pre || last_post) {
            int64_t* dpre = dev_alloc<int64_t>(tmp, e, false);
            int64_t* dpost = dev_alloc<int64_t>(tmp, e, false);
            hipLaunchKernelGGL(k_gather_ticks, g, b, 0, net->stream, net->slot_of, net->S.tpos, net->S.stdp, c2,
                               net->S.codeN, net->S.delay, net->S.tgt, net->S.last_fired, e, dpre, dpost);
            if (last_pre) check(hipMemcpyAsync(last_pre, dpre, e * 8, hipMemcpyDeviceToHost, net->stream), "read ticks");
            if (last_post) check(hipMemcpyAsync(last_post, dpost, e * 8, hipMemcpyDeviceToHost, net->stream), "read ticks");
        }
        if (update_count) {
            std::fill(update_count, update_count + e, 0);
            if (net->n_tpos) {
                const uint64_t nt = net->n_tpos;
                std::vector<uint32_t> upd(nt), ltp(nt);
                std::vector<uint32_t> conn(nt);
                read_stdp_counts(net, upd.data(), ltp.data());
                check(hipMemcpy(conn.data(), net->tpos_conn_dev, nt * 4, hipMemcpyDeviceToHost), "read stdp");
                for (uint64_t t = 0; t < nt; ++t) update_count[conn[t]] = upd[t];
            }
        }
        check(hipStreamSynchronize(net->stream), "read synapses");
        return SNN_OK;
    });
}

int snn_net_read_plasticity(snn_net* net, uint64_t* ltp_updates, uint64_t* ltd_updates) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    return guarded([&]() -> int {
        check(hipSetDevice(net->device), "hipSetDevice");
        const uint64_t e = net->e;
        if (ltp_updates) std::fill(ltp_updates, ltp_updates + e, 0);
        if (ltd_updates) std::fill(ltd_updates, ltd_updates + e, 0);
        if (net->n_tpos == 0 || !(ltp_updates || ltd_updates)) return SNN_OK;
        const uint64_t nt = net->n_tpos;
        std::vector<uint32_t> upd(nt), ltp(nt), conn(nt);
        check(hipStreamSynchronize(net->stream), "sync");
        read_stdp_counts(net, upd.data(), ltp.data());
        check(hipMemcpy(conn.data(), net->tpos_conn_dev, nt * 4, hipMemcpyDeviceToHost), "read stdp");
        for (uint64_t t = 0; t < nt; ++t) {
            if (ltp_updates) ltp_updates[conn[t]] = ltp[t];
            if (ltd_updates) ltd_updates[conn[t]] = upd[t] - ltp[t];
        }
        return SNN_OK;
    });
}

int snn_net_set_profiling(snn_net* net, int enabled) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    net->profiling = enabled != 0;
    net->prof_mask = (enabled & kProfileSelect) ? (uint32_t)(enabled & 0xFFFF) : ~0u;
    net->prof_every = (enabled & kProfileSelect) ? std::max(1u, ((uint32_t)enabled >> 16) & 0xFFu) : 1u;
    for (int k = 0; k < KCount_; ++k) {
        net->kms[k] = 0;
        net->klaunch[k] = 0;
    }
    return SNN_OK;
}

int snn_net_kernel_times(snn_net* net, double* ms_total, uint64_t* launches, int cap, char* names, uint64_t names_cap) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    std::string all;
    const int m = std::min<int>(cap, KCount_);
    for (int k = 0; k < m; ++k) {
        if (ms_total) ms_total[k] = net->kms[k];
        if (launches) launches[k] = net->klaunch[k];
    }
    for (int k = 0; k < KCount_; ++k) {
        if (k) all += ",";
        all += kKernelNames[k];
    }
    if (names && names_cap) {
        const size_t c = std::min<size_t>(names_cap - 1, all.size());
        std::memcpy(names, all.data(), c);
        names[c] = 0;
    }
    return m;
}

}  // extern "C"
