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
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
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
constexpr int kPBlock = 1024;  // partition count / scatter workgroups: 16 waves per CU

// ------------------------------------------------------------------------------------ scans

struct PlainSrc {
    const uint32_t* p;
    const uint32_t* n_dev;
    uint32_t n_static;
    __device__ uint32_t count() const { return n_dev ? *n_dev : n_static; }
    __device__ uint32_t load(uint32_t i) const { return p[i]; }
};

struct PopcSrc {  // popcount of the fire bitmap words covering the tick's events
    const uint32_t* bits;
    const uint32_t* events;
    __device__ uint32_t count() const { return (*events + 31u) >> 5; }
    __device__ uint32_t load(uint32_t i) const { return __popc(bits[i]); }
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

// exclusive scan of the block partials; writes the grand total to *total and, when cap is
// exceeded, raises err_flag and zeroes *zero_on_overflow (so the tick does no further work)
__global__ void __launch_bounds__(kBlock) k_scan_mid(uint32_t* partials, uint32_t nparts, uint32_t* total,
                                                     uint64_t cap, uint32_t err_flag, uint32_t* err,
                                                     uint32_t* zero_on_overflow) {
    uint32_t base = 0;
    for (uint32_t off = 0; off < nparts; off += blockDim.x) {
        const uint32_t i = off + threadIdx.x;
        const uint32_t v = i < nparts ? partials[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, &tot);
        if (i < nparts) partials[i] = base + ex;
        base += tot;
    }
    if (threadIdx.x == 0) {
        if (total) *total = base;
        if (cap && base > cap) {
            atomicOr(err, err_flag);
            if (zero_on_overflow) *zero_on_overflow = 0;
            if (total) *total = 0;
        }
    }
}

template <class Src>
__global__ void __launch_bounds__(kBlock) k_scan_down(Src src, const uint32_t* partials, uint32_t* out) {
    const uint32_t n = src.count();
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = min(n, blockIdx.x * chunk);
    const uint32_t hi = min(n, lo + chunk);
    uint32_t base = partials[blockIdx.x];
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

__global__ void __launch_bounds__(1024) k_plan(Step S) {
    const uint32_t T = S.ctl[0];
    const uint32_t ng = 2 * S.lmax + 1;
    __shared__ uint32_t wsum[1024 / kWave];
    uint32_t base = 0;
    for (uint32_t off = 0; off < ng; off += blockDim.x) {
        const uint32_t gi = off + threadIdx.x;
        Group g;
        g.nseg = 0;
        if (gi < ng) group_info(S, T, gi, g);
        // block exclusive scan of nseg (1024 threads)
        const int lane = threadIdx.x & (kWave - 1);
        const int wid = threadIdx.x / kWave;
        const uint32_t incl = wave_incl_scan(g.nseg);
        if (lane == kWave - 1) wsum[wid] = incl;
        __syncthreads();
        uint32_t wb = 0, tot = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) {
            if (w < wid) wb += wsum[w];
            tot += wsum[w];
        }
        __syncthreads();
        if (gi < ng) {
            g.first = base + wb + incl - g.nseg;
            S.groups[gi] = g;
        }
        base += tot;
    }
    if (threadIdx.x == 0) {
        if (base > S.seg_cap) {
            atomicOr(&S.ctl[1], (uint32_t)kErrSegments);
            base = 0;
        }
        if (S.ctl[1]) base = 0;  // after any error the run is invalid: stop doing work
        S.seg_count[0] = base;
        S.seg_count[1] = 0;
    }
}

__global__ void __launch_bounds__(kBlock) k_segments(Step S) {
    const uint32_t nseg = S.seg_count[0];
    const uint32_t ng = 2 * S.lmax + 1;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nseg; q += gridDim.x * blockDim.x) {
        // last group whose first segment is <= q (groups are in order, empty ones share `first`)
        uint32_t lo = 0, hi = ng;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (S.groups[mid].first <= q) lo = mid + 1; else hi = mid;
        }
        const Group g = S.groups[lo - 1];
        uint64_t v0;
        uint32_t cnt, mult;
        segment_info(S, g, q - g.first, v0, cnt, mult);
        S.seg_v0[q] = v0;
        S.seg_n[q] = cnt;
        S.seg_mult[q] = mult;
    }
}

// Event ranges: workgroup w owns events [D*w/nwg, D*(w+1)/nwg) (a contiguous range of g), caches the
// descriptors of the segments covering it in LDS and lets its lanes walk consecutive events, so the
// CSR reads are coalesced and the load is balanced whatever the fan-out distribution.
constexpr uint32_t kSegChunk = 2048;

struct SegCache {
    uint32_t off[kSegChunk + 1];
    uint64_t v0[kSegChunk];
    uint32_t mult[kSegChunk];
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

// Calls f(v0, g0, mult, i) for every event of this workgroup's range. Every thread of the block
// must call it (it synchronises).
template <bool WITH_MULT, typename F>
__device__ void for_block_events(const Step& S, SegCache& c, F&& f) {
    const uint32_t nseg = S.seg_count[0];
    const uint64_t D = S.seg_count[1];
    const uint32_t g0 = (uint32_t)(D * blockIdx.x / gridDim.x);
    const uint32_t g1 = (uint32_t)(D * (blockIdx.x + 1) / gridDim.x);
    __shared__ uint32_t qstart;
    if (threadIdx.x == 0 && g0 < g1) qstart = seg_find(S.seg_off, nseg, g0);
    __syncthreads();
    if (g0 >= g1) return;
    uint32_t qa = qstart;
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
        for (uint32_t g = ga + threadIdx.x; g < gb; g += blockDim.x) {
            const uint32_t k = seg_find(c.off, nq, g);
            f(c.v0[k], c.off[k], WITH_MULT ? c.mult[k] : 0u, g - c.off[k]);
        }
        __syncthreads();
        if (gb >= g1 || qb >= nseg) break;
        qa = qb;
    }
}

// Per-workgroup histogram of the tick's events over target partitions (LDS atomics only).
__global__ void __launch_bounds__(kPBlock) k_pcount(Step S) {
    extern __shared__ uint32_t h[];
    __shared__ SegCache cache;
    for (uint32_t p = threadIdx.x; p < S.nparts; p += blockDim.x) h[p] = 0;
    __syncthreads();
    const uint32_t shift = S.pshift;
    for_block_events<false>(S, cache, [&](uint64_t v0, uint32_t, uint32_t, uint32_t i) {
        atomicAdd(&h[event_target(S, v0, i) >> shift], 1u);
    });
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < S.nparts; p += blockDim.x) S.hist[(uint64_t)p * S.nwg + blockIdx.x] = h[p];
}

// Place each event in its partition's bucket at (scanned workgroup base + LDS rank). The
// workgroup -> event assignment is the one k_pcount used.
template <bool WIDE, bool CODE_HERE, bool CODE2>
__global__ void __launch_bounds__(kPBlock) k_pscatter(Step S) {
    extern __shared__ uint32_t base[];
    __shared__ SegCache cache;
    for (uint32_t p = threadIdx.x; p < S.nparts; p += blockDim.x) base[p] = S.hist[(uint64_t)p * S.nwg + blockIdx.x];
    __syncthreads();
    const uint32_t T = S.ctl[0];
    const uint32_t shift = S.pshift;
    const uint32_t mask = (1u << shift) - 1;
    typedef typename std::conditional<WIDE, Entry16, Entry8>::type E;
    uint32_t err = 0;
    for_block_events<CODE_HERE>(S, cache, [&](uint64_t v0, uint32_t g0, uint32_t mult, uint32_t i) {
        uint32_t x;
        E e;
        err |= make_event<WIDE, CODE_HERE, CODE2>(S, v0, g0, mult, i, T, x, e);
        const uint32_t pos = atomicAdd(&base[x >> shift], 1u);
        ((E*)S.prec)[pos] = e;
        S.ptl[pos] = (uint16_t)(x & mask);
    });
    if (err) atomicOr(&S.ctl[1], err);
}

constexpr uint32_t kStageBytes = 98304;
constexpr uint32_t kRound = kPBlock;  // targets folded per round (one per thread)

// One workgroup per target partition. The partition's targets are folded in rounds of 1024 (one
// per thread): a counting sort by target over the whole partition gives every target's place, and
// each round stages only its own targets' records in LDS (global scratch if they do not fit), so
// the bucket is read once for its records and once per round for the 2-byte target indices.
template <bool WIDE, bool STDP, bool CODE_HERE, bool CODE2>
__global__ void __launch_bounds__(kPBlock) k_pfold(Step S) {
    typedef typename std::conditional<WIDE, Entry16, Entry8>::type E;
    constexpr uint32_t kCap = kStageBytes / sizeof(E);
    __shared__ E lds_stage[kCap];
    __shared__ uint32_t fill[kRound];
    extern __shared__ uint32_t sh[];
    const uint32_t PS = 1u << S.pshift;
    uint32_t* cnt = sh;
    uint32_t* off = sh + PS;  // PS + 1 entries
    const uint32_t p = blockIdx.x;
    const uint32_t T = S.ctl[0];
    const uint32_t start = S.hist[(uint64_t)p * S.nwg];
    const uint32_t end = p + 1 < S.nparts ? S.hist[(uint64_t)(p + 1) * S.nwg] : S.seg_count[1];
    const uint32_t c = end - start;
    const uint16_t* tl = S.ptl + start;
    const E* src = (const E*)S.prec + start;
    for (uint32_t t = threadIdx.x; t < PS; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) atomicAdd(&cnt[tl[i]], 1u);
    __syncthreads();
    uint32_t run = 0;
    for (uint32_t t0 = 0; t0 < PS; t0 += blockDim.x) {
        const uint32_t v = cnt[t0 + threadIdx.x];
        uint32_t tot;
        const uint32_t ex = block_excl_scan_n<kPBlock>(v, &tot);
        off[t0 + threadIdx.x] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0) off[PS] = run;
    __syncthreads();
    uint32_t err = 0;
    const uint32_t rounds = PS / kRound;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t lo = off[r * kRound], hi = off[(r + 1) * kRound];
        const uint32_t t = r * kRound + threadIdx.x;
        fill[threadIdx.x] = off[t] - lo;
        __syncthreads();
        E* st = hi - lo <= kCap ? lds_stage : (E*)S.stage + start + lo;
        for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) {
            const uint32_t ti = tl[i];
            if (ti / kRound != r) continue;
            const uint32_t pos = atomicAdd(&fill[ti % kRound], 1u);
            st[pos] = src[i];
        }
        __syncthreads();
        const uint32_t x = (p << S.pshift) + t;
        if (x < S.n && cnt[t]) err |= fold_events<WIDE, STDP, CODE_HERE, CODE2>(S, st + (off[t] - lo), cnt[t], x, T);
        __syncthreads();
    }
    if (err) atomicOr(&S.ctl[1], err);
}

// rank-ordered fire list of tick T into the fire ring (and the raster)
__global__ void __launch_bounds__(kBlock) k_fire_write(Step S) {
    const uint32_t T = S.ctl[0];
    const uint32_t F = S.ctl[2];
    const uint32_t nwords = (S.seg_count[1] + 31u) >> 5;
    const uint64_t head = S.totals[2];
    const uint32_t R = S.lmax + 1;
    // entries of ticks older than T+1-lmax may be overwritten
    const uint64_t oldest = (S.lmax == 1) ? head : (T + 1 >= S.lmax ? S.tick_fstart[(T + 1 - S.lmax) % R] : 0);
    const bool fits = head + F - oldest <= S.fire_cap;
    const uint64_t rfill = S.totals[3];
    // the raster is a recording aid: when it is full, recording stops and the run goes on
    const bool rfits = S.raster && !(S.ctl[3] & 1u) && rfill + F <= S.raster_cap;
    for (uint32_t wi = blockIdx.x * blockDim.x + threadIdx.x; wi < nwords; wi += gridDim.x * blockDim.x) {
        uint32_t bits = S.fire_bits[wi];
        if (!bits) continue;
        S.fire_bits[wi] = 0;
        uint32_t pos = S.fire_word_off[wi];
        while (bits) {
            const uint32_t b = __ffs(bits) - 1;
            bits &= bits - 1;
            const uint32_t g = wi * 32 + b;
            if (fits) {
                const uint64_t idx = (head + pos) % S.fire_cap;
                S.firelog[idx] = S.fire_neuron[g];
                S.firemult[idx] = S.fire_mult[g];
            }
            if (rfits) S.raster[rfill + pos] = S.fire_neuron[g];
            ++pos;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (!fits) atomicOr(&S.ctl[1], (uint32_t)kErrFireLog);
        if (S.raster && !rfits) atomicOr(&S.ctl[3], 1u);
        S.tick_fstart[T % R] = head;
        S.tick_fcount[T % R] = fits ? F : 0;
    }
}

// post-synaptic scan: one wave per (first) fire of tick T
__global__ void __launch_bounds__(kBlock) k_post(Step S) {
    const uint32_t T = S.ctl[0];
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
        for (uint64_t t = lo + lane; t < hi; t += kWave) post_scan_record(S, t, T, m);
    }
}

__global__ void k_advance(Step S, uint32_t t0) {
    const uint32_t T = S.ctl[0];
    const uint32_t F = S.ctl[2];
    const uint32_t D = S.seg_count[1];
    S.tick_delivered[T - t0] = D;
    S.tick_fired[T - t0] = F;
    S.totals[0] += D;
    S.totals[1] += F;
    S.totals[2] += S.tick_fcount[T % (S.lmax + 1)];  // fires actually written to the ring
    if (S.raster && !(S.ctl[3] & 1u)) S.totals[3] += F;
    S.ctl[0] = T + 1;
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

__global__ void k_gather_u32(const uint32_t* slot_of, const uint32_t* src, uint64_t e, uint32_t* out) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < e; j += (uint64_t)gridDim.x * blockDim.x)
        out[j] = src[slot_of[j]];
}

// ------------------------------------------------------------------------------------ synthetic E/I networks
// Device generator for the seeded E/I network family of snncuda_amd/workloads.py (ei_network with
// one code per synapse): every draw is the same counter hash, so a generated network is identical
// to the host-built one. Connectome order is source-major (edge j of source s has index
// row[s] + j); slots are sorted by (source, delay, j) exactly as build_host_net sorts them.

struct GenParams {
    uint64_t n, nall, n_exc;
    float w_exc, w_inh;
    uint32_t delay_max, n_comps;
    uint32_t comps[4];
    bool plastic;
    uint64_t stream_key[8];
    const uint64_t* row;
    uint32_t* grp;
    uint32_t* tgt;
    float* w;
    uint8_t* meta;
    uint16_t* delay;
    uint32_t* slot_of;
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

__device__ __forceinline__ uint32_t gen_delay(const GenParams& P, uint64_t s, uint64_t ctr) {
    return 1 + gbelow(ghash(P.stream_key[s < P.n ? 5 : 7], ctr), P.delay_max);
}

__device__ __forceinline__ uint32_t gen_target(const GenParams& P, uint64_t s, uint64_t ctr) {
    if (s < P.n) {
        uint32_t t = gbelow(ghash(P.stream_key[1], ctr), P.n - 1);
        return t + (t >= s ? 1u : 0u);
    }
    return gbelow(ghash(P.stream_key[2], ctr), P.n);
}

__device__ GenEdge gen_edge(const GenParams& P, uint64_t s, uint64_t j) {
    const uint64_t ctr = (s << 24) + j;
    GenEdge g;
    g.tgt = gen_target(P, s, ctr);
    g.delay = gen_delay(P, s, ctr);
    uint32_t comp, rec;
    if (s < P.n) {
        const bool exc = s < P.n_exc;
        const float u = gunif24(ghash(P.stream_key[3], ctr));
        g.w = exc ? u * P.w_exc : u * P.w_inh;
        rec = exc ? kAmpa : kGabaA;
        comp = P.comps[gbelow(ghash(P.stream_key[4], ctr), P.n_comps)];
    } else {
        g.w = gunif24(ghash(P.stream_key[6], ctr)) * 0.5f;
        rec = kAmpa;
        comp = kBasal;
    }
    g.meta = (uint8_t)(rec | (comp << 2) | ((P.plastic ? 1u : 0u) << 4));
    return g;
}

// pass 1: per-source delay histogram -> group table; in-degree counts of plastic synapses
__global__ void __launch_bounds__(kBlock) k_gen_pass1(GenParams P) {
    __shared__ uint32_t hist[kBlock / kWave][64];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wv = threadIdx.x / kWave;
    const uint64_t s = (uint64_t)blockIdx.x * (kBlock / kWave) + wv;
    const bool valid = s < P.nall;
    hist[wv][lane] = 0;
    __syncthreads();
    const uint64_t r0 = valid ? P.row[s] : 0, len = valid ? P.row[s + 1] - r0 : 0;
    for (uint64_t j = lane; j < len; j += kWave) {
        const uint64_t ctr = (s << 24) + j;
        atomicAdd(&hist[wv][gen_delay(P, s, ctr)], 1u);
        if (P.plastic) atomicAdd(&P.incnt[gen_target(P, s, ctr)], 1u);
    }
    __syncthreads();
    if (valid) {
        // lane L-1 holds the count of lag L; exclusive scan gives the group starts
        const uint32_t c = (lane + 1 <= P.delay_max) ? hist[wv][lane + 1] : 0u;
        const uint32_t incl = wave_incl_scan(c);
        const uint32_t W = P.delay_max + 1;
        if (lane < P.delay_max) P.grp[s * W + lane] = incl - c;
        if (lane == 0) P.grp[s * W + P.delay_max] = (uint32_t)len;
    }
}

// pass 2: stable placement of each source's synapses by delay, plus target-major STDP records
__global__ void __launch_bounds__(kBlock) k_gen_pass2(GenParams P) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t s = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (s >= P.nall) return;  // whole wave; no block barriers below
    const uint32_t W = P.delay_max + 1;
    const uint64_t r0 = P.row[s], len = P.row[s + 1] - r0;
    uint32_t running = lane < P.delay_max ? P.grp[s * W + lane] : 0u;  // next free position of lag lane+1
    const uint64_t lt = (1ull << lane) - 1;
    for (uint64_t c = 0; c < len; c += kWave) {
        const uint64_t j = c + lane;
        const bool ok = j < len;
        GenEdge g{};
        if (ok) g = gen_edge(P, s, j);
        uint32_t pos = 0;
        for (uint32_t L = 1; L <= P.delay_max; ++L) {
            const uint64_t m = __ballot(ok && g.delay == L);
            if (!m) continue;
            const uint32_t base = __shfl(running, (int)(L - 1), kWave);
            if (ok && g.delay == L) pos = base + (uint32_t)__popcll(m & lt);
            if (lane == L - 1) running += (uint32_t)__popcll(m);
        }
        if (ok) {
            const uint64_t slot = r0 + pos;
            P.tgt[slot] = g.tgt;
            P.w[slot] = g.w;
            P.meta[slot] = g.meta;
            P.delay[slot] = (uint16_t)g.delay;
            P.slot_of[r0 + j] = (uint32_t)slot;
            if (P.plastic) {
                const uint64_t tp = P.in_ptr[g.tgt] + atomicAdd(&P.incnt[g.tgt], 1u);
                P.tpos[slot] = (uint32_t)tp;
                StdpRec r;
                r.weight = g.w;
                r.last_pre = kNone;
                r.last_post = kNone;
                r.aux = 0;  // wmax table entry 0 = 2.0
                P.stdp[tp] = r;
                P.tpos_conn[tp] = (uint32_t)(r0 + j);
            }
        }
    }
}

__global__ void k_widen(const uint32_t* in, uint64_t n, uint64_t* out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

uint64_t stream_key(uint64_t seed, uint64_t stream) { return mix64(seed * kGold + stream); }

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

template <typename T>
T* dev_upload(std::vector<DevBuf>& owned, const std::vector<T>& v) {
    T* p = dev_alloc<T>(owned, v.size(), false);
    if (!v.empty()) check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return p;
}

struct PendingInject {
    uint32_t neuron;
    uint64_t tick;
    float weight;
    uint64_t stage;
    uint64_t seq;
};

enum KernelId { KPlan, KSegments, KScanSeg, KCount, KScanTgt, KScatter, KFold, KScanFire, KFireWrite, KPost, KAdvance, KCount_ };
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
    // profiling
    bool profiling = false;
    double kms[KCount_] = {0};
    uint64_t klaunch[KCount_] = {0};

    ~snn_net() {
        if (device >= 0) (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& b : run_owned) (void)hipFree(b.p);
        for (auto& b : owned) (void)hipFree(b.p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <class Src>
void launch_scan(snn_net* net, Src src, uint32_t* out, uint32_t* total, uint64_t cap, uint32_t err_flag,
                 uint32_t* zero_on_overflow) {
    hipLaunchKernelGGL(k_scan_reduce<Src>, dim3(kScanBlocks), dim3(kBlock), 0, net->stream, src, net->partials);
    hipLaunchKernelGGL(k_scan_mid, dim3(1), dim3(kBlock), 0, net->stream, net->partials, (uint32_t)kScanBlocks, total, cap,
                       err_flag, &net->S.ctl[1], zero_on_overflow);
    hipLaunchKernelGGL(k_scan_down<Src>, dim3(kScanBlocks), dim3(kBlock), 0, net->stream, src, net->partials, out);
}

template <bool WIDE, bool CODE_HERE, bool CODE2>
void launch_scatter(snn_net* net) {
    hipLaunchKernelGGL((k_pscatter<WIDE, CODE_HERE, CODE2>), dim3(net->S.nwg), dim3(kPBlock), net->S.nparts * 4,
                       net->stream, net->S);
}

template <bool WIDE, bool STDP, bool CODE_HERE, bool CODE2>
void launch_fold(snn_net* net) {
    hipLaunchKernelGGL((k_pfold<WIDE, STDP, CODE_HERE, CODE2>), dim3(net->S.nparts), dim3(kPBlock),
                       (2u * (1u << net->S.pshift) + 1) * 4, net->stream, net->S);
}

void dispatch_scatter(snn_net* net) {
    // (wide, code_here, code2); code_here needs one code per synapse
    if (!net->wide) {
        if (net->code2) launch_scatter<false, true, true>(net); else launch_scatter<false, true, false>(net);
    } else if (net->code_here) {
        if (net->code2) launch_scatter<true, true, true>(net); else launch_scatter<true, true, false>(net);
    } else {
        if (net->code2) launch_scatter<true, false, true>(net); else launch_scatter<true, false, false>(net);
    }
}

void dispatch_fold(snn_net* net) {
    if (!net->wide) {
        if (net->code2) launch_fold<false, false, true, true>(net); else launch_fold<false, false, true, false>(net);
        return;
    }
#define SNN_FOLD(ST, CH)                                                         \
    if (net->stdp == ST && net->code_here == CH) {                             \
        if (net->code2) launch_fold<true, ST, CH, true>(net);                  \
        else launch_fold<true, ST, CH, false>(net);                            \
        return;                                                                \
    }
    SNN_FOLD(true, true)
    SNN_FOLD(true, false)
    SNN_FOLD(false, true)
    SNN_FOLD(false, false)
#undef SNN_FOLD
}

struct Timer {
    snn_net* net;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    int cur = -1;
    hipEvent_t a = nullptr;
    void begin(int id) {
        if (!net->profiling) return;
        cur = id;
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventRecord(a, net->stream), "hipEventRecord");
    }
    void end() {
        if (!net->profiling) return;
        hipEvent_t b;
        check(hipEventCreate(&b), "hipEventCreate");
        check(hipEventRecord(b, net->stream), "hipEventRecord");
        marks.push_back({cur, {a, b}});
    }
    void collect() {
        for (auto& m : marks) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, m.second.first, m.second.second), "hipEventElapsedTime");
            net->kms[m.first] += ms;
            net->klaunch[m.first] += 1;
            (void)hipEventDestroy(m.second.first);
            (void)hipEventDestroy(m.second.second);
        }
        marks.clear();
    }
};

void launch_tick(snn_net* net, uint32_t t0, Timer& tm) {
    Step& S = net->S;
    hipStream_t st = net->stream;
    tm.begin(KPlan);
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, st, S);
    tm.end();
    tm.begin(KSegments);
    hipLaunchKernelGGL(k_segments, dim3(net->grid), dim3(kBlock), 0, st, S);
    tm.end();
    tm.begin(KScanSeg);
    launch_scan(net, PlainSrc{S.seg_n, &S.seg_count[0], 0}, S.seg_off, &S.seg_count[1],
                std::min<uint64_t>(S.ev_cap, kMaxEventsPerTick), kErrEvents, &S.seg_count[0]);
    tm.end();
    tm.begin(KCount);
    hipLaunchKernelGGL(k_pcount, dim3(S.nwg), dim3(kPBlock), S.nparts * 4, st, S);
    tm.end();
    tm.begin(KScanTgt);
    launch_scan(net, PlainSrc{S.hist, nullptr, S.nparts * S.nwg}, S.hist, nullptr, 0, 0, nullptr);
    tm.end();
    tm.begin(KScatter);
    dispatch_scatter(net);
    tm.end();
    tm.begin(KFold);
    dispatch_fold(net);
    tm.end();
    tm.begin(KScanFire);
    launch_scan(net, PopcSrc{S.fire_bits, &S.seg_count[1]}, S.fire_word_off, &S.ctl[2], 0, 0, nullptr);
    tm.end();
    tm.begin(KFireWrite);
    hipLaunchKernelGGL(k_fire_write, dim3(net->grid), dim3(kBlock), 0, st, S);
    tm.end();
    if (net->stdp) {
        tm.begin(KPost);
        hipLaunchKernelGGL(k_post, dim3(net->grid), dim3(kBlock), 0, st, S);
        tm.end();
    }
    tm.begin(KAdvance);
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, st, S, t0);
    tm.end();
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
    if (code_compact) S.code2 = dev_alloc<CodeRec2>(own, e);
    else S.codeN = dev_alloc<CodeRecN>(own, e);
    S.pre_count = o.track_pre_count ? dev_alloc<uint32_t>(own, e) : nullptr;
    S.dyn = dev_alloc<NeuronDyn>(own, n);
    S.th = dev_alloc<TemporalHdr>(own, n);
    S.cand = dev_alloc<Runs>(own, n);
    S.refs = dev_alloc<TemporalRefs>(own, n);
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
    S.seg_n = dev_alloc<uint32_t>(own, seg_cap, false);
    S.seg_off = dev_alloc<uint32_t>(own, seg_cap, false);
    S.seg_mult = dev_alloc<uint32_t>(own, seg_cap, false);
    // Target partitions of 2048 neurons: few enough buckets that each scatter workgroup's runs per
    // bucket stay long (writes coalesce in L2); folded in rounds of 1024 neurons. Networks of at
    // most 1024 neurons use a single 1024-neuron partition.
    uint32_t pshift = n <= 1024 ? 10 : 11;
    if (const char* e = std::getenv("SNNCUDA_PART_SHIFT")) pshift = std::max(10, std::min(12, std::atoi(e)));
    if (((uint64_t)n + (1ull << pshift) - 1) >> pshift > 65535)
        throw std::invalid_argument("too many neurons for 4096-neuron partitions on one device");
    S.pshift = pshift;
    S.nparts = (uint32_t)(((uint64_t)n + (1ull << pshift) - 1) >> pshift);
    int ncu = 256;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, o.device) == hipSuccess) ncu = pr.multiProcessorCount;
    S.nwg = (uint32_t)std::max(1, ncu);  // one scatter workgroup per CU (measured best with 1024-neuron partitions)
    if (const char* e = std::getenv("SNNCUDA_SCATTER_WGS")) S.nwg = (uint32_t)std::max(1, std::atoi(e));
    S.hist = dev_alloc<uint32_t>(own, (uint64_t)S.nparts * S.nwg);
    S.prec = net->wide ? (void*)dev_alloc<Entry16>(own, ev_cap, false) : (void*)dev_alloc<Entry8>(own, ev_cap, false);
    S.ptl = dev_alloc<uint16_t>(own, ev_cap, false);
    S.stage = net->wide ? (void*)dev_alloc<Entry16>(own, ev_cap, false) : (void*)dev_alloc<Entry8>(own, ev_cap, false);
    S.fire_bits = dev_alloc<uint32_t>(own, ev_cap / 32);
    S.fire_neuron = dev_alloc<uint32_t>(own, ev_cap, false);
    S.fire_mult = dev_alloc<uint32_t>(own, ev_cap, false);
    S.fire_word_off = dev_alloc<uint32_t>(own, ev_cap / 32, false);
    S.ctl = dev_alloc<uint32_t>(own, 4);
    S.totals = dev_alloc<uint64_t>(own, 4);
    S.raster_cap = o.raster_capacity ? o.raster_capacity : (1ull << 22);
    S.raster = dev_alloc<uint32_t>(own, S.raster_cap, false);
    if (!net->partials) net->partials = dev_alloc<uint32_t>(own, kScanBlocks);
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, o.device) == hipSuccess) cus = prop.multiProcessorCount;
    net->grid = (uint32_t)std::max(64, cus * 4);
    net->fold_grid = (uint32_t)std::min<uint64_t>((n + kBlock - 1) / kBlock, 65535ull * 4);
    check(hipDeviceSynchronize(), "create");
}

}  // namespace

// ------------------------------------------------------------------------------------ C ABI

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
        NetDesc d;
        d.n_neurons = desc->n_neurons;
        d.threshold = desc->threshold;
        d.window = desc->window;
        d.sim_thr = desc->sim_thr;
        d.max_refs = desc->max_refs;
        d.n_synapses = desc->n_synapses;
        d.source = desc->source;
        d.target = desc->target;
        d.weight = desc->weight;
        d.max_weight = desc->max_weight;
        d.delay = desc->delay;
        d.compartment = desc->compartment;
        d.receptor = desc->receptor;
        d.plastic = desc->plastic;
        d.code_ptr = desc->code_ptr;
        d.codes = desc->codes;
        if (d.n_neurons && (!d.threshold || !d.window || !d.sim_thr || !d.max_refs))
            return fail(SNN_EINVAL, "null neuron array");
        if (d.n_synapses && (!d.source || !d.target || !d.weight || !d.max_weight || !d.delay || !d.compartment ||
                             !d.receptor || !d.plastic))
            return fail(SNN_EINVAL, "null synapse array");
        HostNet h = build_host_net(d);

        std::unique_ptr<snn_net> net(new snn_net());
        open_net(net.get(), o, h.n, h.e, h.lmax, h.any_plastic, h.ve_identity, h.code_compact);
        auto& own = net->owned;
        Step& S = net->S;
        S.row = dev_upload(own, h.row);
        S.grp = dev_upload(own, h.grp);
        S.ve = h.ve_identity ? nullptr : dev_upload(own, h.ve);
        S.tgt = dev_upload(own, h.tgt);
        S.w = dev_upload(own, h.w);
        S.meta = dev_upload(own, h.meta);
        S.delay = dev_upload(own, h.delay);
        S.expected = h.expected.empty() ? nullptr : dev_upload(own, h.expected);
        S.expected_all = h.expected_all ? h.expected_all : 2;
        if (h.any_plastic) {
            S.tpos = dev_upload(own, h.tpos);
            S.in_ptr = dev_upload(own, h.in_ptr);
            S.stdp = dev_upload(own, h.stdp);
            S.upd_count = dev_alloc<uint32_t>(own, h.stdp.size());
            S.ltp_count = dev_alloc<uint32_t>(own, h.stdp.size());
            S.wmax = dev_upload(own, h.wmax);
            net->tpos_conn_dev = dev_upload(own, h.tpos_conn);
            net->n_tpos = h.stdp.size();
        }
        S.np = dev_upload(own, h.params);
        net->slot_of = dev_upload(own, h.slot_of);
        finish_net(net.get(), o, h.nve, h.params);
        *out = net.release();
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
        if (g->n < 2 || !g->fanout) return fail(SNN_EINVAL, "need at least two network neurons and a fan-out array");
        if (g->delay_max < 1 || g->delay_max > 63) return fail(SNN_EINVAL, "delay_max must be in [1, 63]");
        if (g->n_comps < 1 || g->n_comps > 4) return fail(SNN_EINVAL, "n_comps must be in [1, 4]");
        if (g->window > 65535 || g->max_refs < 1) return fail(SNN_EINVAL, "bad temporal parameters");
        const uint64_t nall = g->n + g->n_sources;
        if (nall >= 0xFFFFFFFFull) return fail(SNN_EINVAL, "too many neurons");
        std::vector<uint64_t> row(nall + 1, 0);
        for (uint64_t i = 0; i < g->n; ++i) {
            if (g->fanout[i] >= (1u << 24)) return fail(SNN_EINVAL, "fan-out must be below 2^24");
            row[i + 1] = row[i] + g->fanout[i];
        }
        if (g->source_fanout >= (1u << 24)) return fail(SNN_EINVAL, "fan-out must be below 2^24");
        for (uint64_t i = g->n; i < nall; ++i) row[i + 1] = row[i] + g->source_fanout;
        const uint64_t e = row[nall];
        if (e >= 0xFFFFFFFFull) return fail(SNN_EINVAL, "more than 2^32-2 synapses on one device");
        const uint32_t lmax = g->delay_max;

        std::unique_ptr<snn_net> net(new snn_net());
        const bool plastic = g->plastic != 0;
        const bool compact = g->max_refs == 1;
        open_net(net.get(), o, (uint32_t)nall, e, lmax, plastic, true, compact);
        auto& own = net->owned;
        Step& S = net->S;
        hipStream_t st = net->stream;
        S.row = dev_upload(own, row);
        uint32_t* grp = dev_alloc<uint32_t>(own, nall * (lmax + 1));
        S.grp = grp;
        S.ve = nullptr;
        uint32_t* tgt = dev_alloc<uint32_t>(own, e, false);
        float* w = dev_alloc<float>(own, e, false);
        uint8_t* meta = dev_alloc<uint8_t>(own, e, false);
        uint16_t* delay = dev_alloc<uint16_t>(own, e, false);
        S.tgt = tgt;
        S.w = w;
        S.meta = meta;
        S.delay = delay;
        S.expected = nullptr;
        S.expected_all = 2;
        net->slot_of = dev_alloc<uint32_t>(own, e, false);

        GenParams gp{};
        gp.n = g->n;
        gp.nall = nall;
        gp.n_exc = g->n_exc;
        gp.w_exc = g->w_exc;
        gp.w_inh = g->w_inh;
        gp.delay_max = g->delay_max;
        gp.n_comps = g->n_comps;
        for (int k = 0; k < 4; ++k) gp.comps[k] = g->comps[k];
        gp.plastic = plastic;
        for (uint32_t k = 0; k < 8; ++k) gp.stream_key[k] = stream_key(g->seed, k);
        gp.row = S.row;
        gp.grp = grp;
        gp.tgt = tgt;
        gp.w = w;
        gp.meta = meta;
        gp.delay = delay;
        gp.slot_of = net->slot_of;
        uint32_t* incnt = nullptr;
        if (plastic) {
            incnt = dev_alloc<uint32_t>(own, nall);
            gp.incnt = incnt;
        }
        const uint32_t rows_per_block = kBlock / kWave;
        const dim3 ggrid((uint32_t)((nall + rows_per_block - 1) / rows_per_block));
        hipLaunchKernelGGL(k_gen_pass1, ggrid, dim3(kBlock), 0, st, gp);
        if (plastic) {
            // target-major STDP records
            uint32_t* in32 = dev_alloc<uint32_t>(own, nall + 1);
            net->partials = dev_alloc<uint32_t>(own, kScanBlocks);
            launch_scan(net.get(), PlainSrc{incnt, nullptr, (uint32_t)nall}, in32, &in32[nall], 0, 0, nullptr);
            uint64_t* in_ptr = dev_alloc<uint64_t>(own, nall + 1);
            hipLaunchKernelGGL(k_widen, dim3(1024), dim3(kBlock), 0, st, in32, nall + 1, in_ptr);
            check(hipMemsetAsync(incnt, 0, nall * 4, st), "memset");
            S.in_ptr = in_ptr;
            uint32_t* tpos = dev_alloc<uint32_t>(own, e, false);
            StdpRec* stdp = dev_alloc<StdpRec>(own, e, false);
            uint32_t* tconn = dev_alloc<uint32_t>(own, e, false);
            S.tpos = tpos;
            S.stdp = stdp;
            S.upd_count = dev_alloc<uint32_t>(own, e);
            S.ltp_count = dev_alloc<uint32_t>(own, e);
            std::vector<float> wm(1, 2.0f);  // max_weight 2.0 on every synapse: min(2, 2)
            S.wmax = dev_upload(own, wm);
            net->tpos_conn_dev = tconn;
            net->n_tpos = e;
            gp.in_ptr = in_ptr;
            gp.tpos = tpos;
            gp.stdp = stdp;
            gp.tpos_conn = tconn;
        }
        hipLaunchKernelGGL(k_gen_pass2, ggrid, dim3(kBlock), 0, st, gp);
        std::vector<NeuronParams> params(nall);
        for (uint64_t i = 0; i < nall; ++i) {
            params[i].threshold = i < g->n ? g->threshold : g->source_threshold;
            params[i].window = g->window;
            params[i].sim_thr = g->sim_thr;
            params[i].max_refs = g->max_refs;
        }
        S.np = dev_upload(own, params);
        check(hipStreamSynchronize(st), "generate");
        finish_net(net.get(), o, e, params);
        *out = net.release();
        return SNN_OK;
    });
}

int snn_net_destroy(snn_net* net) {
    delete net;
    return SNN_OK;
}

int snn_net_inject(snn_net* net, const uint32_t* neuron, const uint64_t* tick, const float* weight, uint64_t n) {
    if (!net || (n && (!neuron || !tick || !weight))) return fail(SNN_EINVAL, "null argument");
    for (uint64_t i = 0; i < n; ++i)
        if (neuron[i] >= net->n) return fail(SNN_EINVAL, "inject neuron out of range");
    for (uint64_t i = 0; i < n; ++i) {
        // SpikeScheduler::schedule clamps a late event to the next poll tick
        const uint64_t t = std::max<uint64_t>(tick[i], net->next_tick);
        net->pending.push_back({neuron[i], t, weight[i], net->next_tick, net->inject_seq++});
    }
    return SNN_OK;
}

int snn_net_run_until(snn_net* net, uint64_t inclusive_tick) {
    if (!net) return fail(SNN_EINVAL, "null handle");
    if (inclusive_tick < net->next_tick) return SNN_OK;
    return guarded([&]() -> int {
        if (inclusive_tick + net->S.lmax + 2 >= 0xFFFFFFFFull) return fail(SNN_EINVAL, "tick beyond 2^32");
        check(hipSetDevice(net->device), "hipSetDevice");
        const uint64_t t0 = net->next_tick;
        const uint64_t nt = inclusive_tick - t0 + 1;
        // injects delivered inside this segment, ordered by (tick, stage, call order)
        std::vector<PendingInject> now, later;
        for (const auto& p : net->pending) (p.tick <= inclusive_tick ? now : later).push_back(p);
        std::sort(now.begin(), now.end(), [](const PendingInject& a, const PendingInject& b) {
            if (a.tick != b.tick) return a.tick < b.tick;
            if (a.stage != b.stage) return a.stage < b.stage;
            return a.seq < b.seq;
        });
        net->pending.swap(later);
        std::vector<uint32_t> itgt(now.size()), istage(now.size());
        std::vector<float> iw(now.size());
        std::vector<uint64_t> iptr(nt + 1, 0);
        for (size_t i = 0; i < now.size(); ++i) {
            itgt[i] = now[i].neuron;
            iw[i] = now[i].weight;
            istage[i] = (uint32_t)now[i].stage;
            iptr[now[i].tick - t0 + 1]++;
        }
        for (uint64_t t = 0; t < nt; ++t) iptr[t + 1] += iptr[t];
        for (auto& b : net->run_owned) (void)hipFree(b.p);
        net->run_owned.clear();
        Step& S = net->S;
        S.inj_tgt = dev_upload(net->run_owned, itgt);
        S.inj_w = dev_upload(net->run_owned, iw);
        S.inj_stage = dev_upload(net->run_owned, istage);
        S.inj_ptr = dev_upload(net->run_owned, iptr);
        S.inj_t0 = (uint32_t)t0;
        S.inj_nticks = (uint32_t)nt;
        S.tick_delivered = dev_alloc<uint32_t>(net->run_owned, nt);
        S.tick_fired = dev_alloc<uint32_t>(net->run_owned, nt);
        {
            const uint32_t tick32 = (uint32_t)t0;
            check(hipMemcpy(S.ctl, &tick32, 4, hipMemcpyHostToDevice), "set tick");
        }
        Timer tm{net};
        for (uint64_t t = 0; t < nt; ++t) {
            launch_tick(net, (uint32_t)t0, tm);
            if (net->profiling && tm.marks.size() > 4096) {
                check(hipStreamSynchronize(net->stream), "sync");
                tm.collect();
            }
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
            if (ctl[1] & kErrSort) m += " sort";
            return fail(SNN_ECAPACITY, m);
        }
        return SNN_OK;
    });
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
        if (code_match || code_learned) {
            uint32_t* m = dev_alloc<uint32_t>(tmp, e, false);
            uint32_t* l = dev_alloc<uint32_t>(tmp, e, false);
            if (net->code2)
                hipLaunchKernelGGL(k_gather_codes<true>, g, b, 0, net->stream, net->slot_of, net->S.code2, net->S.codeN, e, m, l);
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
            hipLaunchKernelGGL(k_gather_ticks, g, b, 0, net->stream, net->slot_of, net->S.tpos, net->S.stdp, net->S.code2,
                               net->S.codeN, net->S.delay, net->S.tgt, net->S.last_fired, e, dpre, dpost);
            if (last_pre) check(hipMemcpyAsync(last_pre, dpre, e * 8, hipMemcpyDeviceToHost, net->stream), "read ticks");
            if (last_post) check(hipMemcpyAsync(last_post, dpost, e * 8, hipMemcpyDeviceToHost, net->stream), "read ticks");
        }
        if (update_count) {
            std::fill(update_count, update_count + e, 0);
            if (net->n_tpos) {
                const uint64_t nt = net->n_tpos;
                std::vector<uint32_t> upd(nt);
                std::vector<uint32_t> conn(nt);
                check(hipMemcpy(upd.data(), net->S.upd_count, nt * 4, hipMemcpyDeviceToHost), "read stdp");
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
        check(hipMemcpy(upd.data(), net->S.upd_count, nt * 4, hipMemcpyDeviceToHost), "read stdp");
        check(hipMemcpy(ltp.data(), net->S.ltp_count, nt * 4, hipMemcpyDeviceToHost), "read stdp");
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
