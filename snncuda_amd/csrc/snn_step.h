// The per-tick step as item functions: each __global__ kernel in snn_step.hip is a grid-stride
// loop over one of these, and the CPU model of the kernels in tests/ calls them sequentially.
//
// Tick T, in the reference's terms (src/runtime/NetworkPropagator.cpp:78-132):
//   the due events of T are applied in scheduler insertion order. Insertion order is
//   (stage, kind, position): a fire at tick te schedules its events at stage te+1 (after
//   pop_due advanced the scheduler), in fire order, each fire's outgoing synapses in connectome
//   order and each synapse's codes in order; an inject made while the next poll tick is P has
//   stage P and comes after the events scheduled by the fires of tick P-1
//   (src/runtime/SpikeScheduler.cpp:15-23). Stage-major, the due list of T is therefore
//       [injects with P <= T-lmax], then for L = lmax..1: [fires of T-L, lag-L groups],
//       [injects with P == T-L+1].
//   Every due event gets its global position g in that list. Each target's events are sorted by g
//   and applied sequentially (every write while processing an event for target x touches only x and
//   x's incoming synapses, so targets are independent). Fires are ordered by the g of the event
//   that triggered them, which is their order in the reference.
#pragma once

#include <type_traits>

#include "snn_state.h"

#if defined(__HIPCC__)
#define SNN_ATOMIC_ADD(p, v) atomicAdd((p), (v))
#define SNN_ATOMIC_OR(p, v) atomicOr((p), (v))
#else
template <typename T>
inline T snn_host_atomic_add(T* p, T v) { T o = *p; *p = o + v; return o; }
template <typename T>
inline T snn_host_atomic_or(T* p, T v) { T o = *p; *p = o | v; return o; }
#define SNN_ATOMIC_ADD(p, v) snn_host_atomic_add((p), (v))
#define SNN_ATOMIC_OR(p, v) snn_host_atomic_or((p), (v))
#endif

namespace snn {

constexpr uint64_t kInjFlag = 1ull << 63;
constexpr uint32_t kGroupInj = 0x80000000u;
// g is packed in 28 bits; with g < 2^28 - 1 no entry's gk is all ones, the value that pads the
// per-target sorting networks
constexpr uint32_t kMaxEventsPerTick = (1u << 28) - 1;

SNN_HD uint32_t f2u(float f) {
    union { float f; uint32_t u; } c;
    c.f = f;
    return c.u;
}
SNN_HD float u2f(uint32_t u) {
    union { float f; uint32_t u; } c;
    c.u = u;
    return c.f;
}

// event entries in a target's bin
struct Entry8 {            // no STDP, code patterns updated edge-parallel
    uint32_t gk;           // g << 4 | compartment << 2 | receptor
    float w;               // the receptor current (weight * gain) * sign
};
struct Entry12 {           // STDP, code patterns updated edge-parallel
    uint32_t gk;
    float w;               // the receptor current; for a plastic event (its weight is in the STDP
                           // record) the bits of the synapse's delay
    uint32_t tpos;         // STDP record (kNone: not plastic or external)
};
struct Entry16 {           // code patterns updated per event in the fold
    uint32_t gk;
    float w;               // the receptor current of the static weight (a plastic event's current
                           // comes from its STDP record)
    uint32_t tpos;         // STDP record (kNone: not plastic or external)
    uint32_t slot;         // synapse slot (kNone: external)
};
// the entry of a step variant: WIDE (STDP or per-event code patterns), CODE_HERE (code patterns
// advanced in the scatter)
template <bool WIDE, bool CODE_HERE>
struct EntryOf {
    typedef typename std::conditional<!WIDE, Entry8, typename std::conditional<CODE_HERE, Entry12, Entry16>::type>::type type;
};

struct Group {             // a run of the tick's due list
    uint32_t first;        // first segment index
    uint32_t nseg;
    uint64_t a;            // fire-log index of the first fire, or first inject index
    uint32_t b;            // lag L, or inject count | kGroupInj
    uint32_t pad;
};

// The learned-pattern pool in pages of 2^kRefPageShift neurons when it is larger than the device
// holds: every page's rings live in host memory, and the pages most events reach are copied into a
// device cache between ticks; a page's entry says where its rings are now.
constexpr uint32_t kRefPageShift = 10;
struct RefPage {
    Run* runs;
    RefMeta* meta;
};

// Everything one tick needs. Plain pointers so the same struct is a kernel argument.
struct Step {
    // network
    uint32_t n, lmax;
    const uint64_t* row;
    const uint32_t* grp;
    const uint64_t* gstart;      // optional: row[s] + grp[s * (lmax + 1) + j], the groups' absolute starts
    const uint32_t* ve;          // null when VE == slot
    const uint32_t* tgt;
    const float* w;
    const uint8_t* meta;
    const uint16_t* delay;
    const uint8_t* expected;     // null when uniform
    uint32_t expected_all;
    CodeRec2* code2;
    CodeRecN* codeN;
    uint32_t* pre_count;         // optional
    const uint32_t* tpos;
    const uint64_t* in_ptr;
    StdpRec* stdp;
    const NeuronParams* np;
    uint32_t tp_uniform;         // every neuron has the same window / similarity / max_refs
    uint32_t tp_window, tp_max_refs;
    float tp_sim;
    uint32_t thr_uniform;        // every neuron has the threshold thr_all
    float thr_all;
    NeuronDyn* dyn;
    TemporalHdr* th;
    Run* cand;                   // run_cap runs per neuron
    Run* ref_runs;               // learned patterns: ref_cap slots of run_cap runs per neuron
    uint32_t run_cap;            // runs per temporal pattern (1 .. kRunsMax; kRuns by default)
    RefMeta* ref_meta;
    uint32_t ref_cap;
    RefPage* ref_page;           // paged pool: where each page's rings are (null: ref_runs / ref_meta)
    uint32_t* page_hits;         // paged pool: events per page since the last residency update
    uint32_t* last_fired;
    const float* decay_tab;      // 4 x kDecayTable
    const float* ltp;            // kStdpTable
    const float* ltd;
    // fire ring
    uint32_t* firelog;
    uint32_t* firemult;
    uint64_t fire_cap;
    uint64_t* tick_fstart;       // ring of lmax+1
    uint32_t* tick_fcount;
    // injects of the current run segment
    const uint32_t* inj_tgt;
    const float* inj_w;
    const uint32_t* inj_stage;
    const uint64_t* inj_ptr;     // per tick of the segment, +1
    uint32_t inj_t0, inj_nticks;
    // per-tick work
    Group* groups;               // 2*lmax+1
    uint32_t* seg_count;         // [0] segments, [1] events
    uint64_t seg_cap, ev_cap;
    uint64_t* seg_v0;
    uint32_t* seg_n;
    uint32_t* seg_off;
    uint32_t* seg_loc;           // device: first event relative to the segment's k_segments block
    uint32_t* seg_mult;
    // target partitions: partition p holds neurons [p << pshift, (p+1) << pshift)
    uint32_t pshift, nparts, nwg;
    uint32_t* hist;              // [nparts][nwg] per-workgroup event counts, then their offsets within
                                 // the partition's bucket
    uint32_t* ptotal;            // [nparts] events per partition (bucket p starts at the sum before it)
    uint32_t* pstart;            // [nparts] bucket starts (the scatter's first workgroup writes them)
    uint32_t* partials;          // block partial sums of the two-kernel scans (kScanBlocks)
    // the tick being launched (kNone: read ctl[0], as HIP graph replays do), and whether the first
    // kernel of the tick also finishes the previous tick's counters (stream launches)
    uint32_t tick;
    uint32_t fused_advance;
    uint32_t narrow_max;         // ticks of at most this many events sort by 32-bit keys (2^27)
    // Code patterns kept per source neuron (code_lazy): with one code offset per synapse and the
    // same window / similarity / max_refs on every target, a synapse's code pattern is a function of
    // its source's fire ticks shifted by the synapse's lag, so one record per source, advanced when a
    // fire has reached all of its synapses (lmax ticks later), gives every synapse's record.
    uint32_t code_lazy;
    CodeRec2* src_code;          // per neuron: its fires up to tick T - lmax applied, in fire time
    void* prec;                  // partition buckets: Entry8 / Entry16 events, positions [0, D)
    uint16_t* ptl;               // each bucket event's target index within its partition
    uint32_t chunk;              // scatter write-combining chunk (records per partition buffer; 0: none)
    void* stage;                 // fold scratch for partitions larger than the LDS stage
    uint32_t* fire_bits;
    uint32_t* fire_part;         // fires per block of the fire-list kernel (kFireBlocks), counted by the
                                 // fold as it sets the fire bits (null: the caller ranks the bitmap)
    uint32_t* fire_neuron;
    uint32_t* fire_mult;
    uint32_t* fire_word_off;     // the CPU model's fire ranks (the device ranks fires in k_fire_write)
    // device state
    uint32_t* ctl;               // [0] tick, [1] error flags, [2] fires this tick, [3] bit 0: raster full
    uint64_t* totals;            // [0] delivered, [1] fired, [2] fire-log head, [3] raster fill
    uint32_t* tick_delivered;    // per tick of the run segment
    uint32_t* tick_fired;
    uint32_t* raster;            // optional
    uint64_t raster_cap;
    uint32_t* fire_evt;          // optional: per fire position g, the index of the firing event in the
    uint32_t* raster_evt;        // target's ordered events of the tick; and per raster entry
};

// The fire-list kernel's blocks: the fire bitmap's ceil(D / 32) words of a tick with D events in
// kFireBlocks contiguous ranges of ceil(words / kFireBlocks) words; fire_block is the block holding
// fire position g.
constexpr uint32_t kFireBlocks = 1024;
SNN_HD uint32_t fire_block(uint32_t D, uint32_t g) {
    const uint32_t nwords = (D + 31u) >> 5;
    const uint32_t chunk = (nwords + kFireBlocks - 1) / kFireBlocks;
    return (g >> 5) / chunk;
}

// A fire at position g: its bit in the fire bitmap, its block's fire count, its neuron.
SNN_HD void mark_fire(const Step& S, uint32_t g, uint32_t x) {
    SNN_ATOMIC_OR(&S.fire_bits[g >> 5], 1u << (g & 31u));
    if (S.fire_part) SNN_ATOMIC_ADD(&S.fire_part[fire_block(S.seg_count[1], g)], 1u);
    S.fire_neuron[g] = x;
}

SNN_HD RefsView refs_of(const Step& S, uint32_t x) {
    if (S.ref_page) {
        const RefPage pg = S.ref_page[x >> kRefPageShift];
        const uint64_t b = (uint64_t)(x & ((1u << kRefPageShift) - 1u)) * S.ref_cap;
        return RefsView{pg.runs + b * S.run_cap, pg.meta + b, S.ref_cap, S.run_cap};
    }
    const uint64_t b = (uint64_t)x * S.ref_cap;
    return RefsView{S.ref_runs + b * S.run_cap, S.ref_meta + b, S.ref_cap, S.run_cap};
}

SNN_HD uint32_t lower_bound_u32(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t v) {
    // both ends first (two independent loads): a tick's injects usually share one stage, so the
    // bound is usually one of them and the chain of dependent loads of the search is skipped
    if (lo < hi) {
        const uint32_t first = a[lo], last = a[hi - 1];
        if (last < v) return (uint32_t)hi;
        if (first >= v) return (uint32_t)lo;
    }
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return (uint32_t)lo;
}

// Group gi of tick T (0 <= gi <= 2*lmax): its segment count and payload.
SNN_HD void group_info(const Step& S, uint32_t T, uint32_t gi, Group& g) {
    g.nseg = 0;
    g.a = 0;
    g.b = 0;
    uint64_t ia = 0, ib = 0;
    if (T >= S.inj_t0 && T - S.inj_t0 < S.inj_nticks) {
        ia = S.inj_ptr[T - S.inj_t0];
        ib = S.inj_ptr[T - S.inj_t0 + 1];
    }
    if (gi == 0) {
        // injects with stage P <= T - lmax
        if (ib > ia && T >= S.lmax) {
            const uint32_t hi = lower_bound_u32(S.inj_stage, ia, ib, T - S.lmax + 1);
            if (hi > ia) {
                g.nseg = 1;
                g.a = ia;
                g.b = (uint32_t)(hi - ia) | kGroupInj;
            }
        }
        return;
    }
    const uint32_t L = S.lmax - (gi - 1) / 2;
    if (((gi - 1) & 1u) == 0) {
        if (T >= L) {
            const uint32_t te = T - L;
            const uint32_t slot = te % (S.lmax + 1);
            g.nseg = S.tick_fcount[slot];
            g.a = S.tick_fstart[slot];
            g.b = L;
        }
    } else if (ib > ia && T + 1 >= L) {
        const uint32_t P = T - L + 1;
        const uint32_t lo = lower_bound_u32(S.inj_stage, ia, ib, P);
        const uint32_t hi = lower_bound_u32(S.inj_stage, lo, ib, P + 1);
        if (hi > lo) {
            g.nseg = 1;
            g.a = lo;
            g.b = (uint32_t)(hi - lo) | kGroupInj;
        }
    }
}

// Segment q of the tick: VE start (or inject start | kInjFlag), event count, owner multiplicity.
SNN_HD void segment_info(const Step& S, const Group& g, uint32_t r, uint64_t& v0, uint32_t& cnt,
                         uint32_t& mult) {
    if (g.b & kGroupInj) {
        v0 = g.a | kInjFlag;
        cnt = g.b & ~kGroupInj;
        mult = 0;
        return;
    }
    const uint64_t fi = (g.a + r) % S.fire_cap;
    const uint32_t s = S.firelog[fi];
    mult = S.firemult[fi];
    const uint32_t L = g.b;
    const uint64_t base = (uint64_t)s * (S.lmax + 1);
    if (S.gstart) {
        // one random 16-byte read instead of the row start and the group offsets
        v0 = S.gstart[base + L - 1];
        cnt = (uint32_t)(S.gstart[base + L] - v0);
        return;
    }
    const uint32_t lo = S.grp[base + L - 1];
    const uint32_t hi = S.grp[base + L];
    v0 = S.row[s] + lo;
    cnt = hi - lo;
}

SNN_HD uint32_t event_target(const Step& S, uint64_t v0, uint32_t i) {
    if (v0 & kInjFlag) return S.inj_tgt[(v0 & ~kInjFlag) + i];
    const uint64_t v = v0 + i;
    const uint32_t slot = S.ve ? S.ve[v] : (uint32_t)v;
    return S.tgt[slot];
}

SNN_HD uint32_t code_expected(const Step& S, uint32_t slot) {
    return S.expected ? S.expected[slot] : S.expected_all;
}

// The record of event (q, i): its target x, position key and payload. With CODE_HERE the owner
// segment (first fire of the source this tick) also applies the synapse code pattern `mult` times:
// the code pattern of a synapse depends only on that synapse's own arrival ticks, and same-tick
// arrivals are identical. Returns error flags.
// (v0, g0, mult) are segment q's seg_v0 / seg_off / seg_mult; event i has position g0 + i.
template <bool WIDE, bool CODE_HERE, bool CODE2, typename E>
SNN_HD uint32_t make_event(const Step& S, uint64_t v0, uint32_t g0, uint32_t mult, uint32_t i, uint32_t T,
                           uint32_t& x, E& e) {
    const uint32_t g = g0 + i;
    uint32_t err = 0;
    if (v0 & kInjFlag) {
        const uint64_t k = (v0 & ~kInjFlag) + i;
        x = S.inj_tgt[k];
        e.w = S.inj_w[k] * receptor_gain(kAmpa) * receptor_sign(kAmpa);
        e.gk = (g << 4) | (kSoma << 2) | kAmpa;  // process_external_input: Soma, Ampa
        if constexpr (WIDE) e.tpos = kNone;
        if constexpr (WIDE && !CODE_HERE) e.slot = kNone;
        return 0;
    }
    const uint64_t v = v0 + i;
    const uint32_t slot = S.ve ? S.ve[v] : (uint32_t)v;
    x = S.tgt[slot];
    const uint32_t m = S.meta[slot];
    // the receptor current (w * gain) * sign of process_input (SynapseDendriteProcessor.cpp:183),
    // formed here once per event with the same operations; a plastic event's weight is read from its
    // STDP record in the fold instead
    e.w = S.w[slot] * receptor_gain(m & 3u) * receptor_sign(m & 3u);
    e.gk = (g << 4) | (m & 15u);
    if constexpr (WIDE) {
        e.tpos = (m & 16u) ? S.tpos[slot] : kNone;
        if constexpr (CODE_HERE) {
            if (m & 16u) e.w = u2f(S.delay[slot]);
        } else {
            e.slot = slot;
        }
    }
    if (CODE_HERE) {
        if (mult && !S.code_lazy) {
            uint32_t window = S.tp_window, max_refs = S.tp_max_refs;
            float sim = S.tp_sim;
            if (!S.tp_uniform) {
                const NeuronParams p = S.np[x];
                window = p.window;
                sim = p.sim_thr;
                max_refs = p.max_refs;
            }
            if (CODE2) err |= code_observe2(S.code2[slot], T, mult, window, sim, max_refs);
            else err |= code_observeN(S.codeN[slot], T, mult, code_expected(S, slot), window, sim, max_refs);
        }
        if (mult && S.pre_count) S.pre_count[slot] += mult;
    }
    return err;
}

// Batcher's odd-even merge sort network for the next power of two >= M, keeping only the
// compare-exchanges within the first M places: padded with maximal keys, places M.. hold the padding
// through every stage, so a compare-exchange that reaches them never swaps.
constexpr uint32_t kMaxNetCx = 192;
struct SortNet {
    uint32_t n;
    uint8_t lo[kMaxNetCx], hi[kMaxNetCx];
};
constexpr SortNet odd_even_net(uint32_t M) {
    uint32_t N = 1;
    while (N < M) N <<= 1;
    SortNet s{};
    for (uint32_t p = 1; p < N; p <<= 1)
        for (uint32_t k = p; k >= 1; k >>= 1)
            for (uint32_t j = k % p; j + k < N; j += 2 * k)
                for (uint32_t i = 0; i < k && i + j + k < N; ++i)
                    if ((i + j) / (2 * p) == (i + j + k) / (2 * p) && i + j + k < M) {
                        s.lo[s.n] = (uint8_t)(i + j);
                        s.hi[s.n] = (uint8_t)(i + j + k);
                        ++s.n;
                    }
    return s;
}

// k <= M entries sorted by g in registers: no dependent LDS round trips and no divergence inside
// the network (g is unique within a tick, so comparing whole gk words compares g)
template <uint32_t M, typename E>
SNN_HD void net_sort_by_g(E* ev, uint32_t k) {
    constexpr SortNet net = odd_even_net(M);
    static_assert(net.n <= kMaxNetCx, "sorting network table too small");
    E a[M];
#pragma unroll
    for (uint32_t j = 0; j < M; ++j) {
        if (j < k) {
            a[j] = ev[j];
        } else {
            a[j] = E{};
            a[j].gk = 0xFFFFFFFFu;
        }
    }
#pragma unroll
    for (uint32_t c = 0; c < net.n; ++c) {
        E& x = a[net.lo[c]];
        E& y = a[net.hi[c]];
        const bool sw = x.gk > y.gk;
        const E lo = sw ? y : x;
        const E hi = sw ? x : y;
        x = lo;
        y = hi;
    }
#pragma unroll
    for (uint32_t j = 0; j < M; ++j)
        if (j < k) ev[j] = a[j];
}

// The same with one 32-bit key per entry, g << 5 | its place (g < 2^27): a compare-exchange is a
// min and a max, and the entries are gathered by place afterwards.
template <uint32_t M, typename E>
SNN_HD void key_sort_by_g(E* ev, uint32_t k) {
    static_assert(M <= 32, "the place takes 5 bits of the key");
    constexpr SortNet net = odd_even_net(M);
    static_assert(net.n <= kMaxNetCx, "sorting network table too small");
    uint32_t key[M];
#pragma unroll
    for (uint32_t j = 0; j < M; ++j) key[j] = j < k ? ((ev[j].gk >> 4) << 5) | j : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t c = 0; c < net.n; ++c) {
        const uint32_t x = key[net.lo[c]], y = key[net.hi[c]];
        key[net.lo[c]] = x < y ? x : y;
        key[net.hi[c]] = x < y ? y : x;
    }
    // gathered as 32-bit words: an array of 12- or 16-byte entries would not stay in registers
    constexpr uint32_t W = sizeof(E) / 4;
    uint32_t a[M][W];
#pragma unroll
    for (uint32_t j = 0; j < M; ++j) {
        if (j >= k) continue;
        const char* src = (const char*)(ev + (key[j] & 31u));
#pragma unroll
        for (uint32_t q = 0; q < W; ++q) __builtin_memcpy(&a[j][q], src + 4 * q, 4);
    }
#pragma unroll
    for (uint32_t j = 0; j < M; ++j) {
        if (j >= k) continue;
#pragma unroll
        for (uint32_t q = 0; q < W; ++q) __builtin_memcpy((char*)(ev + j) + 4 * q, &a[j][q], 4);
    }
}

// narrow_g: every g of the tick is below 2^27. SMALL: sorting networks of 8 entries at most (fewer
// registers; larger pieces take the in-place shell sort).
template <bool SMALL = false, typename E>
SNN_HD void sort_by_g(E* ev, uint32_t k, bool narrow_g) {
    if (k <= 1) return;
    if constexpr (SMALL) {
        if (k <= 8) {
            if (narrow_g) key_sort_by_g<8>(ev, k);
            else net_sort_by_g<8>(ev, k);
            return;
        }
    } else if constexpr (sizeof(E) == sizeof(Entry8)) {
        if (k <= 24) {
            if (narrow_g) {
                if (k <= 8) key_sort_by_g<8>(ev, k);
                else key_sort_by_g<24>(ev, k);
            } else {
                if (k <= 8) net_sort_by_g<8>(ev, k);
                else net_sort_by_g<24>(ev, k);
            }
            return;
        }
    } else {
        // wider entries: smaller networks (registers), key and place only
        if (narrow_g && k <= 16) {
            if (k <= 8) key_sort_by_g<8>(ev, k);
            else key_sort_by_g<16>(ev, k);
            return;
        }
    }
    // Shell sort (Ciura gaps) in place for the rest. The gap loop is unrolled so the gap table stays
    // in registers.
#pragma unroll
    for (uint32_t gi = 0; gi < 8; ++gi) {
        const uint32_t gap = gi == 0 ? 701u : gi == 1 ? 301u : gi == 2 ? 132u : gi == 3 ? 57u
                           : gi == 4 ? 23u : gi == 5 ? 10u : gi == 6 ? 4u : 1u;
        if (gap >= k) continue;
        for (uint32_t i = gap; i < k; ++i) {
            const E t = ev[i];
            const uint32_t key = t.gk >> 4;
            uint32_t j = i;
            while (j >= gap && (ev[j - gap].gk >> 4) > key) {
                ev[j] = ev[j - gap];
                j -= gap;
            }
            ev[j] = t;
        }
    }
}

// Bring a record up to `fires` post-synaptic updates of tick T (apply_post_spike_plasticity,
// SynapseDendriteProcessor.cpp:138-173, once per fire of the target at T).
SNN_HD StdpTally post_catch_up(const Step& S, StdpRec& r, uint32_t T, uint32_t fires) {
    uint32_t applied = r.last_post == T ? (r.aux >> 8) : 0u;
    StdpTally changed;
    for (; applied < fires; ++applied) {
        r.last_post = T;
        if (r.last_pre != kNone) changed.add(stdp_apply(r, r.last_pre, T, S.ltp, S.ltd));
    }
    if (r.last_post == T) r.aux = applied << 8;
    return changed;
}

// One target's fold within a tick, kept in registers while its events arrive in pieces: every
// piece holds events of larger g than the pieces before it (the fold reads a partition's bucket
// in runs of whole scatter workgroups, and each scatter workgroup owns a contiguous range of g).
// Everything fold_end writes to the neuron's dynamics is carried (the compartments decayed to T, the
// tau codes, this tick's fires and the stored spike count), so the dynamics are read once per tick.
struct FoldTarget {
    float cur[4];          // compartment currents, decayed to T at fold_begin
    uint32_t taus_fires;   // bits 0-7 receptor code per compartment, bits 8-31 fires this tick
    uint32_t k;            // events applied so far this tick
    uint32_t first_g;      // g of the first fire
    float threshold;
    uint32_t spikes;       // spike count before this tick
};
constexpr uint32_t kFoldMaxFires = (1u << 24) - 1;

// Decaying at fold_begin instead of at the first event gives the same values (decay_to depends only
// on T and the stored state); a target without events this tick is never written back.
SNN_HD void fold_begin(const Step& S, FoldTarget& f, uint32_t x, uint32_t T) {
    NeuronDyn d = S.dyn[x];
    decay_to(d, T, S.decay_tab);
    f.cur[0] = d.cur[0];
    f.cur[1] = d.cur[1];
    f.cur[2] = d.cur[2];
    f.cur[3] = d.cur[3];
    f.taus_fires = d.taus & 0xFFu;
    f.k = 0;
    f.first_g = 0;
    f.threshold = S.thr_uniform ? S.thr_all : S.np[x].threshold;
    f.spikes = d.spike_count;
}

// The next kc events of target x at tick T (ev: this piece, any order; sorted by g here), applied
// in order: receptor current, membrane, threshold / reset and the pre-synaptic side of STDP
// (process_synaptic_input / process_input, SynapseDendriteProcessor.cpp:89-212). Returns error flags.
// narrow_g: every g of the tick is below 2^27 (read once per kernel).
template <bool WIDE, bool STDP, bool CODE_HERE, bool CODE2, typename E>
SNN_HD uint32_t fold_piece(const Step& S, FoldTarget& f, E* ev, uint32_t kc, uint32_t x, uint32_t T, bool narrow_g) {
    if (kc == 0) return 0;
    sort_by_g<true>(ev, kc, narrow_g);
    uint32_t err = 0;
    for (uint32_t i = 0; i < kc; ++i) {
        const E e = ev[i];
        const uint32_t g = e.gk >> 4;
        const uint32_t comp = (e.gk >> 2) & 3u;
        const uint32_t rec = e.gk & 3u;
        float cur = e.w;  // the receptor current, formed in make_event
        if constexpr (WIDE) {
            if (STDP && e.tpos != kNone) {
                StdpRec r = S.stdp[e.tpos];
                StdpTally changed = post_catch_up(S, r, T, f.taus_fires >> 8);
                uint32_t dl;
                if constexpr (CODE_HERE) dl = f2u(e.w);
                else dl = S.delay[e.slot];
                const uint32_t pre = T >= dl ? T - dl : T;
                if (r.last_post != kNone) changed.add(stdp_apply(r, pre, r.last_post, S.ltp, S.ltd));
                r.last_pre = pre;
                changed.store(r);
                S.stdp[e.tpos] = r;
                cur = r.weight * receptor_gain(rec) * receptor_sign(rec);
            }
            if constexpr (!CODE_HERE) {
                if (e.slot != kNone) {
                    const uint32_t slot = e.slot;
                    uint32_t window = S.tp_window, max_refs = S.tp_max_refs;
                    float sim = S.tp_sim;
                    if (!S.tp_uniform) {
                        const NeuronParams p = S.np[x];
                        window = p.window;
                        sim = p.sim_thr;
                        max_refs = p.max_refs;
                    }
                    if (CODE2) err |= code_observe2(S.code2[slot], T, 1, window, sim, max_refs);
                    else err |= code_observeN(S.codeN[slot], T, 1, code_expected(S, slot), window, sim, max_refs);
                    if (S.pre_count) S.pre_count[slot] += 1;
                }
            }
        }
        const float applied_current = cur;
        f.cur[0] = comp == 0 ? f.cur[0] + applied_current : f.cur[0];
        f.cur[1] = comp == 1 ? f.cur[1] + applied_current : f.cur[1];
        f.cur[2] = comp == 2 ? f.cur[2] + applied_current : f.cur[2];
        f.cur[3] = comp == 3 ? f.cur[3] + applied_current : f.cur[3];
        f.taus_fires = (f.taus_fires & ~(3u << (2 * comp))) | (rec << (2 * comp));
        // the membrane is not carried: after every event it is the sum of the compartments (zero
        // after a reset), so fold_end recomputes the last one
        const float membrane = f.cur[0] + f.cur[1] + f.cur[2] + f.cur[3];
        if (membrane >= f.threshold) {
            f.cur[0] = 0.0f;
            f.cur[1] = 0.0f;
            f.cur[2] = 0.0f;
            f.cur[3] = 0.0f;
            if ((f.taus_fires >> 8) == kFoldMaxFires) err |= kErrCounter;
            if ((f.taus_fires >> 8) == 0) f.first_g = g;
            f.taus_fires = ((f.taus_fires & ~0xFFu) | kResetTauCodes) + (1u << 8);
            mark_fire(S, g, x);
            S.fire_mult[g] = 0;
            if (S.fire_evt) S.fire_evt[g] = f.k + i;
        }
    }
    f.k += kc;
    return err;
}

// End of target x's tick: the neuron temporal pattern's k observations (it does not feed the
// membrane, so they are applied together), then the state written back. Nothing is written for a
// target without events. The synapses whose events came before one of this tick's fires are
// brought up to every fire by the post-synaptic scan kernel (post_scan_record). fold_end_load
// reads the target's temporal header (so a thread can issue the loads of all its targets before it
// applies any), fold_end_apply finishes.
SNN_HD void fold_end_load(const Step& S, const FoldTarget& f, uint32_t x, TemporalHdr& h) {
    if (f.k == 0) return;
    h = S.th[x];
}

SNN_HD uint32_t fold_end_apply(const Step& S, const FoldTarget& f, uint32_t x, uint32_t T, TemporalHdr h) {
    if (f.k == 0) return 0;
    NeuronParams p;
    if (S.tp_uniform) {
        p.window = S.tp_window;
        p.sim_thr = S.tp_sim;
        p.max_refs = S.tp_max_refs;
    } else {
        p = S.np[x];
    }
    p.threshold = f.threshold;
    const uint32_t err = temporal_observe_tick(h, S.cand + (uint64_t)x * S.run_cap, refs_of(S, x), T, f.k, p);
    const uint32_t fires = f.taus_fires >> 8;
    if (fires) {
        S.fire_mult[f.first_g] = fires;
        S.last_fired[x] = T;
    }
    NeuronDyn d;
    d.cur[0] = f.cur[0];
    d.cur[1] = f.cur[1];
    d.cur[2] = f.cur[2];
    d.cur[3] = f.cur[3];
    d.membrane = f.cur[0] + f.cur[1] + f.cur[2] + f.cur[3];
    d.last_tick = T;
    d.spike_count = f.spikes + fires;
    d.taus = f.taus_fires & 0xFFu;
    S.dyn[x] = d;
    S.th[x] = h;
    return err;
}

SNN_HD uint32_t fold_end(const Step& S, const FoldTarget& f, uint32_t x, uint32_t T) {
    TemporalHdr h{};
    fold_end_load(S, f, x, h);
    return fold_end_apply(S, f, x, T, h);
}

// The sequential fold of target x's k events at tick T (ev is the target's own array, in any
// order; it is sorted by g here, with networks of at most 8 entries when SMALL; narrow_g: every g of
// the tick is below 2^27). Returns error flags.
template <bool WIDE, bool STDP, bool CODE_HERE, bool CODE2, bool SMALL = false, typename E>
SNN_HD uint32_t fold_events(const Step& S, E* ev, uint32_t k, uint32_t x, uint32_t T, bool narrow_g) {
    if (k == 0) return 0;
    sort_by_g<SMALL>(ev, k, narrow_g);

    const NeuronParams p = S.np[x];
    NeuronDyn d = S.dyn[x];
    TemporalHdr h = S.th[x];
    uint32_t err = 0;
    Run* cand = S.cand + (uint64_t)x * S.run_cap;
    decay_to(d, T, S.decay_tab);
    uint32_t fires = 0, first_g = 0;
    for (uint32_t i = 0; i < k; ++i) {
        const E e = ev[i];
        const uint32_t g = e.gk >> 4;
        const uint32_t comp = (e.gk >> 2) & 3u;
        const uint32_t rec = e.gk & 3u;
        float cur = e.w;  // the receptor current, formed in make_event
        if constexpr (WIDE) {
            if (STDP && e.tpos != kNone) {
                StdpRec r = S.stdp[e.tpos];
                // first the post-synaptic updates of this tick's earlier fires of x
                StdpTally changed = post_catch_up(S, r, T, fires);
                // pre-synaptic side: process_synaptic_input (:95-126); the code offset is ignored
                uint32_t dl;
                if constexpr (CODE_HERE) dl = f2u(e.w);
                else dl = S.delay[e.slot];
                const uint32_t pre = T >= dl ? T - dl : T;
                if (r.last_post != kNone) changed.add(stdp_apply(r, pre, r.last_post, S.ltp, S.ltd));
                r.last_pre = pre;
                changed.store(r);
                S.stdp[e.tpos] = r;
                cur = r.weight * receptor_gain(rec) * receptor_sign(rec);
            }
            if constexpr (!CODE_HERE) {
                if (e.slot != kNone) {
                    const uint32_t slot = e.slot;
                    if (CODE2) err |= code_observe2(S.code2[slot], T, 1, p.window, p.sim_thr, p.max_refs);
                    else err |= code_observeN(S.codeN[slot], T, 1, code_expected(S, slot), p.window, p.sim_thr, p.max_refs);
                    if (S.pre_count) S.pre_count[slot] += 1;
                }
            }
        }
        // process_input (:175-212)
        const float applied_current = cur;
        // cur[comp] += applied, written as selects so the compartments stay in registers
        d.cur[0] = comp == 0 ? d.cur[0] + applied_current : d.cur[0];
        d.cur[1] = comp == 1 ? d.cur[1] + applied_current : d.cur[1];
        d.cur[2] = comp == 2 ? d.cur[2] + applied_current : d.cur[2];
        d.cur[3] = comp == 3 ? d.cur[3] + applied_current : d.cur[3];
        d.taus = (d.taus & ~(3u << (2 * comp))) | (rec << (2 * comp));
        d.membrane = d.cur[0] + d.cur[1] + d.cur[2] + d.cur[3];
        if (d.membrane >= p.threshold) {
            d.membrane = 0.0f;
            d.cur[0] = 0.0f;
            d.cur[1] = 0.0f;
            d.cur[2] = 0.0f;
            d.cur[3] = 0.0f;
            d.taus = kResetTauCodes;
            d.spike_count += 1;
            mark_fire(S, g, x);
            S.fire_mult[g] = 0;
            if (S.fire_evt) S.fire_evt[g] = i;
            if (fires == 0) first_g = g;
            ++fires;
        }
    }
    // the neuron temporal pattern does not feed the membrane: its k observations of this tick are
    // applied together
    err |= temporal_observe_tick(h, cand, refs_of(S, x), T, k, p);
    if (fires) {
        S.fire_mult[first_g] = fires;
        S.last_fired[x] = T;
        if constexpr (WIDE && STDP) {
            // bring the synapses that had events this tick up to all fires; the others are done by
            // the post-synaptic scan kernel
            for (uint32_t i = 0; i < k; ++i) {
                const uint32_t t = ev[i].tpos;
                if (t == kNone) continue;
                StdpRec r = S.stdp[t];
                post_catch_up(S, r, T, fires).store(r);
                S.stdp[t] = r;
            }
        }
    }
    S.dyn[x] = d;
    S.th[x] = h;
    return err;
}

// post-synaptic scan of neuron x after m fires at tick T, for one incoming record r (a register
// copy); returns whether r changed and must be stored
SNN_HD bool post_scan_apply(const Step& S, StdpRec& r, uint32_t T, uint32_t m) {
    if (r.last_post == T && (r.aux >> 8) == m) return false;  // done inside the fold
    post_catch_up(S, r, T, m).store(r);
    return true;
}

SNN_HD void post_scan_record(const Step& S, uint64_t t, uint32_t T, uint32_t m) {
    StdpRec r = S.stdp[t];
    if (post_scan_apply(S, r, T, m)) S.stdp[t] = r;
}

}  // namespace snn
