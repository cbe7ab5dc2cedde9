// Per-target and per-edge state machines of the spiking step, written once for the device
// kernels (hipcc) and for the CPU model of those kernels used by the test suite (g++).
//
// Semantics follow the reference CPU step:
//   decay / process_input      src/runtime/SynapseDendriteProcessor.cpp:175-229
//   receptor tables            src/runtime/SynapseDendriteProcessor.cpp:231-269
//   neuron temporal pattern    src/runtime/SynapseDendriteProcessor.cpp:14-41, 294-316
//   synapse code pattern       src/runtime/SynapseDendriteProcessor.cpp:318-365
//   cosine / matcher           src/snn/TemporalPattern.cpp:12-60
//   STDP                       src/learning/STDP.cpp:20-40, include/snncuda/learning/STDP.h:8-16
//
// Differences in representation (never in results):
//   * all four compartments of a neuron share one last-update tick (every event writes the same
//     tick into all four, so they are always equal);
//   * a compartment's decay constant is stored as a receptor code (taus are always one of
//     5/100/10/150);
//   * the neuron's temporal offsets are run-length encoded (offsets are non-decreasing inside a
//     window), so a window with many events in one tick stays small. The cosine's integer sums are
//     identical to the reference's per-index FP64 sums, which are exact below 2^53.
//   * capacities (runs per pattern, learned patterns) are fixed; reaching one raises an error flag
//     instead of diverging from the reference.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SNN_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define SNN_HD inline
#endif

namespace snn {

// ------------------------------------------------------------------ constants

constexpr uint32_t kNone = 0xFFFFFFFFu;      // unset tick (std::nullopt)
constexpr uint32_t kDecayTable = 16384;      // expf(-dt/tau) is exactly 0 beyond this for tau <= 150
constexpr uint32_t kStdpTable = 4096;        // expf(-dt/20) is exactly 0 beyond this
// run-length capacity of a neuron temporal pattern (the candidate and each learned copy): kRuns
// unless the network asks for more (snn_options.pattern_runs), at most kRunsMax
constexpr uint32_t kRuns = 8;
constexpr uint32_t kRunsMax = 64;
constexpr uint32_t kHeadMask = 0x00FFFFFFu;  // TemporalHdr::head_last: learned-pattern ring head
constexpr uint32_t kLastMatch = 1u << 24;     // TemporalHdr::head_last: the last observation matched
constexpr uint32_t kCodeMax = 8;             // max offsets in a synapse code pattern
constexpr uint32_t kCodeRefs = 4;            // learned code-pattern capacity (general record)

// error flags (OR-ed into a device word)
enum : uint32_t {
    kErrEvents = 1u << 0,        // events in one tick exceed the bin capacity
    kErrSegments = 1u << 1,      // fan-out segments in one tick exceed capacity
    kErrFireLog = 1u << 2,       // fires inside the delay window exceed the fire-log capacity
    kErrTemporal = 1u << 3,      // a neuron temporal pattern exceeded run / learned capacity
    kErrCode = 1u << 4,          // a synapse code pattern exceeded learned capacity
    kErrRaster = 1u << 5,        // raster capacity exceeded
    kErrCounter = 1u << 6,       // a 28-bit match counter overflowed
};

// Dendritic compartment and receptor enums: include/snncuda/snn/Synapse.h:9-21
enum : uint32_t { kSoma = 0, kBasal = 1, kApical = 2, kInhibitory = 3 };
enum : uint32_t { kAmpa = 0, kNmda = 1, kGabaA = 2, kGabaB = 3 };

SNN_HD float receptor_gain(uint32_t r) { return r == kNmda ? 0.55f : (r == kGabaB ? 0.65f : 1.0f); }
SNN_HD float receptor_sign(uint32_t r) { return (r == kGabaA || r == kGabaB) ? -1.0f : 1.0f; }
// receptor code -> tau index: the decay tables are indexed by receptor code directly
// (tau 5 / 100 / 10 / 150 for Ampa / Nmda / GabaA / GabaB).

// reset taus after a fire: soma Ampa, basal Ampa, apical Nmda, inhibitory GabaA
// (src/runtime/SynapseDendriteProcessor.cpp:199-204); packed 2 bits per compartment.
constexpr uint32_t kResetTauCodes = (kAmpa << 0) | (kAmpa << 2) | (kNmda << 4) | (kGabaA << 6);

// ------------------------------------------------------------------ neuron state

struct NeuronParams {      // 16 B, static
    float threshold;
    uint32_t window;       // pattern window in ticks (<= 65535)
    float sim_thr;
    uint32_t max_refs;     // >= 1
};

struct NeuronDyn {         // 32 B, read and written once per touched neuron per tick
    float cur[4];          // soma, basal, apical, inhibitory
    float membrane;
    uint32_t last_tick;    // shared last-update tick of the four compartments
    uint32_t spike_count;
    uint32_t taus;         // 2-bit receptor code per compartment (bits 0-7)
};

// Neuron temporal pattern. Header is touched for every touched neuron; runs and refs only when a
// window starts or a match has to be evaluated.
struct TemporalHdr {       // 32 B
    uint32_t start;        // window start tick
    uint32_t n;            // number of offsets in the window
    uint32_t match;        // match count
    uint32_t maxlen;       // longest learned pattern (0 when none)
    uint32_t nruns;        // stored runs of the candidate
    uint32_t trunc;        // 1: candidate has more runs than stored
    uint32_t nlearned;     // learned patterns held (<= the ring's slots)
    uint32_t head_last;    // bits 0-23 ring head (kHeadMask), bit 24 last_match (kLastMatch)
};

// A temporal pattern is run-length encoded (its offsets are non-decreasing): an array of runs, as
// many as the network's run capacity; a run's value and count share a cache line.
struct Run {
    uint32_t val;          // the offset
    uint32_t cnt;          // how many consecutive offsets equal it
};

struct RefMeta {           // a learned pattern's length and stored runs
    uint32_t len;
    uint32_t nruns;
};

// One neuron's learned patterns (learn_pattern's FIFO of at most max_reference_patterns copies,
// SynapseDendriteProcessor.cpp:31-41) as a ring of `cap` slots in the network's pattern pool:
// neuron x owns slots [x * cap, (x + 1) * cap) of the run and RefMeta arrays.
struct RefsView {
    Run* ref;              // slot s's runs: ref + s * runs
    RefMeta* meta;
    uint32_t cap;
    uint32_t runs;         // run capacity of a pattern (and of the neuron's candidate)
    SNN_HD Run* at(uint32_t slot) const { return ref + (uint64_t)slot * runs; }
};

// cosine over run-length encoded sequences, prefix of k = min(len_l, len_r) entries.
// TemporalPatternMatcher::cosine_similarity (src/snn/TemporalPattern.cpp:35-60): the FP64 sums of
// integer products are exact, so summing per run equals summing per index.
SNN_HD float cosine_runs(const Run* l, uint32_t ln_runs, uint32_t l_len, const Run* r,
                         uint32_t rn_runs, uint32_t r_len) {
    if (l_len == 0 || r_len == 0) return 0.0f;
    uint32_t k = l_len < r_len ? l_len : r_len;
    uint64_t dot = 0, lnorm = 0, rnorm = 0;
    uint32_t i = 0, j = 0, li = 0, rj = 0;  // run index and consumed count within run
    (void)ln_runs;
    (void)rn_runs;
    while (k > 0) {
        const uint32_t la = l[i].cnt - li;
        const uint32_t ra = r[j].cnt - rj;
        uint32_t take = la < ra ? la : ra;
        if (take > k) take = k;
        const uint64_t lv = l[i].val, rv = r[j].val;
        dot += (uint64_t)take * lv * rv;
        lnorm += (uint64_t)take * lv * lv;
        rnorm += (uint64_t)take * rv * rv;
        k -= take;
        li += take;
        rj += take;
        if (li == l[i].cnt) { ++i; li = 0; }
        if (rj == r[j].cnt) { ++j; rj = 0; }
    }
    if (lnorm == 0 || rnorm == 0) return 0.0f;
    const double d = (double)dot, a = (double)lnorm, b = (double)rnorm;
    return (float)(d / (sqrt(a) * sqrt(b)));
}

// cosine_runs with the right sequence's run `ov` counted as `ov_cnt` (a count held in a register)
SNN_HD float cosine_runs_ov(const Run* l, uint32_t l_len, const Run* r, uint32_t r_len, uint32_t ov,
                            uint32_t ov_cnt) {
    if (l_len == 0 || r_len == 0) return 0.0f;
    uint32_t k = l_len < r_len ? l_len : r_len;
    uint64_t dot = 0, lnorm = 0, rnorm = 0;
    uint32_t i = 0, j = 0, li = 0, rj = 0;
    while (k > 0) {
        const uint32_t lc = l[i].cnt;
        const uint32_t rc = j == ov ? ov_cnt : r[j].cnt;
        const uint32_t la = lc - li;
        const uint32_t ra = rc - rj;
        uint32_t take = la < ra ? la : ra;
        if (take > k) take = k;
        const uint64_t lv = l[i].val, rv = r[j].val;
        dot += (uint64_t)take * lv * rv;
        lnorm += (uint64_t)take * lv * lv;
        rnorm += (uint64_t)take * rv * rv;
        k -= take;
        li += take;
        rj += take;
        if (li == lc) { ++i; li = 0; }
        if (rj == rc) { ++j; rj = 0; }
    }
    if (lnorm == 0 || rnorm == 0) return 0.0f;
    const double d = (double)dot, a = (double)lnorm, b = (double)rnorm;
    return (float)(d / (sqrt(a) * sqrt(b)));
}

// Prefix length covered by the stored runs of a candidate.
SNN_HD uint32_t stored_len(const Run* c, uint32_t nruns) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < nruns; ++i) s += c[i].cnt;
    return s;
}

// One observation of update_temporal_pattern (src/runtime/SynapseDendriteProcessor.cpp:294-316)
// with matches_learned_pattern (:14-29) and learn_pattern (:31-41). `cand` and `refs` are only
// dereferenced when needed; they may point to global memory. Returns error flags.
template <typename CandPtr>
SNN_HD uint32_t temporal_observe(TemporalHdr& h, CandPtr& cand, const RefsView& refs, uint32_t tick,
                                 const NeuronParams& p, bool& cand_dirty) {
    uint32_t err = 0;
    const bool reset = h.n == 0 || tick - h.start > p.window;
    if (reset) {
        h.start = tick;
        h.n = 0;
        h.nruns = 0;
        h.trunc = 0;
    }
    const uint32_t v = tick - h.start;
    const bool last_match = (h.head_last & kLastMatch) != 0;
    // With every learned pattern no longer than the offsets seen before this one, the compared
    // prefixes are unchanged, so the previous (matching) result repeats.
    const bool repeat = !reset && last_match && h.n >= h.maxlen;
    // append v
    if (!h.trunc) {
        if (h.nruns > 0 && cand[h.nruns - 1].val == v) {
            cand[h.nruns - 1].cnt += 1;
            cand_dirty = true;
        } else if (h.nruns < refs.runs) {
            cand[h.nruns].val = v;
            cand[h.nruns].cnt = 1;
            h.nruns += 1;
            cand_dirty = true;
        } else {
            h.trunc = 1;
        }
    }
    h.n += 1;
    bool match;
    if (repeat) {
        match = true;
    } else if (h.nlearned == 0 || h.n < 2) {
        match = false;
    } else {
        // every compared prefix must be inside the stored runs
        const uint32_t have = h.trunc ? stored_len(cand, h.nruns) : h.n;
        float best = 0.0f;
        const uint32_t head = h.head_last & kHeadMask;
        for (uint32_t q = 0; q < h.nlearned; ++q) {
            const uint32_t slot = (head + q) % refs.cap;
            const uint32_t rl = refs.meta[slot].len;
            const uint32_t need = rl < h.n ? rl : h.n;
            if (need > have) {
                err |= kErrTemporal;
                continue;
            }
            const float c = cosine_runs(refs.at(slot), refs.meta[slot].nruns, rl, cand, h.nruns, need);
            best = (best < c) ? c : best;
        }
        match = best >= p.sim_thr;
    }
    if (match) {
        h.match += 1;
        h.head_last |= kLastMatch;
        return err;
    }
    h.head_last &= ~kLastMatch;
    if (h.n < 2) return err;
    // learn a copy of the candidate (FIFO eviction at max_refs)
    if (h.trunc) return err | kErrTemporal;
    uint32_t head = h.head_last & kHeadMask;
    if (h.nlearned >= p.max_refs) {
        head = (head + 1) % refs.cap;
        h.nlearned -= 1;
    }
    if (h.nlearned >= refs.cap) return err | kErrTemporal;
    const uint32_t slot = (head + h.nlearned) % refs.cap;
    for (uint32_t i = 0; i < h.nruns; ++i) {
        refs.at(slot)[i].val = cand[i].val;
        refs.at(slot)[i].cnt = cand[i].cnt;
    }
    refs.meta[slot].len = h.n;
    refs.meta[slot].nruns = h.nruns;
    h.nlearned += 1;
    h.head_last = (h.head_last & ~kHeadMask) | head;
    uint32_t ml = 0;
    for (uint32_t q = 0; q < h.nlearned; ++q) {
        const uint32_t s = (head + q) % refs.cap;
        ml = refs.meta[s].len > ml ? refs.meta[s].len : ml;
    }
    h.maxlen = ml;
    return err;
}

// The `times` observations at `tick` of update_temporal_pattern (SynapseDendriteProcessor.cpp:294-316)
// that one neuron receives in one tick, with the same results as `times` calls of temporal_observe:
// they share the tick, so only the first can restart the window and all append the same offset,
// which lands in the candidate's last run. That run's count is kept in a register (written back
// once), and once a match repeats with the candidate at least as long as every learned pattern,
// every remaining observation of the tick is a repeated match and they are applied together.
//
// A tick that restarts the window appends offset 0 `times` times. Every compared candidate prefix is
// then all zeros, so every cosine is 0 (TemporalPattern.cpp:51-52) and, with a similarity threshold
// above 0, no observation matches: the first (n = 1) only clears last_match and learn j = 0, 1, ...
// (observation n = j + 2) learns a copy of the candidate, a single run of j + 2 zeros.
// temporal_restart_tick places those learns in closed form. With max_refs M no larger than the ring:
// learn j first evicts the oldest pattern when nl0 + j >= M, so it lands at ring position
// min(nl0 + j, M - 1) after max(0, nl0 + j + 1 - M) evictions, and a learn followed by M others is
// evicted again, so only the last M are written (slots outside the ring are never read). With M
// above the ring's slots nothing is evicted and the learns beyond the free slots are capacity errors.
template <typename CandPtr>
SNN_HD uint32_t temporal_restart_tick(TemporalHdr& h, CandPtr cand, const RefsView& refs, uint32_t times,
                                      const NeuronParams& p) {
    cand[0].val = 0;
    cand[0].cnt = times;
    h.nruns = 1;
    h.n = times;
    h.head_last &= ~kLastMatch;
    if (times < 2) return 0;
    const uint32_t L = times - 1, C = refs.cap, M = p.max_refs, nl0 = h.nlearned;
    const uint32_t head0 = h.head_last & kHeadMask;
    uint32_t err = 0, head = head0, nl;
    auto learn = [&](uint32_t slot, uint32_t len) {
        refs.at(slot)[0].val = 0;
        refs.at(slot)[0].cnt = len;
        refs.meta[slot].len = len;
        refs.meta[slot].nruns = 1;
    };
    if (M <= C) {
        for (uint32_t j = L > M ? L - M : 0; j < L; ++j) {
            const uint32_t ev = nl0 + j + 1 > M ? nl0 + j + 1 - M : 0;
            const uint32_t pos = nl0 + j < M - 1 ? nl0 + j : M - 1;
            learn((head0 + ev + pos) % C, j + 2);
        }
        head = (head0 + (nl0 + L > M ? nl0 + L - M : 0)) % C;
        nl = nl0 + L < M ? nl0 + L : M;
    } else {
        const uint32_t room = C - nl0, k = L < room ? L : room;
        for (uint32_t j = 0; j < k; ++j) learn((head0 + nl0 + j) % C, j + 2);
        if (L > room) err |= kErrTemporal;
        nl = nl0 + k;
    }
    h.nlearned = nl;
    h.head_last = (h.head_last & ~kHeadMask) | head;
    uint32_t ml = 0;
    for (uint32_t q = 0; q < nl; ++q) {
        const uint32_t l = refs.meta[(head + q) % C].len;
        ml = l > ml ? l : ml;
    }
    h.maxlen = ml;
    return err;
}

template <typename CandPtr>
SNN_HD uint32_t temporal_observe_tick(TemporalHdr& h, CandPtr cand, const RefsView& refs, uint32_t tick, uint32_t times,
                                      const NeuronParams& p) {
    uint32_t err = 0;
    if (times == 0) return 0;
    const bool reset = h.n == 0 || tick - h.start > p.window;
    if (reset) {
        h.start = tick;
        h.n = 0;
        h.nruns = 0;
        h.trunc = 0;
        if (!(0.0f >= p.sim_thr)) return temporal_restart_tick(h, cand, refs, times, p);
    }
    const uint32_t v = tick - h.start;
    // the run this tick's offsets go to (R: none, the candidate is truncated)
    const uint32_t R = refs.runs;
    uint32_t run = R, cnt = 0;
    if (!h.trunc) {
        if (h.nruns > 0 && cand[h.nruns - 1].val == v) {
            run = h.nruns - 1;
            cnt = cand[run].cnt;
        } else if (h.nruns < R) {
            run = h.nruns;
            cand[run].val = v;
            cnt = 0;
            h.nruns += 1;
        } else {
            h.trunc = 1;
        }
    }
    for (uint32_t t = 0; t < times; ++t) {
        const bool last_match = (h.head_last & kLastMatch) != 0;
        const bool repeat = !(reset && t == 0) && last_match && h.n >= h.maxlen;
        if (repeat) {
            // this and every later observation of the tick match
            const uint32_t rest = times - t;
            h.n += rest;
            h.match += rest;
            if (run < R) cnt += rest;
            break;
        }
        if (run < R) cnt += 1;
        h.n += 1;
        bool match = false;
        if (h.nlearned != 0 && h.n >= 2) {
            const uint32_t have = h.trunc ? stored_len(cand, h.nruns) : h.n;
            float best = 0.0f;
            const uint32_t head = h.head_last & kHeadMask;
            for (uint32_t q = 0; q < h.nlearned; ++q) {
                const uint32_t slot = (head + q) % refs.cap;
                const uint32_t rl = refs.meta[slot].len;
                const uint32_t need = rl < h.n ? rl : h.n;
                if (need > have) {
                    err |= kErrTemporal;
                    continue;
                }
                const float c = cosine_runs_ov(refs.at(slot), rl, cand, need, run, cnt);
                best = (best < c) ? c : best;
            }
            match = best >= p.sim_thr;
        }
        if (match) {
            h.match += 1;
            h.head_last |= kLastMatch;
            continue;
        }
        h.head_last &= ~kLastMatch;
        if (h.n < 2) continue;
        // learn a copy of the candidate (FIFO eviction at max_refs)
        if (h.trunc) {
            err |= kErrTemporal;
            continue;
        }
        if (run < R) cand[run].cnt = cnt;
        uint32_t head = h.head_last & kHeadMask;
        if (h.nlearned >= p.max_refs) {
            head = (head + 1) % refs.cap;
            h.nlearned -= 1;
        }
        if (h.nlearned >= refs.cap) {
            err |= kErrTemporal;
            continue;
        }
        const uint32_t slot = (head + h.nlearned) % refs.cap;
        for (uint32_t i = 0; i < h.nruns; ++i) {
            refs.at(slot)[i].val = cand[i].val;
            refs.at(slot)[i].cnt = cand[i].cnt;
        }
        refs.meta[slot].len = h.n;
        refs.meta[slot].nruns = h.nruns;
        h.nlearned += 1;
        h.head_last = (h.head_last & ~kHeadMask) | head;
        uint32_t ml = 0;
        for (uint32_t q = 0; q < h.nlearned; ++q) {
            const uint32_t s2 = (head + q) % refs.cap;
            ml = refs.meta[s2].len > ml ? refs.meta[s2].len : ml;
        }
        h.maxlen = ml;
    }
    if (run < R) cand[run].cnt = cnt;
    return err;
}

// ------------------------------------------------------------------ synapse code pattern

// Compact record for code patterns of 2 offsets with one learned pattern (16 B).
//   w0 window start; w1 bits 0-27 match count, 28-29 observed count, 30 learned count;
//   w2 observed offsets (u16 x2); w3 learned offsets (u16 x2)
struct CodeRec2 {
    uint32_t start;
    uint32_t w1;
    uint32_t obs;
    uint32_t ref;
};

// General record: up to kCodeMax offsets and kCodeRefs learned patterns (96 B).
struct CodeRecN {
    uint32_t start;
    uint32_t match;
    uint32_t flags;        // bits 0-3 observed, 4-7 learned, 8-11 head
    uint32_t pad;
    uint16_t obs[kCodeMax];
    uint16_t ref[kCodeRefs][kCodeMax];
};

// Tick of the last arrival recorded in a code record, or kNone before the first arrival. Every
// synaptic event updates its synapse's code pattern, so this is the synapse's last delivery tick.
SNN_HD uint32_t code_last_arrival2(const CodeRec2& r) {
    const uint32_t nobs = (r.w1 >> 28) & 3u;
    if (nobs == 0) return kNone;
    return r.start + (nobs == 2 ? (r.obs >> 16) : (r.obs & 0xFFFFu));
}
SNN_HD uint32_t code_last_arrival(const CodeRecN& r) {
    const uint32_t nobs = r.flags & 0xFu;
    return nobs == 0 ? kNone : r.start + r.obs[nobs - 1];
}

// FP64 cosine of two equal-length u16 sequences (index order)
SNN_HD float cosine_u16(const uint16_t* l, const uint16_t* r, uint32_t k) {
    double dot = 0.0, ln = 0.0, rn = 0.0;
    for (uint32_t i = 0; i < k; ++i) {
        const double a = (double)l[i], b = (double)r[i];
        dot += a * b;
        ln += a * a;
        rn += b * b;
    }
    if (ln == 0.0 || rn == 0.0) return 0.0f;
    return (float)(dot / (sqrt(ln) * sqrt(rn)));
}

// `times` observations at `tick` of update_synapse_code_pattern (:318-365) on a compact record
// (expected count 2, max one learned pattern held). Returns error flags.
SNN_HD uint32_t code_observe2(CodeRec2& r, uint32_t tick, uint32_t times, uint32_t window,
                              float sim_thr, uint32_t max_refs) {
    uint32_t err = 0;
    uint32_t match = r.w1 & 0x0FFFFFFFu;
    uint32_t nobs = (r.w1 >> 28) & 3u;
    uint32_t nlearned = (r.w1 >> 30) & 1u;
    uint32_t o0 = r.obs & 0xFFFFu, o1 = r.obs >> 16;
    uint32_t f0 = r.ref & 0xFFFFu, f1 = r.ref >> 16;
    for (uint32_t t = 0; t < times; ++t) {
        if (nobs == 0 || tick - r.start > window) {
            r.start = tick;
            nobs = 0;
        }
        const uint32_t v = tick - r.start;
        if (nobs < 2) {
            if (nobs == 0) o0 = v; else o1 = v;
            ++nobs;
        } else {
            o0 = o1;
            o1 = v;
        }
        if (nobs < 2) continue;
        float best = 0.0f;
        if (nlearned) {
            // cosine_u16 of (f0, f1) and (o0, o1), written out (index order, FP64)
            const double a0 = (double)f0, a1 = (double)f1, b0 = (double)o0, b1 = (double)o1;
            double dot = 0.0, ln = 0.0, rn = 0.0;
            dot += a0 * b0;
            ln += a0 * a0;
            rn += b0 * b0;
            dot += a1 * b1;
            ln += a1 * a1;
            rn += b1 * b1;
            const float c = (ln == 0.0 || rn == 0.0) ? 0.0f : (float)(dot / (sqrt(ln) * sqrt(rn)));
            best = (best < c) ? c : best;
        }
        if (best >= sim_thr) {
            ++match;
            continue;
        }
        // learn (FIFO with max_refs; this record holds one pattern)
        if (nlearned >= max_refs) nlearned -= 1;
        if (nlearned >= 1) {
            err |= kErrCode;
            continue;
        }
        f0 = o0;
        f1 = o1;
        nlearned = 1;
    }
    if (match > 0x0FFFFFFFu) err |= kErrCounter;
    r.w1 = (match & 0x0FFFFFFFu) | (nobs << 28) | (nlearned << 30);
    r.obs = o0 | (o1 << 16);
    r.ref = f0 | (f1 << 16);
    return err;
}

// General form: `expected` offsets (2..8), FIFO ring of up to kCodeRefs learned patterns.
SNN_HD uint32_t code_observeN(CodeRecN& r, uint32_t tick, uint32_t times, uint32_t expected,
                              uint32_t window, float sim_thr, uint32_t max_refs) {
    uint32_t err = 0;
    uint32_t nobs = r.flags & 0xFu;
    uint32_t nlearned = (r.flags >> 4) & 0xFu;
    uint32_t head = (r.flags >> 8) & 0xFu;
    for (uint32_t t = 0; t < times; ++t) {
        if (nobs == 0 || tick - r.start > window) {
            r.start = tick;
            nobs = 0;
        }
        const uint16_t v = (uint16_t)(tick - r.start);
        if (nobs < expected) {
            r.obs[nobs++] = v;
        } else {
            for (uint32_t i = 0; i + 1 < expected; ++i) r.obs[i] = r.obs[i + 1];
            r.obs[expected - 1] = v;
        }
        if (nobs < expected) continue;
        float best = 0.0f;
        for (uint32_t q = 0; q < nlearned; ++q) {
            const uint32_t s = (head + q) % kCodeRefs;
            const float c = cosine_u16(r.ref[s], r.obs, expected);
            best = (best < c) ? c : best;
        }
        if (best >= sim_thr) {
            ++r.match;
            continue;
        }
        if (nlearned >= max_refs) {
            head = (head + 1) % kCodeRefs;
            nlearned -= 1;
        }
        if (nlearned >= kCodeRefs) {
            err |= kErrCode;
            continue;
        }
        const uint32_t s = (head + nlearned) % kCodeRefs;
        for (uint32_t i = 0; i < expected; ++i) r.ref[s][i] = r.obs[i];
        nlearned += 1;
    }
    r.flags = nobs | (nlearned << 4) | (head << 8);
    return err;
}

// ------------------------------------------------------------------ STDP

// Plastic-synapse record, stored target-major so a post-synaptic scan is contiguous.
//   aux bits 8-31: post-synaptic updates already applied at tick last_post (only meaningful while
//   last_post is the current tick); bits 0-7 zero.
struct alignas(32) StdpRec {  // 32 B: a synapse's STDP state and counters in one cache-line access
    float weight;
    uint32_t last_pre;     // kNone when unset
    uint32_t last_post;    // kNone when unset
    uint32_t aux;
    uint32_t upd;          // changed updates (plasticity_update_count)
    uint32_t ltp;          // of which with delta > 0
    float hi;              // min(2.0f, max_weight): StdpRule::apply's `maximum`
    uint32_t pad;
};

// StdpRule::apply: clamp(w + delta, 0, min(2, max_w)), std::clamp written out as libstdc++ does.
// ltp[a] = 0.01f*expf(-a/20f), ltd[a] = -0.012f*expf(-a/20f) for a < kStdpTable, 0 beyond.
// Returns 0 when the weight is unchanged, kStdpLtp when it changed with delta > 0, kStdpLtd when
// it changed otherwise (the reference's split of changed updates into LTP and LTD,
// src/backends/cuda/CudaBackend.cu:133-139).
constexpr uint32_t kStdpLtd = 1u;
constexpr uint32_t kStdpLtp = 1u | (1u << 16);

// Changed-update tally of one synapse over several applications.
struct StdpTally {
    uint32_t n = 0, ltp = 0;
    SNN_HD void add(uint32_t c) {
        n += c & 1u;
        ltp += c >> 16;
    }
    SNN_HD void store(StdpRec& r) const {
        r.upd += n;
        r.ltp += ltp;
    }
};
SNN_HD uint32_t stdp_apply(StdpRec& r, uint32_t pre, uint32_t post, const float* ltp, const float* ltd) {
    float delta = 0.0f;
    if (post > pre) {
        const uint32_t a = post - pre;
        delta = a < kStdpTable ? ltp[a] : 0.0f;
    } else if (pre > post) {
        const uint32_t a = pre - post;
        delta = a < kStdpTable ? ltd[a] : 0.0f;
    }
    const float before = r.weight;
    const float v = r.weight + delta;
    const float hi = r.hi;
    r.weight = (v < 0.0f) ? 0.0f : (hi < v) ? hi : v;
    if ((r.weight - before) == 0.0f) return 0u;
    return delta > 0.0f ? kStdpLtp : kStdpLtd;
}

// ------------------------------------------------------------------ decay

// decay_compartments (src/runtime/SynapseDendriteProcessor.cpp:214-229) applied once per tick:
// after the first event of a tick every compartment's last-update tick equals the tick.
SNN_HD void decay_to(NeuronDyn& d, uint32_t tick, const float* decay_tab) {
    if (tick > d.last_tick) {
        const uint32_t dt = tick - d.last_tick;
        // the four factors are loaded unconditionally (an in-bounds index) so the loads overlap
        const uint32_t at = dt < kDecayTable ? dt : 0u;
        float f[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) f[k] = decay_tab[((d.taus >> (2 * k)) & 3u) * kDecayTable + at];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
            if (d.cur[k] != 0.0f) d.cur[k] = d.cur[k] * (dt < kDecayTable ? f[k] : 0.0f);
        d.last_tick = tick;
    }
}

}  // namespace snn
