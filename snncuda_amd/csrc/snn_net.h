// Host-side re-layout of a flattened Connectome for the device step.
//
// Input: neurons and synapses in connectome order (the order ConnectomeBuilder emits, reference
// src/declarative/Connectome.cpp:87-185), as plain arrays.
//
// Output layout (all SoA):
//   * slots: every synapse gets a slot; slots are sorted by (source, lag of its first code,
//     connectome index). lag = max(1, delay + code offset): a delay+offset of 0 is delivered on the
//     next tick because SpikeScheduler clamps late events to the next poll
//     (src/runtime/SpikeScheduler.cpp:15-19).
//   * virtual edges (VE): one per (synapse, code offset), sorted by (source, lag, connectome index,
//     code index). Within one (source, lag) group they are therefore in the order the reference
//     schedules them (src/runtime/NetworkPropagator.cpp:104-130). When every synapse has exactly
//     one code, VE i is slot i and no map is stored.
//   * group table: for source s, grp[s*(lmax+1) + L-1] is the first VE of lag L relative to the
//     source's row start, grp[s*(lmax+1) + lmax] the row length.
//   * plastic synapses get a target-major position (tpos) for their STDP record, so a neuron's
//     post-synaptic scan walks one contiguous range.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "snn_state.h"

namespace snn {

struct NetDesc {
    uint64_t n_neurons = 0;
    const float* threshold = nullptr;
    const uint64_t* window = nullptr;
    const float* sim_thr = nullptr;
    const uint64_t* max_refs = nullptr;
    uint64_t n_synapses = 0;
    const uint32_t* source = nullptr;
    const uint32_t* target = nullptr;
    const float* weight = nullptr;
    const float* max_weight = nullptr;
    const uint32_t* delay = nullptr;
    const uint8_t* compartment = nullptr;
    const uint8_t* receptor = nullptr;
    const uint8_t* plastic = nullptr;
    const uint64_t* code_ptr = nullptr;  // n_synapses + 1, may be null (every synapse has {0})
    const uint32_t* codes = nullptr;
};

constexpr uint32_t kMaxLag = 65535;

// Vectors whose resize leaves new elements default-initialized (no zero fill of a buffer the pack
// overwrites anyway: at 2e9 synapses each fill is a serial pass over gigabytes).
template <typename T>
struct DefaultInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = DefaultInitAlloc<U>;
    };
    DefaultInitAlloc() = default;
    template <typename U>
    DefaultInitAlloc(const DefaultInitAlloc<U>&) noexcept {}
    template <typename U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new ((void*)p) U;
    }
    template <typename U, typename... A>
    void construct(U* p, A&&... a) {
        ::new ((void*)p) U(std::forward<A>(a)...);
    }
};
template <typename T>
using pvec = std::vector<T, DefaultInitAlloc<T>>;

struct HostNet {
    uint32_t n = 0;
    uint64_t e = 0;           // synapses (slots)
    uint64_t nve = 0;         // virtual edges
    uint32_t lmax = 1;
    bool ve_identity = true;  // every synapse has one code
    bool any_plastic = false;
    uint32_t expected_all = 2;  // code pattern length if uniform, else 0
    bool code_compact = true;   // expected == 2 everywhere and every target has max_refs == 1

    pvec<uint64_t> row;      // n+1, VE row starts
    pvec<uint32_t> grp;      // n*(lmax+1)
    pvec<uint32_t> ve;       // nve (empty when identity)
    pvec<uint32_t> tgt;      // per slot
    pvec<float> w;           // per slot
    pvec<uint8_t> meta;      // per slot: receptor | compartment << 2 | plastic << 4
    pvec<uint16_t> delay;    // per slot
    pvec<uint8_t> expected;  // per slot code-pattern length (only when not uniform)
    pvec<uint32_t> slot_of;  // per connectome index
    // STDP
    pvec<uint32_t> tpos;     // per slot (kNone when not plastic)
    pvec<uint64_t> in_ptr;   // n+1 target-major plastic ranges
    pvec<StdpRec> stdp;      // per tpos
    pvec<uint32_t> tpos_conn;  // per tpos: connectome index
    // neurons
    std::vector<NeuronParams> params;
};

// Target partitions on one device. The scatter keeps two counters per partition in 64 KB of dynamic
// LDS (8 B each) and the count kernel one (4 B each), so at most 8,192 partitions fit: 33,554,432
// neurons in 4096-neuron partitions.
constexpr uint32_t kMaxPartitions = 8192;

inline uint64_t partitions_of(uint64_t n, uint32_t pshift) { return (n + (1ull << pshift) - 1) >> pshift; }

// The default partition size for n neurons (2^10, 2^11 from 1,025 neurons, 2^12 from 512k), and the
// error a network too large for one device raises when it is created.
inline uint32_t default_partition_shift(uint64_t n) {
    // partitions of 1,024 neurons up to 2^18 neurons (at most 256 partitions: a network of 10^5
    // neurons with 2,048-neuron partitions leaves most CUs without a fold workgroup), 2,048 up to 2^19
    const uint32_t pshift = n <= (1u << 18) ? 10 : (n < (1u << 19) ? 11 : 12);
    if (partitions_of(n, 12) > kMaxPartitions)
        throw std::invalid_argument("network of " + std::to_string(n) + " neurons needs " +
                                    std::to_string(partitions_of(n, 12)) +
                                    " target partitions of 4096 neurons; one device holds at most " +
                                    std::to_string(kMaxPartitions) + " (" +
                                    std::to_string((uint64_t)kMaxPartitions << 12) + " neurons)");
    return partitions_of(n, pshift) > kMaxPartitions ? 12u : pshift;
}

inline uint32_t lag_of(uint32_t delay, uint32_t off) {
    const uint64_t l = (uint64_t)delay + off;
    return l == 0 ? 1u : (uint32_t)std::min<uint64_t>(l, 0xFFFFFFFFull);
}

// ------------------------------------------------------------------ parallel helpers

// Threads of the host pack: the machine's, at most 16.
inline unsigned pack_threads() {
    const unsigned t = std::thread::hardware_concurrency();
    return t == 0 ? 1u : std::min(t, 16u);
}

// fn(t) for t in [0, T) on T threads (t = 0 on the caller's); fn must not throw.
template <typename F>
void run_threads(unsigned T, F&& fn) {
    std::vector<std::thread> th;
    th.reserve(T);
    for (unsigned t = 1; t < T; ++t) th.emplace_back([&fn, t] { fn(t); });
    fn(0u);
    for (auto& x : th) x.join();
}

// Stable counting sort of items [0, m) by key(i) < nkeys: perm[p] is the item at position p and
// start[k] the first position of key k (nkeys + 1 entries, start[nkeys] = m). Two levels keep the
// per-thread tables small: items first go to buckets of 2^shift consecutive keys (each thread counts
// and places a contiguous item range; bucket-major, thread-minor places keep item order), then
// every bucket is counting-sorted by its own keys. Item order is kept among equal keys.
template <typename Perm, typename KeyF>
void stable_sort_by_key(uint64_t m, uint64_t nkeys, const KeyF& key, unsigned T, pvec<Perm>& perm,
                        pvec<uint64_t>& start) {
    unsigned shift = 0;
    while ((nkeys >> shift) > 65536) ++shift;
    const uint64_t nb = (nkeys >> shift) + 1;
    std::vector<uint64_t> hc((uint64_t)T * nb, 0);
    run_threads(T, [&](unsigned t) {
        uint64_t* h = &hc[(uint64_t)t * nb];
        for (uint64_t i = m * t / T; i < m * (t + 1) / T; ++i) h[key(i) >> shift]++;
    });
    std::vector<uint64_t> bstart(nb + 1);
    uint64_t run = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        bstart[b] = run;
        for (unsigned t = 0; t < T; ++t) {
            const uint64_t c = hc[(uint64_t)t * nb + b];
            hc[(uint64_t)t * nb + b] = run;
            run += c;
        }
    }
    bstart[nb] = run;
    pvec<Perm> tmp(m);
    run_threads(T, [&](unsigned t) {
        uint64_t* h = &hc[(uint64_t)t * nb];
        for (uint64_t i = m * t / T; i < m * (t + 1) / T; ++i) tmp[h[key(i) >> shift]++] = (Perm)i;
    });
    perm.resize(m);
    start.resize(nkeys + 1);  // every key's entry is written below
    std::atomic<uint64_t> next{0};
    run_threads(T, [&](unsigned) {
        std::vector<uint64_t> cnt;
        for (uint64_t b = next.fetch_add(1); b < nb; b = next.fetch_add(1)) {
            const uint64_t k0 = b << shift;
            if (k0 >= nkeys) continue;
            const uint64_t k1 = std::min<uint64_t>(nkeys, (b + 1) << shift);
            cnt.assign(k1 - k0 + 1, 0);
            for (uint64_t q = bstart[b]; q < bstart[b + 1]; ++q) cnt[key(tmp[q]) - k0 + 1]++;
            for (uint64_t k = 1; k <= k1 - k0; ++k) cnt[k] += cnt[k - 1];
            for (uint64_t k = k0; k < k1; ++k) start[k] = bstart[b] + cnt[k - k0];
            for (uint64_t q = bstart[b]; q < bstart[b + 1]; ++q) perm[bstart[b] + cnt[key(tmp[q]) - k0]++] = tmp[q];
        }
    });
    start[nkeys] = m;
}

// First failing index of a parallel validation pass and its message (index ~0: none).
struct PackError {
    uint64_t at = ~0ull;
    const char* msg = nullptr;
    void note(uint64_t j, const char* m) {
        if (j < at) {
            at = j;
            msg = m;
        }
    }
    void merge(const PackError& o) { note(o.at, o.msg); }
    void raise() const {
        if (msg) throw std::invalid_argument(msg);
    }
};

// Points of the pack after which arrays of the HostNet are final, so their upload can start while
// the pack goes on:
//   Slots:  params, slot_of, tgt, w, meta, delay, expected (and n, e, nve, lmax, ve_identity,
//           any_plastic, expected_all)
//   Groups: row, grp, ve
//   Stdp:   tpos, in_ptr, stdp, tpos_conn (and code_compact)
enum class PackStage { Slots, Groups, Stdp };
using PackHook = std::function<void(PackStage, const HostNet&)>;

// Throws std::invalid_argument with a message for inputs outside the device step's limits (the
// message of the first offending neuron or synapse in connectome order). The passes over synapses
// run on `threads` threads (0: pack_threads(), one below 65,536 synapses); every output equals a
// sequential pass's. hook (optional) is called at each PackStage.
inline void build_host_net_into(HostNet& h, const NetDesc& d, unsigned threads = 0, const PackHook& hook = {}) {
    if (d.n_neurons == 0) throw std::invalid_argument("network has no neurons");
    if (d.n_neurons >= 0xFFFFFFFFull) throw std::invalid_argument("more than 2^32-2 neurons");
    if (d.n_synapses >= 0xFFFFFFFFull) throw std::invalid_argument("more than 2^32-2 synapses on one device");
    h.n = (uint32_t)d.n_neurons;
    h.e = d.n_synapses;
    const uint32_t n = h.n;
    const uint64_t e = h.e;
    const unsigned T = threads ? threads : (e < (1u << 16) ? 1u : pack_threads());

    h.params.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (d.max_refs[i] == 0) throw std::invalid_argument("max_reference_patterns must be >= 1");
        if (d.window[i] > 65535) throw std::invalid_argument("pattern_window_ticks above 65535 is not supported");
        h.params[i].threshold = d.threshold[i];
        h.params[i].window = (uint32_t)d.window[i];
        h.params[i].sim_thr = d.sim_thr[i];
        h.params[i].max_refs = (uint32_t)std::min<uint64_t>(d.max_refs[i], 0x7FFFFFFFull);
    }

    // code lists: empty -> {0} (src/runtime/SynapseDendriteProcessor.cpp:68-70)
    static const uint32_t zero = 0;
    auto codes_of = [&](uint64_t j, const uint32_t*& p) -> uint32_t {
        if (d.code_ptr && d.code_ptr[j + 1] > d.code_ptr[j]) {
            p = d.codes + d.code_ptr[j];
            return (uint32_t)(d.code_ptr[j + 1] - d.code_ptr[j]);
        }
        p = &zero;
        return 1;
    };

    // validation and network-wide figures, per thread over a contiguous synapse range
    struct Part {
        PackError err;
        uint32_t lmax = 1;
        uint64_t nve = 0;
        bool identity = true, any_plastic = false, exp_uniform = true, has = false;
        uint32_t exp_first = 0;
    };
    std::vector<Part> part(T);
    run_threads(T, [&](unsigned t) {
        Part& q = part[t];
        for (uint64_t j = e * t / T; j < e * (t + 1) / T; ++j) {
            if (d.source[j] >= n || d.target[j] >= n) {
                q.err.note(j, "synapse endpoint out of range");
                break;
            }
            if (d.compartment[j] > 3 || d.receptor[j] > 3) {
                q.err.note(j, "bad compartment or receptor");
                break;
            }
            const uint32_t* c;
            const uint32_t nc = codes_of(j, c);
            if (nc > kCodeMax) {
                q.err.note(j, "more than 8 spike-code offsets on a synapse");
                break;
            }
            if (nc != 1) q.identity = false;
            const uint32_t ex = nc > 2 ? nc : 2;
            if (!q.has) {
                q.exp_first = ex;
                q.has = true;
            } else if (ex != q.exp_first) {
                q.exp_uniform = false;
            }
            bool bad = false;
            for (uint32_t k = 0; k < nc; ++k) {
                const uint32_t L = lag_of(d.delay[j], c[k]);
                if (L > kMaxLag) bad = true;
                q.lmax = std::max(q.lmax, L);
            }
            if (bad) {
                q.err.note(j, "delay + code offset above 65535 ticks");
                break;
            }
            q.nve += nc;
            if (d.plastic[j]) q.any_plastic = true;
        }
    });
    PackError err;
    uint32_t lmax = 1;
    uint64_t nve = 0;
    bool identity = true, exp_uniform = true, seen = false;
    uint32_t exp_all = 0;
    for (const Part& q : part) {
        err.merge(q.err);
        lmax = std::max(lmax, q.lmax);
        nve += q.nve;
        identity = identity && q.identity;
        h.any_plastic = h.any_plastic || q.any_plastic;
        if (!q.has) continue;
        if (!q.exp_uniform || (seen && q.exp_first != exp_all)) exp_uniform = false;
        if (!seen) exp_all = q.exp_first;
        seen = true;
    }
    err.raise();
    h.lmax = lmax;
    h.nve = nve;
    h.ve_identity = identity;
    h.expected_all = exp_uniform ? (e ? exp_all : 2) : 0;
    if ((uint64_t)n * (lmax + 1) > (1ull << 31))
        throw std::invalid_argument("neurons x (max lag + 1) exceeds the group-table limit");

    const uint32_t W = lmax + 1;
    const uint64_t nkeys = (uint64_t)n * W;
    // slots: stable counting sort by key (source, lag of code 0), connectome order within a key
    auto slot_key = [&](uint64_t j) -> uint64_t {
        const uint32_t* c;
        codes_of(j, c);
        return (uint64_t)d.source[j] * W + lag_of(d.delay[j], c[0]);
    };
    pvec<uint32_t> order;  // slot -> connectome index
    pvec<uint64_t> cnt;    // first slot of each key
    stable_sort_by_key(e, nkeys, slot_key, T, order, cnt);
    h.slot_of.resize(e);
    h.tgt.resize(e);
    h.w.resize(e);
    h.meta.resize(e);
    h.delay.resize(e);
    if (!exp_uniform) h.expected.resize(e);
    std::vector<PackError> perr(T);
    run_threads(T, [&](unsigned t) {
        for (uint64_t s = e * t / T; s < e * (t + 1) / T; ++s) {
            const uint64_t j = order[s];
            h.slot_of[j] = (uint32_t)s;
            h.tgt[s] = d.target[j];
            h.w[s] = d.weight[j];
            h.meta[s] = (uint8_t)(d.receptor[j] | (d.compartment[j] << 2) | ((d.plastic[j] ? 1 : 0) << 4));
            h.delay[s] = (uint16_t)std::min<uint32_t>(d.delay[j], 65535);
            if (!exp_uniform) {
                const uint32_t* c;
                const uint32_t nc = codes_of(j, c);
                h.expected[s] = (uint8_t)(nc > 2 ? nc : 2);
            }
            if (d.plastic[j] && d.delay[j] > 65535) perr[t].note(j, "plastic synapse delay above 65535");
        }
    });
    for (const PackError& q : perr) err.merge(q);
    err.raise();
    pvec<uint32_t>().swap(order);
    if (hook) hook(PackStage::Slots, h);

    // virtual edges and group table
    h.row.assign((uint64_t)n + 1, 0);
    h.grp.assign(nkeys, 0);
    const pvec<uint64_t>* rowkeys = &cnt;
    pvec<uint64_t> vc;
    if (!identity) {
        // one item per (synapse, code) in (connectome index, code index) order
        std::vector<uint64_t> vptr(T + 1, 0);
        run_threads(T, [&](unsigned t) {
            uint64_t c0 = 0;
            for (uint64_t j = e * t / T; j < e * (t + 1) / T; ++j) {
                const uint32_t* c;
                c0 += codes_of(j, c);
            }
            vptr[t + 1] = c0;
        });
        for (unsigned t = 0; t < T; ++t) vptr[t + 1] += vptr[t];
        pvec<uint32_t> vkey(nve);  // n * W <= 2^31
        pvec<uint32_t> vval(nve);
        run_threads(T, [&](unsigned t) {
            uint64_t v = vptr[t];
            for (uint64_t j = e * t / T; j < e * (t + 1) / T; ++j) {
                const uint32_t* c;
                const uint32_t nc = codes_of(j, c);
                for (uint32_t k = 0; k < nc; ++k, ++v) {
                    vkey[v] = (uint32_t)((uint64_t)d.source[j] * W + lag_of(d.delay[j], c[k]));
                    vval[v] = h.slot_of[j];
                }
            }
        });
        h.ve.resize(nve);
        auto place = [&](auto& vorder) {
            stable_sort_by_key(nve, nkeys, [&](uint64_t v) { return (uint64_t)vkey[v]; }, T, vorder, vc);
            run_threads(T, [&](unsigned t) {
                for (uint64_t p = nve * t / T; p < nve * (t + 1) / T; ++p) h.ve[p] = vval[vorder[p]];
            });
        };
        if (nve < 0xFFFFFFFFull) {
            pvec<uint32_t> vorder;
            place(vorder);
        } else {
            pvec<uint64_t> vorder;
            place(vorder);
        }
        rowkeys = &vc;
    }
    {
        const pvec<uint64_t>& rk = *rowkeys;
        std::vector<uint8_t> too_long(T, 0);
        run_threads(T, [&](unsigned t) {
            for (uint32_t s = (uint32_t)((uint64_t)n * t / T); s < (uint32_t)((uint64_t)n * (t + 1) / T); ++s) {
                const uint64_t base = rk[(uint64_t)s * W];
                h.row[s] = base;
                for (uint32_t L = 1; L <= lmax; ++L) h.grp[(uint64_t)s * W + L - 1] = (uint32_t)(rk[(uint64_t)s * W + L] - base);
                const uint64_t len = rk[(uint64_t)(s + 1) * W] - base;
                if (len > 0xFFFFFFFFull) too_long[t] = 1;
                h.grp[(uint64_t)s * W + lmax] = (uint32_t)len;
            }
        });
        for (uint8_t x : too_long)
            if (x) throw std::invalid_argument("a neuron's fan-out x codes exceeds 2^32");
        h.row[n] = identity ? e : nve;
    }
    pvec<uint64_t>().swap(cnt);
    pvec<uint64_t>().swap(vc);
    if (hook) hook(PackStage::Groups, h);

    // compact code record only when every target neuron keeps at most one learned pattern
    std::vector<uint8_t> one(T, 1);
    run_threads(T, [&](unsigned t) {
        for (uint64_t j = e * t / T; j < e * (t + 1) / T && one[t]; ++j) one[t] = h.params[d.target[j]].max_refs == 1;
    });
    const bool all_one = std::all_of(one.begin(), one.end(), [](uint8_t x) { return x != 0; });
    h.code_compact = exp_uniform && exp_all == 2 && all_one;

    // STDP records, target-major in connectome order
    h.tpos.assign(e, kNone);
    if (h.any_plastic) {
        // plastic synapses in connectome order
        std::vector<uint64_t> pptr(T + 1, 0);
        run_threads(T, [&](unsigned t) {
            uint64_t c0 = 0;
            for (uint64_t j = e * t / T; j < e * (t + 1) / T; ++j) c0 += d.plastic[j] ? 1 : 0;
            pptr[t + 1] = c0;
        });
        for (unsigned t = 0; t < T; ++t) pptr[t + 1] += pptr[t];
        const uint64_t np = pptr[T];
        pvec<uint32_t> plist(np);
        run_threads(T, [&](unsigned t) {
            uint64_t q = pptr[t];
            for (uint64_t j = e * t / T; j < e * (t + 1) / T; ++j)
                if (d.plastic[j]) plist[q++] = (uint32_t)j;
        });
        pvec<uint32_t> porder;
        stable_sort_by_key(np, (uint64_t)n, [&](uint64_t q) { return (uint64_t)d.target[plist[q]]; }, T, porder,
                           h.in_ptr);
        h.stdp.resize(np);
        h.tpos_conn.resize(np);
        run_threads(T, [&](unsigned t) {
            for (uint64_t p = np * t / T; p < np * (t + 1) / T; ++p) {
                const uint32_t j = plist[porder[p]];
                h.tpos[h.slot_of[j]] = (uint32_t)p;
                h.tpos_conn[p] = j;
                StdpRec r{};
                r.weight = d.weight[j];
                r.last_pre = kNone;
                r.last_post = kNone;
                r.aux = 0;
                // StdpRule::apply: maximum = std::min(config max 2.0f, synapse max_weight)
                r.hi = (d.max_weight[j] < 2.0f) ? d.max_weight[j] : 2.0f;
                h.stdp[p] = r;
            }
        });
    }
    if (hook) hook(PackStage::Stdp, h);
}

inline HostNet build_host_net(const NetDesc& d, unsigned threads = 0) {
    HostNet h;
    build_host_net_into(h, d, threads);
    return h;
}

}  // namespace snn
