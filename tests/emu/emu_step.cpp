// CPU model of the device step — test infrastructure only.
//
// Runs the same item functions as the HIP kernels (snncuda_amd/csrc/snn_step.h), sequentially, on
// host arrays, with the same launch order as launch_tick in snn_step.hip. Events are counted and
// scattered in a seeded shuffled order, so bins arrive unsorted as they do on the device and the
// per-target sort is exercised. Used by tests/test_emu.py to check the step's logic against the
// oracle on the CPU; the shipped library never contains it.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../snncuda_amd/csrc/snn_net.h"
#include "ref_pack.h"
#include "../../snncuda_amd/csrc/snn_step.h"

using namespace snn;

namespace {

struct Emu {
    HostNet h;
    Step S{};
    bool wide = false, stdp = false, code_here = true, code2 = true;
    std::mt19937_64 rng{12345};
    // owned storage
    std::vector<CodeRec2> code2v;
    std::vector<CodeRecN> codeNv;
    std::vector<uint32_t> pre_count, last_fired;
    std::vector<NeuronDyn> dyn;
    std::vector<TemporalHdr> th;
    std::vector<Run> cand;
    std::vector<Run> ref_runs;
    std::vector<RefMeta> ref_meta;
    std::vector<float> decay, ltp, ltd;
    std::vector<uint32_t> firelog, firemult, tick_fcount;
    std::vector<uint64_t> tick_fstart;
    std::vector<Group> groups;
    std::vector<uint32_t> seg_count;
    std::vector<uint64_t> seg_v0;
    std::vector<uint32_t> seg_n, seg_off, seg_mult;
    std::vector<uint32_t> fire_bits, fire_part, fire_neuron, fire_mult, fire_word_off, ctl;
    std::vector<uint64_t> totals;
    std::vector<uint32_t> raster;
    std::vector<uint32_t> itgt, istage;
    std::vector<float> iw;
    std::vector<uint64_t> iptr;
    std::vector<uint32_t> tdel, tfir;
    // run state
    struct Pend {
        uint32_t neuron;
        uint64_t tick;
        float w;
        uint64_t stage, seq;
    };
    std::vector<Pend> pending;
    uint64_t next_tick = 0, seq = 0;
    std::vector<uint64_t> tick_delivered, tick_fired;
    std::string err;
};

void scan_excl(const std::vector<uint32_t>& in, uint32_t n, std::vector<uint32_t>& out, uint32_t& total) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = s;
        s += in[i];
    }
    total = s;
}

// events of the tick in a shuffled order, grouped by target, then folded target by target
template <bool WIDE, bool ST, bool CH, bool C2>
uint32_t deliver(Emu& m, const std::vector<std::pair<uint32_t, uint32_t>>& order, uint32_t T) {
    typedef typename EntryOf<WIDE, CH>::type E;
    std::vector<uint32_t> xs(order.size());
    std::vector<E> recs(order.size());
    uint32_t err = 0;
    for (size_t k = 0; k < order.size(); ++k) {
        const uint32_t q = order[k].first;
        err |= make_event<WIDE, CH, C2>(m.S, m.S.seg_v0[q], m.S.seg_off[q], m.S.seg_mult[q], order[k].second, T, xs[k], recs[k]);
    }
    const bool narrow = m.S.seg_count[1] <= (1u << 27);
    if (m.rng() & 1u) {
        // the dense fold's chunks: the events in g order cut at random places (whole scatter runs
        // are contiguous g ranges), each chunk shuffled, its events grouped by target and folded
        // into the targets' running state (fold_begin / fold_piece / fold_end)
        std::vector<size_t> idx(recs.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return (recs[a].gk >> 4) < (recs[b].gk >> 4); });
        std::vector<FoldTarget> f(m.S.n);
        for (uint32_t x = 0; x < m.S.n; ++x) fold_begin(m.S, f[x], x, T);
        size_t lo = 0;
        while (lo < idx.size()) {
            const size_t hi = std::min(idx.size(), lo + 1 + (size_t)(m.rng() % 64));
            std::vector<size_t> ch(idx.begin() + lo, idx.begin() + hi);
            std::shuffle(ch.begin(), ch.end(), m.rng);
            std::stable_sort(ch.begin(), ch.end(), [&](size_t a, size_t b) { return xs[a] < xs[b]; });
            std::vector<E> piece;
            for (size_t a = 0; a < ch.size();) {
                size_t b = a;
                piece.clear();
                while (b < ch.size() && xs[ch[b]] == xs[ch[a]]) piece.push_back(recs[ch[b++]]);
                err |= fold_piece<WIDE, ST, CH, C2>(m.S, f[xs[ch[a]]], piece.data(), (uint32_t)piece.size(), xs[ch[a]], T, narrow);
                a = b;
            }
            lo = hi;
        }
        for (uint32_t x = 0; x < m.S.n; ++x) err |= fold_end(m.S, f[x], x, T);
        return err;
    }
    std::vector<uint32_t> cnt(m.S.n + 1, 0);
    for (uint32_t x : xs) cnt[x + 1]++;
    for (uint32_t i = 0; i < m.S.n; ++i) cnt[i + 1] += cnt[i];
    std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
    std::vector<E> by(recs.size());
    for (size_t k = 0; k < recs.size(); ++k) by[fill[xs[k]]++] = recs[k];
    for (uint32_t x = 0; x < m.S.n; ++x)
        err |= fold_events<WIDE, ST, CH, C2>(m.S, by.data() + cnt[x], cnt[x + 1] - cnt[x], x, T, narrow);
    return err;
}

void tick(Emu& m, uint32_t t0) {
    Step& S = m.S;
    const uint32_t T = S.ctl[0];
    // plan
    const uint32_t ng = 2 * S.lmax + 1;
    uint32_t base = 0;
    for (uint32_t gi = 0; gi < ng; ++gi) {
        Group g;
        group_info(S, T, gi, g);
        g.first = base;
        base += g.nseg;
        S.groups[gi] = g;
    }
    if (base > S.seg_cap) {
        S.ctl[1] |= kErrSegments;
        base = 0;
    }
    S.seg_count[0] = base;
    S.seg_count[1] = 0;
    // segments
    const uint32_t nseg = S.seg_count[0];
    for (uint32_t q = 0; q < nseg; ++q) {
        uint32_t lo = 0, hi = ng;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (S.groups[mid].first <= q) lo = mid + 1; else hi = mid;
        }
        const Group g = S.groups[lo - 1];
        segment_info(S, g, q - g.first, S.seg_v0[q], S.seg_n[q], S.seg_mult[q]);
    }
    uint32_t D = 0;
    scan_excl(m.seg_n, nseg, m.seg_off, D);
    if (D > std::min<uint64_t>(S.ev_cap, kMaxEventsPerTick)) {
        S.ctl[1] |= kErrEvents;
        S.seg_count[0] = 0;
        D = 0;
    }
    S.seg_count[1] = D;
    std::fill(m.fire_part.begin(), m.fire_part.end(), 0u);  // k_segments
    // count and scatter in a shuffled order
    std::vector<std::pair<uint32_t, uint32_t>> order;
    order.reserve(D);
    for (uint32_t q = 0; q < S.seg_count[0]; ++q)
        for (uint32_t i = 0; i < S.seg_n[q]; ++i) order.push_back({q, i});
    std::shuffle(order.begin(), order.end(), m.rng);
    uint32_t err = 0;
    if (!m.wide) err |= m.code2 ? deliver<false, false, true, true>(m, order, T) : deliver<false, false, true, false>(m, order, T);
    else if (m.stdp && m.code_here) err |= m.code2 ? deliver<true, true, true, true>(m, order, T) : deliver<true, true, true, false>(m, order, T);
    else if (m.stdp) err |= m.code2 ? deliver<true, true, false, true>(m, order, T) : deliver<true, true, false, false>(m, order, T);
    else if (m.code_here) err |= m.code2 ? deliver<true, false, true, true>(m, order, T) : deliver<true, false, true, false>(m, order, T);
    else err |= m.code2 ? deliver<true, false, false, true>(m, order, T) : deliver<true, false, false, false>(m, order, T);
    S.ctl[1] |= err;
    // fire ranks
    const uint32_t nwords = (D + 31) >> 5;
    std::vector<uint32_t> pc(nwords);
    for (uint32_t w = 0; w < nwords; ++w) pc[w] = __builtin_popcount(S.fire_bits[w]);
    uint32_t F = 0;
    scan_excl(pc, nwords, m.fire_word_off, F);
    {
        // the fold's per-block fire counts equal the bitmap's popcounts over k_fire_write's blocks
        // (an internal inconsistency is reported as an error flag no real error uses)
        const uint32_t chunk = (nwords + kFireBlocks - 1) / kFireBlocks;
        std::vector<uint32_t> want(kFireBlocks, 0);
        for (uint32_t w = 0; w < nwords; ++w) want[w / chunk] += pc[w];
        if (want != m.fire_part) S.ctl[1] |= 1u << 31;
    }
    S.ctl[2] = F;
    // fire write (k_fire_write)
    const uint64_t head = S.totals[2];
    const uint32_t R = S.lmax + 1;
    const uint64_t oldest = (S.lmax == 1) ? head : (T + 1 >= S.lmax ? S.tick_fstart[(T + 1 - S.lmax) % R] : 0);
    const bool fits = head + F - oldest <= S.fire_cap;
    const uint64_t rfill = S.totals[3];
    const bool rfits = !(S.ctl[3] & 1u) && rfill + F <= S.raster_cap;
    for (uint32_t wi = 0; wi < nwords; ++wi) {
        uint32_t bits = S.fire_bits[wi];
        if (!bits) continue;
        S.fire_bits[wi] = 0;
        uint32_t pos = S.fire_word_off[wi];
        while (bits) {
            const uint32_t b = __builtin_ctz(bits);
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
    if (!fits) S.ctl[1] |= kErrFireLog;
    if (!rfits) S.ctl[3] |= 1u;
    S.tick_fstart[T % R] = head;
    S.tick_fcount[T % R] = fits ? F : 0;
    // post-synaptic scans
    if (m.stdp) {
        const uint32_t Fc = S.tick_fcount[T % R];
        for (uint32_t f = 0; f < Fc; ++f) {
            const uint64_t idx = (head + f) % S.fire_cap;
            const uint32_t mult = S.firemult[idx];
            if (!mult) continue;
            const uint32_t x = S.firelog[idx];
            for (uint64_t t = S.in_ptr[x]; t < S.in_ptr[x + 1]; ++t) post_scan_record(S, t, T, mult);
        }
    }
    // advance
    S.tick_delivered[T - t0] = D;
    S.tick_fired[T - t0] = F;
    S.totals[0] += D;
    S.totals[1] += F;
    S.totals[2] += F;
    if (!(S.ctl[3] & 1u)) S.totals[3] += F;
    S.ctl[0] = T + 1;
}

std::vector<float> decay_table() {
    static const float taus[4] = {5.0f, 100.0f, 10.0f, 150.0f};
    std::vector<float> t(4 * (size_t)kDecayTable);
    for (int c = 0; c < 4; ++c)
        for (uint32_t dt = 0; dt < kDecayTable; ++dt) {
            volatile float el = (float)dt;
            t[(size_t)c * kDecayTable + dt] = std::exp(-el / taus[c]);
        }
    return t;
}

}  // namespace

// the per-target sort on E entries held as 32-bit words (small: the 8-entry networks of the
// chunked fold, shell sort past them)
template <typename E>
void sort_words(uint32_t* words, uint32_t k, bool narrow_g, bool small) {
    std::vector<E> ev(k);
    std::memcpy(ev.data(), words, k * sizeof(E));
    if (small) sort_by_g<true>(ev.data(), k, narrow_g);
    else sort_by_g(ev.data(), k, narrow_g);
    std::memcpy(words, ev.data(), k * sizeof(E));
}

extern "C" {

void* emu_create(uint64_t n, const float* thr, const uint64_t* win, const float* sim, const uint64_t* mref, uint64_t e,
                 const uint32_t* src, const uint32_t* tgt, const float* w, const float* mw, const uint32_t* dl,
                 const uint8_t* comp, const uint8_t* rec, const uint8_t* pl, const uint64_t* cptr, const uint32_t* codes,
                 uint64_t ev_cap, uint64_t fire_cap, uint64_t seed, uint64_t ref_cap, uint64_t run_cap,
                 char* errbuf, uint64_t errcap) {
    try {
        NetDesc d{n, thr, win, sim, mref, e, src, tgt, w, mw, dl, comp, rec, pl, cptr, codes};
        Emu* m = new Emu();
        m->rng.seed(seed);
        m->h = build_host_net(d);
        HostNet& h = m->h;
        Step& S = m->S;
        m->stdp = h.any_plastic;
        m->code_here = h.ve_identity;
        m->wide = h.any_plastic || !h.ve_identity;
        m->code2 = h.code_compact;
        S.n = h.n;
        S.lmax = h.lmax;
        S.row = h.row.data();
        S.grp = h.grp.data();
        S.ve = h.ve_identity ? nullptr : h.ve.data();
        S.tgt = h.tgt.data();
        S.w = h.w.data();
        S.meta = h.meta.data();
        S.delay = h.delay.data();
        S.expected = h.expected.empty() ? nullptr : h.expected.data();
        S.expected_all = h.expected_all ? h.expected_all : 2;
        if (h.code_compact) { m->code2v.assign(h.e + 1, CodeRec2{}); S.code2 = m->code2v.data(); }
        else { m->codeNv.assign(h.e + 1, CodeRecN{}); S.codeN = m->codeNv.data(); }
        m->pre_count.assign(h.e + 1, 0);
        S.pre_count = m->pre_count.data();
        if (h.any_plastic) {
            S.tpos = h.tpos.data();
            S.in_ptr = h.in_ptr.data();
            S.stdp = h.stdp.data();
        }
        S.np = h.params.data();
        S.tp_uniform = 1;
        for (const auto& q : h.params)
            if (q.window != h.params[0].window || q.max_refs != h.params[0].max_refs || q.sim_thr != h.params[0].sim_thr)
                S.tp_uniform = 0;
        S.tp_window = h.params[0].window;
        S.tp_max_refs = h.params[0].max_refs;
        S.tp_sim = h.params[0].sim_thr;
        S.thr_uniform = 1;
        for (const auto& q : h.params)
            if (std::memcmp(&q.threshold, &h.params[0].threshold, 4) != 0) S.thr_uniform = 0;
        S.thr_all = h.params[0].threshold;
        m->dyn.assign(h.n, NeuronDyn{});
        m->th.assign(h.n, TemporalHdr{});
        S.run_cap = run_cap ? (uint32_t)run_cap : kRuns;
        if (S.run_cap > kRunsMax) throw std::invalid_argument("pattern_runs above 64");
        m->cand.assign((uint64_t)h.n * S.run_cap, Run{});
        // learned-pattern slots per neuron: the largest max_reference_patterns, or ref_cap
        uint64_t want = 1;
        for (const auto& q : h.params) want = std::max<uint64_t>(want, q.max_refs);
        S.ref_cap = (uint32_t)std::min<uint64_t>(want, ref_cap ? ref_cap : 4096);
        m->ref_runs.assign((uint64_t)h.n * S.ref_cap * S.run_cap, Run{});
        m->ref_meta.assign((uint64_t)h.n * S.ref_cap, RefMeta{});
        m->last_fired.assign(h.n, kNone);
        S.dyn = m->dyn.data();
        S.th = m->th.data();
        S.cand = m->cand.data();
        S.ref_runs = m->ref_runs.data();
        S.ref_meta = m->ref_meta.data();
        S.last_fired = m->last_fired.data();
        m->decay = decay_table();
        S.decay_tab = m->decay.data();
        m->ltp.assign(kStdpTable, 0.0f);
        m->ltd.assign(kStdpTable, 0.0f);
        for (uint32_t a = 1; a < kStdpTable; ++a) {
            volatile float dt = (float)a;
            m->ltp[a] = 0.01f * std::exp(-dt / 20.0f);
            volatile float ndt = -(float)a;
            m->ltd[a] = -0.012f * std::exp(ndt / 20.0f);
        }
        S.ltp = m->ltp.data();
        S.ltd = m->ltd.data();
        S.fire_cap = fire_cap;
        m->firelog.assign(fire_cap, 0);
        m->firemult.assign(fire_cap, 0);
        S.firelog = m->firelog.data();
        S.firemult = m->firemult.data();
        m->tick_fstart.assign(h.lmax + 1, 0);
        m->tick_fcount.assign(h.lmax + 1, 0);
        S.tick_fstart = m->tick_fstart.data();
        S.tick_fcount = m->tick_fcount.data();
        m->groups.assign(2 * h.lmax + 1, Group{});
        S.groups = m->groups.data();
        m->seg_count.assign(4, 0);
        S.seg_count = m->seg_count.data();
        S.seg_cap = fire_cap + h.lmax + 1;
        S.ev_cap = (ev_cap + 31) & ~31ull;
        m->seg_v0.assign(S.seg_cap, 0);
        m->seg_n.assign(S.seg_cap, 0);
        m->seg_off.assign(S.seg_cap, 0);
        m->seg_mult.assign(S.seg_cap, 0);
        S.seg_v0 = m->seg_v0.data();
        S.seg_n = m->seg_n.data();
        S.seg_off = m->seg_off.data();
        S.seg_mult = m->seg_mult.data();
        m->fire_bits.assign(S.ev_cap / 32 + 1, 0);
        m->fire_neuron.assign(S.ev_cap, 0);
        m->fire_mult.assign(S.ev_cap, 0);
        m->fire_word_off.assign(S.ev_cap / 32 + 1, 0);
        S.fire_bits = m->fire_bits.data();
        m->fire_part.assign(kFireBlocks, 0);
        S.fire_part = m->fire_part.data();
        S.fire_neuron = m->fire_neuron.data();
        S.fire_mult = m->fire_mult.data();
        S.fire_word_off = m->fire_word_off.data();
        m->ctl.assign(4, 0);
        m->totals.assign(4, 0);
        S.ctl = m->ctl.data();
        S.totals = m->totals.data();
        S.raster_cap = 1ull << 24;
        m->raster.assign(S.raster_cap, 0);
        S.raster = m->raster.data();
        return m;
    } catch (const std::exception& ex) {
        if (errbuf && errcap) {
            std::strncpy(errbuf, ex.what(), errcap - 1);
            errbuf[errcap - 1] = 0;
        }
        return nullptr;
    }
}

void emu_destroy(void* p) { delete (Emu*)p; }

void emu_inject(void* p, uint32_t neuron, uint64_t tick, float w) {
    Emu* m = (Emu*)p;
    const uint64_t t = std::max<uint64_t>(tick, m->next_tick);
    m->pending.push_back({neuron, t, w, m->next_tick, m->seq++});
}

// returns the error flags of the run
uint32_t emu_run_until(void* p, uint64_t T1) {
    Emu* m = (Emu*)p;
    if (T1 < m->next_tick) return 0;
    const uint64_t t0 = m->next_tick, nt = T1 - t0 + 1;
    std::vector<Emu::Pend> now, later;
    for (auto& q : m->pending) (q.tick <= T1 ? now : later).push_back(q);
    std::sort(now.begin(), now.end(), [](const Emu::Pend& a, const Emu::Pend& b) {
        if (a.tick != b.tick) return a.tick < b.tick;
        if (a.stage != b.stage) return a.stage < b.stage;
        return a.seq < b.seq;
    });
    m->pending.swap(later);
    m->itgt.assign(now.size() + 1, 0);
    m->istage.assign(now.size() + 1, 0);
    m->iw.assign(now.size() + 1, 0.0f);
    m->iptr.assign(nt + 1, 0);
    for (size_t i = 0; i < now.size(); ++i) {
        m->itgt[i] = now[i].neuron;
        m->iw[i] = now[i].w;
        m->istage[i] = (uint32_t)now[i].stage;
        m->iptr[now[i].tick - t0 + 1]++;
    }
    for (uint64_t t = 0; t < nt; ++t) m->iptr[t + 1] += m->iptr[t];
    Step& S = m->S;
    S.inj_tgt = m->itgt.data();
    S.inj_w = m->iw.data();
    S.inj_stage = m->istage.data();
    S.inj_ptr = m->iptr.data();
    S.inj_t0 = (uint32_t)t0;
    S.inj_nticks = (uint32_t)nt;
    m->tdel.assign(nt, 0);
    m->tfir.assign(nt, 0);
    S.tick_delivered = m->tdel.data();
    S.tick_fired = m->tfir.data();
    S.ctl[0] = (uint32_t)t0;
    for (uint64_t t = 0; t < nt; ++t) tick(*m, (uint32_t)t0);
    for (uint64_t t = 0; t < nt; ++t) {
        m->tick_delivered.push_back(m->tdel[t]);
        m->tick_fired.push_back(m->tfir[t]);
    }
    m->next_tick = T1 + 1;
    return S.ctl[1];
}

uint64_t emu_ticks(void* p) { return ((Emu*)p)->tick_delivered.size(); }

void emu_tick_counts(void* p, uint64_t* d, uint64_t* f) {
    Emu* m = (Emu*)p;
    for (size_t i = 0; i < m->tick_delivered.size(); ++i) {
        d[i] = m->tick_delivered[i];
        f[i] = m->tick_fired[i];
    }
}

uint64_t emu_raster(void* p, uint32_t* out, uint64_t cap) {
    Emu* m = (Emu*)p;
    const uint64_t n = m->S.totals[3];
    if (out) std::memcpy(out, m->raster.data(), std::min(n, cap) * 4);
    return n;
}

void emu_neurons(void* p, float* membrane, float* cur, uint64_t* spikes, uint64_t* tmatch, uint64_t* tlearned) {
    Emu* m = (Emu*)p;
    const uint32_t n = m->S.n;
    for (uint32_t i = 0; i < n; ++i) {
        membrane[i] = m->dyn[i].membrane;
        for (int k = 0; k < 4; ++k) cur[(size_t)k * n + i] = m->dyn[i].cur[k];
        spikes[i] = m->dyn[i].spike_count;
        tmatch[i] = m->th[i].match;
        tlearned[i] = m->th[i].nlearned;
    }
}

void emu_synapses(void* p, float* w, uint64_t* pre, uint64_t* cmatch, uint64_t* clearn) {
    Emu* m = (Emu*)p;
    const HostNet& h = m->h;
    for (uint64_t j = 0; j < h.e; ++j) {
        const uint32_t s = h.slot_of[j];
        const uint32_t t = h.tpos[s];
        w[j] = t != kNone ? h.stdp[t].weight : h.w[s];
        pre[j] = m->pre_count[s];
        if (m->code2) {
            cmatch[j] = m->code2v[s].w1 & 0x0FFFFFFFu;
            clearn[j] = (m->code2v[s].w1 >> 30) & 1u;
        } else {
            cmatch[j] = m->codeNv[s].match;
            clearn[j] = (m->codeNv[s].flags >> 4) & 0xFu;
        }
    }
}

// the per-target sort alone, on k entries of `width` 32-bit words (gk first) in place
// The partition shift snn_net_create picks for n neurons, or -1 with the error it raises in msg.
int emu_partition_shift(uint64_t n, char* msg, uint64_t cap) {
    try {
        return (int)default_partition_shift(n);
    } catch (const std::exception& e) {
        std::snprintf(msg, cap, "%s", e.what());
        return -1;
    }
}

// narrow_g: bit 0 every g below 2^27, bit 1 the small networks
void emu_sort_by_g(uint32_t* words, uint32_t k, uint32_t narrow_g, uint32_t width) {
    const bool narrow = narrow_g & 1u, small = narrow_g & 2u;
    if (width == 2) sort_words<Entry8>(words, k, narrow, small);
    else if (width == 3) sort_words<Entry12>(words, k, narrow, small);
    else sort_words<Entry16>(words, k, narrow, small);
}

// One neuron's temporal pattern driven by m ticks with times[i] observations at ticks[i], once
// through temporal_observe_tick and once through times[i] single observations (temporal_observe).
// Returns the index of the first tick after which the two differ (header, error flags, stored
// candidate runs, or any ring slot), or m when they agree throughout.
uint32_t emu_temporal_check(const uint32_t* ticks, const uint32_t* times, uint32_t m, uint32_t window, float sim,
                            uint32_t max_refs, uint32_t cap, uint32_t runs) {
    NeuronParams p{1.0f, window, sim, max_refs};
    TemporalHdr ha{}, hb{};
    std::vector<Run> ca(runs), cb(runs);
    std::vector<Run> rra((uint64_t)cap * runs), rrb((uint64_t)cap * runs);
    std::vector<RefMeta> rma(cap), rmb(cap);
    const RefsView ra{rra.data(), rma.data(), cap, runs}, rb{rrb.data(), rmb.data(), cap, runs};
    for (uint32_t i = 0; i < m; ++i) {
        const uint32_t ea = temporal_observe_tick(ha, ca.data(), ra, ticks[i], times[i], p);
        uint32_t eb = 0;
        Run* pc = cb.data();
        bool dirty = false;
        for (uint32_t t = 0; t < times[i]; ++t) eb |= temporal_observe(hb, pc, rb, ticks[i], p, dirty);
        bool same = ea == eb && std::memcmp(&ha, &hb, sizeof ha) == 0;
        for (uint32_t r = 0; same && r < ha.nruns; ++r) same = ca[r].val == cb[r].val && ca[r].cnt == cb[r].cnt;
        // the learned patterns in ring order (slots outside the ring are never read)
        for (uint32_t q = 0; same && q < ha.nlearned; ++q) {
            const uint32_t k = ((ha.head_last & kHeadMask) + q) % cap;
            same = rma[k].len == rmb[k].len && rma[k].nruns == rmb[k].nruns;
            for (uint32_t r = 0; same && r < rma[k].nruns; ++r)
                same = ra.at(k)[r].val == rb.at(k)[r].val && ra.at(k)[r].cnt == rb.at(k)[r].cnt;
        }
        if (!same) return i;
    }
    return m;
}

// The host pack on `threads` threads against the sequential reference on one network: 0 when every
// output field is equal, or both raise the same message (written to msg); else the 1-based number of
// the first differing field (-1: only one of them raised, -2: different messages).
int emu_pack_check(uint64_t n, const float* thr, const uint64_t* win, const float* sim, const uint64_t* mref,
                   uint64_t e, const uint32_t* src, const uint32_t* tgt, const float* w, const float* mw,
                   const uint32_t* dl, const uint8_t* comp, const uint8_t* rec, const uint8_t* pl,
                   const uint64_t* cptr, const uint32_t* codes, uint32_t threads, char* msg, uint64_t msgcap) {
    NetDesc d{n, thr, win, sim, mref, e, src, tgt, w, mw, dl, comp, rec, pl, cptr, codes};
    HostNet a, b;
    std::string ea, eb;
    try { a = build_host_net(d, threads); } catch (const std::exception& x) { ea = x.what(); }
    try { b = build_host_net_serial(d); } catch (const std::exception& x) { eb = x.what(); }
    std::snprintf(msg, msgcap, "%s", ea.c_str());
    if (!ea.empty() || !eb.empty()) return ea.empty() != eb.empty() ? -1 : (ea == eb ? 0 : -2);
    auto same = [](const auto& x, const auto& y) {
        return x.size() == y.size() && (x.empty() || std::memcmp(x.data(), y.data(), x.size() * sizeof(x[0])) == 0);
    };
    const bool f[] = {a.n == b.n && a.e == b.e && a.nve == b.nve && a.lmax == b.lmax,
                      a.ve_identity == b.ve_identity && a.any_plastic == b.any_plastic,
                      a.expected_all == b.expected_all && a.code_compact == b.code_compact,
                      same(a.row, b.row), same(a.grp, b.grp), same(a.ve, b.ve), same(a.tgt, b.tgt), same(a.w, b.w),
                      same(a.meta, b.meta), same(a.delay, b.delay), same(a.expected, b.expected),
                      same(a.slot_of, b.slot_of), same(a.tpos, b.tpos), same(a.in_ptr, b.in_ptr),
                      same(a.stdp, b.stdp), same(a.tpos_conn, b.tpos_conn),
                      same(a.params, b.params)};
    for (size_t i = 0; i < sizeof f; ++i)
        if (!f[i]) return (int)i + 1;
    return 0;
}

}  // extern "C"
