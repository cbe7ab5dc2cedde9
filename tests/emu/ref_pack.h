// The host pack as one sequential pass per step (the library's first form), kept as the reference
// the parallel pack in snncuda_amd/csrc/snn_net.h is tested against. Test-only.
#pragma once

#include "../../snncuda_amd/csrc/snn_net.h"

namespace snn {

inline HostNet build_host_net_serial(const NetDesc& d) {
    HostNet h;
    if (d.n_neurons == 0) throw std::invalid_argument("network has no neurons");
    if (d.n_neurons >= 0xFFFFFFFFull) throw std::invalid_argument("more than 2^32-2 neurons");
    if (d.n_synapses >= 0xFFFFFFFFull) throw std::invalid_argument("more than 2^32-2 synapses on one device");
    h.n = (uint32_t)d.n_neurons;
    h.e = d.n_synapses;
    const uint32_t n = h.n;
    const uint64_t e = h.e;

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

    uint32_t lmax = 1;
    uint64_t nve = 0;
    bool identity = true;
    uint32_t exp_all = 0;
    bool exp_uniform = true;
    for (uint64_t j = 0; j < e; ++j) {
        if (d.source[j] >= n || d.target[j] >= n) throw std::invalid_argument("synapse endpoint out of range");
        if (d.compartment[j] > 3 || d.receptor[j] > 3) throw std::invalid_argument("bad compartment or receptor");
        const uint32_t* c;
        const uint32_t nc = codes_of(j, c);
        if (nc > kCodeMax) throw std::invalid_argument("more than 8 spike-code offsets on a synapse");
        if (nc != 1) identity = false;
        const uint32_t ex = nc > 2 ? nc : 2;
        if (j == 0) exp_all = ex;
        else if (ex != exp_all) exp_uniform = false;
        for (uint32_t k = 0; k < nc; ++k) {
            const uint32_t L = lag_of(d.delay[j], c[k]);
            if (L > kMaxLag) throw std::invalid_argument("delay + code offset above 65535 ticks");
            lmax = std::max(lmax, L);
        }
        nve += nc;
        if (d.plastic[j]) h.any_plastic = true;
    }
    h.lmax = lmax;
    h.nve = nve;
    h.ve_identity = identity;
    h.expected_all = exp_uniform ? (e ? exp_all : 2) : 0;
    if ((uint64_t)n * (lmax + 1) > (1ull << 31))
        throw std::invalid_argument("neurons x (max lag + 1) exceeds the group-table limit");

    const uint32_t W = lmax + 1;
    // slots: counting sort by key (source, lag of code 0), stable in connectome order
    std::vector<uint64_t> cnt((uint64_t)n * W + 1, 0);
    std::vector<uint32_t> lag0(e);
    for (uint64_t j = 0; j < e; ++j) {
        const uint32_t* c;
        codes_of(j, c);
        lag0[j] = lag_of(d.delay[j], c[0]);
        cnt[(uint64_t)d.source[j] * W + lag0[j] + 1]++;
    }
    for (uint64_t k = 1; k < cnt.size(); ++k) cnt[k] += cnt[k - 1];
    h.slot_of.resize(e);
    h.tgt.resize(e);
    h.w.resize(e);
    h.meta.resize(e);
    h.delay.resize(e);
    if (!exp_uniform) h.expected.resize(e);
    {
        std::vector<uint64_t> fill(cnt.begin(), cnt.end() - 1);
        for (uint64_t j = 0; j < e; ++j) {
            const uint64_t s = fill[(uint64_t)d.source[j] * W + lag0[j]]++;
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
        }
    }
    if (h.any_plastic) {
        for (uint64_t j = 0; j < e; ++j)
            if (d.plastic[j] && d.delay[j] > 65535) throw std::invalid_argument("plastic synapse delay above 65535");
    }

    // virtual edges and group table
    h.row.assign((uint64_t)n + 1, 0);
    h.grp.assign((uint64_t)n * W, 0);
    if (identity) {
        for (uint32_t s = 0; s < n; ++s) {
            const uint64_t base = cnt[(uint64_t)s * W];
            h.row[s] = base;
            for (uint32_t L = 1; L <= lmax; ++L) h.grp[(uint64_t)s * W + L - 1] = (uint32_t)(cnt[(uint64_t)s * W + L] - base);
            h.grp[(uint64_t)s * W + lmax] = (uint32_t)(cnt[(uint64_t)(s + 1) * W] - base);
        }
        h.row[n] = e;
    } else {
        std::vector<uint64_t> vc((uint64_t)n * W + 1, 0);
        for (uint64_t j = 0; j < e; ++j) {
            const uint32_t* c;
            const uint32_t nc = codes_of(j, c);
            for (uint32_t k = 0; k < nc; ++k) vc[(uint64_t)d.source[j] * W + lag_of(d.delay[j], c[k]) + 1]++;
        }
        for (uint64_t k = 1; k < vc.size(); ++k) vc[k] += vc[k - 1];
        h.ve.resize(nve);
        std::vector<uint64_t> fill(vc.begin(), vc.end() - 1);
        for (uint64_t j = 0; j < e; ++j) {
            const uint32_t* c;
            const uint32_t nc = codes_of(j, c);
            for (uint32_t k = 0; k < nc; ++k) h.ve[fill[(uint64_t)d.source[j] * W + lag_of(d.delay[j], c[k])]++] = (uint32_t)h.slot_of[j];
        }
        for (uint32_t s = 0; s < n; ++s) {
            const uint64_t base = vc[(uint64_t)s * W];
            h.row[s] = base;
            for (uint32_t L = 1; L <= lmax; ++L) h.grp[(uint64_t)s * W + L - 1] = (uint32_t)(vc[(uint64_t)s * W + L] - base);
            const uint64_t len = vc[(uint64_t)(s + 1) * W] - base;
            if (len > 0xFFFFFFFFull) throw std::invalid_argument("a neuron's fan-out x codes exceeds 2^32");
            h.grp[(uint64_t)s * W + lmax] = (uint32_t)len;
        }
        h.row[n] = nve;
    }

    // compact code record only when every target neuron keeps at most one learned pattern
    bool all_one = true;
    for (uint64_t j = 0; j < e && all_one; ++j) all_one = h.params[d.target[j]].max_refs == 1;
    h.code_compact = exp_uniform && exp_all == 2 && all_one;

    // STDP records, target-major in connectome order
    h.tpos.assign(e, kNone);
    if (h.any_plastic) {
        h.in_ptr.assign((uint64_t)n + 1, 0);
        for (uint64_t j = 0; j < e; ++j)
            if (d.plastic[j]) h.in_ptr[d.target[j] + 1]++;
        for (uint32_t i = 0; i < n; ++i) h.in_ptr[i + 1] += h.in_ptr[i];
        std::vector<uint64_t> fill(h.in_ptr.begin(), h.in_ptr.end() - 1);
        h.stdp.resize(h.in_ptr[n]);
        h.tpos_conn.resize(h.in_ptr[n]);
        for (uint64_t j = 0; j < e; ++j) {
            if (!d.plastic[j]) continue;
            const uint64_t t = fill[d.target[j]]++;
            h.tpos[h.slot_of[j]] = (uint32_t)t;
            h.tpos_conn[t] = (uint32_t)j;
            // StdpRule::apply: maximum = std::min(config max 2.0f, synapse max_weight)
            StdpRec r{};
            r.weight = d.weight[j];
            r.last_pre = kNone;
            r.last_post = kNone;
            r.aux = 0;
            r.hi = (d.max_weight[j] < 2.0f) ? d.max_weight[j] : 2.0f;
            h.stdp[t] = r;
        }
    }
    return h;
}

}  // namespace snn
