// Driver that runs the reference CPU step (NetworkPropagator) on a SNNNET01 file and writes a
// SNNOUT01 result file. Test infrastructure only: it is compiled by oracle/Makefile against the
// reference's own runtime sources (never copied here) into oracle/_ref/, and used to generate the
// golden fixtures under tests/golden/ and to check oracle/snn_oracle.c.
//
// Usage: ref_harness <net.bin> <out.bin>

#include "snncuda/declarative/Connectome.h"
#include "snncuda/runtime/NetworkPropagator.h"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "net_connectome.hpp"

namespace {

struct Writer {
    FILE* f;
    template <typename T>
    void vec(const std::vector<T>& v) {
        if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    }
    void u64(std::uint64_t x) { std::fwrite(&x, 8, 1, f); }
};

using snncuda::declarative::Connectome;

struct Blob {
    std::vector<std::uint32_t> words;
    void seq(const std::vector<std::uint32_t>& s) {
        words.push_back(static_cast<std::uint32_t>(s.size()));
        words.insert(words.end(), s.begin(), s.end());
    }
};

} // namespace

int main(int argc, char** argv) {
    if (argc != 3 && !(argc == 4 && std::strcmp(argv[3], "--time") == 0)) {
        std::fprintf(stderr, "usage: %s net.bin out.bin [--time]\n", argv[0]);
        return 2;
    }
    const bool time_only = argc == 4;
    try {
        const netfile::Net t = netfile::read(argv[1]);
        const auto n = t.n;
        const auto e = t.e;
        const auto& cmds = t.cmds;
        const Connectome c = netfile::connectome(t);

        snncuda::runtime::MemoryNeuronStateStore store;
        snncuda::runtime::NetworkPropagator prop(c, store, n ? n : 1, 4096);

        std::vector<std::uint64_t> delivered_per_tick;
        std::vector<std::uint64_t> fired_per_tick;
        std::vector<std::uint32_t> fire_records;  // (tick, neuron, count, 0)
        std::vector<std::uint64_t> last_spikes(n, 0);
        std::uint64_t next_tick = 0;
        double run_seconds = 0.0;

        for (const auto& cmd : cmds) {
            if (cmd.kind == SNN_CMD_INJECT) {
                prop.inject(cmd.neuron + 1, cmd.tick, cmd.weight);
                continue;
            }
            if (time_only) {
                // One run_until per command, exactly as a caller would drive it; no per-tick probes.
                const auto t0 = std::chrono::steady_clock::now();
                prop.run_until(cmd.tick);
                run_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                next_tick = cmd.tick + 1;
                continue;
            }
            for (; next_tick <= cmd.tick; ++next_tick) {
                const auto d0 = prop.delivered_spike_count();
                const auto f0 = prop.fired_spike_count();
                const auto t0 = std::chrono::steady_clock::now();
                prop.run_until(next_tick);
                run_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                delivered_per_tick.push_back(prop.delivered_spike_count() - d0);
                const auto df = prop.fired_spike_count() - f0;
                fired_per_tick.push_back(df);
                if (df == 0) continue;
                std::uint64_t seen = 0;
                for (std::uint64_t i = 0; i < n && seen < df; ++i) {
                    const auto sc = prop.spike_count(i + 1);
                    if (sc != last_spikes[i]) {
                        fire_records.push_back(static_cast<std::uint32_t>(next_tick));
                        fire_records.push_back(static_cast<std::uint32_t>(i));
                        fire_records.push_back(static_cast<std::uint32_t>(sc - last_spikes[i]));
                        fire_records.push_back(0);
                        seen += sc - last_spikes[i];
                        last_spikes[i] = sc;
                    }
                }
            }
        }

        if (time_only) {
            std::fprintf(stdout, "ref_harness_time delivered=%llu fired=%llu run_seconds=%.6f\n",
                static_cast<unsigned long long>(prop.delivered_spike_count()),
                static_cast<unsigned long long>(prop.fired_spike_count()), run_seconds);
            return 0;
        }

        std::vector<float> membrane(n), cur(4 * n), tau(4 * n);
        std::vector<std::uint64_t> lastup(4 * n), last_active(n), spikes(n), wstart(n), noff(n), tmatch(n), tlearn(n);
        std::vector<std::int64_t> last_fired(n);
        std::vector<std::uint8_t> tlast(n);
        Blob nblob;
        for (std::uint64_t i = 0; i < n; ++i) {
            const auto st = prop.state(i + 1);
            membrane[i] = st.membrane_potential;
            const snncuda::runtime::NeuronState::CompartmentState* cs[4] = {&st.soma, &st.basal, &st.apical, &st.inhibitory};
            for (int k = 0; k < 4; ++k) {
                cur[k * n + i] = cs[k]->current;
                tau[k * n + i] = cs[k]->decay_tau_ticks;
                lastup[k * n + i] = cs[k]->last_update_tick;
            }
            last_active[i] = st.last_active_tick;
            last_fired[i] = st.last_fired_tick ? static_cast<std::int64_t>(*st.last_fired_tick) : -1;
            spikes[i] = st.spike_count;
            wstart[i] = st.temporal_pattern.window_start_tick;
            noff[i] = st.temporal_pattern.spike_offsets.size();
            tmatch[i] = st.temporal_pattern.match_count;
            tlearn[i] = st.temporal_pattern.learned_patterns.size();
            tlast[i] = st.temporal_pattern.last_match ? 1 : 0;
            nblob.seq(st.temporal_pattern.spike_offsets);
            nblob.words.push_back(static_cast<std::uint32_t>(st.temporal_pattern.learned_patterns.size()));
            for (const auto& p : st.temporal_pattern.learned_patterns) nblob.seq(p.spike_offsets);
        }

        std::vector<float> sw(e), sdelta(e);
        std::vector<std::int64_t> lpre(e), lpost(e);
        std::vector<std::uint64_t> prec(e), updc(e), cws(e), cno(e), cmatch(e), clearn(e);
        std::vector<std::uint8_t> clast(e);
        Blob sblob;
        for (std::uint64_t j = 0; j < e; ++j) {
            const auto* s = prop.synapse_state(j + 1);
            sw[j] = s->weight;
            sdelta[j] = s->last_weight_delta;
            lpre[j] = s->last_pre_spike_tick ? static_cast<std::int64_t>(*s->last_pre_spike_tick) : -1;
            lpost[j] = s->last_post_spike_tick ? static_cast<std::int64_t>(*s->last_post_spike_tick) : -1;
            prec[j] = s->pre_spike_count;
            updc[j] = s->plasticity_update_count;
            cws[j] = s->code_window_start_tick;
            cno[j] = s->code_offsets.size();
            cmatch[j] = s->code_match_count;
            clearn[j] = s->learned_code_patterns.size();
            clast[j] = s->last_code_match ? 1 : 0;
            sblob.seq(s->code_offsets);
            sblob.words.push_back(static_cast<std::uint32_t>(s->learned_code_patterns.size()));
            for (const auto& p : s->learned_code_patterns) sblob.seq(p.spike_offsets);
        }

        FILE* out = std::fopen(argv[2], "wb");
        if (!out) throw std::runtime_error("cannot open output");
        std::fwrite(SNN_OUT_MAGIC, 1, 8, out);
        Writer wr{out};
        wr.u64(n);
        wr.u64(e);
        wr.u64(delivered_per_tick.size());
        wr.u64(fire_records.size() / 4);
        wr.vec(delivered_per_tick);
        wr.vec(fired_per_tick);
        wr.vec(fire_records);
        wr.vec(membrane);
        wr.vec(cur);
        wr.vec(tau);
        wr.vec(lastup);
        wr.vec(last_active);
        wr.vec(last_fired);
        wr.vec(spikes);
        wr.vec(wstart);
        wr.vec(noff);
        wr.vec(tmatch);
        wr.vec(tlearn);
        wr.vec(tlast);
        wr.vec(sw);
        wr.vec(sdelta);
        wr.vec(lpre);
        wr.vec(lpost);
        wr.vec(prec);
        wr.vec(updc);
        wr.vec(cws);
        wr.vec(cno);
        wr.vec(cmatch);
        wr.vec(clearn);
        wr.vec(clast);
        wr.u64(nblob.words.size());
        wr.vec(nblob.words);
        wr.u64(sblob.words.size());
        wr.vec(sblob.words);
        wr.u64(prop.delivered_spike_count());
        wr.u64(prop.fired_spike_count());
        std::fclose(out);
        std::fprintf(stdout, "ref_harness ticks=%zu delivered=%llu fired=%llu run_seconds=%.6f\n",
            delivered_per_tick.size(),
            static_cast<unsigned long long>(prop.delivered_spike_count()),
            static_cast<unsigned long long>(prop.fired_spike_count()), run_seconds);
    } catch (const std::exception& err) {
        std::fprintf(stderr, "ref_harness failed: %s\n", err.what());
        return 1;
    }
    return 0;
}
