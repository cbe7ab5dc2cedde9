// Checks the drop-in CudaBackend (snncuda_amd/cpp/CudaBackendHip.cpp, over libsnncuda_hip.so)
// against the reference CPU NetworkPropagator on one SNNNET01 network. Test infrastructure only:
// compiled by oracle/Makefile against the reference's headers and runtime sources into oracle/_ref/.
//
//   backend_check run <net.bin> <max_steps>
//       initially_active = the neurons of the file's inject commands at tick 0, in file order.
//       CPU side, as the reference's own CUDA tests drive it (reference
//       tests/cuda_synapse_dendrite_tests.cpp:122-130): inject(id, 0, 1.0) per distinct id, then
//       run_until(max_steps). Device side: CudaBackend::run_resident_propagation(connectome,
//       initially_active, max_steps). Every CudaPropagationResult vector must equal the CPU state
//       bit for bit, and the counters must agree. Exit 0 on agreement, 1 on a mismatch, 3 without a
//       device.
//   backend_check pack <net.bin>
//       Host-only: snncuda_hip::pack of the Connectome must reproduce the file's arrays exactly
//       (ids mapped back to indices, empty code lists kept empty), and a synapse naming an unknown
//       neuron id must raise snncuda_hip::Error with SNN_EINVAL.

#include "snncuda/backends/CudaBackend.h"
#include "snncuda/runtime/NetworkPropagator.h"
#include "snncuda_hip.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "net_connectome.hpp"

namespace {

int mismatches = 0;

template <typename A, typename B>
void expect_eq(const char* what, std::size_t i, const A& got, const B& want) {
    if (!(got == want)) {
        if (mismatches < 20) {
            std::fprintf(stderr, "mismatch %s[%zu]: device %s, cpu %s\n", what, i, std::to_string(got).c_str(),
                         std::to_string(want).c_str());
        }
        ++mismatches;
    }
}

void expect_bits(const char* what, std::size_t i, float got, float want) {
    std::uint32_t a, b;
    std::memcpy(&a, &got, 4);
    std::memcpy(&b, &want, 4);
    if (a != b) {
        if (mismatches < 20) std::fprintf(stderr, "mismatch %s[%zu]: device %.9g, cpu %.9g\n", what, i, got, want);
        ++mismatches;
    }
}

template <typename T>
bool same(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

int cmd_pack(const netfile::Net& t) {
    const auto c = netfile::connectome(t);
    const auto p = snncuda_hip::pack(c);
    bool ok = same(p.threshold, t.thr) && same(p.window, t.win) && same(p.sim_thr, t.sim) &&
              same(p.max_refs, t.mref) && same(p.source, t.src) && same(p.target, t.tgt) && same(p.weight, t.w) &&
              same(p.max_weight, t.mw) && same(p.delay, t.dl) && same(p.compartment, t.comp) &&
              same(p.receptor, t.rec) && same(p.plastic, t.pl) && same(p.code_ptr, t.cptr) &&
              same(p.codes, t.codes);
    const auto d = p.desc();
    ok = ok && d.n_neurons == t.n && d.n_synapses == t.e;
    if (!ok) {
        std::fprintf(stderr, "pack differs from the file's arrays\n");
        return 1;
    }
    if (t.e > 0) {
        auto bad = c;
        bad.synapses[0].target = 0;  // invalid_id: no neuron carries it
        try {
            (void)snncuda_hip::pack(bad);
            std::fprintf(stderr, "pack accepted a synapse with an unknown target\n");
            return 1;
        } catch (const snncuda_hip::Error& err) {
            if (err.code() != SNN_EINVAL) {
                std::fprintf(stderr, "pack raised code %d, expected SNN_EINVAL\n", err.code());
                return 1;
            }
        }
    }
    std::printf("backend_check pack ok: %llu neurons, %llu synapses\n", static_cast<unsigned long long>(t.n),
                static_cast<unsigned long long>(t.e));
    return 0;
}

int cmd_run(const netfile::Net& t, std::uint32_t max_steps) {
    const snncuda::backends::CudaBackend backend;
    if (!backend.available()) {
        std::printf("backend_check: %s\n", backend.availability_status().c_str());
        return 3;
    }
    const auto c = netfile::connectome(t);
    std::vector<snncuda::core::NeuronId> active;
    for (const auto& cmd : t.cmds) {
        if (cmd.kind == SNN_CMD_INJECT && cmd.tick == 0) active.push_back(cmd.neuron + 1);
    }

    snncuda::runtime::MemoryNeuronStateStore store;
    snncuda::runtime::NetworkPropagator cpu(c, store, t.n ? t.n : 1, 4096);
    std::unordered_set<snncuda::core::NeuronId> seen;
    std::uint64_t n_inject = 0;
    for (const auto id : active) {
        if (seen.insert(id).second) {
            cpu.inject(id, 0, 1.0F);
            ++n_inject;
        }
    }
    cpu.run_until(max_steps);

    const auto gpu = backend.run_resident_propagation(c, active, max_steps);
    expect_eq("executed", 0, gpu.executed, true);
    expect_eq("delivered", 0, gpu.delivered_spikes, cpu.delivered_spike_count());
    expect_eq("fired", 0, gpu.fired_spikes, cpu.fired_spike_count());
    if (active.empty()) {
        std::printf("backend_check run: empty active set\n");
        return mismatches ? 1 : 0;
    }
    expect_eq("n_weights", 0, gpu.final_synapse_weights.size(), t.e);
    expect_eq("n_membrane", 0, gpu.final_membrane_potentials.size(), t.n);
    if (mismatches) return 1;

    std::uint64_t t_match = 0, t_learned = 0, updates = 0;
    for (std::uint64_t i = 0; i < t.n; ++i) {
        const auto st = cpu.state(i + 1);
        expect_bits("membrane", i, gpu.final_membrane_potentials[i], st.membrane_potential);
        expect_eq("spike_count", i, gpu.final_neuron_spike_counts[i], st.spike_count);
        expect_eq("temporal_match", i, gpu.final_temporal_match_counts[i], st.temporal_pattern.match_count);
        expect_eq("temporal_learned", i, static_cast<std::uint64_t>(gpu.final_temporal_learned_pattern_counts[i]),
                  static_cast<std::uint64_t>(st.temporal_pattern.learned_patterns.size()));
        t_match += st.temporal_pattern.match_count;
        t_learned += st.temporal_pattern.learned_patterns.size();
    }
    for (std::uint64_t j = 0; j < t.e; ++j) {
        const auto* s = cpu.synapse_state(j + 1);
        expect_bits("weight", j, gpu.final_synapse_weights[j], s->weight);
        expect_eq("code_match", j, gpu.final_synapse_code_match_counts[j], s->code_match_count);
        expect_eq("code_learned", j, static_cast<std::uint64_t>(gpu.final_synapse_code_learned_pattern_counts[j]),
                  static_cast<std::uint64_t>(s->learned_code_patterns.size()));
        updates += s->plasticity_update_count;
    }
    const auto synaptic = cpu.delivered_spike_count() - n_inject;
    expect_eq("debug.ticks_processed", 0, gpu.debug.ticks_processed, static_cast<std::uint64_t>(max_steps) + 1);
    expect_eq("debug.scheduled_events", 0, gpu.debug.scheduled_events, synaptic);
    expect_eq("debug.processed_events", 0, gpu.debug.processed_events, synaptic);
    expect_eq("debug.fired_appends", 0, gpu.debug.fired_appends, cpu.fired_spike_count());
    expect_eq("debug.stdp_updates", 0, gpu.debug.stdp_updates, updates);
    expect_eq("debug.stdp_ltp+ltd", 0, gpu.debug.stdp_ltp + gpu.debug.stdp_ltd, updates);
    expect_eq("debug.temporal_matches", 0, gpu.debug.temporal_matches, t_match);
    expect_eq("debug.temporal_patterns_learned", 0, gpu.debug.temporal_patterns_learned, t_learned);
    expect_eq("debug.dropped_events", 0, gpu.debug.dropped_events, std::uint64_t{0});
    if (gpu.debug.scheduled_event_requests < synaptic) {
        std::fprintf(stderr, "scheduled_event_requests %llu below delivered synaptic events %llu\n",
                     static_cast<unsigned long long>(gpu.debug.scheduled_event_requests),
                     static_cast<unsigned long long>(synaptic));
        ++mismatches;
    }
    std::printf("backend_check run %s: max_steps=%u active=%zu delivered=%llu fired=%llu stdp_ltp=%llu stdp_ltd=%llu "
                "requests=%llu mismatches=%d\n",
                mismatches ? "FAILED" : "ok", max_steps, active.size(),
                static_cast<unsigned long long>(gpu.delivered_spikes), static_cast<unsigned long long>(gpu.fired_spikes),
                static_cast<unsigned long long>(gpu.debug.stdp_ltp), static_cast<unsigned long long>(gpu.debug.stdp_ltd),
                static_cast<unsigned long long>(gpu.debug.scheduled_event_requests), mismatches);
    return mismatches ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc == 3 && std::strcmp(argv[1], "pack") == 0) return cmd_pack(netfile::read(argv[2]));
        if (argc == 4 && std::strcmp(argv[1], "run") == 0) {
            return cmd_run(netfile::read(argv[2]), static_cast<std::uint32_t>(std::stoul(argv[3])));
        }
    } catch (const std::exception& err) {
        std::fprintf(stderr, "backend_check failed: %s\n", err.what());
        return 1;
    }
    std::fprintf(stderr, "usage: %s run <net.bin> <max_steps> | pack <net.bin>\n", argv[0]);
    return 2;
}
