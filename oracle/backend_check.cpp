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
//   backend_check session <net.bin> <max_steps> <samples>
//       CudaInferenceSession::run_batch: the reference's session tests restated on the LTP pair, then
//       samples built from the file's injects against fresh CPU NetworkPropagator runs (stateless)
//       and against one CPU propagator with a continuing clock (stateful).
//   backend_check pack <net.bin>
//       Host-only: snncuda_hip::pack of the Connectome must reproduce the file's arrays exactly
//       (ids mapped back to indices, empty code lists kept empty), and a synapse naming an unknown
//       neuron id must raise snncuda_hip::Error with SNN_EINVAL.

#include "snncuda/backends/CudaBackend.h"
#include "snncuda/runtime/NetworkPropagator.h"
#include "snncuda_hip.hpp"

#include <chrono>
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
    // the automatic thread count and forced ones (several threads over a few synapses each)
    for (const unsigned threads : {0u, 1u, 3u, 16u}) {
        const auto p = snncuda_hip::pack(c, threads);
        bool ok = same(p.threshold, t.thr) && same(p.window, t.win) && same(p.sim_thr, t.sim) &&
                  same(p.max_refs, t.mref) && same(p.source, t.src) && same(p.target, t.tgt) &&
                  same(p.weight, t.w) && same(p.max_weight, t.mw) && same(p.delay, t.dl) &&
                  same(p.compartment, t.comp) && same(p.receptor, t.rec) && same(p.plastic, t.pl) &&
                  same(p.code_ptr, t.cptr) && same(p.codes, t.codes);
        const auto d = p.desc();
        ok = ok && d.n_neurons == t.n && d.n_synapses == t.e;
        if (!ok) {
            std::fprintf(stderr, "pack on %u threads differs from the file's arrays\n", threads);
            return 1;
        }
    }
    if (t.e > 0) {
        // the first unknown id in connectome order is the one reported, at any thread count
        auto bad = c;
        const std::size_t first = t.e / 2;
        bad.synapses[first].target = 0;  // invalid_id: no neuron carries it
        bad.synapses[t.e - 1].source = 0;
        const std::string want = "synapse " + std::to_string(first) + " names";
        for (const unsigned threads : {1u, 3u, 16u}) {
            try {
                (void)snncuda_hip::pack(bad, threads);
                std::fprintf(stderr, "pack accepted a synapse with an unknown target\n");
                return 1;
            } catch (const snncuda_hip::Error& err) {
                if (err.code() != SNN_EINVAL || std::string(err.what()).find(want) == std::string::npos) {
                    std::fprintf(stderr, "pack raised code %d (%s), expected SNN_EINVAL for %s\n", err.code(),
                                 err.what(), want.c_str());
                    return 1;
                }
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

// The LTP pair of the reference's CUDA tests (tests/cuda_synapse_dendrite_tests.cpp:20-80): input
// (threshold 1.0) -> target (threshold 0.4), weight 0.5, delay 1, Basal, AMPA, plastic.
snncuda::declarative::Connectome ltp_pair() {
    snncuda::declarative::Connectome c;
    c.neurons.resize(2);
    c.neurons[0].id = 1;
    c.neurons[0].params.threshold = 1.0F;
    c.neurons[1].id = 2;
    c.neurons[1].params.threshold = 0.4F;
    c.synapses.resize(1);
    c.synapses[0].id = 1;
    c.synapses[0].source = 1;
    c.synapses[0].target = 2;
    c.synapses[0].weight = 0.5F;
    return c;
}

void require(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "failed: %s\n", what);
        ++mismatches;
    }
}

// The reference's session tests restated (cuda_synapse_dendrite_tests.cpp:238-303,421-463), then
// every sample of a batch built from a network file against a fresh CPU NetworkPropagator run, and
// the stateful mode against one CPU propagator whose clock continues across samples.
int cmd_session(const netfile::Net& t, std::uint32_t max_steps, std::uint32_t n_samples) {
    using namespace snncuda::backends;
    const CudaBackend backend;
    if (!backend.available()) {
        std::printf("backend_check: %s\n", backend.availability_status().c_str());
        return 3;
    }
    {
        const auto c = ltp_pair();
        const auto first = backend.run_resident_propagation(c, {1}, 1);
        const auto second = backend.run_resident_propagation(c, {1}, 1);
        CudaInferenceSession session(c);
        CudaInferenceOptions options;
        options.max_steps = 1;
        options.readout_neurons = {2};
        const auto batch = session.run_batch({{.inputs = {{.neuron = 1}}}, {.inputs = {{.neuron = 1}}}, {.inputs = {}}},
                                             options);
        require(batch.executed && batch.samples.size() == 3, "session returns every sample");
        require(batch.samples[0].delivered_spikes == first.delivered_spikes, "sample 0 delivered = single run");
        require(batch.samples[0].fired_spikes == first.fired_spikes, "sample 0 fired = single run");
        require(batch.samples[0].readout_spike_counts.size() == 1 &&
                    batch.samples[0].readout_spike_counts[0] == first.final_neuron_spike_counts[1],
                "sample 0 readout = single run");
        require(batch.samples[1].delivered_spikes == second.delivered_spikes &&
                    batch.samples[1].readout_spike_counts[0] == second.final_neuron_spike_counts[1],
                "sample 1 = single run");
        require(batch.samples[2].delivered_spikes == 0 && batch.samples[2].fired_spikes == 0 &&
                    batch.samples[2].readout_spike_counts[0] == 0,
                "empty sample is empty");
        const auto weighted = session.run_batch(
            {{.inputs = {{.neuron = 1, .weight = 0.5F}}},
             {.inputs = {{.neuron = 1, .weight = 1.0F}}},
             {.inputs = {{.neuron = 1, .weight = 0.4F}, {.neuron = 1, .weight = 0.6F}}}},
            options);
        require(weighted.samples[0].delivered_spikes == 1 && weighted.samples[0].fired_spikes == 0 &&
                    weighted.samples[0].readout_spike_counts[0] == 0,
                "subthreshold weighted input");
        require(weighted.samples[1].fired_spikes == 2 && weighted.samples[1].readout_spike_counts[0] == 1,
                "threshold weighted input fires input and readout");
        require(weighted.samples[2].fired_spikes == 2 && weighted.samples[2].readout_spike_counts[0] == 1,
                "accumulated weighted inputs fire input and readout");
        std::vector<CudaInferenceSample> many(100, CudaInferenceSample{.inputs = {{.neuron = 1}}});
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < 100; ++k) require(backend.run_resident_propagation(c, {1}, 1).executed, "repeated run");
        const double repeated = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const auto persistent = session.run_batch(many, options);
        require(persistent.samples.size() == 100 && persistent.elapsed_seconds > 0.0, "100-sample batch");
        require(persistent.elapsed_seconds < repeated, "one batch beats 100 separate runs");
        std::printf("session: 100 separate runs %.4f s, one batch %.4f s\n", repeated, persistent.elapsed_seconds);
    }
    // samples from the file's injects: inject q goes to sample q % n_samples (ticks <= max_steps)
    const auto c = netfile::connectome(t);
    std::vector<CudaInferenceSample> samples(n_samples);
    for (std::size_t q = 0; q < t.cmds.size(); ++q) {
        const auto& cmd = t.cmds[q];
        if (cmd.kind != SNN_CMD_INJECT) continue;
        samples[q % n_samples].inputs.push_back(
            {.neuron = cmd.neuron + 1ull, .weight = cmd.weight, .tick = static_cast<std::uint32_t>(cmd.tick)});
    }
    CudaInferenceOptions options;
    options.max_steps = max_steps;
    for (std::uint64_t i = 0; i < t.n; ++i) options.readout_neurons.push_back(i + 1);
    CudaInferenceSession session(c);
    const auto batch = session.run_batch(samples, options);
    for (std::uint32_t s = 0; s < n_samples; ++s) {
        snncuda::runtime::MemoryNeuronStateStore store;
        snncuda::runtime::NetworkPropagator cpu(c, store, t.n ? t.n : 1, 4096);
        for (const auto& in : samples[s].inputs)
            if (in.tick <= max_steps) cpu.inject(in.neuron, in.tick, in.weight);
        cpu.run_until(max_steps);
        const auto& r = batch.samples[s];
        expect_eq("session delivered", s, r.delivered_spikes, cpu.delivered_spike_count());
        expect_eq("session fired", s, r.fired_spikes, cpu.fired_spike_count());
        for (std::uint64_t i = 0; i < t.n; ++i) expect_eq("session readout", i, r.readout_spike_counts[i], cpu.spike_count(i + 1));
    }
    // stateful: one propagator, the clock continues across samples
    options.reset_state_between_samples = false;
    const auto seq = session.run_batch(samples, options);
    {
        snncuda::runtime::MemoryNeuronStateStore store;
        snncuda::runtime::NetworkPropagator cpu(c, store, t.n ? t.n : 1, 4096);
        std::uint64_t base = 0;
        std::vector<std::uint64_t> before(t.n);
        for (std::uint32_t s = 0; s < n_samples; ++s) {
            const auto d0 = cpu.delivered_spike_count(), f0 = cpu.fired_spike_count();
            for (std::uint64_t i = 0; i < t.n; ++i) before[i] = cpu.spike_count(i + 1);
            for (const auto& in : samples[s].inputs)
                if (in.tick <= max_steps) cpu.inject(in.neuron, base + in.tick, in.weight);
            cpu.run_until(base + max_steps);
            const auto& r = seq.samples[s];
            expect_eq("stateful delivered", s, r.delivered_spikes, cpu.delivered_spike_count() - d0);
            expect_eq("stateful fired", s, r.fired_spikes, cpu.fired_spike_count() - f0);
            for (std::uint64_t i = 0; i < t.n; ++i)
                expect_eq("stateful readout", i, r.readout_spike_counts[i], cpu.spike_count(i + 1) - before[i]);
            base += max_steps + 1ull;
        }
    }
    std::printf("backend_check session %s: %u samples, max_steps=%u, mismatches=%d\n", mismatches ? "FAILED" : "ok",
                n_samples, max_steps, mismatches);
    return mismatches ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc == 3 && std::strcmp(argv[1], "pack") == 0) return cmd_pack(netfile::read(argv[2]));
        if (argc == 4 && std::strcmp(argv[1], "run") == 0) {
            return cmd_run(netfile::read(argv[2]), static_cast<std::uint32_t>(std::stoul(argv[3])));
        }
        if (argc == 5 && std::strcmp(argv[1], "session") == 0) {
            return cmd_session(netfile::read(argv[2]), static_cast<std::uint32_t>(std::stoul(argv[3])),
                               static_cast<std::uint32_t>(std::stoul(argv[4])));
        }
    } catch (const std::exception& err) {
        std::fprintf(stderr, "backend_check failed: %s\n", err.what());
        return 1;
    }
    std::fprintf(stderr, "usage: %s run <net.bin> <max_steps> | pack <net.bin> | session <net.bin> <max_steps> <samples>\n",
                 argv[0]);
    return 2;
}
