// C++ layer over the C ABI in snncuda_hip.h: an RAII handle that rethrows status codes as
// exceptions, a packer from any Connectome-shaped type to snn_net_desc, and the body of
// CudaBackend::run_resident_propagation with the reference CPU semantics.
//
// Header-only and free of the reference's headers: the packer and the propagation function are
// templates over the reference types (reference include/snncuda/declarative/Connectome.h:13-37,
// include/snncuda/backends/CudaBackend.h:14-53), so the same code compiles inside the reference
// tree (snncuda_amd/cpp/CudaBackendHip.cpp) and anywhere else with structs of the same shape.
#ifndef SNNCUDA_HIP_HPP
#define SNNCUDA_HIP_HPP

#include "snncuda_hip.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

namespace snncuda_hip {

// Status codes from the C ABI become exceptions here, as the reference's check_cuda raises
// std::runtime_error (reference src/backends/cuda/CudaBackend.cu:47-51).
class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    [[nodiscard]] int code() const noexcept { return code_; }

private:
    int code_;
};

inline void check(int rc, const char* call) {
    if (rc != SNN_OK) {
        throw Error(rc, std::string(call) + ": " + snn_last_error());
    }
}

// One device-resident network (move-only). The state persists across run_until calls, like
// runtime::NetworkPropagator (reference include/snncuda/runtime/NetworkPropagator.h:14-47).
class Network {
public:
    Network() = default;
    explicit Network(const snn_net_desc& desc, const snn_options* opts = nullptr) {
        check(snn_net_create(&desc, opts, &handle_), "snn_net_create");
    }
    ~Network() { reset(); }
    Network(const Network&) = delete;
    Network& operator=(const Network&) = delete;
    Network(Network&& o) noexcept : handle_(std::exchange(o.handle_, nullptr)) {}
    Network& operator=(Network&& o) noexcept {
        if (this != &o) {
            reset();
            handle_ = std::exchange(o.handle_, nullptr);
        }
        return *this;
    }

    void inject(std::uint32_t neuron, std::uint64_t tick, float weight) {
        check(snn_net_inject(handle_, &neuron, &tick, &weight, 1), "snn_net_inject");
    }
    void inject(const std::vector<std::uint32_t>& neuron, const std::vector<std::uint64_t>& tick,
                const std::vector<float>& weight) {
        if (neuron.size() != tick.size() || neuron.size() != weight.size()) {
            throw Error(SNN_EINVAL, "inject: neuron, tick and weight differ in length");
        }
        check(snn_net_inject(handle_, neuron.data(), tick.data(), weight.data(), neuron.size()),
              "snn_net_inject");
    }
    void run_until(std::uint64_t inclusive_tick) {
        check(snn_net_run_until(handle_, inclusive_tick), "snn_net_run_until");
    }
    [[nodiscard]] std::uint64_t delivered() const { return counts().delivered; }
    [[nodiscard]] std::uint64_t fired() const { return counts().fired; }

    struct Counts {
        std::uint64_t delivered, fired, next_tick;
    };
    [[nodiscard]] Counts counts() const {
        Counts c{};
        check(snn_net_counts(handle_, &c.delivered, &c.fired, &c.next_tick), "snn_net_counts");
        return c;
    }
    [[nodiscard]] snn_net* get() const noexcept { return handle_; }

private:
    void reset() noexcept {
        if (handle_) {
            snn_net_destroy(handle_);
            handle_ = nullptr;
        }
    }
    snn_net* handle_{nullptr};
};

// A Connectome flattened into the SoA arrays snn_net_desc points into, in connectome order.
// Neuron ids are mapped to dense indices in neuron order; a synapse whose source or target is not
// a neuron of the connectome is rejected (the device layout has no slot for it).
struct PackedConnectome {
    std::vector<float> threshold, sim_thr;
    std::vector<std::uint64_t> window, max_refs;
    std::vector<std::uint32_t> source, target, delay, codes;
    std::vector<float> weight, max_weight;
    std::vector<std::uint8_t> compartment, receptor, plastic;
    std::vector<std::uint64_t> code_ptr;
    std::unordered_map<std::uint64_t, std::uint32_t> index;  // NeuronId -> dense index

    [[nodiscard]] snn_net_desc desc() const {
        snn_net_desc d{};
        d.n_neurons = threshold.size();
        d.threshold = threshold.data();
        d.window = window.data();
        d.sim_thr = sim_thr.data();
        d.max_refs = max_refs.data();
        d.n_synapses = source.size();
        d.source = source.data();
        d.target = target.data();
        d.weight = weight.data();
        d.max_weight = max_weight.data();
        d.delay = delay.data();
        d.compartment = compartment.data();
        d.receptor = receptor.data();
        d.plastic = plastic.data();
        d.code_ptr = code_ptr.data();
        d.codes = codes.data();
        return d;
    }
};

// Connectome: .neurons[i].{id, params.{threshold, pattern_window_ticks, similarity_threshold,
// max_reference_patterns}}, .synapses[j].{source, target, weight, max_weight, delay_ticks,
// spike_code_offsets, compartment, receptor, plasticity_enabled}.
// The synapse passes run on `threads` threads (0: the machine's, at most 16; one below 65,536
// synapses), each over a contiguous range of synapses; the arrays and the error raised (that of the
// first offending synapse) are the same at any thread count.
template <class Connectome>
PackedConnectome pack(const Connectome& c, unsigned threads = 0) {
    PackedConnectome p;
    const std::size_t n = c.neurons.size();
    const std::size_t e = c.synapses.size();
    if (n > 0xFFFFFFFEull) {
        throw Error(SNN_EINVAL, "pack: more than 2^32-1 neurons");
    }
    p.threshold.resize(n);
    p.window.resize(n);
    p.sim_thr.resize(n);
    p.max_refs.resize(n);
    p.index.reserve(n);
    for (std::size_t i = 0; i < n; ++i) {
        const auto& nr = c.neurons[i];
        p.index.emplace(static_cast<std::uint64_t>(nr.id), static_cast<std::uint32_t>(i));
        p.threshold[i] = nr.params.threshold;
        p.window[i] = nr.params.pattern_window_ticks;
        p.sim_thr[i] = nr.params.similarity_threshold;
        p.max_refs[i] = static_cast<std::uint64_t>(nr.params.max_reference_patterns);
    }
    p.source.resize(e);
    p.target.resize(e);
    p.weight.resize(e);
    p.max_weight.resize(e);
    p.delay.resize(e);
    p.compartment.resize(e);
    p.receptor.resize(e);
    p.plastic.resize(e);
    p.code_ptr.resize(e + 1);
    p.code_ptr[0] = 0;
    unsigned T = threads;
    if (T == 0) {
        const unsigned hw = std::thread::hardware_concurrency();
        T = e < (1u << 16) ? 1u : std::max(1u, std::min(hw, 16u));
    }
    auto run = [T](auto&& fn) {
        std::vector<std::thread> th;
        for (unsigned t = 1; t < T; ++t) th.emplace_back([&fn, t] { fn(t); });
        fn(0u);
        for (auto& x : th) x.join();
    };
    // per synapse: dense endpoints and fields, and the code count in code_ptr[j + 1]; the index map
    // is only read here
    std::vector<std::size_t> bad(T, e);
    std::vector<std::uint64_t> ncodes(T + 1, 0);
    run([&](unsigned t) {
        std::uint64_t cnt = 0;
        for (std::size_t j = e * t / T; j < e * (t + 1) / T; ++j) {
            const auto& s = c.synapses[j];
            const auto src = p.index.find(static_cast<std::uint64_t>(s.source));
            const auto tgt = p.index.find(static_cast<std::uint64_t>(s.target));
            if (src == p.index.end() || tgt == p.index.end()) {
                bad[t] = j;
                break;
            }
            p.source[j] = src->second;
            p.target[j] = tgt->second;
            p.weight[j] = s.weight;
            p.max_weight[j] = s.max_weight;
            p.delay[j] = s.delay_ticks;
            p.compartment[j] = static_cast<std::uint8_t>(s.compartment);
            p.receptor[j] = static_cast<std::uint8_t>(s.receptor);
            p.plastic[j] = s.plasticity_enabled ? 1 : 0;
            p.code_ptr[j + 1] = s.spike_code_offsets.size();
            cnt += s.spike_code_offsets.size();
        }
        ncodes[t + 1] = cnt;
    });
    const std::size_t first_bad = *std::min_element(bad.begin(), bad.end());
    if (first_bad < e) {
        throw Error(SNN_EINVAL, "pack: synapse " + std::to_string(first_bad) + " names a neuron id not in the connectome");
    }
    for (unsigned t = 0; t < T; ++t) ncodes[t + 1] += ncodes[t];
    p.codes.resize(ncodes[T]);
    run([&](unsigned t) {
        std::uint64_t at = ncodes[t];
        for (std::size_t j = e * t / T; j < e * (t + 1) / T; ++j) {
            for (const auto off : c.synapses[j].spike_code_offsets) p.codes[at++] = static_cast<std::uint32_t>(off);
            p.code_ptr[j + 1] = at;
        }
    });
    return p;
}

// CudaBackend::run_resident_propagation (reference include/snncuda/backends/CudaBackend.h:134-137)
// with the reference CPU semantics, which the reference's own CUDA tests compare against
// (reference tests/cuda_synapse_dendrite_tests.cpp:122-130): every distinct, known id of
// initially_active receives inject(id, 0, 1.0) in list order, then ticks 0..max_steps run.
// Unknown ids are skipped and duplicates count once, as in the reference (CudaBackend.cu:3480-3491);
// an empty connectome or an empty active set gives executed=true with empty results
// (CudaBackend.cu:3378-3380, 3489-3491).
//
// Result is CudaPropagationResult-shaped (CudaBackend.h:40-53). Its debug block holds the reference
// diagnostic counters that have a meaning under exact CPU semantics, computed from the final state:
// ticks, scheduled requests (emissions from fires before the last tick) and in-window events,
// processed synaptic events, fires, STDP weight changes, events per receptor, dendritic
// integrations, post-spike plasticity scans, temporal observations / matches / learned patterns and
// the per-tick maxima. Lock, drop and overflow counters stay 0: the step takes no locks and drops
// nothing (a capacity overflow throws Error).
template <class Result, class Connectome, class Id>
Result run_resident_propagation(const Connectome& connectome, const std::vector<Id>& initially_active,
                                std::uint32_t max_steps, const snn_options* opts = nullptr) {
    Result result{};
    if (connectome.neurons.empty() || initially_active.empty()) {
        result.executed = true;
        return result;
    }
    const auto start = std::chrono::steady_clock::now();
    const PackedConnectome p = pack(connectome);
    std::vector<std::uint32_t> active;
    {
        std::unordered_set<std::uint64_t> seen;
        for (const auto id : initially_active) {
            const auto it = p.index.find(static_cast<std::uint64_t>(id));
            if (it != p.index.end() && seen.insert(static_cast<std::uint64_t>(id)).second) {
                active.push_back(it->second);
            }
        }
    }
    if (active.empty()) {
        result.executed = true;
        return result;
    }
    const std::size_t n = p.threshold.size();
    const std::size_t e = p.source.size();
    snn_options o{};
    if (opts) {
        o = *opts;
    }
    o.track_pre_count = 1;
    Network net(p.desc(), &o);
    net.inject(active, std::vector<std::uint64_t>(active.size(), 0), std::vector<float>(active.size(), 1.0F));

    // Fires before the last tick emit into the schedule; the reference counts those emissions as
    // scheduled requests, so the spike counts are read once more just before the last tick.
    std::vector<std::uint64_t> spikes_before_last(n, 0);
    if (max_steps > 0) {
        net.run_until(max_steps - 1);
        check(snn_net_read_neurons(net.get(), nullptr, nullptr, nullptr, spikes_before_last.data(), nullptr,
                                   nullptr, nullptr),
              "snn_net_read_neurons");
    }
    net.run_until(max_steps);
    const auto counts = net.counts();

    result.executed = true;
    result.delivered_spikes = counts.delivered;
    result.fired_spikes = counts.fired;
    result.final_synapse_weights.resize(e);
    result.final_membrane_potentials.resize(n);
    result.final_neuron_spike_counts.resize(n);
    std::vector<std::uint64_t> t_match(n), t_learned(n), c_match(e), c_learned(e), pre_count(e), ltp(e), ltd(e);
    check(snn_net_read_neurons(net.get(), result.final_membrane_potentials.data(), nullptr, nullptr,
                               result.final_neuron_spike_counts.data(), nullptr, t_match.data(), t_learned.data()),
          "snn_net_read_neurons");
    check(snn_net_read_synapses(net.get(), result.final_synapse_weights.data(), pre_count.data(), nullptr,
                                nullptr, nullptr, c_match.data(), c_learned.data()),
          "snn_net_read_synapses");
    check(snn_net_read_plasticity(net.get(), ltp.data(), ltd.data()), "snn_net_read_plasticity");
    result.final_temporal_match_counts.assign(t_match.begin(), t_match.end());
    result.final_temporal_learned_pattern_counts.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        result.final_temporal_learned_pattern_counts[i] = static_cast<std::uint32_t>(t_learned[i]);
    }
    result.final_synapse_code_match_counts.assign(c_match.begin(), c_match.end());
    result.final_synapse_code_learned_pattern_counts.resize(e);
    for (std::size_t j = 0; j < e; ++j) {
        result.final_synapse_code_learned_pattern_counts[j] = static_cast<std::uint32_t>(c_learned[j]);
    }

    // Diagnostic counters.
    const std::uint64_t ticks = static_cast<std::uint64_t>(max_steps) + 1;
    std::vector<std::uint64_t> dpt(ticks), fpt(ticks);
    check(snn_net_read_tick_counts(net.get(), 0, ticks, dpt.data(), fpt.data()), "snn_net_read_tick_counts");
    const std::uint64_t synaptic = counts.delivered - active.size();
    std::uint64_t requests = 0, scans = 0, receptor[4] = {0, 0, 0, 0}, stdp_ltp = 0, stdp_ltd = 0;
    std::vector<std::uint64_t> emit(n, 0), indeg(n, 0);
    for (std::size_t j = 0; j < e; ++j) {
        emit[p.source[j]] += std::max<std::uint64_t>(1, p.code_ptr[j + 1] - p.code_ptr[j]);
        indeg[p.target[j]] += 1;
        receptor[p.receptor[j] & 3] += pre_count[j];
        stdp_ltp += ltp[j];
        stdp_ltd += ltd[j];
    }
    std::uint64_t t_matches = 0, t_learned_total = 0;
    for (std::size_t i = 0; i < n; ++i) {
        requests += spikes_before_last[i] * emit[i];
        scans += result.final_neuron_spike_counts[i] * indeg[i];
        t_matches += t_match[i];
        t_learned_total += t_learned[i];
    }
    auto& d = result.debug;
    d.ticks_processed = ticks;
    d.scheduled_event_requests = requests;
    d.scheduled_events = synaptic;
    d.processed_events = synaptic;
    d.fired_appends = counts.fired;
    d.stdp_updates = stdp_ltp + stdp_ltd;
    d.stdp_ltp = stdp_ltp;
    d.stdp_ltd = stdp_ltd;
    d.receptor_ampa_events = receptor[0];
    d.receptor_nmda_events = receptor[1];
    d.receptor_gaba_a_events = receptor[2];
    d.receptor_gaba_b_events = receptor[3];
    d.dendritic_integrations = synaptic;
    d.post_plasticity_scans = scans;
    d.temporal_observations = counts.delivered + synaptic;
    d.temporal_patterns_learned = t_learned_total;
    d.temporal_matches = t_matches;
    d.max_scheduled_events_per_tick = *std::max_element(dpt.begin(), dpt.end());
    d.max_fired_neurons_per_tick = *std::max_element(fpt.begin(), fpt.end());
    result.elapsed_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return result;
}

// CudaInferenceSession::run_batch (reference include/snncuda/backends/CudaBackend.h:97-126,
// src/backends/cuda/CudaBackend.cu:1692-1723) with the reference CPU semantics.
//
// Stateless (options.reset_state_between_samples, the reference default): every sample runs on a
// fresh copy of the network, all copies stepped together as one batch (snn_batch_*): sample s's
// inputs are injects (neuron, tick, weight) in input order (ticks above max_steps and unknown ids
// skipped, as the reference skips them, CudaBackend.cu:2316-2324), then ticks 0..max_steps run.
// Two inputs to one neuron at one tick stay two injects, as on the CPU (the reference CUDA path sums
// them into one, :2325-2330). Per sample: delivered and fired totals (injects included, as the CPU
// counts them), the readout neurons' spike counts, ticks_processed, processed/scheduled synaptic
// events and fired_appends. readout_scores stays empty, as on the reference's stateless path.
//
// Stateful (reset_state_between_samples false): one network keeps its state across samples and
// calls, and the clock continues: a sample's input tick t lands at (first tick of the sample) + t and
// the sample runs max_steps + 1 ticks; its counts are the increments over the sample. (The
// reference CUDA path restarts its tick counter at 0 for every sample while keeping state.)
template <class BatchResult, class Sample, class Options>
class InferenceRunner {
public:
    explicit InferenceRunner(PackedConnectome p) : p_(std::move(p)) {}

    BatchResult run(const std::vector<Sample>& samples, const Options& options) {
        BatchResult batch{};
        batch.executed = true;
        const auto start = std::chrono::steady_clock::now();
        if (!samples.empty() && !p_.threshold.empty()) {
            if (options.reset_state_between_samples) run_stateless(samples, options, batch);
            else run_stateful(samples, options, batch);
        } else {
            batch.samples.resize(samples.size());
            for (auto& r : batch.samples) r.readout_spike_counts.assign(options.readout_neurons.size(), 0);
        }
        batch.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        return batch;
    }

    void reset() { state_ = Network(); }

    // CudaInferenceSession::run_feedforward_batch (reference src/backends/cuda/CudaBackend.cu:1725-1849)
    // on snn_graph_feedforward: inputs merged per neuron in order of first appearance (unknown ids
    // skipped), delivered = merged inputs, scores and 0/1 spike counts in readout order, fired = the
    // spike counts' sum.
    BatchResult feedforward(const std::vector<Sample>& samples, const Options& options) {
        const auto start = std::chrono::steady_clock::now();
        BatchResult batch{};
        batch.executed = true;
        batch.samples.resize(samples.size());
        if (samples.empty() || options.readout_neurons.empty()) return batch;
        Merged m = merge(samples);
        const std::vector<std::uint32_t> readout = dense_list(options.readout_neurons);
        const std::size_t S = samples.size(), R = readout.size();
        std::vector<float> scores(S * R);
        std::vector<std::uint64_t> spikes(S * R);
        check(snn_graph_feedforward(graph(), S, m.ptr.data(), m.neuron.data(), m.weight.data(), R, readout.data(),
                                    scores.data(), spikes.data()),
              "snn_graph_feedforward");
        fill_readouts(batch, m, scores, spikes, R, 0);
        batch.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        return batch;
    }

    // CudaInferenceSession::run_recurrent_feedforward_batch (:1851-2100) on snn_graph_recurrent. With
    // use_full_spike_timing the samples run through the spiking step (run) for max_timing_steps, or
    // max_delay * (1 + 2 * iterations) + 1 ticks when that is 0. Hidden neurons default to every
    // neuron that is not a readout, in connectome order (the reference takes them in the iteration
    // order of its id hash map, :2237-2251).
    template <class RecurrentOptions>
    BatchResult recurrent(const std::vector<Sample>& samples, const RecurrentOptions& options) {
        if (options.use_full_spike_timing) {
            Options timing{};
            timing.max_steps = options.max_timing_steps == 0
                                   ? static_cast<std::uint32_t>(static_cast<std::size_t>(max_delay()) *
                                                                    (1u + 2u * static_cast<std::size_t>(options.iterations)) +
                                                                1u)
                                   : options.max_timing_steps;
            timing.reset_state_between_samples = true;
            timing.readout_neurons = options.readout_neurons;
            return run(samples, timing);
        }
        const auto start = std::chrono::steady_clock::now();
        BatchResult batch{};
        batch.executed = true;
        batch.samples.resize(samples.size());
        if (samples.empty() || options.readout_neurons.empty()) return batch;
        std::vector<std::uint32_t> hidden;
        if (!options.hidden_neurons.empty()) {
            hidden = dense_list(options.hidden_neurons);
        } else {
            std::unordered_set<std::uint64_t> ro;
            for (const auto id : options.readout_neurons) ro.insert(static_cast<std::uint64_t>(id));
            std::vector<std::uint64_t> id_of(p_.threshold.size());
            for (const auto& [id, x] : p_.index) id_of[x] = id;
            for (std::uint32_t x = 0; x < id_of.size(); ++x)
                if (!ro.count(id_of[x])) hidden.push_back(x);
        }
        if (hidden.empty()) return batch;
        Merged m = merge(samples);
        const std::vector<std::uint32_t> readout = dense_list(options.readout_neurons);
        const std::size_t S = samples.size(), R = readout.size();
        std::vector<float> scores(S * R);
        std::vector<std::uint64_t> spikes(S * R);
        snn_recurrent_opts o{};
        o.iterations = options.iterations;
        o.feedback_decay = options.feedback_decay;
        o.top_k_feedback_readouts = static_cast<std::int32_t>(options.top_k_feedback_readouts);
        o.top_k_hidden = static_cast<std::int32_t>(options.top_k_hidden);
        o.use_scores_not_spikes = options.use_scores_not_spikes ? 1 : 0;
        std::uint64_t ticks = 0;
        check(snn_graph_recurrent(graph(), S, m.ptr.data(), m.neuron.data(), m.weight.data(), R, readout.data(),
                                  hidden.size(), hidden.data(), &o, scores.data(), spikes.data(), &ticks),
              "snn_graph_recurrent");
        fill_readouts(batch, m, scores, spikes, R, ticks);
        batch.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        return batch;
    }

    // the largest max(1, delay_ticks) over the connectome's synapses
    std::uint32_t max_delay() {
        std::uint32_t d = 0;
        check(snn_graph_max_delay(graph(), &d), "snn_graph_max_delay");
        return d;
    }

private:
    template <class Result>
    void fill_debug(Result& r, const Options& options, std::uint64_t synaptic, std::uint64_t fired) const {
        r.debug.ticks_processed = static_cast<std::uint64_t>(options.max_steps) + 1;
        r.debug.processed_events = synaptic;
        r.debug.scheduled_events = synaptic;
        r.debug.dendritic_integrations = synaptic;
        r.debug.fired_appends = fired;
    }

    void run_stateless(const std::vector<Sample>& samples, const Options& options, BatchResult& batch) {
        const std::size_t n = p_.threshold.size(), e = p_.source.size();
        // replicas per launch group: the union stays below the device's 2^32 neuron / synapse limits
        const std::size_t per = std::max<std::size_t>(
            1, std::min<std::size_t>((std::size_t)0xF0000000ull / std::max<std::size_t>(n, 1),
                                     (std::size_t)0xF0000000ull / std::max<std::size_t>(e, 1)));
        batch.samples.resize(samples.size());
        const snn_net_desc d = p_.desc();
        for (std::size_t first = 0; first < samples.size(); first += per) {
            const std::size_t m = std::min(per, samples.size() - first);
            std::vector<snn_net_desc> descs(m, d);
            snn_options o{};
            o.track_pre_count = 1;
            snn_batch* b = nullptr;
            check(snn_batch_create(descs.data(), m, &o, &b), "snn_batch_create");
            struct Guard {
                snn_batch* b;
                ~Guard() { snn_batch_destroy(b); }
            } guard{b};
            std::vector<std::uint32_t> net, neuron;
            std::vector<std::uint64_t> tick;
            std::vector<float> weight;
            std::vector<std::uint64_t> injected(m, 0);
            for (std::size_t s = 0; s < m; ++s)
                for (const auto& in : samples[first + s].inputs) {
                    if (in.tick > options.max_steps) continue;
                    const auto it = p_.index.find(static_cast<std::uint64_t>(in.neuron));
                    if (it == p_.index.end()) continue;
                    net.push_back(static_cast<std::uint32_t>(s));
                    neuron.push_back(it->second);
                    tick.push_back(in.tick);
                    weight.push_back(in.weight);
                    ++injected[s];
                }
            check(snn_batch_inject(b, net.data(), neuron.data(), tick.data(), weight.data(), net.size()),
                  "snn_batch_inject");
            check(snn_batch_run_until(b, options.max_steps), "snn_batch_run_until");
            std::vector<std::uint64_t> delivered(m), fired(m), spikes(m * n);
            check(snn_batch_net_counts(b, delivered.data(), fired.data()), "snn_batch_net_counts");
            snn_net* u = nullptr;
            check(snn_batch_network(b, &u), "snn_batch_network");
            check(snn_net_read_neurons(u, nullptr, nullptr, nullptr, spikes.data(), nullptr, nullptr, nullptr),
                  "snn_net_read_neurons");
            for (std::size_t s = 0; s < m; ++s) {
                auto& r = batch.samples[first + s];
                r.delivered_spikes = delivered[s];
                r.fired_spikes = fired[s];
                for (const auto id : options.readout_neurons) {
                    const auto it = p_.index.find(static_cast<std::uint64_t>(id));
                    r.readout_spike_counts.push_back(it == p_.index.end() ? 0 : spikes[s * n + it->second]);
                }
                fill_debug(r, options, delivered[s] - injected[s], fired[s]);
            }
        }
    }

    void run_stateful(const std::vector<Sample>& samples, const Options& options, BatchResult& batch) {
        if (!state_.get()) {
            snn_options o{};
            o.track_pre_count = 1;
            state_ = Network(p_.desc(), &o);
        }
        const std::size_t n = p_.threshold.size();
        std::vector<std::uint64_t> before(n), after(n);
        for (const auto& sample : samples) {
            const Counts c0 = state_.counts();
            check(snn_net_read_neurons(state_.get(), nullptr, nullptr, nullptr, before.data(), nullptr, nullptr, nullptr),
                  "snn_net_read_neurons");
            std::uint64_t injected = 0;
            for (const auto& in : sample.inputs) {
                if (in.tick > options.max_steps) continue;
                const auto it = p_.index.find(static_cast<std::uint64_t>(in.neuron));
                if (it == p_.index.end()) continue;
                state_.inject(it->second, c0.next_tick + in.tick, in.weight);
                ++injected;
            }
            state_.run_until(c0.next_tick + options.max_steps);
            const Counts c1 = state_.counts();
            check(snn_net_read_neurons(state_.get(), nullptr, nullptr, nullptr, after.data(), nullptr, nullptr, nullptr),
                  "snn_net_read_neurons");
            typename decltype(batch.samples)::value_type r{};
            r.delivered_spikes = c1.delivered - c0.delivered;
            r.fired_spikes = c1.fired - c0.fired;
            for (const auto id : options.readout_neurons) {
                const auto it = p_.index.find(static_cast<std::uint64_t>(id));
                r.readout_spike_counts.push_back(it == p_.index.end() ? 0 : after[it->second] - before[it->second]);
            }
            fill_debug(r, options, r.delivered_spikes - injected, r.fired_spikes);
            batch.samples.push_back(std::move(r));
        }
    }

    struct Merged {
        std::vector<std::uint64_t> ptr{0};
        std::vector<std::uint32_t> neuron;
        std::vector<float> weight;
    };

    // per sample: one entry per distinct known neuron, weights summed in input order
    Merged merge(const std::vector<Sample>& samples) const {
        Merged m;
        std::unordered_map<std::uint32_t, std::size_t> at;
        for (const auto& sample : samples) {
            at.clear();
            for (const auto& in : sample.inputs) {
                const auto it = p_.index.find(static_cast<std::uint64_t>(in.neuron));
                if (it == p_.index.end()) continue;
                const auto [pos, fresh] = at.emplace(it->second, m.neuron.size());
                if (fresh) {
                    m.neuron.push_back(it->second);
                    m.weight.push_back(in.weight);
                } else {
                    m.weight[pos->second] += in.weight;
                }
            }
            m.ptr.push_back(m.neuron.size());
        }
        return m;
    }

    template <class Ids>
    std::vector<std::uint32_t> dense_list(const Ids& ids) const {
        std::vector<std::uint32_t> out;
        out.reserve(ids.size());
        for (const auto id : ids) {
            const auto it = p_.index.find(static_cast<std::uint64_t>(id));
            out.push_back(it == p_.index.end() ? SNN_GRAPH_NONE : it->second);
        }
        return out;
    }

    void fill_readouts(BatchResult& batch, const Merged& m, const std::vector<float>& scores,
                       const std::vector<std::uint64_t>& spikes, std::size_t R, std::uint64_t ticks) const {
        for (std::size_t s = 0; s < batch.samples.size(); ++s) {
            auto& r = batch.samples[s];
            r.delivered_spikes = m.ptr[s + 1] - m.ptr[s];
            r.readout_scores.assign(scores.begin() + s * R, scores.begin() + (s + 1) * R);
            r.readout_spike_counts.assign(spikes.begin() + s * R, spikes.begin() + (s + 1) * R);
            for (const auto c : r.readout_spike_counts) r.fired_spikes += c;
            r.debug.ticks_processed = ticks;
            r.debug.processed_events = r.delivered_spikes;
            r.debug.dendritic_integrations = r.delivered_spikes;
        }
    }

    snn_graph* graph() {
        if (!graph_) {
            const snn_net_desc d = p_.desc();
            snn_graph* g = nullptr;
            check(snn_graph_create(&d, 0, &g), "snn_graph_create");
            graph_.reset(g);
        }
        return graph_.get();
    }

    struct GraphFree {
        void operator()(snn_graph* g) const noexcept { snn_graph_destroy(g); }
    };

    using Counts = Network::Counts;
    PackedConnectome p_;
    Network state_;
    std::unique_ptr<snn_graph, GraphFree> graph_;
};

}  // namespace snncuda_hip

#endif
