/*
 * C ABI of the MI355X (gfx950) spiking step. Plain pointers and sizes; no C++ or torch types.
 *
 * Status codes: 0 ok, negative on error (see SNN_E*); snn_last_error() returns a thread-local
 * message for the last failing call on this thread. No C++ exception crosses this boundary.
 *
 * Threading: one host thread per handle at a time; distinct handles are independent. Each handle
 * owns a HIP stream on its device (default priority).
 */
#ifndef SNNCUDA_HIP_H
#define SNNCUDA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SNN_OK 0
#define SNN_EINVAL (-1)     /* invalid argument or unsupported network shape */
#define SNN_EDEVICE (-2)    /* HIP runtime error */
#define SNN_ECAPACITY (-3)  /* a device capacity was exceeded during a run; results are invalid */
#define SNN_ENODEVICE (-5)  /* no HIP device */

/* Unchanged reference symbols (reference include/snncuda/c/SNNCudaC.h:7-8, src/c/SNNCudaC.cpp:5-11):
 * the library version string (static storage) and the compile-time accelerator flag (1: this
 * library is built with its device backend). */
const char* snncuda_version(void);
int snncuda_cuda_available(void);

/* Number of visible HIP devices (0 when none); a device probe, unlike snncuda_cuda_available. */
int snn_device_count(void);
const char* snn_last_error(void);

/* A flattened Connectome in connectome order (reference include/snncuda/declarative/Connectome.h:13-37
 * and NetworkIR.h:14-19). Neuron i is NeuronId i+1, synapse j is SynapseId j+1; source/target are
 * neuron indices. code_ptr (n_synapses+1 entries, may be NULL) indexes codes; an empty list means
 * {0}, as the reference's initialize_synapse does. Compartment 0..3 = Soma, Basal, Apical,
 * Inhibitory; receptor 0..3 = Ampa, Nmda, GabaA, GabaB. All arrays are copied; the caller keeps
 * ownership. */
typedef struct snn_net_desc {
    uint64_t n_neurons;
    const float* threshold;
    const uint64_t* window;
    const float* sim_thr;
    const uint64_t* max_refs;
    uint64_t n_synapses;
    const uint32_t* source;
    const uint32_t* target;
    const float* weight;
    const float* max_weight;
    const uint32_t* delay;
    const uint8_t* compartment;
    const uint8_t* receptor;
    const uint8_t* plastic;
    const uint64_t* code_ptr;
    const uint32_t* codes;
} snn_net_desc;

/* Capacities (0 = default). Exceeding one during a run returns SNN_ECAPACITY: events are never
 * dropped silently. */
typedef struct snn_options {
    uint64_t max_events_per_tick;   /* delivered events in one tick (< 2^28) */
    uint64_t fire_log_capacity;     /* fires held for the delay window */
    uint64_t raster_capacity;       /* fires recorded for snn_net_read_raster over the handle's life */
    int32_t track_pre_count;        /* keep per-synapse delivered counts (pre_spike_count) */
    int32_t device;                 /* HIP device ordinal */
    uint64_t learned_pattern_slots; /* learned temporal patterns held per neuron (0: the network's
                                       largest max_reference_patterns); a neuron whose
                                       max_reference_patterns is larger fails the run once it holds
                                       that many */
    uint64_t pattern_device_pages;  /* the learned-pattern pool in pages of 1,024 neurons: 0 keeps it
                                       on the device when a quarter of device memory holds it, else
                                       in pinned host memory behind a device cache of as many pages
                                       as that quarter holds; n > 0 gives a device cache of n pages
                                       (the whole pool on the device when n covers the network).
                                       Results do not depend on where the pages are. */
    uint64_t pattern_runs;          /* runs (distinct offsets, run-length encoded) a neuron's temporal
                                       pattern holds, its candidate and each learned copy (0: 8; at
                                       most 64). A neuron that learns a candidate with more distinct
                                       offsets in its window fails the run with SNN_ECAPACITY; each
                                       run costs 8 B per neuron and per learned-pattern slot. */
} snn_options;

typedef struct snn_net snn_net;

/* Replaces CudaBackend::run_resident_propagation's host pack + upload
 * (reference src/backends/cuda/CudaBackend.cu:3382-3512) and NetworkPropagator's constructor
 * (src/runtime/NetworkPropagator.cpp:5-29): re-lays the connectome out on the device. */
int snn_net_create(const snn_net_desc* desc, const snn_options* opts, snn_net** out);
int snn_net_destroy(snn_net* net);

/* No counterpart in the reference: builds a seeded synthetic E/I network directly on the device
 * (the network the reference would receive as a Connectome; used for benchmarks too large to pack
 * on the host). Network neurons [0, n) with fan-out fanout[i], then n_sources input neurons with
 * source_fanout each; the same counter hash as snncuda_amd/workloads.py's ei_network, so the result
 * equals snn_net_create on that function's arrays. */
typedef struct snn_ei_desc {
    uint64_t n;
    uint64_t n_sources;
    const uint32_t* fanout;
    uint32_t source_fanout;
    uint64_t seed;
    uint64_t n_exc;
    float w_exc;
    float w_inh;
    uint32_t delay_max;         /* <= 63 */
    uint32_t n_comps;
    uint8_t comps[4];
    int32_t plastic;            /* 0: no plastic synapses, 1: every synapse, 2: every synapse except
                                   those of inhibitory neurons [n_exc, n) */
    float threshold;
    float source_threshold;
    uint32_t window;
    float sim_thr;
    uint32_t max_refs;
    uint32_t n_codes;           /* code offsets per synapse, 1..8 (0: one code {0}) */
    uint32_t code_max;          /* offsets drawn from U{0..code_max} (0: every offset is 0); <= 63 */
} snn_ei_desc;
int snn_net_create_ei(const snn_ei_desc* desc, const snn_options* opts, snn_net** out);

/* NetworkPropagator::inject (src/runtime/NetworkPropagator.cpp:31-37), n events in call order. */
int snn_net_inject(snn_net* net, const uint32_t* neuron, const uint64_t* tick, const float* weight, uint64_t n);

/* NetworkPropagator::run_until (src/runtime/NetworkPropagator.cpp:39-43): process every tick up to
 * and including inclusive_tick. Synchronous; the per-tick loop itself never waits on the host. */
int snn_net_run_until(snn_net* net, uint64_t inclusive_tick);

/* No counterpart in the reference. snn_net_run_until for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * ticks and results, with the run segment's inject tables and per-tick counters staged in a fixed
 * 4 MiB device array that each device holds until the process ends, instead of buffers allocated
 * and freed per call; no HIP graph replay, autotuning or profiling events inside the call. A segment
 * that needs more staging than the array holds (about 262,000 ticks without injects, 12 bytes per
 * pending inject) fails with SNN_EINVAL and changes nothing. All calls on one device share its
 * array: calls that overlap on one device race. */
int snn_net_run_until_static(snn_net* net, uint64_t inclusive_tick);

/* No counterpart in the reference (its NeuronStateCache pages neuron state on the CPU,
 * src/runtime/NeuronStateCache.cpp:5-43): the learned-pattern pool's pages (of 1,024 neurons), how
 * many of them the device cache holds (0: the whole pool is on the device) and the page copies made
 * so far. Any pointer may be NULL. */
int snn_net_pattern_paging(snn_net* net, uint64_t* pages, uint64_t* device_pages, uint64_t* moves);

/* delivered_spike_count / fired_spike_count (NetworkPropagator.h:27-28) and the next tick. */
int snn_net_counts(snn_net* net, uint64_t* delivered, uint64_t* fired, uint64_t* next_tick);

/* Per-tick delivered / fired counts for ticks [first_tick, first_tick + n). */
int snn_net_read_tick_counts(snn_net* net, uint64_t first_tick, uint64_t n, uint64_t* delivered, uint64_t* fired);

/* Fires in processing order (tick-major; within a tick in the reference's order). *n_out receives
 * the number recorded; neurons (may be NULL) receives min(n_out, cap) neuron indices. Recording stops
 * (the run does not) once raster_capacity is reached; the call then returns SNN_ECAPACITY after
 * filling what was recorded. */
int snn_net_read_raster(snn_net* net, uint64_t* n_out, uint32_t* neurons, uint64_t cap);

/* Final neuron state (CudaPropagationResult fields, reference include/snncuda/backends/CudaBackend.h:40-53,
 * plus the compartment currents/taus of NeuronState). Any pointer may be NULL. currents and taus are
 * [4][n] (soma, basal, apical, inhibitory). last_fired is -1 when the neuron never fired. */
int snn_net_read_neurons(snn_net* net, float* membrane, float* currents, float* taus, uint64_t* spike_count,
                         int64_t* last_fired, uint64_t* temporal_match, uint64_t* temporal_learned);

/* Final synapse state in connectome order (SynapseRuntimeState fields). Any pointer may be NULL.
 * last_pre / last_post are -1 when unset. pre_count needs track_pre_count. */
int snn_net_read_synapses(snn_net* net, float* weight, uint64_t* pre_count, uint64_t* update_count,
                          int64_t* last_pre, int64_t* last_post, uint64_t* code_match, uint64_t* code_learned);

/* Changed STDP updates per synapse in connectome order, split by the sign of the applied delta
 * (delta > 0: LTP; otherwise LTD), as the reference's stdp_ltp / stdp_ltd diagnostic counters split
 * them (reference src/backends/cuda/CudaBackend.cu:133-139). ltp + ltd equals update_count of
 * snn_net_read_synapses. Either pointer may be NULL. */
int snn_net_read_plasticity(snn_net* net, uint64_t* ltp_updates, uint64_t* ltd_updates);

/* Launch schedule: target partition size (2^partition_shift neurons, 10..12), workgroups of the
 * partition count / scatter kernels, and ticks per HIP graph replay (0: one stream launch per kernel;
 * n > 0: the launches of n ticks are captured once per snn_net_run_until call and replayed). Results
 * are identical under every schedule. No counterpart in the reference (its launch shapes are fixed,
 * src/backends/cuda/CudaBackend.cu:3672-3799). */
typedef struct snn_schedule {
    uint32_t partition_shift;
    uint32_t scatter_workgroups;
    uint32_t graph_ticks;
    uint32_t fold_threads;      /* fold workgroup: 1024, or 256 (small LDS stage, more workgroups per
                                   CU: faster when few events reach each partition); 0 on set: 1024 */
} snn_schedule;
int snn_net_get_schedule(snn_net* net, snn_schedule* out);
int snn_net_set_schedule(snn_net* net, const snn_schedule* schedule);

/* Online autotuner: the next snn_net_run_until call that is long enough spends its first ticks
 * (ticks of the run itself, not extra work) on chunks of trial_ticks under candidate schedules, timed
 * with HIP events on the handle's stream, and keeps the fastest for the rest of that call and later
 * calls. trial_ticks 0 cancels a pending tune. snn_net_tune_log copies the trial table (text) and
 * returns its length. */
int snn_net_autotune(snn_net* net, uint32_t trial_ticks);
int snn_net_tune_log(snn_net* net, char* buf, uint64_t cap);

/* Batched form: independent networks of any sizes stepped together by one set of launches per tick
 * (one graph replay per graph_ticks ticks). Replaces the reference's batched step
 * (src/backends/cuda/CudaBackend.cu:837-1258, run_stateless_device_batch :2296-2578), which steps
 * replicas of one connectome; here every network may differ. Network k's neurons and synapses occupy
 * [neuron_off[k], neuron_off[k+1]) and [synapse_off[k], synapse_off[k+1]) of one union network, and
 * its results equal a run of network k alone. snn_batch_network returns that union network
 * (borrowed, valid until snn_batch_destroy) for the snn_net_read_* readers, schedules and the
 * autotuner. Injects go through snn_batch_inject (network-local neuron indices). */
typedef struct snn_batch snn_batch;
int snn_batch_create(const snn_net_desc* nets, uint64_t n_nets, const snn_options* opts, snn_batch** out);
/* No counterpart in the reference: generates n_nets seeded E/I networks (snn_net_create_ei's family,
 * one descriptor each) on the device as a batch; equal to snn_batch_create on the host-built arrays. */
int snn_batch_create_ei(const snn_ei_desc* nets, uint64_t n_nets, const snn_options* opts, snn_batch** out);
int snn_batch_destroy(snn_batch* batch);
int snn_batch_network(snn_batch* batch, snn_net** net);
int snn_batch_offsets(snn_batch* batch, uint64_t* n_nets, uint64_t* neuron_off, uint64_t* synapse_off);
int snn_batch_inject(snn_batch* batch, const uint32_t* net, const uint32_t* neuron, const uint64_t* tick,
                     const float* weight, uint64_t n);
int snn_batch_run_until(snn_batch* batch, uint64_t inclusive_tick);
/* Per-network delivered (external injects included, as delivered_spike_count counts them) and fired
 * totals, arrays of n_nets. delivered needs track_pre_count; either pointer may be NULL. */
int snn_batch_net_counts(snn_batch* batch, uint64_t* delivered, uint64_t* fired);

/* Non-spiking readouts of a fixed-weight connectome: CudaInferenceSession::run_feedforward_batch and
 * run_recurrent_feedforward_batch (reference include/snncuda/backends/CudaBackend.h:113-119, kernels
 * src/backends/cuda/CudaBackend.cu:1260-1528, host sequence :1725-2100). A graph holds the
 * connectome's out-edges (target, weight, delay max(1, delay_ticks)) and neuron thresholds on the
 * device; the Connectome's plastic / code / compartment fields are not used. */
typedef struct snn_graph snn_graph;
#define SNN_GRAPH_NONE 0xFFFFFFFFu /* a readout / hidden entry that is not a neuron of the connectome */
int snn_graph_create(const snn_net_desc* desc, int device, snn_graph** out);
int snn_graph_destroy(snn_graph* graph);
/* The largest max(1, delay_ticks) over the synapses (1 without synapses). */
int snn_graph_max_delay(snn_graph* graph, uint32_t* max_delay);

/* Samples are given merged: sample s's inputs are entries [input_ptr[s], input_ptr[s+1]) of
 * input_neuron / input_weight (neuron indices), one entry per distinct neuron, its weights summed in
 * sample order, in order of first appearance (the reference merges them the same way,
 * CudaBackend.cu:1757-1775). readout[r] (and hidden[h]) is a neuron index, or SNN_GRAPH_NONE for an id
 * that is not a neuron of the connectome (score 0, threshold 1.0, no edges). A neuron listed more
 * than once maps to its last listing. Sums run in a fixed order: inputs (or sources) in list order,
 * each source's edges in connectome order.
 *
 * Feedforward (:1260-1310): scores[s*R + r] = sum of input_weight * synapse weight over the edges
 * from sample s's inputs to readout r; spikes[s*R + r] = scores >= the readout's threshold. */
int snn_graph_feedforward(snn_graph* graph, uint64_t n_samples, const uint64_t* input_ptr,
                          const uint32_t* input_neuron, const float* input_weight, uint64_t n_readouts,
                          const uint32_t* readout, float* scores, uint64_t* spikes);

/* Recurrent readout / feedback timeline (:1312-1528, :1851-2100) over time_count =
 * max_delay * (1 + 2 * iterations) + 1 steps (written to *time_count). top_k_feedback_readouts and
 * top_k_hidden are the reference options cast to int as its kernels receive them: a readout top-k
 * <= 0 gates every readout whose step score reaches its threshold; the hidden top-k acts when
 * 0 < top_k_hidden < n_hidden. Fails with SNN_EINVAL when the timelines
 * (n_samples * time_count * (n_hidden + n_readouts) floats) exceed a quarter of free device memory. */
typedef struct snn_recurrent_opts {
    uint32_t iterations;
    float feedback_decay;
    int32_t top_k_feedback_readouts;
    int32_t top_k_hidden;
    int32_t use_scores_not_spikes;
} snn_recurrent_opts;
int snn_graph_recurrent(snn_graph* graph, uint64_t n_samples, const uint64_t* input_ptr,
                        const uint32_t* input_neuron, const float* input_weight, uint64_t n_readouts,
                        const uint32_t* readout, uint64_t n_hidden, const uint32_t* hidden,
                        const snn_recurrent_opts* opts, float* scores, uint64_t* spikes, uint64_t* time_count);

/* Recording of a run for the fixed-spike adjoint (no counterpart in the reference). Switched on
 * before the first tick, the step keeps, for every fire of the raster, the index of the event the
 * neuron fired on among its ordered events of that tick, and keeps every delivered inject (neuron,
 * tick, scheduler poll tick at the inject call, weight) in delivery order. */
int snn_net_record_fire_events(snn_net* net, int enabled);
int snn_net_read_fire_events(snn_net* net, uint64_t* n_out, uint32_t* event_index, uint64_t cap);
int snn_net_read_inject_log(snn_net* net, uint64_t* n_out, uint32_t* neuron, uint64_t* tick, uint64_t* stage,
                            float* weight, uint64_t cap);

/* Fixed-spike adjoint (no counterpart in the reference, which has no gradients). The subthreshold
 * receptor / dendrite / soma dynamics of the step (reference src/runtime/SynapseDendriteProcessor.cpp:175-229)
 * with every event time and every reset held at those of a recorded run (ticks [0, ticks)):
 *   c_k(t) = exp(-1/tau[r_k]) c_k(t-1), then each event of tick t in processing order:
 *   c_comp += w * gain[r] * sign[r], r_comp = r, and c = 0 on the event the neuron fired on;
 *   trace v(t) = sum_k alpha[k] c_k(t)   (alpha = 1: the reference membrane sum, :287-292)
 * Parameters: synapse weights (connectome order) and params[8] = alpha[4] (soma, basal, apical,
 * inhibitory), tau[4] (AMPA, NMDA, GABA_A, GABA_B; reference values 5, 100, 10, 150).
 * snn_adjoint_create builds every neuron's event list from a run recorded with
 * snn_net_record_fire_events and the same connectome. Device pointers (d_*) live on the device that
 * was current at create; calls synchronize the adjoint's stream before returning.
 * snn_adjoint_forward writes the trace [ticks][n_neurons] and keeps checkpoints every 64 ticks;
 * snn_adjoint_backward takes dL/dtrace [ticks][n_neurons] for the last forward's parameters,
 * recomputes each 64-tick segment from its checkpoint and writes dL/dweight (device, n_synapses) and
 * dL/dparams (host, 8 doubles). The backward reads the weights through the last forward's d_weight,
 * so that buffer must stay valid and unchanged until the next snn_adjoint_forward or
 * snn_adjoint_destroy. */
typedef struct snn_adjoint snn_adjoint;
int snn_adjoint_create(snn_net* run, const snn_net_desc* desc, uint64_t ticks, snn_adjoint** out);
int snn_adjoint_destroy(snn_adjoint* adj);
int snn_adjoint_sizes(snn_adjoint* adj, uint64_t* n_neurons, uint64_t* n_synapses, uint64_t* ticks,
                      uint64_t* n_events);
/* The event lists (host copies): ptr[n+1]; per event tick, synapse (UINT32_MAX for an inject),
 * inject weight, reset flag. Any pointer may be NULL. */
int snn_adjoint_read_events(snn_adjoint* adj, uint64_t* ptr, uint32_t* tick, uint32_t* synapse, float* weight,
                            uint8_t* reset);
int snn_adjoint_forward(snn_adjoint* adj, const float* d_weight, const float* params, float* d_trace);
int snn_adjoint_backward(snn_adjoint* adj, const float* d_grad_trace, float* d_grad_weight, double* grad_params);

/* Per-kernel time of the step, measured with HIP events on the handle's stream while profiling is
 * enabled (no counterpart in the reference). snn_net_set_profiling(enabled): 0 off, 1 every kernel,
 * 0x40000000 | every << 16 | mask only the kernels whose bit is set (bit k = the k-th name that
 * snn_net_kernel_times reports, mask < 2^16) and only on ticks that are a multiple of every (1..255;
 * 0: every tick), so a timed run carries events around few launches; it also clears the totals. snn_net_kernel_times: total ms and launches per kernel since then; names
 * receives a comma-separated list. Returns the number of kernels written. */
int snn_net_set_profiling(snn_net* net, int enabled);
int snn_net_kernel_times(snn_net* net, double* ms_total, uint64_t* launches, int cap, char* names, uint64_t names_cap);

#ifdef __cplusplus
}
#endif

#endif
