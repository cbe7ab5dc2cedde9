/*
 * ORACLE — test infrastructure, not product code.
 *
 * A plain-C restatement of the reference SNNCuda CPU per-tick step (NetworkPropagator +
 * SpikeScheduler + SynapseDendriteProcessor + StdpRule + TemporalPatternMatcher). It is the checker
 * for the HIP step: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The shipped library never links or calls it.
 *
 * Neurons and synapses are addressed by dense index (NeuronId - 1, SynapseId - 1).
 */
#ifndef SNN_ORACLE_H
#define SNN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct snno snno;

/* Build the oracle network. Arrays are in connectome order and copied. code_ptr has n_syn + 1
 * entries indexing codes; an empty code list means {0}. Returns NULL on invalid input (an index out
 * of range, max_refs == 0). */
snno* snno_create(uint64_t n_neurons, const float* threshold, const uint64_t* window,
                  const float* sim_thr, const uint64_t* max_refs, uint64_t n_syn,
                  const uint32_t* src, const uint32_t* tgt, const float* weight,
                  const float* max_weight, const uint32_t* delay, const uint8_t* comp,
                  const uint8_t* receptor, const uint8_t* plastic, const uint64_t* code_ptr,
                  const uint32_t* codes);
void snno_destroy(snno* o);

/* NetworkPropagator::inject / run_until (inclusive). */
void snno_inject(snno* o, uint32_t neuron, uint64_t tick, float weight);
void snno_run_until(snno* o, uint64_t inclusive_tick);

uint64_t snno_delivered(const snno* o);
uint64_t snno_fired(const snno* o);
uint64_t snno_ticks_recorded(const snno* o);
/* Per-tick delivered / fired counts for ticks [0, ticks_recorded). */
void snno_read_tick_counts(const snno* o, uint64_t* delivered, uint64_t* fired);
/* Fires in processing order: n = snno_fired(); tick[i], neuron[i]. */
void snno_read_fires(const snno* o, uint64_t* tick, uint32_t* neuron);

/* Neuron fields: membrane, currents[4][N], taus[4][N], spike_count, last_fired (-1 = none),
 * temporal match count and learned-pattern count. */
void snno_read_neurons(const snno* o, float* membrane, float* currents, float* taus,
                       uint64_t* spike_count, int64_t* last_fired, uint64_t* t_match,
                       uint64_t* t_learned);
/* Synapse fields in connectome order. last_pre / last_post are -1 when unset. */
void snno_read_synapses(const snno* o, float* weight, uint64_t* pre_count, uint64_t* upd_count,
                        int64_t* last_pre, int64_t* last_post, uint64_t* code_match,
                        uint64_t* code_learned);
/* Log of every processed event in processing order (off by default; switch on before running):
 * tick, target, synapse (UINT32_MAX for an external inject), the event's weight and whether the
 * target fired on it. */
void snno_set_event_log(snno* o, int on);
uint64_t snno_event_log_size(const snno* o);
void snno_read_event_log(const snno* o, uint64_t* tick, uint32_t* target, uint32_t* synapse, float* weight,
                         uint8_t* fired);
/* Changed STDP updates with a positive delta, per synapse (the rest of upd_count had delta <= 0). */
void snno_read_ltp(const snno* o, uint64_t* ltp_count);

/* Run a SNNNET01 file's commands and write a SNNOUT01 result file (same layout as the reference
 * harness writes). Returns 0 on success. If seconds is non-NULL it receives the time spent inside
 * run_until. */
int snno_run_file(const char* net_path, const char* out_path, double* seconds);

#ifdef __cplusplus
}
#endif

#endif
