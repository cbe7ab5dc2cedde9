/*
 * Binary network/run description shared by the oracle, the reference harness and the tests.
 *
 * Test infrastructure only: nothing in the shipped library reads this format.
 *
 * File "SNNNET01" (little endian, no padding between sections):
 *   char magic[8]; u64 n_neurons, n_synapses, n_codes, n_commands;
 *   neurons  (SoA): f32 threshold[N]; u64 window[N]; f32 sim_thr[N]; u64 max_refs[N]
 *   synapses (SoA): u32 src[E]; u32 tgt[E]; f32 weight[E]; f32 max_weight[E]; u32 delay[E];
 *                   u8 comp[E]; u8 receptor[E]; u8 plastic[E]; u64 code_ptr[E+1]; u32 codes[n_codes]
 *   commands: n_commands x { u32 kind (0 inject, 1 run_until); u32 neuron; u64 tick; f32 weight; u32 pad }
 *
 * Neuron index i is NeuronId i+1 and synapse index j is SynapseId j+1, the ids ConnectomeBuilder
 * assigns (reference src/declarative/Connectome.cpp:93-95,141-158). A synapse with an empty code
 * list (code_ptr[j] == code_ptr[j+1]) means "spike_code_offsets = {}", which the reference turns
 * into {0} (src/runtime/SynapseDendriteProcessor.cpp:68-70).
 *
 * Result file "SNNOUT01": see oracle/snn_oracle.c (snno_write_results) and ref_harness.cpp.
 */
#ifndef SNN_NETFILE_H
#define SNN_NETFILE_H

#include <stdint.h>

#define SNN_NET_MAGIC "SNNNET01"
#define SNN_OUT_MAGIC "SNNOUT01"

enum { SNN_CMD_INJECT = 0, SNN_CMD_RUN_UNTIL = 1 };

typedef struct {
    uint32_t kind;
    uint32_t neuron;
    uint64_t tick;
    float weight;
    uint32_t pad;
} snn_command;

#endif
