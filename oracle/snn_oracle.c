/*
 * ORACLE — test infrastructure, not product code (see snn_oracle.h).
 *
 * Plain-C restatement of the reference CPU step. Each function names the reference file:line it
 * follows. Everything the reference keeps in unbounded std::vectors (temporal offsets, learned
 * patterns) is unbounded here too, so this oracle has none of the device kernels' capacity limits.
 *
 * Pinned by (tests/test_oracle.py): the reference's own known-answer tests restated as cases, the
 * golden fixtures written by the reference itself (oracle/_ref/ref_harness, tests/golden/), closed
 * forms (decay, STDP delta, cosine) and brute-force invariants on tiny nets.
 */
#define _POSIX_C_SOURCE 199309L
#include "snn_oracle.h"
#include "netfile.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------- growable arrays */

typedef struct {
    uint32_t* v;
    size_t n, cap;
} u32vec;

static void u32vec_push(u32vec* a, uint32_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 4;
        a->v = (uint32_t*)realloc(a->v, a->cap * sizeof(uint32_t));
    }
    a->v[a->n++] = x;
}

static void u32vec_copy(u32vec* dst, const u32vec* src) {
    if (dst->cap < src->n) {
        dst->cap = src->n;
        dst->v = (uint32_t*)realloc(dst->v, dst->cap * sizeof(uint32_t));
    }
    if (src->n) memcpy(dst->v, src->v, src->n * sizeof(uint32_t));
    dst->n = src->n;
}

/* A FIFO list of learned patterns (reference: std::vector<TemporalPattern>). */
typedef struct {
    u32vec* p;
    size_t n, cap;
} patlist;

/* learn with FIFO eviction: TemporalPatternMatcher::learn (src/snn/TemporalPattern.cpp:21-26),
 * learn_pattern (src/runtime/SynapseDendriteProcessor.cpp:31-41) and the code-pattern tail
 * (:361-364) all erase the first entry when size >= max, then push a copy. */
static void patlist_learn(patlist* l, const u32vec* pat, uint64_t max_refs) {
    if (l->n >= max_refs) {
        u32vec first = l->p[0];
        memmove(l->p, l->p + 1, (l->n - 1) * sizeof(u32vec));
        l->p[l->n - 1] = first; /* recycle the buffer of the evicted pattern */
        l->n--;
    }
    if (l->n == l->cap) {
        size_t old = l->cap;
        l->cap = l->cap ? l->cap * 2 : 2;
        l->p = (u32vec*)realloc(l->p, l->cap * sizeof(u32vec));
        memset(l->p + old, 0, (l->cap - old) * sizeof(u32vec));
    }
    u32vec_copy(&l->p[l->n], pat);
    l->n++;
}

static void patlist_free(patlist* l) {
    for (size_t i = 0; i < l->cap; ++i) free(l->p[i].v);
    free(l->p);
}

/* ---------------------------------------------------------------- model state */

enum { SOMA = 0, BASAL = 1, APICAL = 2, INHIBITORY = 3 };
enum { AMPA = 0, NMDA = 1, GABA_A = 2, GABA_B = 3 };

typedef struct {
    /* NeuronState (include/snncuda/runtime/NeuronStateCache.h:15-45) */
    float membrane;
    float threshold;
    uint64_t window;
    float sim_thr;
    uint64_t max_refs;
    uint64_t last_active;
    int64_t last_fired; /* -1: std::nullopt */
    uint64_t spike_count;
    float cur[4];
    uint64_t last_up[4];
    float tau[4];
    uint64_t t_start;
    u32vec t_offsets;
    patlist t_learned;
    int t_last_match;
    uint64_t t_match;
} neuron_t;

typedef struct {
    /* SynapseRuntimeState (include/snncuda/runtime/SynapseDendriteProcessor.h:14-35) */
    uint32_t source, target;
    float weight, max_weight;
    uint32_t delay;
    uint32_t ncodes;
    const uint32_t* codes; /* points into snno.codes; ncodes >= 1 */
    uint8_t comp, receptor, plastic;
    int64_t last_pre, last_post; /* -1: std::nullopt */
    uint64_t pre_count, upd_count, ltp_count;
    float last_delta;
    uint64_t c_start;
    u32vec c_offsets;
    patlist c_learned;
    uint64_t c_match;
    int c_last_match;
} synapse_t;

/* SpikeEvent (include/snncuda/snn/SpikeEvent.h:9-15); synapse == UINT32_MAX is an external inject. */
typedef struct {
    uint64_t delivery;
    uint32_t target;
    uint32_t synapse;
    float weight;
} event_t;

typedef struct {
    event_t* e;
    size_t n, cap;
} bucket_t;

#define WHEEL_SLOTS 4096u

struct snno {
    uint64_t n_neurons, n_syn;
    neuron_t* neurons;
    synapse_t* syn;
    uint32_t* codes;
    uint32_t zero_code;
    /* outgoing_/incoming_ lists in connectome order (src/runtime/NetworkPropagator.cpp:24-28) */
    uint64_t *out_ptr, *in_ptr;
    uint32_t *out_idx, *in_idx;
    /* SpikeScheduler (src/runtime/SpikeScheduler.cpp) */
    bucket_t wheel[WHEEL_SLOTS];
    uint64_t next_pop;
    event_t* due;
    size_t due_cap;
    /* counters and per-tick records */
    uint64_t delivered, fired;
    uint64_t* tick_delivered;
    uint64_t* tick_fired;
    size_t ticks_cap, ticks_n;
    uint64_t* fire_tick;
    uint32_t* fire_neuron;
    size_t fires_cap;
    /* optional log of every processed event, in processing order */
    int log_on;
    size_t log_n, log_cap;
    uint64_t* log_tick;
    uint32_t* log_target;
    uint32_t* log_syn;
    float* log_weight;
    uint8_t* log_fired;
};

static void bucket_push(bucket_t* b, event_t ev) {
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 16;
        b->e = (event_t*)realloc(b->e, b->cap * sizeof(event_t));
    }
    b->e[b->n++] = ev;
}

/* SpikeScheduler::schedule (src/runtime/SpikeScheduler.cpp:15-23): late events are clamped to the
 * next poll tick; the slot is tick % wheel size; bucket order is insertion order. */
static void schedule(snno* o, event_t ev) {
    if (ev.delivery < o->next_pop) ev.delivery = o->next_pop;
    bucket_push(&o->wheel[ev.delivery % WHEEL_SLOTS], ev);
}

/* SpikeScheduler::pop_due (src/runtime/SpikeScheduler.cpp:25-47): every slot from next_pop to tick,
 * due events in bucket order, later-rotation events retained in order. */
static size_t pop_due(snno* o, uint64_t tick) {
    size_t nd = 0;
    while (o->next_pop <= tick) {
        bucket_t* b = &o->wheel[o->next_pop % WHEEL_SLOTS];
        size_t kept = 0;
        for (size_t i = 0; i < b->n; ++i) {
            if (b->e[i].delivery <= tick) {
                if (nd == o->due_cap) {
                    o->due_cap = o->due_cap ? o->due_cap * 2 : 1024;
                    o->due = (event_t*)realloc(o->due, o->due_cap * sizeof(event_t));
                }
                o->due[nd++] = b->e[i];
            } else {
                b->e[kept++] = b->e[i];
            }
        }
        b->n = kept;
        o->next_pop++;
    }
    return nd;
}

/* ---------------------------------------------------------------- arithmetic */

/* receptor tables: SynapseDendriteProcessor::receptor_gain/decay_tau/sign
 * (src/runtime/SynapseDendriteProcessor.cpp:231-269) */
static float receptor_gain(uint8_t r) {
    switch (r) {
    case AMPA: return 1.0f;
    case NMDA: return 0.55f;
    case GABA_A: return 1.0f;
    case GABA_B: return 0.65f;
    }
    return 1.0f;
}
static float receptor_tau(uint8_t r) {
    switch (r) {
    case AMPA: return 5.0f;
    case NMDA: return 100.0f;
    case GABA_A: return 10.0f;
    case GABA_B: return 150.0f;
    }
    return 5.0f;
}
static float receptor_sign(uint8_t r) { return (r == GABA_A || r == GABA_B) ? -1.0f : 1.0f; }

/* TemporalPatternMatcher::cosine_similarity (src/snn/TemporalPattern.cpp:35-60): FP64 over the
 * common prefix, index order, 0 for an empty or zero-norm side, cast to float. */
static float cosine(const u32vec* left, const u32vec* right) {
    if (left->n == 0 || right->n == 0) return 0.0f;
    size_t count = left->n < right->n ? left->n : right->n;
    double dot = 0.0, ln = 0.0, rn = 0.0;
    for (size_t i = 0; i < count; ++i) {
        double l = (double)left->v[i];
        double r = (double)right->v[i];
        dot += l * r;
        ln += l * l;
        rn += r * r;
    }
    if (ln == 0.0 || rn == 0.0) return 0.0f;
    return (float)(dot / (sqrt(ln) * sqrt(rn)));
}

/* best_similarity + matches (src/snn/TemporalPattern.cpp:12-19): best starts at 0.0f and is
 * std::max(best, cos) with the reference pattern as the left argument. */
static int matches(const patlist* refs, const u32vec* cand, float thr) {
    float best = 0.0f;
    for (size_t i = 0; i < refs->n; ++i) {
        float c = cosine(&refs->p[i], cand);
        best = (best < c) ? c : best;
    }
    return best >= thr;
}

/* StdpRule::weight_delta + apply (src/learning/STDP.cpp:20-40) with the defaults of
 * include/snncuda/learning/STDP.h:8-16; std::min and std::clamp written out as libstdc++ defines
 * them. */
static float stdp_delta(uint64_t pre, uint64_t post) {
    const float dt = (float)((long long)post - (long long)pre);
    if (dt > 0.0f) return 0.01f * expf(-dt / 20.0f);
    if (dt < 0.0f) return -0.012f * expf(dt / 20.0f);
    return 0.0f;
}
/* A changed update counts as LTP when its delta is positive, else as LTD: the split of the
 * reference's stdp_ltp / stdp_ltd diagnostic counters (src/backends/cuda/CudaBackend.cu:133-139). */
static void stdp_apply(synapse_t* s, uint64_t pre, uint64_t post) {
    const float before = s->weight;
    const float maximum = (s->max_weight < 2.0f) ? s->max_weight : 2.0f;
    const float delta = stdp_delta(pre, post);
    const float v = s->weight + delta;
    s->weight = (v < 0.0f) ? 0.0f : (maximum < v) ? maximum : v;
    s->last_delta = s->weight - before;
    if (s->last_delta != 0.0f) {
        s->upd_count++;
        if (delta > 0.0f) s->ltp_count++;
    }
}

/* decay_compartments (src/runtime/SynapseDendriteProcessor.cpp:214-229) */
static void decay_compartments(neuron_t* t, uint64_t tick) {
    for (int k = 0; k < 4; ++k) {
        if (tick <= t->last_up[k] || t->cur[k] == 0.0f) {
            t->last_up[k] = tick;
            continue;
        }
        const float elapsed = (float)(tick - t->last_up[k]);
        t->cur[k] = t->cur[k] * expf(-elapsed / t->tau[k]);
        t->last_up[k] = tick;
    }
}

/* update_temporal_pattern + matches_learned_pattern + learn_pattern
 * (src/runtime/SynapseDendriteProcessor.cpp:294-316, :14-41) */
static void update_temporal_pattern(neuron_t* t, uint64_t tick) {
    if (t->t_offsets.n == 0 || tick - t->t_start > t->window) {
        t->t_start = tick;
        t->t_offsets.n = 0;
    }
    u32vec_push(&t->t_offsets, (uint32_t)(tick - t->t_start));
    t->t_last_match = (t->t_learned.n != 0 && t->t_offsets.n >= 2)
        ? matches(&t->t_learned, &t->t_offsets, t->sim_thr) : 0;
    if (t->t_last_match) {
        t->t_match++;
        return;
    }
    if (t->t_offsets.n < 2) return;
    patlist_learn(&t->t_learned, &t->t_offsets, t->max_refs);
}

/* process_input (src/runtime/SynapseDendriteProcessor.cpp:175-212); returns 1 when the neuron fired */
static int process_input(neuron_t* t, uint8_t comp, uint8_t receptor, float weight, uint64_t tick) {
    decay_compartments(t, tick);
    const float applied = weight * receptor_gain(receptor) * receptor_sign(receptor);
    t->cur[comp] += applied;
    t->last_up[comp] = tick;
    t->tau[comp] = receptor_tau(receptor);
    t->last_active = tick;
    update_temporal_pattern(t, tick);
    t->membrane = t->cur[SOMA] + t->cur[BASAL] + t->cur[APICAL] + t->cur[INHIBITORY];
    if (t->membrane >= t->threshold) {
        static const float reset_tau[4] = {5.0f, 5.0f, 100.0f, 10.0f}; /* Ampa, Ampa, Nmda, GabaA */
        t->membrane = 0.0f;
        for (int k = 0; k < 4; ++k) {
            t->cur[k] = 0.0f;
            t->last_up[k] = tick;
            t->tau[k] = reset_tau[k];
        }
        t->last_fired = (int64_t)tick;
        t->spike_count++;
        return 1;
    }
    return 0;
}

/* update_synapse_code_pattern (src/runtime/SynapseDendriteProcessor.cpp:318-365); the target's
 * window, threshold and max_refs apply. */
static void update_code_pattern(synapse_t* s, const neuron_t* t, uint64_t tick) {
    if (s->c_offsets.n == 0 || tick - s->c_start > t->window) {
        s->c_start = tick;
        s->c_offsets.n = 0;
    }
    u32vec_push(&s->c_offsets, (uint32_t)(tick - s->c_start));
    s->c_last_match = 0;
    const size_t expected = s->ncodes > 2 ? s->ncodes : 2;
    if (s->c_offsets.n < expected) return;
    if (s->c_offsets.n > expected) {
        const size_t drop = s->c_offsets.n - expected;
        memmove(s->c_offsets.v, s->c_offsets.v + drop, expected * sizeof(uint32_t));
        s->c_offsets.n = expected;
    }
    s->c_last_match = matches(&s->c_learned, &s->c_offsets, t->sim_thr);
    if (s->c_last_match) {
        s->c_match++;
        return;
    }
    patlist_learn(&s->c_learned, &s->c_offsets, t->max_refs);
}

/* process_synaptic_input (src/runtime/SynapseDendriteProcessor.cpp:89-136) */
static int process_synaptic(neuron_t* t, synapse_t* s, uint64_t tick) {
    const uint64_t pre_tick = tick >= s->delay ? tick - s->delay : tick;
    if (s->plastic && s->last_post >= 0) stdp_apply(s, pre_tick, (uint64_t)s->last_post);
    s->last_pre = (int64_t)pre_tick;
    s->pre_count++;
    const int fired = process_input(t, s->comp, s->receptor, s->weight, tick);
    update_code_pattern(s, t, tick);
    return fired;
}

/* apply_post_spike_plasticity (src/runtime/SynapseDendriteProcessor.cpp:138-173) */
static void post_spike_plasticity(synapse_t* s, uint64_t tick) {
    s->last_post = (int64_t)tick;
    if (!s->plastic || s->last_pre < 0) return;
    stdp_apply(s, (uint64_t)s->last_pre, tick);
}

static void record_tick(snno* o, uint64_t tick, uint64_t delivered, uint64_t fired) {
    if (tick >= o->ticks_cap) {
        size_t cap = o->ticks_cap ? o->ticks_cap : 1024;
        while (cap <= tick) cap *= 2;
        o->tick_delivered = (uint64_t*)realloc(o->tick_delivered, cap * sizeof(uint64_t));
        o->tick_fired = (uint64_t*)realloc(o->tick_fired, cap * sizeof(uint64_t));
        memset(o->tick_delivered + o->ticks_cap, 0, (cap - o->ticks_cap) * sizeof(uint64_t));
        memset(o->tick_fired + o->ticks_cap, 0, (cap - o->ticks_cap) * sizeof(uint64_t));
        o->ticks_cap = cap;
    }
    o->tick_delivered[tick] = delivered;
    o->tick_fired[tick] = fired;
    if (tick + 1 > o->ticks_n) o->ticks_n = tick + 1;
}

static void record_fire(snno* o, uint64_t tick, uint32_t neuron) {
    if (o->fired == o->fires_cap) {
        o->fires_cap = o->fires_cap ? o->fires_cap * 2 : 1024;
        o->fire_tick = (uint64_t*)realloc(o->fire_tick, o->fires_cap * sizeof(uint64_t));
        o->fire_neuron = (uint32_t*)realloc(o->fire_neuron, o->fires_cap * sizeof(uint32_t));
    }
    o->fire_tick[o->fired] = tick;
    o->fire_neuron[o->fired] = neuron;
}

static void log_event(snno* o, uint64_t tick, const event_t* ev, int fired) {
    if (o->log_n == o->log_cap) {
        o->log_cap = o->log_cap ? o->log_cap * 2 : 4096;
        o->log_tick = (uint64_t*)realloc(o->log_tick, o->log_cap * sizeof(uint64_t));
        o->log_target = (uint32_t*)realloc(o->log_target, o->log_cap * sizeof(uint32_t));
        o->log_syn = (uint32_t*)realloc(o->log_syn, o->log_cap * sizeof(uint32_t));
        o->log_weight = (float*)realloc(o->log_weight, o->log_cap * sizeof(float));
        o->log_fired = (uint8_t*)realloc(o->log_fired, o->log_cap);
    }
    o->log_tick[o->log_n] = tick;
    o->log_target[o->log_n] = ev->target;
    o->log_syn[o->log_n] = ev->synapse;
    o->log_weight[o->log_n] = ev->weight;
    o->log_fired[o->log_n] = (uint8_t)fired;
    o->log_n++;
}

/* NetworkPropagator::process_tick (src/runtime/NetworkPropagator.cpp:78-132) */
static void process_tick(snno* o, uint64_t tick) {
    const size_t nd = pop_due(o, tick);
    uint64_t fired_here = 0;
    for (size_t i = 0; i < nd; ++i) {
        const event_t ev = o->due[i];
        o->delivered++;
        neuron_t* t = &o->neurons[ev.target];
        int fired;
        if (ev.synapse == UINT32_MAX) {
            /* process_external_input (:77-87): Soma, Ampa, the event's weight */
            fired = process_input(t, SOMA, AMPA, ev.weight, tick);
        } else {
            fired = process_synaptic(t, &o->syn[ev.synapse], tick);
        }
        if (o->log_on) log_event(o, tick, &ev, fired);
        if (!fired) continue;
        record_fire(o, tick, ev.target);
        o->fired++;
        fired_here++;
        for (uint64_t k = o->in_ptr[ev.target]; k < o->in_ptr[ev.target + 1]; ++k)
            post_spike_plasticity(&o->syn[o->in_idx[k]], tick);
        for (uint64_t k = o->out_ptr[ev.target]; k < o->out_ptr[ev.target + 1]; ++k) {
            const uint32_t sid = o->out_idx[k];
            const synapse_t* s = &o->syn[sid];
            for (uint32_t c = 0; c < s->ncodes; ++c) {
                event_t out = {tick + s->delay + s->codes[c], s->target, sid, s->weight};
                schedule(o, out);
            }
        }
    }
    record_tick(o, tick, nd, fired_here);
}

/* ---------------------------------------------------------------- API */

snno* snno_create(uint64_t n_neurons, const float* threshold, const uint64_t* window,
                  const float* sim_thr, const uint64_t* max_refs, uint64_t n_syn,
                  const uint32_t* src, const uint32_t* tgt, const float* weight,
                  const float* max_weight, const uint32_t* delay, const uint8_t* comp,
                  const uint8_t* receptor, const uint8_t* plastic, const uint64_t* code_ptr,
                  const uint32_t* codes) {
    for (uint64_t i = 0; i < n_neurons; ++i)
        if (max_refs[i] == 0) return NULL; /* the reference erases begin() of an empty vector */
    for (uint64_t j = 0; j < n_syn; ++j)
        if (src[j] >= n_neurons || tgt[j] >= n_neurons || comp[j] > 3 || receptor[j] > 3) return NULL;

    snno* o = (snno*)calloc(1, sizeof(snno));
    o->n_neurons = n_neurons;
    o->n_syn = n_syn;
    o->neurons = (neuron_t*)calloc(n_neurons ? n_neurons : 1, sizeof(neuron_t));
    o->syn = (synapse_t*)calloc(n_syn ? n_syn : 1, sizeof(synapse_t));
    const uint64_t ncodes = n_syn ? code_ptr[n_syn] : 0;
    o->codes = (uint32_t*)malloc((ncodes ? ncodes : 1) * sizeof(uint32_t));
    if (ncodes) memcpy(o->codes, codes, ncodes * sizeof(uint32_t));
    o->zero_code = 0;

    for (uint64_t i = 0; i < n_neurons; ++i) {
        neuron_t* t = &o->neurons[i];
        t->threshold = threshold[i];
        t->window = window[i];
        t->sim_thr = sim_thr[i];
        t->max_refs = max_refs[i];
        t->last_fired = -1;
        for (int k = 0; k < 4; ++k) t->tau[k] = 5.0f; /* CompartmentState default */
    }
    for (uint64_t j = 0; j < n_syn; ++j) {
        synapse_t* s = &o->syn[j];
        s->source = src[j];
        s->target = tgt[j];
        s->weight = weight[j];
        s->max_weight = max_weight[j];
        s->delay = delay[j];
        /* initialize_synapse: an empty code list becomes {0} (:68-70) */
        if (code_ptr[j + 1] > code_ptr[j]) {
            s->ncodes = (uint32_t)(code_ptr[j + 1] - code_ptr[j]);
            s->codes = o->codes + code_ptr[j];
        } else {
            s->ncodes = 1;
            s->codes = &o->zero_code;
        }
        s->comp = comp[j];
        s->receptor = receptor[j];
        s->plastic = plastic[j] ? 1 : 0;
        s->last_pre = -1;
        s->last_post = -1;
    }
    /* outgoing_/incoming_ in connectome order (src/runtime/NetworkPropagator.cpp:24-28) */
    o->out_ptr = (uint64_t*)calloc(n_neurons + 1, sizeof(uint64_t));
    o->in_ptr = (uint64_t*)calloc(n_neurons + 1, sizeof(uint64_t));
    o->out_idx = (uint32_t*)malloc((n_syn ? n_syn : 1) * sizeof(uint32_t));
    o->in_idx = (uint32_t*)malloc((n_syn ? n_syn : 1) * sizeof(uint32_t));
    for (uint64_t j = 0; j < n_syn; ++j) {
        o->out_ptr[src[j] + 1]++;
        o->in_ptr[tgt[j] + 1]++;
    }
    for (uint64_t i = 0; i < n_neurons; ++i) {
        o->out_ptr[i + 1] += o->out_ptr[i];
        o->in_ptr[i + 1] += o->in_ptr[i];
    }
    uint64_t* fo = (uint64_t*)malloc((n_neurons + 1) * sizeof(uint64_t));
    uint64_t* fi = (uint64_t*)malloc((n_neurons + 1) * sizeof(uint64_t));
    memcpy(fo, o->out_ptr, (n_neurons + 1) * sizeof(uint64_t));
    memcpy(fi, o->in_ptr, (n_neurons + 1) * sizeof(uint64_t));
    for (uint64_t j = 0; j < n_syn; ++j) {
        o->out_idx[fo[src[j]]++] = (uint32_t)j;
        o->in_idx[fi[tgt[j]]++] = (uint32_t)j;
    }
    free(fo);
    free(fi);
    return o;
}

void snno_destroy(snno* o) {
    if (!o) return;
    for (uint64_t i = 0; i < o->n_neurons; ++i) {
        free(o->neurons[i].t_offsets.v);
        patlist_free(&o->neurons[i].t_learned);
    }
    for (uint64_t j = 0; j < o->n_syn; ++j) {
        free(o->syn[j].c_offsets.v);
        patlist_free(&o->syn[j].c_learned);
    }
    for (unsigned k = 0; k < WHEEL_SLOTS; ++k) free(o->wheel[k].e);
    free(o->neurons);
    free(o->syn);
    free(o->codes);
    free(o->out_ptr);
    free(o->in_ptr);
    free(o->out_idx);
    free(o->in_idx);
    free(o->due);
    free(o->log_tick);
    free(o->log_target);
    free(o->log_syn);
    free(o->log_weight);
    free(o->log_fired);
    free(o->tick_delivered);
    free(o->tick_fired);
    free(o->fire_tick);
    free(o->fire_neuron);
    free(o);
}

/* NetworkPropagator::inject (src/runtime/NetworkPropagator.cpp:31-37): an event with no synapse */
void snno_inject(snno* o, uint32_t neuron, uint64_t tick, float weight) {
    event_t ev = {tick, neuron, UINT32_MAX, weight};
    schedule(o, ev);
}

/* NetworkPropagator::run_until (src/runtime/NetworkPropagator.cpp:39-43): it loops from tick 0, but
 * pop_due returns nothing for ticks below next_pop, so only ticks >= next_pop do any work. */
void snno_run_until(snno* o, uint64_t inclusive_tick) {
    for (uint64_t tick = o->next_pop; tick <= inclusive_tick; ++tick) process_tick(o, tick);
}

uint64_t snno_delivered(const snno* o) { return o->delivered; }
uint64_t snno_fired(const snno* o) { return o->fired; }
uint64_t snno_ticks_recorded(const snno* o) { return o->ticks_n; }

void snno_read_tick_counts(const snno* o, uint64_t* delivered, uint64_t* fired) {
    if (o->ticks_n) {
        memcpy(delivered, o->tick_delivered, o->ticks_n * sizeof(uint64_t));
        memcpy(fired, o->tick_fired, o->ticks_n * sizeof(uint64_t));
    }
}

void snno_read_fires(const snno* o, uint64_t* tick, uint32_t* neuron) {
    if (o->fired) {
        memcpy(tick, o->fire_tick, o->fired * sizeof(uint64_t));
        memcpy(neuron, o->fire_neuron, o->fired * sizeof(uint32_t));
    }
}

void snno_read_neurons(const snno* o, float* membrane, float* currents, float* taus,
                       uint64_t* spike_count, int64_t* last_fired, uint64_t* t_match,
                       uint64_t* t_learned) {
    const uint64_t n = o->n_neurons;
    for (uint64_t i = 0; i < n; ++i) {
        const neuron_t* t = &o->neurons[i];
        if (membrane) membrane[i] = t->membrane;
        for (int k = 0; k < 4; ++k) {
            if (currents) currents[k * n + i] = t->cur[k];
            if (taus) taus[k * n + i] = t->tau[k];
        }
        if (spike_count) spike_count[i] = t->spike_count;
        if (last_fired) last_fired[i] = t->last_fired;
        if (t_match) t_match[i] = t->t_match;
        if (t_learned) t_learned[i] = t->t_learned.n;
    }
}

void snno_read_synapses(const snno* o, float* weight, uint64_t* pre_count, uint64_t* upd_count,
                        int64_t* last_pre, int64_t* last_post, uint64_t* code_match,
                        uint64_t* code_learned) {
    for (uint64_t j = 0; j < o->n_syn; ++j) {
        const synapse_t* s = &o->syn[j];
        if (weight) weight[j] = s->weight;
        if (pre_count) pre_count[j] = s->pre_count;
        if (upd_count) upd_count[j] = s->upd_count;
        if (last_pre) last_pre[j] = s->last_pre;
        if (last_post) last_post[j] = s->last_post;
        if (code_match) code_match[j] = s->c_match;
        if (code_learned) code_learned[j] = s->c_learned.n;
    }
}

void snno_set_event_log(snno* o, int on) { o->log_on = on; }

uint64_t snno_event_log_size(const snno* o) { return o->log_n; }

void snno_read_event_log(const snno* o, uint64_t* tick, uint32_t* target, uint32_t* synapse, float* weight,
                         uint8_t* fired) {
    for (size_t i = 0; i < o->log_n; ++i) {
        tick[i] = o->log_tick[i];
        target[i] = o->log_target[i];
        synapse[i] = o->log_syn[i];
        weight[i] = o->log_weight[i];
        fired[i] = o->log_fired[i];
    }
}

void snno_read_ltp(const snno* o, uint64_t* ltp_count) {
    for (uint64_t j = 0; j < o->n_syn; ++j) ltp_count[j] = o->syn[j].ltp_count;
}

/* ---------------------------------------------------------------- file driver */

static int read_all(FILE* f, void* dst, size_t size, size_t n) {
    return n == 0 || fread(dst, size, n, f) == n;
}

static void put_seq(uint32_t** words, size_t* n, size_t* cap, const u32vec* s) {
    if (*n + s->n + 1 > *cap) {
        while (*n + s->n + 1 > *cap) *cap = *cap ? *cap * 2 : 1024;
        *words = (uint32_t*)realloc(*words, *cap * sizeof(uint32_t));
    }
    (*words)[(*n)++] = (uint32_t)s->n;
    if (s->n) memcpy(*words + *n, s->v, s->n * sizeof(uint32_t));
    *n += s->n;
}

static void put_word(uint32_t** words, size_t* n, size_t* cap, uint32_t w) {
    if (*n + 1 > *cap) {
        *cap = *cap ? *cap * 2 : 1024;
        *words = (uint32_t*)realloc(*words, *cap * sizeof(uint32_t));
    }
    (*words)[(*n)++] = w;
}

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int snno_run_file(const char* net_path, const char* out_path, double* seconds) {
    FILE* f = fopen(net_path, "rb");
    if (!f) return -1;
    char magic[8];
    uint64_t hdr[4];
    if (!read_all(f, magic, 1, 8) || memcmp(magic, SNN_NET_MAGIC, 8) != 0 || !read_all(f, hdr, 8, 4)) {
        fclose(f);
        return -2;
    }
    const uint64_t n = hdr[0], e = hdr[1], nc = hdr[2], ncmd = hdr[3];
    float* thr = malloc((n + 1) * 4);
    uint64_t* win = malloc((n + 1) * 8);
    float* sim = malloc((n + 1) * 4);
    uint64_t* mref = malloc((n + 1) * 8);
    uint32_t* src = malloc((e + 1) * 4);
    uint32_t* tgt = malloc((e + 1) * 4);
    float* w = malloc((e + 1) * 4);
    float* mw = malloc((e + 1) * 4);
    uint32_t* dl = malloc((e + 1) * 4);
    uint8_t* comp = malloc(e + 1);
    uint8_t* rec = malloc(e + 1);
    uint8_t* pl = malloc(e + 1);
    uint64_t* cptr = malloc((e + 1) * 8);
    uint32_t* codes = malloc((nc + 1) * 4);
    snn_command* cmds = malloc((ncmd + 1) * sizeof(snn_command));
    int ok = read_all(f, thr, 4, n) && read_all(f, win, 8, n) && read_all(f, sim, 4, n)
        && read_all(f, mref, 8, n) && read_all(f, src, 4, e) && read_all(f, tgt, 4, e)
        && read_all(f, w, 4, e) && read_all(f, mw, 4, e) && read_all(f, dl, 4, e)
        && read_all(f, comp, 1, e) && read_all(f, rec, 1, e) && read_all(f, pl, 1, e)
        && read_all(f, cptr, 8, e + 1) && read_all(f, codes, 4, nc)
        && read_all(f, cmds, sizeof(snn_command), ncmd);
    fclose(f);
    snno* o = ok ? snno_create(n, thr, win, sim, mref, e, src, tgt, w, mw, dl, comp, rec, pl, cptr, codes) : NULL;
    free(thr); free(win); free(sim); free(mref); free(src); free(tgt); free(w); free(mw); free(dl);
    free(comp); free(rec); free(pl); free(cptr); free(codes);
    if (!o) {
        free(cmds);
        return -3;
    }
    double run_s = 0.0;
    for (uint64_t c = 0; c < ncmd; ++c) {
        if (cmds[c].kind == SNN_CMD_INJECT) {
            snno_inject(o, cmds[c].neuron, cmds[c].tick, cmds[c].weight);
        } else {
            const double t0 = now_seconds();
            snno_run_until(o, cmds[c].tick);
            run_s += now_seconds() - t0;
        }
    }
    free(cmds);
    if (seconds) *seconds = run_s;
    if (!out_path) {
        snno_destroy(o);
        return 0;
    }

    FILE* out = fopen(out_path, "wb");
    if (!out) {
        snno_destroy(o);
        return -4;
    }
    /* raster as (tick, neuron, count) sorted by (tick, neuron), as the harness writes it */
    uint32_t* rec4 = NULL;
    size_t nrec = 0;
    {
        uint32_t* count = calloc(n ? n : 1, sizeof(uint32_t));
        uint32_t* touched = malloc((n ? n : 1) * sizeof(uint32_t));
        rec4 = malloc((o->fired + 1) * 4 * sizeof(uint32_t));
        size_t i = 0;
        while (i < o->fired) {
            const uint64_t tk = o->fire_tick[i];
            size_t nt = 0;
            size_t j = i;
            for (; j < o->fired && o->fire_tick[j] == tk; ++j) {
                if (count[o->fire_neuron[j]]++ == 0) touched[nt++] = o->fire_neuron[j];
            }
            /* sort touched ascending (insertion sort; fires per tick are few in fixtures) */
            for (size_t a = 1; a < nt; ++a) {
                uint32_t x = touched[a];
                size_t b = a;
                while (b > 0 && touched[b - 1] > x) { touched[b] = touched[b - 1]; --b; }
                touched[b] = x;
            }
            for (size_t a = 0; a < nt; ++a) {
                rec4[4 * nrec + 0] = (uint32_t)tk;
                rec4[4 * nrec + 1] = touched[a];
                rec4[4 * nrec + 2] = count[touched[a]];
                rec4[4 * nrec + 3] = 0;
                count[touched[a]] = 0;
                ++nrec;
            }
            i = j;
        }
        free(count);
        free(touched);
    }
    fwrite(SNN_OUT_MAGIC, 1, 8, out);
    const uint64_t h[4] = {n, e, o->ticks_n, nrec};
    fwrite(h, 8, 4, out);
    fwrite(o->tick_delivered, 8, o->ticks_n, out);
    fwrite(o->tick_fired, 8, o->ticks_n, out);
    fwrite(rec4, 16, nrec, out);
    free(rec4);

    float* fbuf = malloc((4 * n + 1) * 4);
    uint64_t* ubuf = malloc((4 * n + 1) * 8);
    uint8_t* bbuf = malloc(n + 1);
    for (uint64_t i = 0; i < n; ++i) fbuf[i] = o->neurons[i].membrane;
    fwrite(fbuf, 4, n, out);
    for (int k = 0; k < 4; ++k) for (uint64_t i = 0; i < n; ++i) fbuf[k * n + i] = o->neurons[i].cur[k];
    fwrite(fbuf, 4, 4 * n, out);
    for (int k = 0; k < 4; ++k) for (uint64_t i = 0; i < n; ++i) fbuf[k * n + i] = o->neurons[i].tau[k];
    fwrite(fbuf, 4, 4 * n, out);
    for (int k = 0; k < 4; ++k) for (uint64_t i = 0; i < n; ++i) ubuf[k * n + i] = o->neurons[i].last_up[k];
    fwrite(ubuf, 8, 4 * n, out);
#define WRITE_NEURON_U64(expr) do { for (uint64_t i = 0; i < n; ++i) { const neuron_t* t = &o->neurons[i]; ubuf[i] = (uint64_t)(expr); } fwrite(ubuf, 8, n, out); } while (0)
    WRITE_NEURON_U64(t->last_active);
    WRITE_NEURON_U64(t->last_fired);
    WRITE_NEURON_U64(t->spike_count);
    WRITE_NEURON_U64(t->t_start);
    WRITE_NEURON_U64(t->t_offsets.n);
    WRITE_NEURON_U64(t->t_match);
    WRITE_NEURON_U64(t->t_learned.n);
#undef WRITE_NEURON_U64
    for (uint64_t i = 0; i < n; ++i) bbuf[i] = (uint8_t)o->neurons[i].t_last_match;
    fwrite(bbuf, 1, n, out);
    free(fbuf); free(ubuf); free(bbuf);

    fbuf = malloc((e + 1) * 4);
    ubuf = malloc((e + 1) * 8);
    bbuf = malloc(e + 1);
    for (uint64_t j = 0; j < e; ++j) fbuf[j] = o->syn[j].weight;
    fwrite(fbuf, 4, e, out);
    for (uint64_t j = 0; j < e; ++j) fbuf[j] = o->syn[j].last_delta;
    fwrite(fbuf, 4, e, out);
#define WRITE_SYN_U64(expr) do { for (uint64_t j = 0; j < e; ++j) { const synapse_t* s = &o->syn[j]; ubuf[j] = (uint64_t)(expr); } fwrite(ubuf, 8, e, out); } while (0)
    WRITE_SYN_U64(s->last_pre);
    WRITE_SYN_U64(s->last_post);
    WRITE_SYN_U64(s->pre_count);
    WRITE_SYN_U64(s->upd_count);
    WRITE_SYN_U64(s->c_start);
    WRITE_SYN_U64(s->c_offsets.n);
    WRITE_SYN_U64(s->c_match);
    WRITE_SYN_U64(s->c_learned.n);
#undef WRITE_SYN_U64
    for (uint64_t j = 0; j < e; ++j) bbuf[j] = (uint8_t)o->syn[j].c_last_match;
    fwrite(bbuf, 1, e, out);
    free(fbuf); free(ubuf); free(bbuf);

    uint32_t* words = NULL;
    size_t nw = 0, cw = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const neuron_t* t = &o->neurons[i];
        put_seq(&words, &nw, &cw, &t->t_offsets);
        put_word(&words, &nw, &cw, (uint32_t)t->t_learned.n);
        for (size_t k = 0; k < t->t_learned.n; ++k) put_seq(&words, &nw, &cw, &t->t_learned.p[k]);
    }
    uint64_t nw64 = nw;
    fwrite(&nw64, 8, 1, out);
    fwrite(words, 4, nw, out);
    nw = 0;
    for (uint64_t j = 0; j < e; ++j) {
        const synapse_t* s = &o->syn[j];
        put_seq(&words, &nw, &cw, &s->c_offsets);
        put_word(&words, &nw, &cw, (uint32_t)s->c_learned.n);
        for (size_t k = 0; k < s->c_learned.n; ++k) put_seq(&words, &nw, &cw, &s->c_learned.p[k]);
    }
    nw64 = nw;
    fwrite(&nw64, 8, 1, out);
    fwrite(words, 4, nw, out);
    free(words);
    const uint64_t tot[2] = {o->delivered, o->fired};
    fwrite(tot, 8, 2, out);
    fclose(out);
    snno_destroy(o);
    return 0;
}
