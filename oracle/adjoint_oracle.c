/* ORACLE — test infrastructure, not product code.
 *
 * FP64 CPU restatement of the fixed-spike adjoint: the subthreshold receptor / dendrite / soma
 * dynamics of the reference step (src/runtime/SynapseDendriteProcessor.cpp:175-229) with the
 * presynaptic event times and the reset events held fixed, a per-tick somatic trace, its MSE loss and
 * the gradient by reverse-mode recursion. The reference has no gradients, no compartment coupling and
 * no per-tick trace; this file defines them so that at the default parameters (coupling 1, receptor
 * time constants 5 / 100 / 10 / 150) the trace at a neuron's event ticks equals the reference
 * membrane (soma + basal + apical + inhibitory after the event, :287-292).
 *
 * Model, per neuron i, compartments k = 0..3, for ticks t = 0..T-1:
 *   decay   c_k <- c_k * exp(-1 / tau[r_k])   (t > 0; r_k = receptor of compartment k's last input;
 *                                              exp(-1/tau) per tick composes to the reference's
 *                                              exp(-dt/tau) over a gap)
 *   events  for each event of tick t, in processing order: c_comp += u; r_comp = receptor;
 *           u = weight * gain[receptor] * sign[receptor] (:183), external injects on the soma with
 *           AMPA (:77-87); when the neuron fired on the event, every c_k <- 0 (:199-209)
 *   trace   v(t) = sum_k alpha_k * c_k
 * Loss = sum_{t,i} (v_i(t) - target_i(t))^2 / (N * T). Parameters: synaptic weights, alpha[4], tau[4]
 * (indexed by receptor: AMPA, NMDA, GABA_A, GABA_B).
 *
 * The event lists come from snn_oracle.c's event log (the reference order and fire positions).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ADJ_NONE 0xFFFFFFFFu

static const double kGain[4] = {1.0, 0.55, 1.0, 0.65};
static const double kSign[4] = {1.0, 1.0, -1.0, -1.0};

typedef struct adjo {
    uint64_t n, e, T, nev;
    uint64_t* ptr;    /* n + 1: neuron i's events are [ptr[i], ptr[i+1]) in processing order */
    uint32_t* tick;
    uint32_t* syn;    /* ADJ_NONE: external inject */
    double* extw;     /* weight of an external inject */
    uint8_t* comp;
    uint8_t* rec;
    uint8_t* fired;
    uint32_t* syn_target;
    double* cs;       /* T x 4 states after each tick (scratch for one neuron) */
    uint8_t* rks;     /* T x 4 receptor of each compartment's last input after each tick */
    double* v;        /* T */
} adjo;

/* Events in processing order (tick-major): target, synapse (ADJ_NONE for external), weight of an
 * external inject, fired flag. Synapse compartments and receptors give each synaptic event's
 * compartment and receptor. */
adjo* adjo_create(uint64_t n, uint64_t e, uint64_t T, uint64_t nev, const uint64_t* ev_tick, const uint32_t* ev_target,
                  const uint32_t* ev_syn, const float* ev_weight, const uint8_t* ev_fired, const uint8_t* syn_comp,
                  const uint8_t* syn_rec, const uint32_t* syn_target) {
    adjo* a = (adjo*)calloc(1, sizeof(adjo));
    a->n = n;
    a->e = e;
    a->T = T;
    a->ptr = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
    for (uint64_t q = 0; q < nev; ++q)
        if (ev_tick[q] < T) a->ptr[ev_target[q] + 1]++;
    for (uint64_t i = 0; i < n; ++i) a->ptr[i + 1] += a->ptr[i];
    a->nev = a->ptr[n];
    a->tick = (uint32_t*)malloc((a->nev + 1) * sizeof(uint32_t));
    a->syn = (uint32_t*)malloc((a->nev + 1) * sizeof(uint32_t));
    a->extw = (double*)malloc((a->nev + 1) * sizeof(double));
    a->comp = (uint8_t*)malloc(a->nev + 1);
    a->rec = (uint8_t*)malloc(a->nev + 1);
    a->fired = (uint8_t*)malloc(a->nev + 1);
    uint64_t* fill = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    memcpy(fill, a->ptr, (n + 1) * sizeof(uint64_t));
    for (uint64_t q = 0; q < nev; ++q) {
        if (ev_tick[q] >= T) continue;
        const uint64_t p = fill[ev_target[q]]++;
        a->tick[p] = (uint32_t)ev_tick[q];
        a->syn[p] = ev_syn[q];
        a->extw[p] = ev_weight[q];
        a->comp[p] = ev_syn[q] == ADJ_NONE ? 0 : syn_comp[ev_syn[q]];
        a->rec[p] = ev_syn[q] == ADJ_NONE ? 0 : syn_rec[ev_syn[q]];
        a->fired[p] = ev_fired[q];
    }
    free(fill);
    a->syn_target = (uint32_t*)malloc((e + 1) * sizeof(uint32_t));
    if (e) memcpy(a->syn_target, syn_target, e * sizeof(uint32_t));
    a->cs = (double*)malloc(T * 4 * sizeof(double));
    a->rks = (uint8_t*)malloc(T * 4);
    a->v = (double*)malloc(T * sizeof(double));
    return a;
}

void adjo_destroy(adjo* a) {
    if (!a) return;
    free(a->ptr);
    free(a->tick);
    free(a->syn);
    free(a->extw);
    free(a->comp);
    free(a->rec);
    free(a->fired);
    free(a->syn_target);
    free(a->cs);
    free(a->rks);
    free(a->v);
    free(a);
}

uint64_t adjo_event_count(const adjo* a) { return a->nev; }

/* neuron i's trajectory into a->cs / a->rks / a->v */
static void forward_one(adjo* a, uint64_t i, const double* w, const double* alpha, const double* tau) {
    double c[4] = {0, 0, 0, 0};
    uint8_t rk[4] = {0, 0, 0, 0};
    double A[4];
    for (int r = 0; r < 4; ++r) A[r] = exp(-1.0 / tau[r]);
    uint64_t p = a->ptr[i];
    const uint64_t end = a->ptr[i + 1];
    for (uint64_t t = 0; t < a->T; ++t) {
        if (t > 0)
            for (int k = 0; k < 4; ++k) c[k] *= A[rk[k]];
        for (; p < end && a->tick[p] == t; ++p) {
            const int k = a->comp[p], r = a->rec[p];
            const double wt = a->syn[p] == ADJ_NONE ? a->extw[p] : w[a->syn[p]];
            c[k] += wt * kGain[r] * kSign[r];
            rk[k] = (uint8_t)r;
            if (a->fired[p]) c[0] = c[1] = c[2] = c[3] = 0.0;
        }
        double v = 0.0;
        for (int k = 0; k < 4; ++k) {
            v += alpha[k] * c[k];
            a->cs[t * 4 + k] = c[k];
            a->rks[t * 4 + k] = rk[k];
        }
        a->v[t] = v;
    }
}

/* trace[t * n + i] */
void adjo_forward(adjo* a, const double* w, const double* alpha, const double* tau, double* trace) {
    for (uint64_t i = 0; i < a->n; ++i) {
        forward_one(a, i, w, alpha, tau);
        for (uint64_t t = 0; t < a->T; ++t) trace[t * a->n + i] = a->v[t];
    }
}

/* neuron i's share of the loss */
double adjo_loss_target(adjo* a, uint64_t i, const double* w, const double* alpha, const double* tau,
                        const double* target) {
    forward_one(a, i, w, alpha, tau);
    double s = 0.0;
    for (uint64_t t = 0; t < a->T; ++t) {
        const double d = a->v[t] - target[t * a->n + i];
        s += d * d;
    }
    return s / ((double)a->n * (double)a->T);
}

double adjo_loss(adjo* a, const double* w, const double* alpha, const double* tau, const double* target) {
    double s = 0.0;
    for (uint64_t i = 0; i < a->n; ++i) s += adjo_loss_target(a, i, w, alpha, tau, target);
    return s;
}

/* Loss and its gradient. gw has e entries (synapses without events get 0). */
double adjo_grad(adjo* a, const double* w, const double* alpha, const double* tau, const double* target, double* gw,
                 double* galpha, double* gtau) {
    const double scale = 2.0 / ((double)a->n * (double)a->T);
    double loss = 0.0;
    memset(gw, 0, a->e * sizeof(double));
    for (int k = 0; k < 4; ++k) galpha[k] = gtau[k] = 0.0;
    double A[4];
    for (int r = 0; r < 4; ++r) A[r] = exp(-1.0 / tau[r]);
    for (uint64_t i = 0; i < a->n; ++i) {
        forward_one(a, i, w, alpha, tau);
        double lam[4] = {0, 0, 0, 0};
        uint64_t p = a->ptr[i + 1];
        const uint64_t begin = a->ptr[i];
        for (uint64_t tt = a->T; tt-- > 0;) {
            const double d = a->v[tt] - target[tt * a->n + i];
            loss += d * d;
            const double g = scale * d;
            for (int k = 0; k < 4; ++k) {
                galpha[k] += g * a->cs[tt * 4 + k];
                lam[k] += g * alpha[k];
            }
            for (; p > begin && a->tick[p - 1] == tt; --p) {
                const uint64_t q = p - 1;
                if (a->fired[q]) lam[0] = lam[1] = lam[2] = lam[3] = 0.0;
                const int k = a->comp[q], r = a->rec[q];
                if (a->syn[q] != ADJ_NONE) gw[a->syn[q]] += lam[k] * kGain[r] * kSign[r];
            }
            if (tt > 0)
                for (int k = 0; k < 4; ++k) {
                    const int r = a->rks[(tt - 1) * 4 + k];
                    gtau[r] += lam[k] * a->cs[(tt - 1) * 4 + k] * A[r] / (tau[r] * tau[r]);
                    lam[k] *= A[r];
                }
        }
    }
    return loss / ((double)a->n * (double)a->T);
}
