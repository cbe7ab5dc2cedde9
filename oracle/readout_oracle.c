/*
 * ORACLE — test infrastructure, not product code. Used only by tests/ as the checker of the
 * device readouts.
 *
 * Plain-C restatement of the reference's non-spiking readouts, one sample and one edge at a time:
 *   feedforward  src/backends/cuda/CudaBackend.cu:1260-1310 (accumulate_feedforward_readout_kernel,
 *                threshold_feedforward_readout_kernel), host sequence :1725-1849
 *   recurrent    :1312-1528 (initialize_recurrent_hidden_events_kernel,
 *                propagate_hidden_events_to_readouts_kernel, select_recurrent_readout_events_kernel,
 *                propagate_readout_feedback_events_kernel, apply_topk_hidden_events_kernel,
 *                sum_recurrent_readout_events_kernel), host sequence :1851-2100
 * The reference adds each edge's contribution with atomicAdd, in no defined order; here every sum
 * runs in one fixed order: sources (inputs, hidden or readout entries) in list order, each source's
 * out-edges in connectome order (the order of the reference's outgoing lists, :2137-2182), steps in
 * ascending order. Inputs arrive merged per neuron (:1757-1775). A readout or hidden entry that is
 * not a neuron of the connectome (RO_NONE) has no edges (the reference points such an entry at
 * dense neuron 0, :1892-1913, which only matters for a readout gated at 1.0 in top-k mode).
 *
 * Pinned by tests/test_readout_cpu.py: the reference's own known answers
 * (tests/cuda_synapse_dendrite_tests.cpp:306-383), dense matrix restatements in float64 on random
 * graphs, brute-force top-k and threshold gating cases.
 */
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define RO_NONE 0xFFFFFFFFu

typedef struct {
    uint64_t* ptr;
    uint32_t* tgt;
    float* w;
    uint32_t* delay;
    uint32_t max_delay;
} Csr;

static void csr_free(Csr* c) {
    free(c->ptr);
    free(c->tgt);
    free(c->w);
    free(c->delay);
}

/* out-edges by source, connectome order within a source; delay max(1, d) (:2146) */
static int csr_build(Csr* c, uint64_t n, uint64_t e, const uint32_t* src, const uint32_t* tgt, const float* w,
                     const uint32_t* delay) {
    memset(c, 0, sizeof(*c));
    c->ptr = calloc(n + 1, sizeof(uint64_t));
    c->tgt = malloc((e ? e : 1) * sizeof(uint32_t));
    c->w = malloc((e ? e : 1) * sizeof(float));
    c->delay = malloc((e ? e : 1) * sizeof(uint32_t));
    uint64_t* at = malloc((n ? n : 1) * sizeof(uint64_t));
    if (!c->ptr || !c->tgt || !c->w || !c->delay || !at) {
        free(at);
        csr_free(c);
        return -1;
    }
    c->max_delay = 1;
    for (uint64_t j = 0; j < e; ++j) {
        ++c->ptr[src[j] + 1];
        const uint32_t d = delay ? (delay[j] > 1 ? delay[j] : 1) : 1;
        if (d > c->max_delay) c->max_delay = d;
    }
    for (uint64_t x = 0; x < n; ++x) c->ptr[x + 1] += c->ptr[x];
    for (uint64_t x = 0; x < n; ++x) at[x] = c->ptr[x];
    for (uint64_t j = 0; j < e; ++j) {
        const uint64_t q = at[src[j]]++;
        c->tgt[q] = tgt[j];
        c->w[q] = w[j];
        c->delay[q] = delay ? (delay[j] > 1 ? delay[j] : 1) : 1;
    }
    free(at);
    return 0;
}

/* dense neuron -> list index + 1, the last listing winning (:1745-1754) */
static uint32_t* dense_map(uint64_t n, uint64_t m, const uint32_t* list) {
    uint32_t* d = calloc(n ? n : 1, sizeof(uint32_t));
    if (!d) return NULL;
    for (uint64_t j = 0; j < m; ++j)
        if (list[j] != RO_NONE) d[list[j]] = (uint32_t)(j + 1);
    return d;
}

static float threshold_of(const float* threshold, uint32_t x) { return x == RO_NONE ? 1.0f : threshold[x]; }

int roo_feedforward(uint64_t n, const float* threshold, uint64_t e, const uint32_t* src, const uint32_t* tgt,
                    const float* w, uint64_t S, const uint64_t* in_ptr, const uint32_t* in_x, const float* in_w,
                    uint64_t R, const uint32_t* readout, float* scores, uint64_t* spikes) {
    Csr c;
    if (csr_build(&c, n, e, src, tgt, w, NULL)) return -1;
    uint32_t* d2r = dense_map(n, R, readout);
    if (!d2r) {
        csr_free(&c);
        return -1;
    }
    for (uint64_t s = 0; s < S; ++s) {
        float* row = scores + s * R;
        for (uint64_t r = 0; r < R; ++r) row[r] = 0.0f;
        for (uint64_t i = in_ptr[s]; i < in_ptr[s + 1]; ++i) {
            const uint32_t x = in_x[i];
            for (uint64_t q = c.ptr[x]; q < c.ptr[x + 1]; ++q) {
                const uint32_t r1 = d2r[c.tgt[q]];
                if (r1) row[r1 - 1] += in_w[i] * c.w[q];
            }
        }
        for (uint64_t r = 0; r < R; ++r) spikes[s * R + r] = row[r] >= threshold_of(threshold, readout[r]) ? 1u : 0u;
    }
    free(d2r);
    csr_free(&c);
    return 0;
}

/* k passes of "largest value not yet selected, ties to the lower index, above -FLT_MAX" (:1405-1425) */
static int64_t best_unselected(const float* row, const uint8_t* sel, uint64_t m, float* best_score) {
    int64_t best = -1;
    float bs = -FLT_MAX;
    for (uint64_t j = 0; j < m; ++j)
        if (!sel[j] && row[j] > bs) {
            bs = row[j];
            best = (int64_t)j;
        }
    *best_score = bs;
    return best;
}

int roo_recurrent(uint64_t n, const float* threshold, uint64_t e, const uint32_t* src, const uint32_t* tgt,
                  const float* w, const uint32_t* delay, uint64_t S, const uint64_t* in_ptr, const uint32_t* in_x,
                  const float* in_w, uint64_t R, const uint32_t* readout, uint64_t H, const uint32_t* hidden,
                  uint32_t iterations, float decay, int32_t topk_ro, int32_t topk_hid, int32_t use_scores,
                  float* scores, uint64_t* spikes, uint64_t* time_count) {
    Csr c;
    if (csr_build(&c, n, e, src, tgt, w, delay)) return -1;
    const uint64_t T = (uint64_t)c.max_delay * (1u + 2u * (uint64_t)iterations) + 1u;
    if (time_count) *time_count = T;
    int rc = 0;
    uint32_t* d2r = dense_map(n, R, readout);
    uint32_t* d2h = dense_map(n, H, hidden);
    float* hid = malloc((T && H ? T * H : 1) * sizeof(float));
    float* ro = malloc((T && R ? T * R : 1) * sizeof(float));
    float* gate = malloc((R ? R : 1) * sizeof(float));
    uint8_t* rsel = malloc(R ? R : 1);
    uint8_t* hsel = malloc(H ? H : 1);
    if (!d2r || !d2h || !hid || !ro || !gate || !rsel || !hsel) {
        rc = -1;
        goto out;
    }
    for (uint64_t s = 0; s < S; ++s) {
        memset(hid, 0, T * H * sizeof(float));
        memset(ro, 0, T * R * sizeof(float));
        /* initialize_recurrent_hidden_events_kernel (:1312-1340) */
        for (uint64_t i = in_ptr[s]; i < in_ptr[s + 1]; ++i) {
            const uint32_t h1 = d2h[in_x[i]];
            if (h1) hid[h1 - 1] += in_w[i];
        }
        for (uint64_t t = 0; t < T; ++t) {
            float* hrow = hid + t * H;
            float* rrow = ro + t * R;
            memset(gate, 0, R * sizeof(float));
            memset(rsel, 0, R);
            memset(hsel, 0, H);
            /* apply_topk_hidden_events_kernel (:1472-1507): only when 0 < k < H */
            if (topk_hid > 0 && (uint64_t)topk_hid < H) {
                for (int32_t pass = 0; pass < topk_hid; ++pass) {
                    float bs;
                    const int64_t b = best_unselected(hrow, hsel, H, &bs);
                    if (b >= 0) hsel[b] = 1;
                }
                for (uint64_t h = 0; h < H; ++h)
                    if (!hsel[h]) hrow[h] = 0.0f;
            }
            /* propagate_hidden_events_to_readouts_kernel (:1342-1381) */
            for (uint64_t h = 0; h < H; ++h) {
                const float v = hrow[h];
                if (v == 0.0f || hidden[h] == RO_NONE) continue;
                const uint32_t x = hidden[h];
                for (uint64_t q = c.ptr[x]; q < c.ptr[x + 1]; ++q) {
                    const uint32_t r1 = d2r[c.tgt[q]];
                    if (r1 && t + c.delay[q] < T) ro[(t + c.delay[q]) * R + r1 - 1] += v * c.w[q];
                }
            }
            /* select_recurrent_readout_events_kernel (:1383-1428) */
            if (topk_ro <= 0) {
                for (uint64_t r = 0; r < R; ++r)
                    if (rrow[r] >= threshold_of(threshold, readout[r])) {
                        gate[r] = use_scores ? rrow[r] : 1.0f;
                        rsel[r] = 1;
                    }
            } else {
                const uint64_t limit = (uint64_t)topk_ro < R ? (uint64_t)topk_ro : R;
                for (uint64_t pass = 0; pass < limit; ++pass) {
                    float bs;
                    const int64_t b = best_unselected(rrow, rsel, R, &bs);
                    if (b >= 0) {
                        rsel[b] = 1;
                        gate[b] = use_scores ? bs : 1.0f;
                    }
                }
            }
            /* propagate_readout_feedback_events_kernel (:1430-1470) */
            for (uint64_t r = 0; r < R; ++r) {
                const float g = gate[r];
                if (g == 0.0f || readout[r] == RO_NONE) continue;
                const uint32_t x = readout[r];
                for (uint64_t q = c.ptr[x]; q < c.ptr[x + 1]; ++q) {
                    const uint32_t h1 = d2h[c.tgt[q]];
                    if (h1 && t + c.delay[q] < T) hid[(t + c.delay[q]) * H + h1 - 1] += decay * g * c.w[q];
                }
            }
        }
        /* sum_recurrent_readout_events_kernel (:1509-1528), then the threshold kernel (:1296-1310) */
        for (uint64_t r = 0; r < R; ++r) {
            float sum = 0.0f;
            for (uint64_t t = 0; t < T; ++t) sum += ro[t * R + r];
            scores[s * R + r] = sum;
            spikes[s * R + r] = sum >= threshold_of(threshold, readout[r]) ? 1u : 0u;
        }
    }
out:
    free(d2r);
    free(d2h);
    free(hid);
    free(ro);
    free(gate);
    free(rsel);
    free(hsel);
    csr_free(&c);
    return rc;
}
