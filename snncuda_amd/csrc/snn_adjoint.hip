// Fixed-spike adjoint of the per-tick step's subthreshold dynamics (no counterpart in the
// reference, which has no gradients).
//
// Input: a recorded run of the plain step (snn_net_record_fire_events): its raster in processing
// order, the index of each firing event within its target's ordered events of the tick, and every
// delivered inject. From these and the connectome the host builds each neuron's events of ticks
// [0, T) in the exact processing order of the step (reference src/runtime/NetworkPropagator.cpp:78-132
// with SpikeScheduler's insertion order): within tick T, injects of stage P sit after the fires of
// tick P-1 and before those of tick P, and the events of one fire follow its outgoing synapses in
// connectome order and each synapse's code offsets in list order. The event a neuron fired on is
// its reset event.
//
// Model per neuron, compartments k = 0..3 (reference src/runtime/SynapseDendriteProcessor.cpp:175-229
// with the reset at the recorded events instead of a threshold test):
//   c_k(t) = A[r_k] * c_k(t-1), A[r] = exp(-1/tau_r), r_k the receptor of compartment k's last input
//   each event of tick t in order: c_comp += w * gain[r] * sign[r]; r_comp = r; reset event: c = 0
//   v(t) = sum_k alpha_k c_k(t)                       (the somatic trace; alpha = 1 is the reference
//                                                      membrane sum, :287-292)
// Every step of this recursion is affine in the state, with coefficients fixed by the event lists
// (the receptor a compartment decays with is that of its last input, known from the events alone).
// Each neuron is one wave and each 64-tick window one pass of it, lane j holding tick t0 + j:
//   forward: lane j composes its tick's events into c -> ea*c + eb per compartment, a wave scan of
//   "last receptor written" gives the receptor each tick decays with, and a wave scan of the affine
//   maps gives c after every tick of the window from the state at its start. The trace is written
//   and the state at every window start (every 64 ticks) is the checkpoint.
//   backward: windows from the last; the window's c(t) is recomputed from its checkpoint the same
//   way, the adjoint lambda enters from the later window and is carried back through the window by a
//   reverse wave scan of its (also affine) per-tick maps:
//     lambda_k += dL/dv(t) * alpha_k; events of tick t in reverse: reset -> lambda = 0;
//     dL/dw_syn += lambda_comp * gain * sign; dL/dalpha_k += dL/dv(t) * c_k(t);
//     dL/dtau_r += lambda_k c_k(t-1) A[r] / tau_r^2 over k with r_k(t-1) = r; lambda_k *= A[r_k(t-1)]
//   Each synaptic event's weight-gradient term is written at its event position; a last kernel sums
//   each synapse's terms in event order (fixed order, no atomics).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/snncuda_hip.h"

void snn_internal_set_error(const std::string& msg);

namespace {

struct AdjError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void acheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw AdjError(std::string(what) + ": " + hipGetErrorString(e));
}

int afail(int code, const std::string& msg) {
    snn_internal_set_error(msg);
    return code;
}

template <typename F>
int aguarded(F&& f) {
    try {
        return f();
    } catch (const AdjError& e) {
        return afail(SNN_EDEVICE, e.what());
    } catch (const std::invalid_argument& e) {
        return afail(SNN_EINVAL, e.what());
    } catch (const std::bad_alloc&) {
        return afail(SNN_EINVAL, "host allocation failed");
    } catch (const std::exception& e) {
        return afail(SNN_EDEVICE, e.what());
    }
}

constexpr uint32_t kSeg = 64;     // window = checkpoint interval (ticks): one lane per tick
constexpr uint32_t kWave = 64;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kBlock = 256;       // reduction block
constexpr int kWavesPerBlock = 4; // neurons per workgroup, one wave each

// One event: the synapse (connectome index) or, for an external inject, the weight's bits; tick;
// meta = compartment | receptor << 2 | external << 4 | reset << 5.
struct AdjEvent {
    uint32_t ref;
    uint16_t tick;
    uint8_t meta;
    uint8_t pad;
};

struct AdjParams {
    float alpha[4];
    float A[4];   // exp(-1/tau_r)
    float dA[4];  // d A / d tau_r = A / tau_r^2
};

__constant__ float kGainSign[4] = {1.0f, 0.55f, -1.0f, -0.65f};

struct AdjArgs {
    uint32_t n, T, nseg;
    const uint64_t* ptr;    // n + 1
    const AdjEvent* ev;
    const float* w;         // synapse weights, connectome order
    float* trace;           // [T][n]
    float* ck;              // [nseg][4][n] state at the start of window s
    uint32_t* ckr;          // [nseg][n] receptor codes (2 bits per compartment) at the start of window s
    uint64_t* ckp;          // [nseg + 1][n] event position at the start of window s
    const float* gtrace;    // [T][n]
    float* gev;             // per event: its weight-gradient term (synaptic events)
    const uint64_t* sptr;   // e + 1: synapse s's events are sev[sptr[s], sptr[s+1]), in event order
    const uint32_t* sev;
    float* gw;              // synapse weight gradients, connectome order
    double* part;           // [n][8] per-neuron dL/dalpha, dL/dtau
};

__device__ __forceinline__ float ev_input(const AdjEvent& e, const float* w) {
    const uint32_t r = (e.meta >> 2) & 3u;
    const float wt = (e.meta & 16u) ? __uint_as_float(e.ref) : w[e.ref];
    return wt * kGainSign[r];
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// bit k of wr set -> the 2-bit receptor field of compartment k
__device__ __forceinline__ uint32_t field_mask(uint32_t wr) {
    return ((wr & 1u) ? 3u : 0u) | ((wr & 2u) ? 12u : 0u) | ((wr & 4u) ? 48u : 0u) | ((wr & 8u) ? 192u : 0u);
}

__device__ __forceinline__ float sel4(const float v[4], uint32_t k) {
    return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : v[3];
}

struct WaveLds {                 // one wave's LDS
    float u[kWave];              // a batch of events: input, meta | tick offset << 8
    uint32_t md[kWave];
    uint32_t d[kWave];
    uint32_t lo[kWave], hi[kWave];  // per tick offset: its events' range within the batch
    float lam[kWave][4];         // backward: per tick, lambda seen by events after its last reset
    uint32_t lastr[kWave];       // backward: 1 + window position of the tick's last reset (0: none)
};

// The events of tick t0 + lane as c_k -> ea_k * c_k + eb_k (decay excluded), the compartments they
// write (wr) with their last receptors (lr), and the window position of the last reset (+1, 0: none).
struct TickMap {
    float ea[4], eb[4];
    uint32_t wr, lr, lastr;
};

// Reads the window's events from position p (at most up to end) in batches of 64, one event per
// lane; lane j applies the batch's events of tick t0 + j in order. kPrefetch batches (and their
// weights) are requested at once, so a window costs one or two round trips to memory rather than
// one per batch; batches past the window's end are read and not used. Returns the position after
// the window's events.
constexpr uint32_t kPrefetch = 4;
__device__ uint64_t window_maps(const AdjArgs& a, uint64_t p, uint64_t end, uint32_t t0, uint32_t tw, uint32_t lane,
                                WaveLds& L, TickMap& m) {
    for (int k = 0; k < 4; ++k) {
        m.ea[k] = 1.0f;
        m.eb[k] = 0.0f;
    }
    m.wr = 0;
    m.lr = 0;
    m.lastr = 0;
    const uint64_t pw = p;
    bool more = true;
    while (more) {
        AdjEvent eb[kPrefetch];
        float ub[kPrefetch];
#pragma unroll
        for (uint32_t b = 0; b < kPrefetch; ++b) {
            const uint64_t q = p + b * kWave + lane;
            eb[b] = AdjEvent{};
            if (q < end) eb[b] = a.ev[q];
        }
#pragma unroll
        for (uint32_t b = 0; b < kPrefetch; ++b) ub[b] = p + b * kWave + lane < end ? ev_input(eb[b], a.w) : 0.0f;
#pragma unroll
        for (uint32_t b = 0; b < kPrefetch; ++b) {
            if (!more) break;  // uniform: cnt below is a wave-wide count
            // every batch before this one was full, so p is the round's start + b * 64
            const uint64_t q = p + lane;
            uint32_t d = kWave;  // not in the window
            if (q < end) d = (uint32_t)eb[b].tick - t0;  // events before the window were consumed: tick >= t0
            const bool in = d < tw;
            const uint32_t cnt = (uint32_t)__popcll(__ballot(in));  // a prefix: events are sorted by tick
            L.d[lane] = in ? d : kWave;
            L.lo[lane] = 0;
            L.hi[lane] = 0;
            if (in) {
                L.u[lane] = ub[b];
                L.md[lane] = eb[b].meta;
            }
            wave_sync();
            if (in) {
                if (lane == 0 || L.d[lane - 1] != d) L.lo[d] = lane;
                if (lane + 1 == kWave || L.d[lane + 1] != d) L.hi[d] = lane + 1;
            }
            wave_sync();
            const uint32_t lo = L.lo[lane], hi = L.hi[lane];
            for (uint32_t x = lo; x < hi; ++x) {
                const uint32_t md = L.md[x];
                const float u = L.u[x];
                const uint32_t k = md & 3u, r = (md >> 2) & 3u;
                m.eb[0] = k == 0 ? m.eb[0] + u : m.eb[0];
                m.eb[1] = k == 1 ? m.eb[1] + u : m.eb[1];
                m.eb[2] = k == 2 ? m.eb[2] + u : m.eb[2];
                m.eb[3] = k == 3 ? m.eb[3] + u : m.eb[3];
                m.wr |= 1u << k;
                m.lr = (m.lr & ~(3u << (2 * k))) | (r << (2 * k));
                if (md & 32u) {
                    for (int c = 0; c < 4; ++c) {
                        m.ea[c] = 0.0f;
                        m.eb[c] = 0.0f;
                    }
                    m.lastr = (uint32_t)(p + x - pw) + 1;
                }
            }
            wave_sync();
            p += cnt;
            if (cnt < kWave) more = false;
        }
    }
    return p;
}

// receptor codes before tick t0 + lane: r_start updated by the writes of the ticks before it
__device__ __forceinline__ uint32_t receptors_in(const TickMap& m, uint32_t r_start, uint32_t lane) {
    uint32_t swr = m.wr, slr = m.lr;
    for (uint32_t off = 1; off < kWave; off <<= 1) {
        const uint32_t owr = __shfl_up(swr, off, kWave), olr = __shfl_up(slr, off, kWave);
        if (lane >= off) {
            const uint32_t M = field_mask(swr);
            slr = (olr & ~M) | (slr & M);
            swr |= owr;
        }
    }
    uint32_t pwr = __shfl_up(swr, 1, kWave), plr = __shfl_up(slr, 1, kWave);
    if (lane == 0) pwr = plr = 0;
    const uint32_t M = field_mask(pwr);
    return (r_start & ~M) | (plr & M);
}

// c after tick t0 + lane, from the state c0 at the window start
__device__ __forceinline__ void states_after(const TickMap& m, const AdjParams& P, uint32_t r_in, bool decays,
                                             const float c0[4], uint32_t lane, float c[4]) {
    float am[4], bm[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float D = decays ? P.A[(r_in >> (2 * k)) & 3u] : 1.0f;
        am[k] = m.ea[k] * D;
        bm[k] = m.eb[k];
    }
    for (uint32_t off = 1; off < kWave; off <<= 1) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float oa = __shfl_up(am[k], off, kWave), ob = __shfl_up(bm[k], off, kWave);
            if (lane >= off) {
                bm[k] = am[k] * ob + bm[k];
                am[k] = am[k] * oa;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = am[k] * c0[k] + bm[k];
}

__global__ void __launch_bounds__(kWavesPerBlock * kWave) k_adj_forward(AdjArgs a, AdjParams P) {
    __shared__ WaveLds lds[kWavesPerBlock];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wv = threadIdx.x / kWave;
    const uint32_t i = blockIdx.x * kWavesPerBlock + wv;
    if (i >= a.n) return;  // whole wave; no block barriers below
    WaveLds& L = lds[wv];
    float c0[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t r0 = 0;
    uint64_t p = a.ptr[i];
    const uint64_t end = a.ptr[i + 1];
    for (uint32_t s = 0; s < a.nseg; ++s) {
        const uint32_t t0 = s * kSeg, tw = min(kSeg, a.T - t0);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) a.ck[((uint64_t)s * 4 + k) * a.n + i] = c0[k];
            a.ckr[(uint64_t)s * a.n + i] = r0;
            a.ckp[(uint64_t)s * a.n + i] = p;
        }
        TickMap m;
        p = window_maps(a, p, end, t0, tw, lane, L, m);
        const uint32_t r_in = receptors_in(m, r0, lane);
        float c[4];
        states_after(m, P, r_in, t0 + lane > 0, c0, lane, c);
        if (lane < tw)
            a.trace[(uint64_t)(t0 + lane) * a.n + i] =
                P.alpha[0] * c[0] + P.alpha[1] * c[1] + P.alpha[2] * c[2] + P.alpha[3] * c[3];
        const uint32_t r_out = (r_in & ~field_mask(m.wr)) | (m.lr & field_mask(m.wr));
#pragma unroll
        for (int k = 0; k < 4; ++k) c0[k] = __shfl(c[k], (int)(tw - 1), kWave);
        r0 = __shfl(r_out, (int)(tw - 1), kWave);
    }
    if (lane == 0) a.ckp[(uint64_t)a.nseg * a.n + i] = p;
}

__global__ void __launch_bounds__(kWavesPerBlock * kWave) k_adj_backward(AdjArgs a, AdjParams P) {
    __shared__ WaveLds lds[kWavesPerBlock];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wv = threadIdx.x / kWave;
    const uint32_t i = blockIdx.x * kWavesPerBlock + wv;
    if (i >= a.n) return;  // whole wave; no block barriers below
    WaveLds& L = lds[wv];
    float carry[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // lambda entering the window's last tick
    float galpha[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float gtau[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t s = a.nseg; s-- > 0;) {
        const uint32_t t0 = s * kSeg, tw = min(kSeg, a.T - t0);
        const uint32_t t = t0 + lane;
        const bool live = lane < tw;
        float c0[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c0[k] = a.ck[((uint64_t)s * 4 + k) * a.n + i];
        const uint32_t r0 = a.ckr[(uint64_t)s * a.n + i];
        const uint64_t pbeg = a.ckp[(uint64_t)s * a.n + i], pend = a.ckp[(uint64_t)(s + 1) * a.n + i];
        // the window's forward states
        TickMap m;
        window_maps(a, pbeg, pend, t0, tw, lane, L, m);
        const uint32_t r_in = receptors_in(m, r0, lane);
        float c[4], cp[4];
        states_after(m, P, r_in, t > 0, c0, lane, c);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cp[k] = __shfl_up(c[k], 1, kWave);
            if (lane == 0) cp[k] = c0[k];
        }
        // lambda maps: lambda after tick t = mm * (lambda entering tick t) + qq
        const float g = live ? a.gtrace[(uint64_t)t * a.n + i] : 0.0f;
        const bool reset = m.lastr != 0;
        float mm[4], qq[4], D[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            D[k] = t > 0 ? P.A[(r_in >> (2 * k)) & 3u] : 1.0f;
            mm[k] = !live ? 1.0f : reset ? 0.0f : D[k];
            qq[k] = (live && !reset) ? D[k] * (g * P.alpha[k]) : 0.0f;
        }
        for (uint32_t off = 1; off < kWave; off <<= 1) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float om = __shfl_down(mm[k], off, kWave), oq = __shfl_down(qq[k], off, kWave);
                if (lane + off < kWave) {
                    qq[k] = mm[k] * oq + qq[k];
                    mm[k] = mm[k] * om;
                }
            }
        }
        float lin[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float lout = mm[k] * carry[k] + qq[k];
            lin[k] = __shfl_down(lout, 1, kWave);
            if (lane + 1 == kWave) lin[k] = carry[k];
            const float l0 = __shfl(lout, 0, kWave);
            carry[k] = l0;
        }
        if (live) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float pre = lin[k] + g * P.alpha[k];
                L.lam[lane][k] = pre;
                galpha[k] += g * c[k];
                if (t > 0) {
                    const uint32_t r = (r_in >> (2 * k)) & 3u;
                    const float gt = (reset ? 0.0f : pre) * cp[k] * P.dA[r];
                    gtau[0] += r == 0 ? gt : 0.0f;
                    gtau[1] += r == 1 ? gt : 0.0f;
                    gtau[2] += r == 2 ? gt : 0.0f;
                    gtau[3] += r == 3 ? gt : 0.0f;
                }
            }
            L.lastr[lane] = m.lastr;
        }
        wave_sync();
        // weight-gradient terms of the window's synaptic events
        for (uint64_t b = pbeg; b < pend; b += kWave) {
            const uint64_t q = b + lane;
            if (q < pend) {
                const AdjEvent e = a.ev[q];
                if (!(e.meta & 16u)) {
                    const uint32_t d = (uint32_t)e.tick - t0;
                    const uint32_t k = e.meta & 3u, r = (e.meta >> 2) & 3u;
                    const float lam = (uint32_t)(q - pbeg) + 1 > L.lastr[d] ? sel4(L.lam[d], k) : 0.0f;
                    a.gev[q] = lam * kGainSign[r];
                }
            }
        }
        wave_sync();
    }
    // per-neuron partials: lane sums in FP64, fixed order
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double v = k < 4 ? (double)galpha[k] : (double)gtau[k - 4];
        for (uint32_t off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
        if (lane == 0) a.part[(uint64_t)i * 8 + k] = v;
    }
}

// each synapse's weight gradient: the sum of its events' terms in event order
__global__ void __launch_bounds__(kBlock) k_adj_weight_grad(const uint64_t* sptr, const uint32_t* sev, const float* gev,
                                                            uint64_t e, float* gw) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < e; j += (uint64_t)gridDim.x * blockDim.x) {
        float acc = 0.0f;
        for (uint64_t q = sptr[j]; q < sptr[j + 1]; ++q) acc += gev[sev[q]];
        gw[j] = acc;
    }
}

// sum of the per-neuron partials in FP64, one block, fixed order: thread t sums neurons t, t + 1024,
// ... in order, then a pairwise tree over the threads (the same shape every call)
constexpr int kReduceBlock = 1024;
__global__ void __launch_bounds__(kReduceBlock) k_adj_reduce(const double* part, uint32_t n, double* out) {
    __shared__ double sm[kReduceBlock][8];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        for (int q = 0; q < 8; ++q) acc[q] += part[(uint64_t)i * 8 + q];
    for (int q = 0; q < 8; ++q) sm[threadIdx.x][q] = acc[q];
    __syncthreads();
    for (uint32_t h = kReduceBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int q = 0; q < 8; ++q) sm[threadIdx.x][q] += sm[threadIdx.x + h][q];
        __syncthreads();
    }
    if (threadIdx.x < 8) out[threadIdx.x] = sm[0][threadIdx.x];
}

struct DBuf {
    void* p = nullptr;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
};

template <typename T>
T* dalloc(std::vector<std::unique_ptr<DBuf>>& own, size_t count) {
    auto b = std::make_unique<DBuf>();
    acheck(hipMalloc(&b->p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
    T* p = (T*)b->p;
    own.push_back(std::move(b));
    return p;
}

}  // namespace

struct snn_adjoint {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0, T = 0, nseg = 0;
    uint64_t e = 0, nev = 0;
    std::vector<std::unique_ptr<DBuf>> own;
    AdjArgs a{};
    AdjParams last{};
    double* sums = nullptr;  // the eight parameter gradients
    bool have_forward = false;
    std::vector<uint64_t> h_ptr;
    std::vector<AdjEvent> h_ev;
    ~snn_adjoint() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
            own.clear();
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace {

AdjParams make_params(const float* params) {
    AdjParams P{};
    for (int k = 0; k < 4; ++k) {
        P.alpha[k] = params[k];
        const double tau = params[4 + k];
        if (!(tau > 0.0)) throw std::invalid_argument("time constants must be positive");
        const double A = std::exp(-1.0 / tau);
        P.A[k] = (float)A;
        P.dA[k] = (float)(A / (tau * tau));
    }
    return P;
}

}  // namespace

extern "C" {

int snn_adjoint_create(snn_net* run, const snn_net_desc* desc, uint64_t ticks, snn_adjoint** out) {
    if (!run || !desc || !out) return afail(SNN_EINVAL, "null argument");
    *out = nullptr;
    return aguarded([&]() -> int {
        const uint64_t n = desc->n_neurons, e = desc->n_synapses;
        if (ticks == 0 || ticks > 65535) return afail(SNN_EINVAL, "ticks must be in [1, 65535]");
        if (n == 0 || n >= 0xFFFFFFFFull) return afail(SNN_EINVAL, "bad neuron count");
        uint64_t delivered = 0, fired = 0, next = 0;
        int rc = snn_net_counts(run, &delivered, &fired, &next);
        if (rc) return rc;
        if (next < ticks) return afail(SNN_EINVAL, "the run has not processed the requested ticks");
        // raster, firing-event indices, per-tick fire counts, injects
        uint64_t nf = 0;
        rc = snn_net_read_raster(run, &nf, nullptr, 0);
        if (rc) return rc;
        std::vector<uint32_t> fire_neuron(nf), fire_evt(nf);
        rc = snn_net_read_raster(run, &nf, fire_neuron.data(), nf);
        if (rc) return rc;
        uint64_t nf2 = 0;
        rc = snn_net_read_fire_events(run, &nf2, fire_evt.data(), nf);
        if (rc) return rc;
        if (nf2 != nf) return afail(SNN_EINVAL, "raster and fire events differ in length");
        std::vector<uint64_t> tick_fired(next);
        rc = snn_net_read_tick_counts(run, 0, next, nullptr, tick_fired.data());
        if (rc) return rc;
        uint64_t ni = 0;
        rc = snn_net_read_inject_log(run, &ni, nullptr, nullptr, nullptr, nullptr, 0);
        if (rc) return rc;
        std::vector<uint32_t> inj_neuron(ni);
        std::vector<uint64_t> inj_tick(ni), inj_stage(ni);
        std::vector<float> inj_w(ni);
        rc = snn_net_read_inject_log(run, &ni, inj_neuron.data(), inj_tick.data(), inj_stage.data(), inj_w.data(), ni);
        if (rc) return rc;

        // outgoing synapses per source, connectome order
        std::vector<uint64_t> optr(n + 1, 0);
        for (uint64_t j = 0; j < e; ++j) {
            if (desc->source[j] >= n || desc->target[j] >= n) throw std::invalid_argument("synapse endpoint out of range");
            optr[desc->source[j] + 1]++;
        }
        for (uint64_t s = 0; s < n; ++s) optr[s + 1] += optr[s];
        std::vector<uint32_t> oidx(e);
        {
            std::vector<uint64_t> fill(optr.begin(), optr.end() - 1);
            for (uint64_t j = 0; j < e; ++j) oidx[fill[desc->source[j]]++] = (uint32_t)j;
        }
        static const uint32_t zero = 0;
        auto codes_of = [&](uint64_t j, const uint32_t*& c) -> uint32_t {
            if (desc->code_ptr && desc->code_ptr[j + 1] > desc->code_ptr[j]) {
                c = desc->codes + desc->code_ptr[j];
                return (uint32_t)(desc->code_ptr[j + 1] - desc->code_ptr[j]);
            }
            c = &zero;
            return 1;
        };

        // events in emission order: [injects of stage P][fires of tick P] for P = 0, 1, ...
        struct Raw {
            uint32_t target, tick;
            AdjEvent ev;
        };
        std::vector<Raw> raw;
        std::vector<uint64_t> order(ni);
        for (uint64_t q = 0; q < ni; ++q) order[q] = q;
        std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return inj_stage[x] < inj_stage[y]; });
        uint64_t iq = 0, f = 0;
        std::vector<uint64_t> fire_tick(nf);
        for (uint64_t P = 0; P < next && P < ticks; ++P) {
            for (; iq < ni && inj_stage[order[iq]] <= P; ++iq) {
                const uint64_t q = order[iq];
                if (inj_tick[q] >= ticks) continue;
                AdjEvent ev{};
                std::memcpy(&ev.ref, &inj_w[q], 4);
                ev.tick = (uint16_t)inj_tick[q];
                ev.meta = (uint8_t)(0u | (0u << 2) | 16u);  // soma, AMPA, external
                raw.push_back({inj_neuron[q], (uint32_t)inj_tick[q], ev});
            }
            const uint64_t fend = f + tick_fired[P];
            if (fend > nf) return afail(SNN_EINVAL, "raster shorter than the fire counts");
            for (; f < fend; ++f) {
                fire_tick[f] = P;
                const uint32_t s = fire_neuron[f];
                for (uint64_t q = optr[s]; q < optr[s + 1]; ++q) {
                    const uint32_t j = oidx[q];
                    const uint32_t* c;
                    const uint32_t nc = codes_of(j, c);
                    for (uint32_t k = 0; k < nc; ++k) {
                        const uint64_t lag = std::max<uint64_t>(1, (uint64_t)desc->delay[j] + c[k]);
                        const uint64_t T = P + lag;
                        if (T >= ticks) continue;
                        AdjEvent ev{};
                        ev.ref = j;
                        ev.tick = (uint16_t)T;
                        ev.meta = (uint8_t)((desc->compartment[j] & 3u) | ((desc->receptor[j] & 3u) << 2));
                        raw.push_back({desc->target[j], (uint32_t)T, ev});
                    }
                }
            }
        }
        // per neuron, ordered by tick, emission order within a tick
        std::unique_ptr<snn_adjoint> adj(new snn_adjoint());
        adj->h_ptr.assign(n + 1, 0);
        for (const Raw& r : raw) adj->h_ptr[r.target + 1]++;
        for (uint64_t i = 0; i < n; ++i) adj->h_ptr[i + 1] += adj->h_ptr[i];
        adj->h_ev.resize(raw.size());
        std::vector<uint32_t> ev_tick(raw.size());
        {
            std::vector<uint64_t> fill(adj->h_ptr.begin(), adj->h_ptr.end() - 1);
            for (const Raw& r : raw) {
                const uint64_t at = fill[r.target]++;
                adj->h_ev[at] = r.ev;
                ev_tick[at] = r.tick;
            }
        }
        raw.clear();
        raw.shrink_to_fit();
        std::vector<AdjEvent> tmp;
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t lo = adj->h_ptr[i], hi = adj->h_ptr[i + 1];
            std::vector<uint64_t> idx(hi - lo);
            for (uint64_t q = 0; q < hi - lo; ++q) idx[q] = lo + q;
            std::stable_sort(idx.begin(), idx.end(), [&](uint64_t x, uint64_t y) { return ev_tick[x] < ev_tick[y]; });
            tmp.resize(hi - lo);
            for (uint64_t q = 0; q < hi - lo; ++q) tmp[q] = adj->h_ev[idx[q]];
            std::copy(tmp.begin(), tmp.end(), adj->h_ev.begin() + lo);
        }
        // reset events: fire f of neuron x at tick P happened on x's event fire_evt[f] of tick P
        {
            for (uint64_t q = 0; q < nf; ++q) {
                if (fire_tick[q] >= ticks) continue;
                const uint32_t x = fire_neuron[q];
                const uint64_t lo = adj->h_ptr[x], hi = adj->h_ptr[x + 1];
                // first event of tick P (binary search on the sorted ticks)
                uint64_t a = lo, b = hi;
                while (a < b) {
                    const uint64_t m = (a + b) / 2;
                    if (adj->h_ev[m].tick < fire_tick[q]) a = m + 1; else b = m;
                }
                const uint64_t at = a + fire_evt[q];
                if (at >= hi || adj->h_ev[at].tick != fire_tick[q])
                    return afail(SNN_EINVAL, "a recorded fire does not match its neuron's events");
                adj->h_ev[at].meta |= 32u;
            }
        }
        // device side
        adj->n = (uint32_t)n;
        adj->e = e;
        adj->T = (uint32_t)ticks;
        adj->nseg = (uint32_t)((ticks + kSeg - 1) / kSeg);
        adj->nev = adj->h_ev.size();
        int dev = 0;
        acheck(hipGetDevice(&dev), "hipGetDevice");
        adj->device = dev;
        acheck(hipStreamCreate(&adj->stream), "hipStreamCreate");
        AdjArgs& A = adj->a;
        A.n = adj->n;
        A.T = adj->T;
        A.nseg = adj->nseg;
        uint64_t* dptr = dalloc<uint64_t>(adj->own, n + 1);
        AdjEvent* dev_ev = dalloc<AdjEvent>(adj->own, adj->nev);
        acheck(hipMemcpy(dptr, adj->h_ptr.data(), (n + 1) * 8, hipMemcpyHostToDevice), "upload");
        if (adj->nev) acheck(hipMemcpy(dev_ev, adj->h_ev.data(), adj->nev * sizeof(AdjEvent), hipMemcpyHostToDevice), "upload");
        A.ptr = dptr;
        A.ev = dev_ev;
        A.ck = dalloc<float>(adj->own, (uint64_t)adj->nseg * 4 * n);
        A.ckr = dalloc<uint32_t>(adj->own, (uint64_t)adj->nseg * n);
        A.ckp = dalloc<uint64_t>(adj->own, (uint64_t)(adj->nseg + 1) * n);
        A.part = dalloc<double>(adj->own, (uint64_t)n * 8);
        A.gev = dalloc<float>(adj->own, adj->nev);
        // each synapse's events, in event order
        {
            std::vector<uint64_t> sptr(e + 1, 0);
            for (const AdjEvent& ev : adj->h_ev)
                if (!(ev.meta & 16u)) sptr[ev.ref + 1]++;
            for (uint64_t j = 0; j < e; ++j) sptr[j + 1] += sptr[j];
            std::vector<uint32_t> sev(sptr[e]);
            std::vector<uint64_t> fill(sptr.begin(), sptr.end() - 1);
            for (uint64_t q = 0; q < adj->nev; ++q)
                if (!(adj->h_ev[q].meta & 16u)) sev[fill[adj->h_ev[q].ref]++] = (uint32_t)q;
            uint64_t* dsp = dalloc<uint64_t>(adj->own, e + 1);
            uint32_t* dse = dalloc<uint32_t>(adj->own, sev.size());
            acheck(hipMemcpy(dsp, sptr.data(), (e + 1) * 8, hipMemcpyHostToDevice), "upload");
            if (!sev.empty()) acheck(hipMemcpy(dse, sev.data(), sev.size() * 4, hipMemcpyHostToDevice), "upload");
            A.sptr = dsp;
            A.sev = dse;
        }
        adj->sums = dalloc<double>(adj->own, 8);
        *out = adj.release();
        return SNN_OK;
    });
}

int snn_adjoint_destroy(snn_adjoint* adj) {
    delete adj;
    return SNN_OK;
}

int snn_adjoint_sizes(snn_adjoint* adj, uint64_t* n_neurons, uint64_t* n_synapses, uint64_t* ticks, uint64_t* n_events) {
    if (!adj) return afail(SNN_EINVAL, "null handle");
    if (n_neurons) *n_neurons = adj->n;
    if (n_synapses) *n_synapses = adj->e;
    if (ticks) *ticks = adj->T;
    if (n_events) *n_events = adj->nev;
    return SNN_OK;
}

int snn_adjoint_read_events(snn_adjoint* adj, uint64_t* ptr, uint32_t* tick, uint32_t* synapse, float* weight,
                            uint8_t* reset) {
    if (!adj) return afail(SNN_EINVAL, "null handle");
    if (ptr) std::copy(adj->h_ptr.begin(), adj->h_ptr.end(), ptr);
    for (uint64_t q = 0; q < adj->nev; ++q) {
        const AdjEvent& e = adj->h_ev[q];
        const bool ext = (e.meta & 16u) != 0;
        if (tick) tick[q] = e.tick;
        if (synapse) synapse[q] = ext ? kNone : e.ref;
        if (weight) {
            float w = 0.0f;
            if (ext) std::memcpy(&w, &e.ref, 4);
            weight[q] = w;
        }
        if (reset) reset[q] = (e.meta & 32u) ? 1 : 0;
    }
    return SNN_OK;
}

int snn_adjoint_forward(snn_adjoint* adj, const float* d_weight, const float* params, float* d_trace) {
    if (!adj || !params || !d_trace || (adj->e && !d_weight)) return afail(SNN_EINVAL, "null argument");
    return aguarded([&]() -> int {
        acheck(hipSetDevice(adj->device), "hipSetDevice");
        const AdjParams P = make_params(params);
        AdjArgs a = adj->a;
        a.w = d_weight;
        a.trace = d_trace;
        const dim3 grid((adj->n + kWavesPerBlock - 1) / kWavesPerBlock);
        hipLaunchKernelGGL(k_adj_forward, grid, dim3(kWavesPerBlock * kWave), 0, adj->stream, a, P);
        acheck(hipGetLastError(), "k_adj_forward");
        acheck(hipStreamSynchronize(adj->stream), "adjoint forward");
        adj->last = P;
        adj->a.w = d_weight;
        adj->have_forward = true;
        return SNN_OK;
    });
}

int snn_adjoint_backward(snn_adjoint* adj, const float* d_grad_trace, float* d_grad_weight, double* grad_params) {
    if (!adj || !d_grad_trace || !grad_params || (adj->e && !d_grad_weight)) return afail(SNN_EINVAL, "null argument");
    if (!adj->have_forward) return afail(SNN_EINVAL, "backward needs a forward first");
    return aguarded([&]() -> int {
        acheck(hipSetDevice(adj->device), "hipSetDevice");
        AdjArgs a = adj->a;
        a.gtrace = d_grad_trace;
        a.gw = d_grad_weight;
        const dim3 grid((adj->n + kWavesPerBlock - 1) / kWavesPerBlock);
        hipLaunchKernelGGL(k_adj_backward, grid, dim3(kWavesPerBlock * kWave), 0, adj->stream, a, adj->last);
        acheck(hipGetLastError(), "k_adj_backward");
        if (adj->e) {
            const uint64_t nb = std::min<uint64_t>((adj->e + kBlock - 1) / kBlock, 8192);
            hipLaunchKernelGGL(k_adj_weight_grad, dim3((uint32_t)nb), dim3(kBlock), 0, adj->stream, a.sptr, a.sev,
                               (const float*)a.gev, adj->e, d_grad_weight);
            acheck(hipGetLastError(), "k_adj_weight_grad");
        }
        hipLaunchKernelGGL(k_adj_reduce, dim3(1), dim3(kReduceBlock), 0, adj->stream, (const double*)a.part, adj->n,
                           adj->sums);
        acheck(hipMemcpyAsync(grad_params, adj->sums, 8 * sizeof(double), hipMemcpyDeviceToHost, adj->stream), "read");
        acheck(hipStreamSynchronize(adj->stream), "adjoint backward");
        return SNN_OK;
    });
}

}  // extern "C"
