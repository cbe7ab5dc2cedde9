// Non-spiking readouts of a connectome: the feedforward projection and the recurrent readout /
// feedback timeline of CudaInferenceSession::run_feedforward_batch and
// run_recurrent_feedforward_batch (reference src/backends/cuda/CudaBackend.cu:1260-1528 kernels,
// :1725-2100 host sequence).
//
// The connectome is held once per graph handle as a source-major CSR in connectome order within a
// source (target, weight, delay = max(1, delay_ticks)), plus neuron thresholds.
//
// One wave per sample. A sample's work is a sequence of (source, coefficient) lists walked edge by
// edge: for every source in list order, its out-edges in connectome order, i.e. a sequential sum
// in the order a single thread would take. The wave flattens the (source, edge) pairs of 64 sources
// at a time (a wave scan of their out-degrees) and takes 64 consecutive pairs per step, one per
// lane, so short rows (a pixel with ten class edges) do not leave lanes idle. Adds land in the
// sample's own rows (LDS for a feedforward score row of at most kLdsRow readouts, else device memory),
// written only by that wave, so no two samples touch one cell and the sums need no cross-wave order.
//
// Feedforward: score[r] = sum over merged inputs (x, w_in) and edges x -> readout r of w_in * w_edge;
// spike = score >= threshold of r (1.0 for a readout that is not a neuron of the connectome).
//
// Recurrent, time_count T = max_delay * (1 + 2 * iterations) + 1 steps; per step t:
//   1. top-k hidden (0 < k < H): keep the k largest hidden scores of row t (ties: lower index),
//      zero the rest;
//   2. every hidden h with a nonzero row-t score s: for each edge to readout r with t + d < T,
//      readout[t + d][r] += s * w;
//   3. gates of row t: top-k readouts (k > 0; ties: lower index) with gate = score (or 1), or every
//      readout whose score reaches its threshold (k <= 0);
//   4. every readout with a nonzero gate g: for each edge to hidden h with t + d < T,
//      hidden[t + d][h] += (decay * g) * w.
// Final score[r] = sum over t of readout[t][r]; spike = score >= threshold. Row 0 of the hidden
// timeline holds the merged input weights of the hidden neurons among the inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/snncuda_hip.h"

void snn_internal_set_error(const std::string& msg);

namespace {

struct RoError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void rcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw RoError(std::string(what) + ": " + hipGetErrorString(e));
}

int rfail(int code, const std::string& msg) {
    snn_internal_set_error(msg);
    return code;
}

template <typename F>
int rguarded(F&& f) {
    try {
        return f();
    } catch (const RoError& e) {
        return rfail(SNN_EDEVICE, e.what());
    } catch (const std::invalid_argument& e) {
        return rfail(SNN_EINVAL, e.what());
    } catch (const std::bad_alloc&) {
        return rfail(SNN_EINVAL, "host allocation failed");
    } catch (const std::exception& e) {
        return rfail(SNN_EDEVICE, e.what());
    }
}

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// feedforward score rows of at most kLdsRow readouts accumulate in LDS (4 rows of 8 KB per block)
constexpr uint32_t kLdsRow = 2048;

struct GraphDev {
    uint32_t n;
    uint64_t e;
    const uint64_t* out_ptr;  // n + 1
    const uint32_t* tgt;      // e, by source, connectome order within a source
    const float* w;
    const uint32_t* delay;    // max(1, delay_ticks)
    const float* threshold;   // n
};

// per wave: the current 64 sources' first edge, pair offset and coefficient
struct WaveStage {
    uint64_t e0[kWave];
    uint32_t off[kWave + 1];
    float c[kWave];
};

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

// LDS written by some lanes of the wave is read by others: order the wave's LDS accesses
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rows that receive device-memory atomics are read around the per-CU vector cache.
__device__ __forceinline__ float load_row(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_row(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t load_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_u32(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// the wave's earlier device-memory writes and atomics complete (every counter drained) before its
// next reads
__device__ __forceinline__ void wave_mem_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
}

// Walk the (source, edge) pairs of n sources in order, 64 pairs per step. get(i, x, c) gives
// source i's neuron (kNone: no edges) and coefficient; visit(c, e, ok) is called by every lane of
// every step (ok: the lane holds pair (i, e) with coefficient c).
template <typename Get, typename Visit>
__device__ void wave_pairs(const GraphDev& G, uint32_t n, WaveStage& L, Get&& get, Visit&& visit) {
    const uint32_t lane = lane_id();
    for (uint32_t i0 = 0; i0 < n; i0 += kWave) {
        const uint32_t i = i0 + lane;
        uint32_t x = kNone;
        float c = 0.0f;
        if (i < n) get(i, x, c);
        uint64_t e0 = 0;
        uint32_t deg = 0;
        if (x != kNone) {
            e0 = G.out_ptr[x];
            deg = (uint32_t)(G.out_ptr[x + 1] - e0);
        }
        const uint32_t incl = wave_incl_scan(deg);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        L.off[lane] = incl - deg;
        L.e0[lane] = e0;
        L.c[lane] = c;
        if (lane == 0) L.off[kWave] = total;
        wave_lds_sync();
        for (uint32_t p0 = 0; p0 < total; p0 += kWave) {
            const uint32_t p = p0 + lane;
            const bool ok = p < total;
            // the covering source: the last k with off[k] <= p (sources without edges share the
            // offset of the next one and are passed over)
            uint32_t k = 0;
#pragma unroll
            for (uint32_t s = kWave / 2; s > 0; s >>= 1)
                if (L.off[k + s] <= p) k += s;
            visit(L.c[k], ok ? L.e0[k] + (p - L.off[k]) : 0ull, ok);
        }
        wave_lds_sync();
    }
}

struct FFArgs {
    uint32_t S, R;
    const uint64_t* in_ptr;  // S + 1
    const uint32_t* in_x;
    const float* in_w;
    const uint32_t* readout;  // R: dense neuron or kNone
    const uint32_t* d2r;      // n: readout index + 1 (0: not a readout)
    float* scores;            // S * R (zeroed when rows live in device memory)
    uint64_t* spikes;         // S * R
};

template <bool LDS_ROW>
__global__ void __launch_bounds__(kBlock) k_feedforward(GraphDev G, FFArgs A) {
    __shared__ WaveStage stage[kWavesPerBlock];
    __shared__ float rows[LDS_ROW ? kWavesPerBlock * kLdsRow : 1];
    const uint32_t wv = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint64_t s = (uint64_t)blockIdx.x * kWavesPerBlock + wv;
    if (s >= A.S) return;  // whole waves leave; no block barrier follows
    float* row = LDS_ROW ? rows + wv * kLdsRow : A.scores + s * A.R;
    if (LDS_ROW) {
        for (uint32_t r = lane; r < A.R; r += kWave) row[r] = 0.0f;
        wave_lds_sync();
    }
    const uint64_t a = A.in_ptr[s];
    const uint32_t n = (uint32_t)(A.in_ptr[s + 1] - a);
    wave_pairs(
        G, n, stage[wv],
        [&](uint32_t i, uint32_t& x, float& c) {
            x = A.in_x[a + i];
            c = A.in_w[a + i];
        },
        [&](float c, uint64_t e, bool ok) {
            if (!ok) return;
            const uint32_t r1 = A.d2r[G.tgt[e]];
            if (!r1) return;
            const float v = c * G.w[e];
            if (LDS_ROW) atomicAdd(&row[r1 - 1], v);
            else unsafeAtomicAdd(&row[r1 - 1], v);
        });
    if (LDS_ROW) wave_lds_sync();
    else wave_mem_sync();
    for (uint32_t r = lane; r < A.R; r += kWave) {
        const float v = LDS_ROW ? row[r] : load_row(&row[r]);
        const uint32_t x = A.readout[r];
        const float thr = x != kNone ? G.threshold[x] : 1.0f;
        if (LDS_ROW) A.scores[s * A.R + r] = v;
        A.spikes[s * A.R + r] = v >= thr ? 1ull : 0ull;
    }
}

struct RecArgs {
    uint32_t S, R, H, T;
    const uint64_t* in_ptr;
    const uint32_t* in_x;
    const float* in_w;
    const uint32_t* readout;  // R: dense neuron or kNone
    const uint32_t* hidden;   // H: dense neuron or kNone
    const uint32_t* d2r;      // n: readout index + 1
    const uint32_t* d2h;      // n: hidden index + 1
    float* hid;               // S * T * H, zeroed
    float* ro;                // S * T * R, zeroed
    float* gate;              // S * R
    uint32_t* hsel;           // S * H: step + 1 when selected at that step, zeroed
    uint32_t* rsel;           // S * R, zeroed
    int32_t topk_ro, topk_hid;
    int32_t use_scores;
    float decay;
    float* scores;            // S * R
    uint64_t* spikes;         // S * R
};

// Index of the largest value among the entries [0, m) not yet selected at this step (stamp), ties to
// the lower index, and value > -FLT_MAX as the reference's passes start from; kNone when there is
// none. Every lane returns the same answer.
__device__ uint32_t wave_argmax(const float* row, const uint32_t* sel, uint32_t m, uint32_t stamp, float& best_v) {
    const uint32_t lane = lane_id();
    float bv = -FLT_MAX;
    uint32_t bi = kNone;
    for (uint32_t j = lane; j < m; j += kWave) {
        const float v = load_row(&row[j]);
        if (load_u32(&sel[j]) != stamp && v > bv) {
            bv = v;
            bi = j;
        }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o, kWave);
        const uint32_t oi = __shfl_xor(bi, o, kWave);
        // a candidate beats another with a larger value, or the same value at a lower index
        const bool take = oi != kNone && (bi == kNone || ov > bv || (ov == bv && oi < bi));
        if (take) {
            bv = ov;
            bi = oi;
        }
    }
    best_v = bv;
    return bi;
}

__global__ void __launch_bounds__(kBlock) k_recurrent(GraphDev G, RecArgs A) {
    __shared__ WaveStage stage[kWavesPerBlock];
    const uint32_t wv = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint64_t s = (uint64_t)blockIdx.x * kWavesPerBlock + wv;
    if (s >= A.S) return;
    const uint32_t R = A.R, H = A.H, T = A.T;
    float* hid = A.hid + s * T * H;
    float* ro = A.ro + s * T * R;
    float* gate = A.gate + s * R;
    uint32_t* hsel = A.hsel + s * H;
    uint32_t* rsel = A.rsel + s * R;
    // row 0 of the hidden timeline: the merged inputs that are hidden neurons (distinct neurons, so
    // distinct cells)
    {
        const uint64_t a = A.in_ptr[s];
        const uint32_t n = (uint32_t)(A.in_ptr[s + 1] - a);
        for (uint32_t i = lane; i < n; i += kWave) {
            const uint32_t h1 = A.d2h[A.in_x[a + i]];
            if (h1) unsafeAtomicAdd(&hid[h1 - 1], A.in_w[a + i]);
        }
        wave_mem_sync();
    }
    for (uint32_t t = 0; t < T; ++t) {
        const uint32_t stamp = t + 1;
        float* hrow = hid + (uint64_t)t * H;
        float* rrow = ro + (uint64_t)t * R;
        for (uint32_t r = lane; r < R; r += kWave) store_row(&gate[r], 0.0f);
        // 1. the k strongest hidden scores of row t
        if (A.topk_hid > 0 && (uint32_t)A.topk_hid < H) {
            for (int32_t pass = 0; pass < A.topk_hid; ++pass) {
                float bv;
                const uint32_t bi = wave_argmax(hrow, hsel, H, stamp, bv);
                if (bi != kNone && lane == 0) store_u32(&hsel[bi], stamp);
                wave_mem_sync();
            }
            for (uint32_t h = lane; h < H; h += kWave)
                if (load_u32(&hsel[h]) != stamp) store_row(&hrow[h], 0.0f);
            wave_mem_sync();
        }
        // 2. hidden -> readout events of later rows
        wave_pairs(
            G, H, stage[wv],
            [&](uint32_t h, uint32_t& x, float& c) {
                c = load_row(&hrow[h]);
                x = c != 0.0f ? A.hidden[h] : kNone;
            },
            [&](float c, uint64_t e, bool ok) {
                if (!ok) return;
                const uint32_t d = G.delay[e];
                if ((uint64_t)t + d >= T) return;
                const uint32_t r1 = A.d2r[G.tgt[e]];
                if (r1) unsafeAtomicAdd(&ro[(uint64_t)(t + d) * R + r1 - 1], c * G.w[e]);
            });
        wave_mem_sync();
        // 3. gates of row t
        if (A.topk_ro <= 0) {
            for (uint32_t r = lane; r < R; r += kWave) {
                const float v = load_row(&rrow[r]);
                const uint32_t x = A.readout[r];
                const float thr = x != kNone ? G.threshold[x] : 1.0f;
                if (v >= thr) {
                    store_row(&gate[r], A.use_scores ? v : 1.0f);
                    store_u32(&rsel[r], stamp);
                }
            }
        } else {
            const uint32_t limit = min((uint32_t)A.topk_ro, R);
            for (uint32_t pass = 0; pass < limit; ++pass) {
                float bv;
                const uint32_t bi = wave_argmax(rrow, rsel, R, stamp, bv);
                if (bi != kNone && lane == 0) {
                    store_u32(&rsel[bi], stamp);
                    store_row(&gate[bi], A.use_scores ? bv : 1.0f);
                }
                wave_mem_sync();
            }
        }
        wave_mem_sync();
        // 4. readout feedback to hidden rows of later steps
        wave_pairs(
            G, R, stage[wv],
            [&](uint32_t r, uint32_t& x, float& c) {
                const float g = load_row(&gate[r]);
                c = A.decay * g;
                x = g != 0.0f ? A.readout[r] : kNone;
            },
            [&](float c, uint64_t e, bool ok) {
                if (!ok) return;
                const uint32_t d = G.delay[e];
                if ((uint64_t)t + d >= T) return;
                const uint32_t h1 = A.d2h[G.tgt[e]];
                if (h1) unsafeAtomicAdd(&hid[(uint64_t)(t + d) * H + h1 - 1], c * G.w[e]);
            });
        wave_mem_sync();
    }
    for (uint32_t r = lane; r < R; r += kWave) {
        float sum = 0.0f;
        for (uint32_t t = 0; t < T; ++t) sum += load_row(&ro[(uint64_t)t * R + r]);
        const uint32_t x = A.readout[r];
        const float thr = x != kNone ? G.threshold[x] : 1.0f;
        A.scores[s * R + r] = sum;
        A.spikes[s * R + r] = sum >= thr ? 1ull : 0ull;
    }
}

// d2x[list[j]] = j + 1 for every listed neuron, the last listing of a neuron winning (as the
// reference's host loop assigns them in list order)
__global__ void k_mark(const uint32_t* list, uint32_t m, uint32_t* d2x) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m && list[j] != kNone) atomicMax(&d2x[list[j]], j + 1);
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t n) {
        if (n) rcheck(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        std::swap(p, o.p);
        return *this;
    }
};

template <typename T>
DevBuf<T> upload(const T* h, size_t n, hipStream_t st) {
    DevBuf<T> d(n);
    if (n) rcheck(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    return d;
}

uint32_t grid_of(uint64_t waves) { return (uint32_t)((waves + kWavesPerBlock - 1) / kWavesPerBlock); }

}  // namespace

struct snn_graph {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0;
    uint64_t e = 0;
    uint32_t max_delay = 1;
    DevBuf<uint64_t> out_ptr;
    DevBuf<uint32_t> tgt, delay;
    DevBuf<float> w, threshold;
    GraphDev dev() const { return GraphDev{n, e, out_ptr.p, tgt.p, w.p, delay.p, threshold.p}; }
    ~snn_graph() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Merged sample inputs and readout / hidden lists as device arrays, with the dense -> index maps.
struct Lists {
    DevBuf<uint64_t> in_ptr;
    DevBuf<uint32_t> in_x;
    DevBuf<float> in_w;
    DevBuf<uint32_t> readout, d2r, hidden, d2h;
};

void check_inputs(const snn_graph* g, uint64_t S, const uint64_t* in_ptr, const uint32_t* in_x, const float* in_w) {
    if (!in_ptr) throw std::invalid_argument("input_ptr is NULL");
    if (in_ptr[0] != 0) throw std::invalid_argument("input_ptr[0] must be 0");
    for (uint64_t s = 0; s < S; ++s)
        if (in_ptr[s + 1] < in_ptr[s]) throw std::invalid_argument("input_ptr is not non-decreasing");
    const uint64_t m = in_ptr[S];
    if (m && (!in_x || !in_w)) throw std::invalid_argument("input_neuron / input_weight is NULL");
    for (uint64_t i = 0; i < m; ++i)
        if (in_x[i] >= g->n) throw std::invalid_argument("input neuron " + std::to_string(in_x[i]) + " is not a neuron of the graph");
    if (S >= (1ull << 31)) throw std::invalid_argument("too many samples");
}

void check_list(const snn_graph* g, uint64_t m, const uint32_t* list, const char* what) {
    if (m && !list) throw std::invalid_argument(std::string(what) + " is NULL");
    if (m >= (1ull << 31)) throw std::invalid_argument(std::string(what) + ": too many entries");
    for (uint64_t j = 0; j < m; ++j)
        if (list[j] != kNone && list[j] >= g->n)
            throw std::invalid_argument(std::string(what) + " entry " + std::to_string(list[j]) + " is not a neuron of the graph");
}

DevBuf<uint32_t> dense_map(const snn_graph* g, const DevBuf<uint32_t>& list, uint64_t m) {
    DevBuf<uint32_t> d2x(g->n);
    rcheck(hipMemsetAsync(d2x.p, 0, (size_t)g->n * sizeof(uint32_t), g->stream), "hipMemsetAsync");
    if (m) {
        hipLaunchKernelGGL(k_mark, dim3((uint32_t)((m + 255) / 256)), dim3(256), 0, g->stream, list.p, (uint32_t)m, d2x.p);
        rcheck(hipGetLastError(), "k_mark launch");
    }
    return d2x;
}

Lists upload_lists(const snn_graph* g, uint64_t S, const uint64_t* in_ptr, const uint32_t* in_x, const float* in_w,
                   uint64_t R, const uint32_t* readout, uint64_t H, const uint32_t* hidden) {
    Lists L;
    L.in_ptr = upload(in_ptr, S + 1, g->stream);
    L.in_x = upload(in_x, in_ptr[S], g->stream);
    L.in_w = upload(in_w, in_ptr[S], g->stream);
    L.readout = upload(readout, R, g->stream);
    L.d2r = dense_map(g, L.readout, R);
    if (hidden) {
        L.hidden = upload(hidden, H, g->stream);
        L.d2h = dense_map(g, L.hidden, H);
    }
    return L;
}

// device memory the recurrent timelines may take: a quarter of what is free
void check_budget(int device, uint64_t bytes, const char* what) {
    size_t free_b = 0, total_b = 0;
    rcheck(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    (void)device;
    if (bytes > free_b / 4)
        throw std::invalid_argument(std::string(what) + " needs " + std::to_string(bytes) + " bytes of device memory, more than a quarter of the " +
                                    std::to_string(free_b) + " bytes free");
}

}  // namespace

extern "C" {

int snn_graph_create(const snn_net_desc* d, int device, snn_graph** out) {
    if (!d || !out) return rfail(SNN_EINVAL, "snn_graph_create: NULL argument");
    *out = nullptr;
    return rguarded([&]() -> int {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return rfail(SNN_ENODEVICE, "no HIP device");
        if (device < 0 || device >= count) throw std::invalid_argument("device ordinal out of range");
        if (d->n_neurons >= kNone) throw std::invalid_argument("too many neurons");
        if (d->n_neurons && !d->threshold) throw std::invalid_argument("threshold is NULL");
        if (d->n_synapses && (!d->source || !d->target || !d->weight || !d->delay))
            throw std::invalid_argument("source / target / weight / delay is NULL");
        const uint32_t n = (uint32_t)d->n_neurons;
        const uint64_t e = d->n_synapses;
        // source-major CSR, connectome order within a source (the reference's outgoing lists,
        // CudaBackend.cu:2137-2182): a stable counting sort by source
        std::vector<uint64_t> ptr((size_t)n + 1, 0);
        uint32_t max_delay = 1;
        for (uint64_t j = 0; j < e; ++j) {
            if (d->source[j] >= n || d->target[j] >= n)
                throw std::invalid_argument("synapse " + std::to_string(j) + " names a neuron outside the graph");
            ++ptr[d->source[j] + 1];
            max_delay = std::max(max_delay, std::max(1u, d->delay[j]));
        }
        for (uint32_t x = 0; x < n; ++x) ptr[x + 1] += ptr[x];
        std::vector<uint32_t> tgt(e), dl(e);
        std::vector<float> w(e);
        {
            std::vector<uint64_t> at(ptr.begin(), ptr.end() - 1);
            for (uint64_t j = 0; j < e; ++j) {
                const uint64_t q = at[d->source[j]]++;
                tgt[q] = d->target[j];
                w[q] = d->weight[j];
                dl[q] = std::max(1u, d->delay[j]);
            }
        }
        auto g = std::make_unique<snn_graph>();
        g->device = device;
        g->n = n;
        g->e = e;
        g->max_delay = max_delay;
        rcheck(hipSetDevice(device), "hipSetDevice");
        rcheck(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate");
        g->out_ptr = upload(ptr.data(), ptr.size(), g->stream);
        g->tgt = upload(tgt.data(), e, g->stream);
        g->w = upload(w.data(), e, g->stream);
        g->delay = upload(dl.data(), e, g->stream);
        g->threshold = upload(d->threshold, n, g->stream);
        rcheck(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
        *out = g.release();
        return SNN_OK;
    });
}

int snn_graph_destroy(snn_graph* g) {
    if (!g) return SNN_OK;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    delete g;
    return SNN_OK;
}

int snn_graph_max_delay(snn_graph* g, uint32_t* max_delay) {
    if (!g || !max_delay) return rfail(SNN_EINVAL, "snn_graph_max_delay: NULL argument");
    *max_delay = g->max_delay;
    return SNN_OK;
}

int snn_graph_feedforward(snn_graph* g, uint64_t n_samples, const uint64_t* input_ptr, const uint32_t* input_neuron,
                          const float* input_weight, uint64_t n_readouts, const uint32_t* readout, float* scores,
                          uint64_t* spikes) {
    if (!g) return rfail(SNN_EINVAL, "snn_graph_feedforward: NULL graph");
    return rguarded([&]() -> int {
        const uint64_t S = n_samples, R = n_readouts;
        check_inputs(g, S, input_ptr, input_neuron, input_weight);
        check_list(g, R, readout, "readout");
        if (S == 0 || R == 0) return SNN_OK;
        if (!scores || !spikes) throw std::invalid_argument("scores / spikes is NULL");
        rcheck(hipSetDevice(g->device), "hipSetDevice");
        Lists L = upload_lists(g, S, input_ptr, input_neuron, input_weight, R, readout, 0, nullptr);
        DevBuf<float> sc(S * R);
        DevBuf<uint64_t> sp(S * R);
        const bool lds = R <= kLdsRow;
        if (!lds) rcheck(hipMemsetAsync(sc.p, 0, S * R * sizeof(float), g->stream), "hipMemsetAsync");
        FFArgs A{(uint32_t)S, (uint32_t)R, L.in_ptr.p, L.in_x.p, L.in_w.p, L.readout.p, L.d2r.p, sc.p, sp.p};
        if (lds) hipLaunchKernelGGL(k_feedforward<true>, dim3(grid_of(S)), dim3(kBlock), 0, g->stream, g->dev(), A);
        else hipLaunchKernelGGL(k_feedforward<false>, dim3(grid_of(S)), dim3(kBlock), 0, g->stream, g->dev(), A);
        rcheck(hipGetLastError(), "k_feedforward launch");
        rcheck(hipMemcpyAsync(scores, sc.p, S * R * sizeof(float), hipMemcpyDeviceToHost, g->stream), "hipMemcpyAsync");
        rcheck(hipMemcpyAsync(spikes, sp.p, S * R * sizeof(uint64_t), hipMemcpyDeviceToHost, g->stream), "hipMemcpyAsync");
        rcheck(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
        return SNN_OK;
    });
}

int snn_graph_recurrent(snn_graph* g, uint64_t n_samples, const uint64_t* input_ptr, const uint32_t* input_neuron,
                        const float* input_weight, uint64_t n_readouts, const uint32_t* readout, uint64_t n_hidden,
                        const uint32_t* hidden, const snn_recurrent_opts* opts, float* scores, uint64_t* spikes,
                        uint64_t* time_count) {
    if (!g || !opts) return rfail(SNN_EINVAL, "snn_graph_recurrent: NULL argument");
    return rguarded([&]() -> int {
        const uint64_t S = n_samples, R = n_readouts, H = n_hidden;
        check_inputs(g, S, input_ptr, input_neuron, input_weight);
        check_list(g, R, readout, "readout");
        check_list(g, H, hidden, "hidden");
        const uint64_t T = (uint64_t)g->max_delay * (1ull + 2ull * opts->iterations) + 1ull;
        if (time_count) *time_count = T;
        if (S == 0 || R == 0 || H == 0) return SNN_OK;
        if (!scores || !spikes) throw std::invalid_argument("scores / spikes is NULL");
        if (T >= (1ull << 31)) throw std::invalid_argument("time_count " + std::to_string(T) + " is too large");
        rcheck(hipSetDevice(g->device), "hipSetDevice");
        const uint64_t cells = S * T * (H + R);
        check_budget(g->device, cells * 4 + S * (H + 2 * R) * 4 + S * R * 12, "the recurrent timelines");
        Lists L = upload_lists(g, S, input_ptr, input_neuron, input_weight, R, readout, H, hidden);
        DevBuf<float> hid(S * T * H), ro(S * T * R), gate(S * R), sc(S * R);
        DevBuf<uint32_t> hsel(S * H), rsel(S * R);
        DevBuf<uint64_t> sp(S * R);
        rcheck(hipMemsetAsync(hid.p, 0, S * T * H * sizeof(float), g->stream), "hipMemsetAsync");
        rcheck(hipMemsetAsync(ro.p, 0, S * T * R * sizeof(float), g->stream), "hipMemsetAsync");
        rcheck(hipMemsetAsync(hsel.p, 0, S * H * sizeof(uint32_t), g->stream), "hipMemsetAsync");
        rcheck(hipMemsetAsync(rsel.p, 0, S * R * sizeof(uint32_t), g->stream), "hipMemsetAsync");
        RecArgs A{};
        A.S = (uint32_t)S;
        A.R = (uint32_t)R;
        A.H = (uint32_t)H;
        A.T = (uint32_t)T;
        A.in_ptr = L.in_ptr.p;
        A.in_x = L.in_x.p;
        A.in_w = L.in_w.p;
        A.readout = L.readout.p;
        A.hidden = L.hidden.p;
        A.d2r = L.d2r.p;
        A.d2h = L.d2h.p;
        A.hid = hid.p;
        A.ro = ro.p;
        A.gate = gate.p;
        A.hsel = hsel.p;
        A.rsel = rsel.p;
        A.topk_ro = opts->top_k_feedback_readouts;
        A.topk_hid = opts->top_k_hidden;
        A.use_scores = opts->use_scores_not_spikes;
        A.decay = opts->feedback_decay;
        A.scores = sc.p;
        A.spikes = sp.p;
        hipLaunchKernelGGL(k_recurrent, dim3(grid_of(S)), dim3(kBlock), 0, g->stream, g->dev(), A);
        rcheck(hipGetLastError(), "k_recurrent launch");
        rcheck(hipMemcpyAsync(scores, sc.p, S * R * sizeof(float), hipMemcpyDeviceToHost, g->stream), "hipMemcpyAsync");
        rcheck(hipMemcpyAsync(spikes, sp.p, S * R * sizeof(uint64_t), hipMemcpyDeviceToHost, g->stream), "hipMemcpyAsync");
        rcheck(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
        return SNN_OK;
    });
}

}  // extern "C"
