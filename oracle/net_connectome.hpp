// Reading a SNNNET01 file into a reference Connectome. Test infrastructure only: included by the
// drivers under oracle/ that are compiled against the reference's headers into oracle/_ref/.
//
// The Connectome is built directly (not through NetworkIR) with ids i+1 / j+1, which is what
// ConnectomeBuilder produces for the same neuron and synapse order
// (reference src/declarative/Connectome.cpp:26-36,93-95).
#pragma once

#include "snncuda/declarative/Connectome.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

extern "C" {
#include "netfile.h"
}

namespace netfile {

struct Net {
    std::uint64_t n = 0, e = 0;
    std::vector<float> thr, sim, w, mw;
    std::vector<std::uint64_t> win, mref, cptr;
    std::vector<std::uint32_t> src, tgt, dl, codes;
    std::vector<std::uint8_t> comp, rec, pl;
    std::vector<snn_command> cmds;
};

struct Reader {
    FILE* f;
    template <typename T>
    std::vector<T> vec(std::size_t n) {
        std::vector<T> v(n);
        if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short read");
        return v;
    }
    std::uint64_t u64() {
        std::uint64_t x;
        if (std::fread(&x, 8, 1, f) != 1) throw std::runtime_error("short read");
        return x;
    }
};

inline Net read(const char* path) {
    FILE* in = std::fopen(path, "rb");
    if (!in) throw std::runtime_error("cannot open input");
    char magic[8];
    if (std::fread(magic, 1, 8, in) != 8 || std::memcmp(magic, SNN_NET_MAGIC, 8) != 0) {
        std::fclose(in);
        throw std::runtime_error("bad magic");
    }
    Reader r{in};
    Net t;
    t.n = r.u64();
    t.e = r.u64();
    const auto ncodes = r.u64();
    const auto ncmd = r.u64();
    t.thr = r.vec<float>(t.n);
    t.win = r.vec<std::uint64_t>(t.n);
    t.sim = r.vec<float>(t.n);
    t.mref = r.vec<std::uint64_t>(t.n);
    t.src = r.vec<std::uint32_t>(t.e);
    t.tgt = r.vec<std::uint32_t>(t.e);
    t.w = r.vec<float>(t.e);
    t.mw = r.vec<float>(t.e);
    t.dl = r.vec<std::uint32_t>(t.e);
    t.comp = r.vec<std::uint8_t>(t.e);
    t.rec = r.vec<std::uint8_t>(t.e);
    t.pl = r.vec<std::uint8_t>(t.e);
    t.cptr = r.vec<std::uint64_t>(t.e + 1);
    t.codes = r.vec<std::uint32_t>(ncodes);
    t.cmds = r.vec<snn_command>(ncmd);
    std::fclose(in);
    return t;
}

inline snncuda::declarative::Connectome connectome(const Net& t) {
    snncuda::declarative::Connectome c;
    c.neurons.resize(t.n);
    for (std::uint64_t i = 0; i < t.n; ++i) {
        c.neurons[i].id = i + 1;
        c.neurons[i].params.threshold = t.thr[i];
        c.neurons[i].params.pattern_window_ticks = t.win[i];
        c.neurons[i].params.similarity_threshold = t.sim[i];
        c.neurons[i].params.max_reference_patterns = t.mref[i];
    }
    c.synapses.resize(t.e);
    for (std::uint64_t j = 0; j < t.e; ++j) {
        auto& s = c.synapses[j];
        s.id = j + 1;
        s.source = t.src[j] + 1;
        s.target = t.tgt[j] + 1;
        s.weight = t.w[j];
        s.max_weight = t.mw[j];
        s.delay_ticks = t.dl[j];
        s.spike_code_offsets.assign(t.codes.begin() + t.cptr[j], t.codes.begin() + t.cptr[j + 1]);
        s.compartment = static_cast<snncuda::snn::DendriticCompartment>(t.comp[j]);
        s.receptor = static_cast<snncuda::snn::ReceptorType>(t.rec[j]);
        s.plasticity_enabled = t.pl[j] != 0;
    }
    return c;
}

}  // namespace netfile
