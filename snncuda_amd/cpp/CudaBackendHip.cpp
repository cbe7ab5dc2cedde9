// Drop-in definitions of snncuda::backends::CudaBackend's members (reference
// include/snncuda/backends/CudaBackend.h:128-138) on the MI355X step behind libsnncuda_hip.so.
//
// Built inside the reference tree in place of the CudaBackend:: member definitions of
// src/backends/cuda/CudaBackend.cu (:3350-3800), with include/ of this repository on the include path
// and libsnncuda_hip.so linked. The class declaration, CudaPropagationResult and the callers stay
// unchanged. Errors raised by the device library surface as snncuda_hip::Error, a
// std::runtime_error, as the reference's check_cuda does.

#include "snncuda/backends/CudaBackend.h"
#include "snncuda_hip.hpp"

#include <string>

namespace snncuda::backends {

std::string_view CudaBackend::name() const noexcept {
    return "cuda";
}

bool CudaBackend::available() const noexcept {
    return snn_device_count() > 0;
}

std::string CudaBackend::availability_status() const {
    const int count = snn_device_count();
    if (count <= 0) {
        return "no HIP devices reported by the runtime";
    }
    return std::to_string(count) + " HIP device(s) available";
}

CudaPropagationResult CudaBackend::run_resident_propagation(
    const declarative::Connectome& connectome,
    const std::vector<core::NeuronId>& initially_active,
    std::uint32_t max_steps) const {
    if (!available()) {
        return {};
    }
    return snncuda_hip::run_resident_propagation<CudaPropagationResult>(connectome, initially_active, max_steps);
}

} // namespace snncuda::backends
