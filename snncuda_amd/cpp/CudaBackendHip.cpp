// Drop-in definitions of snncuda::backends::CudaBackend's and CudaInferenceSession's members
// (reference include/snncuda/backends/CudaBackend.h:97-138) on the MI355X step behind
// libsnncuda_hip.so.
//
// Built inside the reference tree in place of src/backends/cuda/CudaBackend.cu, with include/ of
// this repository on the include path and libsnncuda_hip.so linked. The class declarations,
// result types and callers stay unchanged. Errors raised by the device library surface as
// snncuda_hip::Error, a std::runtime_error, as the reference's check_cuda does. The spiking
// session runs on the step (snn_batch_*), the non-spiking readouts (run_feedforward_batch,
// run_recurrent_feedforward_batch, CudaBackend.cu:1260-1528) on snn_graph_*.

#include "snncuda/backends/CudaBackend.h"
#include "snncuda_hip.hpp"

#include <memory>
#include <stdexcept>
#include <string>

namespace snncuda::backends {

class CudaInferenceSession::Impl {
public:
    explicit Impl(const declarative::Connectome& connectome)
        : available_(snn_device_count() > 0), runner_(snncuda_hip::pack(connectome)) {}

    bool available_;
    snncuda_hip::InferenceRunner<CudaInferenceBatchResult, CudaInferenceSample, CudaInferenceOptions> runner_;
};

CudaInferenceSession::CudaInferenceSession(const declarative::Connectome& connectome)
    : impl_(std::make_unique<Impl>(connectome)) {}
CudaInferenceSession::~CudaInferenceSession() = default;
CudaInferenceSession::CudaInferenceSession(CudaInferenceSession&&) noexcept = default;
CudaInferenceSession& CudaInferenceSession::operator=(CudaInferenceSession&&) noexcept = default;

bool CudaInferenceSession::available() const noexcept {
    return impl_->available_;
}

CudaInferenceBatchResult CudaInferenceSession::run_batch(
    const std::vector<CudaInferenceSample>& samples,
    const CudaInferenceOptions& options) {
    if (!impl_->available_) {
        return {};
    }
    return impl_->runner_.run(samples, options);
}

CudaInferenceBatchResult CudaInferenceSession::run_feedforward_batch(
    const std::vector<CudaInferenceSample>& samples,
    const CudaInferenceOptions& options) {
    if (!impl_->available_) {
        return {};
    }
    return impl_->runner_.feedforward(samples, options);
}

CudaInferenceBatchResult CudaInferenceSession::run_recurrent_feedforward_batch(
    const std::vector<CudaInferenceSample>& samples,
    const CudaRecurrentFeedforwardOptions& options) {
    if (!impl_->available_) {
        return {};
    }
    return impl_->runner_.recurrent(samples, options);
}

void CudaInferenceSession::reset_state() {
    impl_->runner_.reset();
}

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
