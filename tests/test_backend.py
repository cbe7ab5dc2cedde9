"""The drop-in CudaBackend (snncuda_amd/cpp/CudaBackendHip.cpp over libsnncuda_hip.so) against the
reference CPU runtime.

oracle/_ref/backend_check and oracle/_ref/cuda_learning_consistency_experiment are built by
oracle/Makefile from the reference's headers and sources where they lie; the tests skip when they
were not built (no reference sources at build time). backend_check drives the reference
NetworkPropagator as the reference's own CUDA tests do (inject(id, 0, 1.0) per active neuron, then
run_until(max_steps)) and requires every CudaPropagationResult vector to equal the CPU state bit for
bit; the experiment is the reference's own tests/cuda_learning_consistency_experiment.cpp linked
against the drop-in backend.
"""
import dataclasses
import pathlib
import subprocess

import pytest

import oracle
from snncuda_amd import workloads as W
from tests import kat_nets as K

ROOT = pathlib.Path(__file__).resolve().parents[1]
CHECK = ROOT / "oracle" / "_ref" / "backend_check"
EXPERIMENT = ROOT / "oracle" / "_ref" / "cuda_learning_consistency_experiment"
THROUGHPUT = ROOT / "oracle" / "_ref" / "spike_throughput_benchmark"
needs_check = pytest.mark.skipif(not CHECK.exists(), reason="oracle/_ref/backend_check not built")


def _feature(seed, threshold):
    net = W.feature_net(seed, threshold=threshold)
    return dataclasses.replace(net, commands=[(0, i, 0, 1.0) for i in range(0, net.n_neurons, 3)])


def _cases():
    out = {}
    for name, fn in sorted(K.ALL.items()):
        net = fn()
        last = max((t for k, _, t, _ in net.commands if k == 1), default=10)
        out[name] = (fn, int(last))
    for seed in (1, 2, 3):
        for thr in (0.5, 0.6):
            out[f"feature{seed}_{thr}"] = ((lambda s=seed, t=thr: _feature(s, t)), 150)
    return out


CASES = _cases()


def _run(args, timeout=120):
    r = subprocess.run([str(CHECK)] + args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    return r


@needs_check
@pytest.mark.parametrize("name", sorted(CASES))
def test_pack_reproduces_connectome_arrays(name, tmp_path):
    net = CASES[name][0]()
    path = tmp_path / "net.bin"
    oracle.write_net(path, net)
    assert _run(["pack", str(path)]).returncode == 0


@pytest.mark.gpu
@needs_check
@pytest.mark.parametrize("name", sorted(CASES))
def test_run_resident_propagation_matches_reference_cpu(name, tmp_path):
    fn, max_steps = CASES[name]
    print(f"case {name} max_steps {max_steps}", flush=True)
    path = tmp_path / "net.bin"
    oracle.write_net(path, fn())
    r = _run(["run", str(path), str(max_steps)])
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.skipif(not EXPERIMENT.exists(), reason="oracle/_ref experiment not built")
def test_reference_learning_consistency_experiment_passes():
    print("reference tests/cuda_learning_consistency_experiment.cpp on the drop-in backend", flush=True)
    r = subprocess.run([str(EXPERIMENT)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "cuda_learning_consistency_experiment=passed" in r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not THROUGHPUT.exists(), reason="oracle/_ref benchmark not built")
def test_reference_spike_throughput_benchmark_passes():
    # the reference's tests/spike_throughput_benchmark.cpp (27,648-neuron chain net, 6 ticks) on the
    # drop-in: it exits 0 only when the device's delivered and fired counts equal the reference CPU's
    print("reference tests/spike_throughput_benchmark.cpp on the drop-in backend", flush=True)
    r = subprocess.run([str(THROUGHPUT)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    kv = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    assert kv["cuda_executed"] == "yes"
    assert kv["cuda_delivered_spikes"] == kv["cpu_delivered_spikes"] and int(kv["cpu_delivered_spikes"]) > 0
    assert kv["cuda_fired_spikes"] == kv["cpu_fired_spikes"]


@pytest.mark.gpu
@needs_check
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_inference_session_run_batch_matches_reference_cpu(seed, tmp_path):
    """CudaInferenceSession::run_batch on the batched form: the reference's session tests restated,
    then 5 samples drawn from a feature network's injects, each against a fresh reference CPU run
    (stateless) and against one CPU propagator with a continuing clock (stateful)."""
    print(f"session feature{seed}", flush=True)
    path = tmp_path / "net.bin"
    oracle.write_net(path, W.feature_net(seed))
    r = _run(["session", str(path), "80", "5"])
    assert r.returncode == 0, r.stderr


SESSION_TESTS = ROOT / "oracle" / "_ref" / "cuda_synapse_dendrite_tests"


@pytest.mark.gpu
@pytest.mark.skipif(not SESSION_TESTS.exists(), reason="oracle/_ref session tests not built")
def test_reference_cuda_synapse_dendrite_tests_pass():
    # the reference's tests/cuda_synapse_dendrite_tests.cpp, unchanged, on the drop-in: LTP / LTD /
    # inhibitory / coded synapses against its CPU runtime, the inference session, the feedforward and
    # recurrent readouts (:306-421) and the persistent-session timing check
    print("reference tests/cuda_synapse_dendrite_tests.cpp on the drop-in backend", flush=True)
    r = subprocess.run([str(SESSION_TESTS)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "All CUDA synapse/dendrite tests passed" in r.stdout
