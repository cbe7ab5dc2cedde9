"""The C-ABI library loads and exports every symbol include/snncuda_hip.h declares."""
import ctypes
import pathlib
import re

import pytest

import snncuda_amd as sa

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "snncuda_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sn[a-z_]+)\s*\(", text)) - {"snn_net_desc", "snn_options"})


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("snncuda_version", "snncuda_cuda_available", "snn_net_create", "snn_net_run_until",
                 "snn_net_inject", "snn_net_read_neurons", "snn_net_read_synapses", "snn_net_read_raster"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(sa.library_path()))
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_reference_symbols_unchanged():
    assert sa.version() == "0.1.0"
    assert sa.cuda_available() is True


def test_create_without_device_or_with_bad_input_fails_cleanly():
    from snncuda_amd import workloads as W
    net = W.feature_net(0)
    if sa.device_count() == 0:
        with pytest.raises(sa.SNNError) as e:
            sa.Network(net)
        assert e.value.code == sa.SNN_ENODEVICE
    bad = W.feature_net(0)
    bad.tgt[0] = bad.n_neurons + 5
    with pytest.raises(sa.SNNError):
        sa.Network(bad)


def test_batch_create_without_device_fails_cleanly():
    from snncuda_amd import workloads as W
    if sa.device_count() != 0:
        pytest.skip("a device is visible")
    with pytest.raises(sa.SNNError) as e:
        sa.Batch([W.feature_net(0), W.feature_net(1)])
    assert e.value.code == sa.SNN_ENODEVICE


def test_graph_create_without_device_fails_cleanly():
    from snncuda_amd import workloads as W
    if sa.device_count() != 0:
        pytest.skip("a device is visible")
    with pytest.raises(sa.SNNError) as e:
        sa.Graph(W.feature_net(0))
    assert e.value.code == sa.SNN_ENODEVICE
