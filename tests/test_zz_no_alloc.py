"""snn_net_run_until_static (the run_until variant that allocates no device memory) against
snn_net_run_until, through the C ABI: two networks built from the same connectome get the same
inject calls; one steps with each entry point. Two calls in a row, the second a smaller segment with
fewer injects; every result must be identical after each call. Also: a segment larger than the
static staging array is refused and changes nothing."""
import ctypes

import numpy as np
import pytest

import snncuda_amd as sa
from snncuda_amd import workloads as W

pytestmark = pytest.mark.gpu


def _inject(net, cmds):
    if cmds:
        net.inject([c[1] for c in cmds], [c[2] for c in cmds], [c[3] for c in cmds])


def _same(a, b, where):
    assert a.counts() == b.counts(), where
    da, fa = a.tick_counts()
    db, fb = b.tick_counts()
    assert np.array_equal(da, db) and np.array_equal(fa, fb), where
    assert np.array_equal(a.raster(), b.raster()), where
    na, nb = a.neurons(), b.neurons()
    for k in na:
        assert np.array_equal(na[k], nb[k]), (where, k)
    sa_, sb = a.synapses(), b.synapses()
    for k in sa_:
        assert np.array_equal(sa_[k], sb[k]), (where, k)


def test_static_run_matches_run_until_twice_in_a_row():
    print("no-alloc: config0 (1,000 neurons, STDP on, threshold 1.1), ticks 0-399 then 400-449", flush=True)
    net = W.config0(0, threshold=1.1, ticks=1000)
    inj = [c for c in net.commands if c[0] == 0]
    first = [c for c in inj if c[2] < 400]
    second = [c for c in inj if 400 <= c[2] < 450]
    assert len(second) < len(first)
    L = sa.lib()
    with sa.Network(net) as ref, sa.Network(net) as dev:
        for cmds, last in ((first, 399), (second, 449)):
            _inject(ref, cmds)
            _inject(dev, cmds)
            assert L.snn_net_run_until(ref._h, ctypes.c_uint64(last)) == 0
            rc = L.snn_net_run_until_static(dev._h, ctypes.c_uint64(last))
            assert rc == 0, sa.lib().snn_last_error().decode()
            _same(ref, dev, f"after run_until({last})")
        assert dev.delivered > 0 and dev.fired > 0


def test_static_run_refuses_a_segment_beyond_its_staging_and_changes_nothing():
    print("no-alloc: a segment of 10^6 ticks (beyond the 4 MiB staging) is refused", flush=True)
    net = W.config0(0, threshold=1.1, ticks=200)
    with sa.Network(net) as dev:
        dev.run_commands(net.commands)
        before = dev.counts()
        with pytest.raises(sa.SNNError) as e:
            dev.run_until_static(before[2] + 1_000_000)
        assert e.value.code == sa.SNN_EINVAL
        assert dev.counts() == before
        dev.run_until_static(before[2] + 50)  # a short segment still runs
        assert dev.counts()[2] == before[2] + 51
