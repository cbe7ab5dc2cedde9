"""The host pack (Connectome arrays -> source-sorted, lag-grouped SoA for the device) runs its passes
over synapses on several threads; every output must equal the sequential pack's (kept as the
reference in tests/emu/ref_pack.h), and an invalid network must raise the same message (that of
the first offending synapse in connectome order), at any thread count."""
import numpy as np
import pytest

from snncuda_amd import workloads as W
from tests.emu import pack_check


def _random_net(rng, n, e, codes="one", plastic=0.5, max_refs=(1,), mw_values=(0.7, 1.5, 2.0, 3.0),
                delay_hi=40):
    src = rng.integers(0, n, e).astype(np.uint32)
    if rng.random() < 0.5:
        src.sort(kind="stable")  # builder order (source-major) as well as arbitrary order
    tgt = rng.integers(0, n, e).astype(np.uint32)
    if codes == "one":
        cps = None
    else:
        lens = rng.integers(0 if codes == "mixed" else 1, 5, e)
        cps = [list(rng.integers(0, 8, k)) for k in lens]
    thr, win, sim, _ = W.neurons(n)
    mref = rng.choice(np.asarray(max_refs, np.uint64), n)
    return W.make_net(
        (thr, win, sim, mref), src, tgt, rng.random(e).astype(np.float32),
        rng.choice(np.asarray(mw_values, np.float32), e), rng.integers(0, delay_hi, e).astype(np.uint32),
        rng.integers(0, 4, e).astype(np.uint8), rng.integers(0, 4, e).astype(np.uint8),
        (rng.random(e) < plastic).astype(np.uint8), codes_per_syn=cps)


@pytest.mark.parametrize("threads", [1, 2, 3, 8, 16])
def test_parallel_pack_equals_sequential_pack(threads):
    rng = np.random.default_rng(threads)
    cases = [
        _random_net(rng, 50, 3000),
        _random_net(rng, 1000, 40000, plastic=1.0),
        _random_net(rng, 300, 20000, codes="multi", plastic=0.3, max_refs=(1, 2)),
        _random_net(rng, 300, 20000, codes="mixed", plastic=0.0),
        _random_net(rng, 7, 5000, plastic=0.9, delay_hi=3),          # few keys, long runs per key
        _random_net(rng, 70000, 90000, plastic=0.2, delay_hi=2),     # more keys than synapses
        _random_net(rng, 2000, 30000, mw_values=np.linspace(0.1, 1.9, 200)),
        _random_net(rng, 2000, 30000, plastic=1.0, mw_values=np.linspace(0.01, 2.5, 30000)),  # a bound per synapse
    ]
    for i, net in enumerate(cases):
        rc, msg = pack_check(net, threads)
        assert rc == 0 and not msg, (i, rc, msg)


def _with(net, **kw):
    for k, v in kw.items():
        getattr(net, k)[...] = v
    return net


@pytest.mark.parametrize("threads", [1, 4, 16])
def test_invalid_networks_raise_the_sequential_packs_message(threads):
    rng = np.random.default_rng(11)

    def base():
        return _random_net(rng, 200, 20000, plastic=0.5)

    bad = []
    n1 = base(); n1.tgt[15000] = 200; n1.comp[19000] = 7; bad.append((n1, "synapse endpoint out of range"))
    n2 = base(); n2.comp[9000] = 7; n2.src[15000] = 999; bad.append((n2, "bad compartment or receptor"))
    n3 = base(); n3.delay[12345] = 70000; bad.append((n3, "delay + code offset above 65535 ticks"))
    n5 = base(); n5.max_refs[3] = 0; bad.append((n5, "max_reference_patterns must be >= 1"))
    for net, want in bad:
        rc, msg = pack_check(net, threads)
        assert rc == 0 and msg == want, (rc, msg, want)
