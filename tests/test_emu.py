"""The device step's logic, run on the CPU through the same item functions as the kernels
(tests/emu), against the reference fixtures — including the per-target sort (events arrive in a
shuffled order) and every code path the templates select."""
import numpy as np
import pytest

from tests import golden_util as G
from tests.emu import Emu


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_step_model_matches_reference_fixture(name):
    net = G.CASES[name]()
    g = G.load(name)
    m = Emu(net, ev_cap=1 << 22, fire_cap=1 << 20)
    m.run_commands(net.commands)
    assert m.flags == 0
    d, f = m.tick_counts()
    bad = G.mismatches(g, d, f, m.raster(), m.neurons(), m.synapses())
    assert not bad, f"{name}: step model differs from the reference in {bad}"


def test_capacity_overflow_is_an_error_not_a_drop():
    net = G.CASES["config0_burst"]()
    m = Emu(net, ev_cap=1 << 22, fire_cap=1 << 12)
    m.run_commands(net.commands)
    assert m.flags & 4  # fire-log capacity
    m = Emu(net, ev_cap=1 << 12, fire_cap=1 << 20)
    m.run_commands(net.commands)
    assert m.flags & 1  # events per tick
