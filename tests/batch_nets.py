"""Small batches of uneven networks for the batched-form tests (all sizes finish in the oracle in
seconds). Every network's command list ends with the same run_until, so each network runs exactly
as many ticks inside the batch as it does alone."""
from snncuda_amd import workloads as W


def equalize(nets):
    """Append run_until(last) to every network's commands, last = the latest tick of any network."""
    last = max(max(t for k, _, t, _ in n.commands if k == 1) for n in nets)
    for n in nets:
        if n.commands[-1] != (1, 0, last, 0.0):
            n.commands = list(n.commands) + [(1, 0, last, 0.0)]
    return nets, last


def mixed_batch():
    """Feature networks (codes, delay 0, every compartment and receptor, mixed plasticity and
    max_refs, injects between run_until calls), a config-0 network and small networks of the
    batched workload's law (2 codes from U{0..7})."""
    nets = [W.feature_net(s) for s in (1, 2, 3)]
    nets.append(W.config0(4, threshold=1.1, ticks=250))
    nets += [W.config3_net(s, ticks=250, threshold=3.0, n_range=(500, 1500), fan_range=(50, 100)) for s in (5, 6, 7)]
    return equalize(nets)


def code3_batch():
    """Only networks of the batched workload's law (compact code records, VE order by lag)."""
    nets = [W.config3_net(s, ticks=200, threshold=3.0, n_range=(500, 3000), fan_range=(50, 150)) for s in range(6)]
    return equalize(nets)
