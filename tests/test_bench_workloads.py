"""The bench's synthetic workloads match their configuration text: network sizes, topology,
the served stream's model mix, size and duration distributions, the batched sweep."""
import math

import numpy as np

import bench


def _synapses(pops, conns):
    return sum(pops[s] * pops[d] * p for s, d, p in conns)


def test_headline_and_reference_networks():
    pops, conns = bench.workload_layout("brunel_stdp_1e9")
    assert pops == [20_000, 64_000, 16_000] and len(conns) == 9 and all(p == 0.1 for *_, p in conns)
    assert abs(_synapses(pops, conns) - 1e9) < 1e6
    pops, conns = bench.workload_layout("ref_brunel_1e9")  # samples/main.cpp: {N/2, N/2}, 0->1, 1->1
    assert pops == [70_710, 70_710] and conns == [(0, 1, 0.1), (1, 1, 0.1)]
    assert abs(_synapses(pops, conns) - 1e9) / 1e9 < 1e-3
    pops, conns = bench.workload_layout("ref_vogels_1e9")
    assert pops == [223_606] and abs(_synapses(pops, conns) - 1e9) / 1e9 < 1e-3
    assert bench.WORKLOADS["ref_vogels_1e9"][3] == 8 and bench.WORKLOADS["ref_brunel_1e9"][3] == 15
    pops, conns = bench.workload_layout("brunel_1e10")
    assert abs(_synapses(pops, conns) - 1e10) / 1e10 < 1e-3
    assert set(bench.PUBLISHED) <= set(bench.WORKLOADS)


def test_served_stream_distribution():
    jobs = bench.serve_stream(4000, 2.0, 3)
    models = [j[1] for j in jobs]
    frac = {m: models.count(m) / len(jobs) for m in set(models)}
    assert abs(frac["brunel_3pop"] - 0.4) < 0.03 and abs(frac["vogels_abbott"] - 0.4) < 0.03
    assert abs(frac["brunel_3pop_stdp"] - 0.2) < 0.03
    syn = np.array([_synapses(j[2], j[3]) for j in jobs])
    assert syn.min() >= 0.9e5 and syn.max() <= 2.1e9
    lg = np.log(syn)  # log-uniform: uniform in log space
    assert abs(np.mean(lg) - (math.log(1e5) + math.log(2e9)) / 2) < 0.2
    bio = np.array([j[5] for j in jobs]) * 1e-4
    assert bio.min() >= 0.1 - 1e-4 and bio.max() <= 10 + 1e-4 and abs(bio.mean() - 5.05) < 0.3
    gaps = np.diff([0.0] + [j[0] for j in jobs])
    assert abs(gaps.mean() - 0.5) < 0.05  # Poisson arrivals at 2 jobs/s
    assert all(j[4] == (1 if j[1] == "vogels_abbott" else 15) for j in jobs)


def test_batched_sweep():
    nets = bench.sweep_nets(512, 7)
    assert sum(1 for n in nets if n[0] == "brunel_3pop") == 256
    syn = np.array([_synapses(n[1], n[2]) for n in nets])
    assert syn.min() > 0.8e4 and syn.max() < 1.2e8
    rates = [n[5][0] for n in nets if n[0] == "brunel_3pop"]
    assert min(rates) >= 5 and max(rates) <= 30
    assert len({n[4] for n in nets}) == 512  # per-net seeds


def test_reference_synth_networks():
    """samples/main.cpp:1285-1291: synth is one population, N = sqrt(S / p); the simulation row
    at p = 0.00156 (delay 1), the setup-time row at p = 0.05."""
    for name, p in (("ref_synth_1e9", 0.00156), ("ref_synth_setup_1e9", 0.05)):
        model, n, pp, delay = bench.WORKLOADS[name]
        assert model == "synth" and pp == p and delay == 1 and n == int(math.sqrt(1e9 / p))
        pops, conns = bench.workload_layout(name)
        assert pops == [n] and conns == [(0, 0, p)]
    assert bench.PUBLISHED["ref_synth_1e9"] == 1.39947 and bench.PUBLISHED_SETUP["ref_synth_setup_1e9"] == 0.253356


def test_statistics_populations():
    assert bench.stat_pops("vogels_abbott", [70_710]) == [56_568, 14_142]
    assert bench.stat_pops("brunel_3pop", [20, 64, 16]) == [20, 64, 16]


def test_algorithmic_bytes_model():
    """The per-launch byte model on hand-made counters (one step): delivery reads 4 B per
    spike id + 16 B of row offsets per spike + the target id (and, with synapse state, the
    weight) per event; plasticity reads 4 B per index word, the state of every synapse brought
    up to date, and writes back only the synapses it changed; static models have no passes."""
    st = {"spikes": 100, "events": 1000, "plastic_events": 10, "flush_slots": 200, "flush_synapses": 50,
          "flush_written": 30}
    b = bench.algorithmic_bytes("brunel_3pop_stdp", 1000, st, 1)
    assert b["receive"] == 100 * 20 + 1000 * 8 + 10 * 4 + 200 * 4 + (50 + 30) * 4
    s = bench.algorithmic_bytes("brunel_3pop", 1000, dict(st, plastic_events=0), 1)
    assert s["receive"] == 100 * 20 + 1000 * 4
    # update: state (8 B) read + write, the V accumulator read + clear, 4 B per spike
    assert s["update"] == 1000 * (2 * 8 + 2 * 4) + 100 * 4
    # the overlapped step kernel: one update, one delivery and one step's plasticity work
    assert b["step"] == b["receive"] + b["update"] and s["step"] == s["receive"] + s["update"]


def test_cpu_runnable_configuration():
    """brunel_1e6: three populations (Poisson 20 %, E, I) of N = 3,162, all pairs at p = 0.1,
    ~1e6 synapses."""
    pops, conns = bench.workload_layout("brunel_1e6")
    assert sum(pops) == 3_162 and pops[0] == 632 and len(conns) == 9
    assert abs(_synapses(pops, conns) - 1e6) / 1e6 < 1e-3
