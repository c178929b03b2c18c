"""Batched stepping (many independent networks, one launch per step phase): every network's
result equals the same network stepped alone (bit for bit in fixed-point mode), for uneven
sizes, delays, seeds and per-network model parameters; checked through the C ABI."""
import numpy as np
import pytest

from oracle import oracle as O
from spice_amd import Batch, Layout, Snn
from tests.helpers import assert_state_equal, three_pop_layout

pytestmark = pytest.mark.gpu


def _mixed_nets(model, count, seed0, delays=(1, 2, 5, 15)):
    rng = np.random.default_rng(seed0)
    nets = []
    for k in range(count):
        if model == "vogels_abbott":
            n = int(rng.integers(300, 4000))
            lay = ([n], [(0, 0, 0.05)])
        else:
            n = int(rng.integers(300, 5000))
            lay = three_pop_layout(n)
        delay = int(rng.choice(list(delays)))
        params = [float(rng.uniform(5, 30))] if model == "brunel_3pop" else None
        nets.append((Layout(*lay), delay, seed0 * 1000 + k, params))
    return nets


@pytest.mark.parametrize("model,delays", [("brunel_3pop", (1, 2, 5, 15)), ("brunel_3pop", (2, 5, 15)),
                                          ("vogels_abbott", (1, 2, 5, 15))],
                         ids=["brunel_3pop-mixed", "brunel_3pop-overlapped", "vogels_abbott"])
def test_batch_equals_networks_alone(gpu, model, delays):
    """With every delay >= 2 a batch of brunel_3pop nets runs overlapped steps (one kernel per
    step boundary over all nets); with a delay-1 net among them, two kernels per step."""
    nets = _mixed_nets(model, 9, 3, delays)
    steps = 300  # graph replays plus eager remainder steps
    b = Batch(model, nets, 1e-4, fixed_point=True)
    b.run(steps)
    b.synchronize()
    for k, (lay, delay, seed, params) in enumerate(nets):
        alone = Snn(model, lay, 1e-4, delay, seed=seed, fixed_point=True, params=params)
        alone.run(steps)
        bn = b.net(k)
        assert bn.clock() == alone.clock()
        assert_state_equal(bn.neurons(), alone.neurons(), f"net {k}")
        sa, sb = alone.stats(), bn.stats()
        assert sa["spikes"] == sb["spikes"] and sa["events"] == sb["events"], (k, sa, sb)
        np.testing.assert_array_equal(np.sort(bn.step_spikes_of_last()), np.sort(alone.step_spikes_of_last()))
        alone.close()
    b.close()


@pytest.mark.parametrize("model,delays", [("brunel_3pop_stdp", (1, 4, 15)), ("brunel_3pop_stdp", (2, 4, 15)),
                                          ("brunel_3pop_stdp_deferred", (2, 4, 15)), ("brunel_with_plasticity", (1, 4, 15))],
                         ids=["stdp-mixed", "stdp-overlapped", "stdp_deferred", "brunel_with_plasticity"])
def test_plastic_batch_equals_networks_alone(gpu, model, delays):
    """Plastic nets in one batch (batched receive, update and plasticity launches: eager passes
    for the brunel_3pop_stdp models, lazy flushes for brunel_with_plasticity; overlapped steps
    when every net of brunel_3pop_stdp has a delay >= 2): spikes, neuron state, synaptic
    weights and counters equal each net run alone, over enough steps for every lazy row to be
    flushed several times (window 255 steps)."""
    rng = np.random.default_rng(11)
    nets = []
    for k in range(5):
        n = int(rng.integers(400, 2500))
        lay = three_pop_layout(n) if model.startswith("brunel_3pop_stdp") else ([n // 2, n - n // 2], [(0, 1, 0.1), (1, 1, 0.1)])
        nets.append((Layout(*lay), int(rng.choice(list(delays))), 500 + k, None))
    steps = 700
    b = Batch(model, nets, 1e-4, fixed_point=True)
    b.run(steps)
    b.synchronize()
    for k, (lay, delay, seed, _) in enumerate(nets):
        alone = Snn(model, lay, 1e-4, delay, seed=seed, fixed_point=True)
        alone.run(steps)
        bn = b.net(k)
        assert_state_equal(bn.neurons(), alone.neurons(), f"net {k}")
        sa, sb = alone.stats(), bn.stats()
        for key in ("spikes", "events", "plastic_events", "flush_rows", "flush_synapses"):
            assert sa[key] == sb[key], (k, key, sa, sb)
        wa, wb = alone.synapses(ell=False), bn.synapses(ell=False)
        assert wa.keys() == wb.keys()
        for key in wa:
            np.testing.assert_array_equal(wa[key], wb[key], err_msg=f"net {k} synapse attribute {key}")
        alone.close()
    b.close()


@pytest.mark.parametrize("model,delays", [("brunel_3pop_stdp", (1, 4, 15)), ("brunel_3pop_stdp", (2, 4, 15)),
                                          ("brunel_with_plasticity", (1, 4, 15))],
                         ids=["stdp-mixed", "stdp-overlapped", "brunel_with_plasticity"])
def test_plastic_batch_export_mid_run_equals_alone(gpu, model, delays):
    """A batched net's weights exported mid-run (the export first brings that net's synapses up
    to date: for eager plasticity it applies the last step's passes once, which the batch's next
    delivery kernel then skips for that net only), then the batch goes on: every net equals the
    same net run alone with the same exports, at both exports."""
    rng = np.random.default_rng(12)
    nets = []
    for k, d in enumerate(delays):
        n = int(rng.integers(500, 1500))
        lay = three_pop_layout(n) if model == "brunel_3pop_stdp" else ([n // 2, n - n // 2], [(0, 1, 0.1), (1, 1, 0.1)])
        nets.append((Layout(*lay), d, 700 + k, None))
    b = Batch(model, nets, 1e-4, fixed_point=True)
    alone = [Snn(model, lay, 1e-4, d, seed=s, fixed_point=True) for lay, d, s, _ in nets]
    exported = [0, 2]  # nets exported after the first half; net 1 only at the end
    for half in range(2):
        b.run(333)
        b.synchronize()
        for k, a in enumerate(alone):
            a.run(333)
            if half == 0 and k not in exported:
                continue
            wa, wb = a.synapses(ell=False), b.net(k).synapses(ell=False)
            for key in wa:
                np.testing.assert_array_equal(wa[key], wb[key], err_msg=f"half {half} net {k} synapse attribute {key}")
            assert_state_equal(b.net(k).neurons(), a.neurons(), f"half {half} net {k}")
    for a in alone:
        a.close()
    b.close()


def test_batch_net_parameters_match_checker(gpu):
    """A per-network Poisson rate (snn_info::param[0]) reaches the model identically on the GPU
    (inside a batch) and in the CPU checker."""
    lay = three_pop_layout(1500)
    b = Batch("brunel_3pop", [(Layout(*lay), 15, 77, [7.5]), (Layout(*lay), 15, 78, [28.0])], 1e-4, fixed_point=True)
    for k, (seed, rate) in enumerate([(77, 7.5), (78, 28.0)]):
        h = O.CpuSnn("brunel_3pop", lay, 1e-4, 15, seed=seed, fixed_point=True)
        h.set_params([rate])
        h.run(200)
        if k == 0:
            b.run(200)
        assert_state_equal(b.net(k).neurons(), h.neurons(), f"net {k}")
    b.close()


def test_borrowed_nets_cannot_be_stepped_or_destroyed(gpu):
    b = Batch("brunel", [(Layout([200, 200], [(0, 1, 0.1), (1, 1, 0.1)]), 1, 1)], 1e-4)
    n = b.net(0)
    with pytest.raises(ValueError):
        n.step()
    with pytest.raises(ValueError):
        n.run(5)
    b.run(10)
    assert n.clock()[0] == 10
    b.close()
