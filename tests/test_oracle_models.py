"""The checker's model semantics against closed forms and invariants that the mathematics
fixes (independent of how the checker or the models are written)."""
import numpy as np
import pytest

from oracle import oracle as O

DT = 1e-4


def test_vogels_isolated_neuron_follows_closed_form_and_period():
    """No synapses: V_k = Vinf + (V0 - Vinf)(1 - a)^k with Vinf = Vrest + Ibg = -40 mV and
    a = dt / 20 ms; first spike at s1 = min{k : V_k > -50 mV}; afterwards spikes are periodic
    with period s1 + 50 (50 refractory steps)."""
    h = O.CpuSnn("vogels_abbott", ([5], [(0, 0, 0.0)]), DT, 1, seed=1)
    a = DT / 0.02
    vinf, v0 = -0.04, -0.06
    k = np.arange(400)
    vk = vinf + (v0 - vinf) * (1 - a) ** k
    s1 = int(np.argmax(vk > -0.05))
    spikes = []
    for i in range(3 * (s1 + 50) + 10):
        s = h.step(True)
        if i < s1:
            np.testing.assert_allclose(h.neurons()["V"], vk[i + 1], rtol=2e-6)
        if s.size:
            assert s.size == 5
            spikes.append(i)
    assert spikes[0] == s1
    assert np.all(np.diff(spikes) == s1 + 50)


def test_vogels_delivery_delay_and_conductance_decay():
    """A spike emitted at step t is received at step t + delay (before that step's update),
    then decays by (1 - dt / 5 ms) in the update: observable on a target with no other input."""
    for delay in (1, 3, 8):
        h = O.CpuSnn("vogels_abbott", ([1, 1], [(0, 1, 1.0)]), DT, delay, seed=1)
        wex = 0.4 * 16e6 / 4  # 0.4 * 16e6 / N^2 siemens, N = 2
        t_spike = None
        for i in range(400):
            s = h.step(True)
            g = h.neurons()["Gex"]
            if t_spike is None and 0 in s:
                t_spike = i
            if t_spike is not None and i == t_spike + delay:
                np.testing.assert_allclose(g[1], wex * (1 - DT / 0.005), rtol=1e-6)
                break
            assert g[1] == 0.0
        assert t_spike is not None


@pytest.mark.parametrize("model,npois", [("brunel", 500), ("brunel_3pop", 200)])
def test_poisson_neurons_use_philox_keyed_on_neuron_and_step(model, npois):
    """Poisson neuron i fires at step t iff u(seed, i, t) < 20 Hz * dt, u the top 24 bits of
    word 0 of Philox4x32-10(ctr = (i, t, 0, tag 4), key = seed) scaled to [0, 1)."""
    n = 1000
    lay = ([500, 500], [(1, 1, 0.0)]) if model == "brunel" else ([200, 640, 160], [])
    seed = 0x1234_5678_9ABC
    h = O.CpuSnn(model, lay, DT, 1, seed=seed)
    key = [seed & 0xffffffff, seed >> 32]
    thr = np.float32(np.float32(20.0) * np.float32(DT))
    for t in range(50):
        s = h.step(True)
        got = np.sort(s[s < npois])
        u = np.array([(O.philox([i, t, 0, 4], key)[0] >> 8) / 16777216.0 for i in range(npois)], np.float32)
        np.testing.assert_array_equal(got, np.nonzero(u < thr)[0])
    assert n == sum(lay[0])


def test_poisson_rate_statistics():
    h = O.CpuSnn("brunel", ([2000, 2000], [(1, 1, 0.0)]), DT, 1, seed=9)
    total = 0
    steps = 2000
    for _ in range(steps):
        s = h.step(True)
        total += np.count_nonzero(s < 2000)
    p = 20 * DT
    mean = 2000 * steps * p
    assert abs(total - mean) < 4 * np.sqrt(mean * (1 - p))


def test_lif_refractory_period():
    """brunel LIF: after a spike no spike for Tref = 20 steps."""
    h = O.CpuSnn("brunel_3pop", ([200, 640, 160], [(0, 1, 0.5), (0, 2, 0.5), (1, 1, 0.1), (1, 2, 0.1),
                                                    (2, 1, 0.1), (2, 2, 0.1)]), DT, 1, seed=4)
    last = np.full(1000, -1000)
    nspk = 0
    for t in range(1500):
        s = h.step(True)
        lif = s[s >= 200]
        assert np.all(t - last[lif] > 20), t
        last[lif] = t
        nspk += lif.size
    assert nspk > 100


def test_stdp_traces_match_spike_trains_and_weights_stay_bounded():
    """Trace-synapse model: the pre trace of every neuron equals the decayed sum over its
    spike arrivals (emission + delay) and the post trace the decayed sum over its spikes
    (closed form from the emitted spike trains); weights stay in [0, wmax]; non-plastic
    weights never move; plasticity moves some weights."""
    n = 1000
    npois = n // 5
    ne = (n - npois) * 4 // 5
    lay = ([npois, ne, n - npois - ne], [(s, d, 0.1) for s in range(3) for d in range(3)])
    D = 15
    h = O.CpuSnn("brunel_3pop_stdp", lay, DT, D, seed=2)
    steps = 400
    spk = np.zeros((steps, n), bool)
    for t in range(steps):
        spk[t, h.step(True)] = True
    x, y = h.traces()
    d = 1 - DT / 0.02
    decay = d ** (steps - 1 - np.arange(steps))
    arr = np.zeros_like(spk)
    arr[D:] = spk[:-D]
    np.testing.assert_allclose(x, (arr * decay[:, None]).sum(0), rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(y, (spk * decay[:, None]).sum(0), rtol=1e-4, atol=1e-6)
    rp, col = h.csr()
    w = h.synapses(ell=False)["W"]
    src = np.repeat(np.arange(n), np.diff(rp))
    plastic = (src >= npois) & (src < npois + ne) & (col >= npois) & (col < npois + ne)
    assert np.all((w[plastic] >= 0) & (w[plastic] <= np.float32(3e-4)))
    w0 = np.where(src < npois + ne, np.float32(1e-4), np.float32(-5e-4))
    np.testing.assert_array_equal(w[~plastic], w0[~plastic])
    assert np.any(w[plastic] != np.float32(1e-4))


def test_stdp_weight_change_single_pair_closed_form():
    """Additive pair rule on a tiny net: every plastic weight equals its initial value plus
    A+ wmax times the sum over post spikes of the pre trace minus A- wmax times the sum over
    arrivals of the post trace, clipped to [0, wmax] (no clipping occurs at these rates, so
    the unclipped closed form applies), with traces taken from the spike trains."""
    n = 400
    npois = n // 5
    ne = (n - npois) * 4 // 5
    lay = ([npois, ne, n - npois - ne], [(0, 1, 0.3), (1, 1, 0.2), (2, 1, 0.1)])
    D = 4
    h = O.CpuSnn("brunel_3pop_stdp", lay, DT, D, seed=8)
    steps = 300
    spk = np.zeros((steps, n), bool)
    for t in range(steps):
        spk[t, h.step(True)] = True
    arr = np.zeros_like(spk)
    arr[D:] = spk[:-D]
    d = 1 - DT / 0.02
    # traces after each step t: x_t = sum_{u<=t} arr[u] d^(t-u)
    X = np.zeros((steps, n))
    Y = np.zeros((steps, n))
    xs = np.zeros(n)
    ys = np.zeros(n)
    for t in range(steps):
        xs = xs * d + arr[t]
        ys = ys * d + spk[t]
        X[t], Y[t] = xs, ys
    rp, col = h.csr()
    w = h.synapses(ell=False)["W"].astype(np.float64)
    src = np.repeat(np.arange(n), np.diff(rp))
    plastic = (src >= npois) & (src < npois + ne) & (col >= npois) & (col < npois + ne)
    Ap, Am = 0.01 * 3e-4, 0.0105 * 3e-4
    s_idx, d_idx = src[plastic], col[plastic]
    pot = (spk[:, d_idx] * X[:, s_idx]).sum(0)
    dep = (arr[:, s_idx] * Y[:, d_idx]).sum(0)
    expect = 1e-4 + Ap * pot - Am * dep
    assert np.all((expect > 0) & (expect < 3e-4))
    np.testing.assert_allclose(w[plastic], expect, rtol=1e-5, atol=1e-10)
    assert np.any(pot > 0) and np.any(dep > 0)


@pytest.mark.parametrize("model,fixed", [("brunel_3pop", True), ("brunel_3pop", False), ("vogels_abbott", True)])
def test_checker_resumes_from_state_clock_and_spikes_in_flight(model, fixed):
    """A fresh checker given another's neuron state, clock and the spikes of the last `delay`
    steps continues exactly as the original for the next steps (what the full-size GPU tests
    use to check single steps of a network too large to simulate on the CPU from the start)."""
    lay = ([200, 640, 160], [(s, d, 0.1) for s in range(3) for d in range(3)]) if model == "brunel_3pop" \
        else ([1000], [(0, 0, 0.1)])
    D = 15 if model == "brunel_3pop" else 1
    h = O.CpuSnn(model, lay, DT, D, seed=3, fixed_point=fixed)
    sp = [h.step(True) for _ in range(300)]
    assert sum(len(x) for x in sp[200:]) > 50
    g = O.CpuSnn(model, lay, DT, D, seed=3, fixed_point=fixed)
    g.set_neurons(h.neurons())
    g.set_clock(*h.clock())
    g.set_in_flight(sp[-D:])
    for t in range(3 * D):
        np.testing.assert_array_equal(np.sort(g.step(True)), np.sort(h.step(True)), err_msg=f"step {t}")
    for k, v in h.neurons().items():
        np.testing.assert_array_equal(g.neurons()[k], v, err_msg=k)
    with pytest.raises(ValueError):
        g.set_in_flight(sp[-D - 1:])  # more than `delay` steps
