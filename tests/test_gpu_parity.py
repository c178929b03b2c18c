"""GPU parity: the gfx950 path against the CPU checker (oracle/), through the C ABI.

Bars: graphs, initial state and integer state are bit-exact; in fixed-point mode every
state value and every spike is bit-exact; in fp32 mode state agrees within the tolerance
written in each test (the reference's own bar is 1e-5 mean absolute difference,
spice_test/cuda/snn.d.cpp:26-44, 120).
"""
import numpy as np
import pytest

from oracle import oracle as O
from spice_amd import Layout, Snn
from tests.helpers import (VOGELS_4000, assert_state_close, assert_state_equal, brunel_ref_layout,
                           three_pop_layout)

pytestmark = pytest.mark.gpu

GRAPH_CASES = [
    ("p1_identity", ([100], [(0, 0, 1.0)]), 3),
    ("p0_empty", ([100], [(0, 0, 0.0)]), 3),
    ("three_pop_ranges", ([10, 20, 30], [(0, 0, 0.5), (0, 1, 0.1), (0, 2, 0.5), (1, 0, 1.0), (1, 2, 0.5)]), 5),
    ("vogels_4000", VOGELS_4000, 1),
    ("brunel_ref_2000", brunel_ref_layout(2000), 7),
    ("three_pop_3162", three_pop_layout(3162), 1),
    ("wide_rows_20000", ([20000], [(0, 0, 0.3)]), 11),
]


@pytest.mark.parametrize("name,layout,seed", GRAPH_CASES, ids=[c[0] for c in GRAPH_CASES])
def test_graph_bit_exact_vs_host_build(gpu, name, layout, seed):
    """Device generator == host build of the same generator, bit for bit."""
    rp_h, col_h, w_h = O.generate_csr(layout, seed)
    net = Snn("brunel", Layout(*layout), 1e-4, 1, seed=seed)
    rp_d, col_d = net.csr()
    assert net.width() == w_h
    np.testing.assert_array_equal(rp_d, rp_h)
    np.testing.assert_array_equal(col_d, col_h)
    # ELL export keeps the reference's format: rows padded with INT_MAX to the layout width
    ell, w = net.adj()
    assert ell.size == net.num_neurons() * w == net.num_synapses()


@pytest.mark.parametrize("model", ["vogels_abbott", "brunel", "brunel_with_plasticity", "synth", "brunel_3pop",
                                   "brunel_3pop_stdp"])
def test_init_state_exact(gpu, model):
    lay = three_pop_layout(1000) if "3pop" in model else ([1000], [(0, 0, 0.1)])
    h = O.CpuSnn(model, lay, 1e-4, 15, seed=3)
    d = Snn(model, Layout(*lay), 1e-4, 15, seed=3)
    assert_state_equal(d.neurons(), h.neurons(), "neurons")
    assert_state_equal(d.synapses(ell=False), h.synapses(ell=False), "synapses")


def _run_pair(model, layout, delay, steps, fixed, seed=5, check_every=1, rtol=1e-5, atol=1e-9, exact=None):
    h = O.CpuSnn(model, layout, 1e-4, delay, seed=seed, fixed_point=fixed)
    d = Snn(model, Layout(*layout), 1e-4, delay, seed=seed, fixed_point=fixed)
    total = 0
    for i in range(steps):
        sh = np.sort(h.step(True))
        sd = np.sort(d.step(True))
        np.testing.assert_array_equal(sd, sh, err_msg=f"{model} step {i}: spike sets differ")
        total += sh.size
        if check_every and (i % check_every == 0 or i == steps - 1):
            if fixed if exact is None else exact:
                assert_state_equal(d.neurons(), h.neurons(), f"step {i}")
            else:
                assert_state_close(d.neurons(), h.neurons(), rtol, atol, f"step {i}")
    return h, d, total


@pytest.mark.parametrize("delay", [1, 15])
def test_tiles_without_incoming_synapses_get_no_delivery(gpu, delay):
    """The reference's Brunel topology targets only population 1, so with 64-neuron tiles the
    first half of the tiles has no incoming synapse and gets no delivery work items (the
    delivery runs over the other tiles only). Fixed point, against the checker: spikes every
    step and state bit for bit, graph-replayed (overlapped at delay 15) and single steps; and a
    network without any synapse (every tile idle: one delivery work item remains) steps
    exactly as the checker's."""
    lay = brunel_ref_layout(3000)
    h = O.CpuSnn("brunel", lay, 1e-4, delay, seed=31, fixed_point=True)
    d = Snn("brunel", Layout(*lay), 1e-4, delay, seed=31, fixed_point=True)
    d.set_launch(64, 0, 0, 256, 2)
    for n in (1, 37, 128, 2, 200):
        if n == 1:
            np.testing.assert_array_equal(np.sort(d.step(True)), np.sort(h.step(True)))
        else:
            d.run(n)
            h.run(n)
        assert_state_equal(d.neurons(), h.neurons(), f"after {n}")
    assert d.stats()["events"] > 0
    d.close(), h.close()
    empty = ([600, 600], [(0, 1, 0.0)])
    _run_pair("brunel", empty, delay, 60, True, seed=32, check_every=20)


@pytest.mark.parametrize("delay", [1, 8])
def test_vogels_reference_scenario_fixed_point_bit_exact(gpu, delay):
    """Reference scenario (vogels, N=4000, p=0.02, 1000 steps, delay 1 and 8): per-step spike
    sets equal, state bit-exact in fixed-point mode."""
    h, d, total = _run_pair("vogels_abbott", VOGELS_4000, delay, 1000, True, check_every=50)
    assert total > 1000


@pytest.mark.parametrize("delay", [1, 8])
def test_vogels_reference_scenario_fp32(gpu, delay):
    """Same scenario in fp32 atomics mode: spike sets equal every step, final state within
    the reference's bar (mean absolute difference per neuron <= 1e-5) and 1e-5 relative."""
    h, d, total = _run_pair("vogels_abbott", VOGELS_4000, delay, 1000, False, check_every=100, rtol=1e-5, atol=1e-7)
    a, b = d.neurons(), h.neurons()
    mad = np.mean(sum(np.abs(a[k].astype(np.float64) - b[k]) for k in a) / len(a))
    assert mad <= 1e-5


@pytest.mark.parametrize("model", ["brunel", "brunel_3pop", "synth"])
def test_poisson_models_bit_exact_fixed_point(gpu, model):
    """Models with random (Poisson) neurons: Philox keyed on (seed, neuron, step) makes the GPU
    draw the checker's numbers, so spikes and state are bit-exact (the reference could not
    test brunel at all, spice_test/cuda/snn.d.cpp:102-103)."""
    lay = three_pop_layout(2000) if "3pop" in model else brunel_ref_layout(2000)
    if model == "synth":
        lay = ([2000], [(0, 0, 0.05)])
    _run_pair(model, lay, 15 if model != "synth" else 1, 400, True, check_every=40)



@pytest.mark.parametrize("case", ["empty_segments", "bitmap_overflow"])
def test_delivery_segment_lookup_paths_bit_exact(gpu, case):
    """The delivery finds the spike (segment) of each delivered entry from an LDS bitmap of
    segment starts, after compacting away the spikes with no target in the tile, and by binary
    search when a chunk holds more than 32 x recv_block entries. Both paths against the checker,
    bit-exact in fixed-point mode: a sparse network with 64-target tiles (most (spike, tile)
    segments empty) and a dense one with 1024-target tiles (~2e4 entries per chunk)."""
    if case == "empty_segments":
        model, lay, delay, launch, steps = "brunel_3pop", three_pop_layout(20000, 0.002), 15, (64, 0, 0, 256, 1), 150
    else:
        model, lay, delay, launch, steps = "synth", ([20000], [(0, 0, 0.2)]), 1, (1024, 0, 0, 256, 1), 60
    h = O.CpuSnn(model, lay, 1e-4, delay, seed=7, fixed_point=True)
    d = Snn(model, Layout(*lay), 1e-4, delay, seed=7, fixed_point=True)
    d.set_launch(*launch)
    total = 0
    for i in range(steps):
        sh = np.sort(h.step(True))
        sd = np.sort(d.step(True))
        np.testing.assert_array_equal(sd, sh, err_msg=f"{case} step {i}: spike sets differ")
        total += sh.size
        if i % 30 == 29:
            assert_state_equal(d.neurons(), h.neurons(), f"{case} step {i}")
    assert total > 500, "the network must deliver spikes for the test to mean anything"
    d.close()

def test_one_step_fp32_within_1e5(gpu):
    """One step of update + propagation from identical state matches within 1e-5 relative:
    before every step the GPU neuron state is reset to the checker's, so each comparison
    covers exactly one step (spikes in flight are identical because spike sets are)."""
    lay = three_pop_layout(3162)
    h = O.CpuSnn("brunel_3pop", lay, 1e-4, 15, seed=1)
    d = Snn("brunel_3pop", Layout(*lay), 1e-4, 15, seed=1)
    for i in range(120):
        d.set_neurons(h.neurons())
        sh = np.sort(h.step(True))
        sd = np.sort(d.step(True))
        np.testing.assert_array_equal(sd, sh, err_msg=f"step {i}")
        assert_state_close(d.neurons(), h.neurons(), 1e-5, 1e-9, f"step {i}")


def test_lazy_stdp_generic_contract_matches_dense(gpu):
    """Lazy plasticity for the generic synapse contract (reference model brunel+) vs the
    checker's dense per-step updates. Replay merges event-free steps into one update(n) whose
    trace decay is pow(d, n) (one rounding) where the dense checker multiplies by d once per
    step (n roundings of up to 1/2 ulp each), so synapse traces drift apart by up to about
    (steps / 2) ulp: 2e-4 relative over 1200 steps (several flush windows), not bit for bit.
    Spike sets identical."""
    model = "brunel_with_plasticity"
    lay = brunel_ref_layout(1500)
    steps = 1200
    h, d, _ = _run_pair(model, lay, 15, steps, True, check_every=100, rtol=1e-5, atol=1e-9, exact=False)
    sd, sh = d.synapses(ell=False), h.synapses(ell=False)
    rtol = 3 * steps * 2.0 ** -24  # 2.1e-4
    for k in sd:
        np.testing.assert_allclose(sd[k], sh[k], rtol=rtol, atol=1e-9, err_msg=k)


@pytest.mark.parametrize("model", ["brunel_3pop_stdp", "brunel_3pop_stdp_deferred"])
@pytest.mark.parametrize("delay", [1, 15, 255])
def test_eager_stdp_neuron_traces_bit_exact(gpu, model, delay):
    """Trace-synapse plasticity, applied eagerly -- pre pass over arriving rows, post pass over
    spiking targets through the by-target index -- equals the checker's dense per-step
    plasticity bit for bit (fixed-point mode): spikes every step, neuron state, and every weight
    mid-run (an export applies the pending passes once and the run goes on) and at the end;
    delays 1, 15 and the largest supported (255). brunel_3pop_stdp promises refractory plastic
    endpoints, so its passes apply every event and its deliveries only read weights;
    brunel_3pop_stdp_deferred (the same model without the promise) leaves the events of rows
    arriving next to that delivery."""
    lay = three_pop_layout(1500)
    h = O.CpuSnn(model, lay, 1e-4, delay, seed=5, fixed_point=True)
    d = Snn(model, Layout(*lay), 1e-4, delay, seed=5, fixed_point=True)
    for i in range(900):
        sh = np.sort(h.step(True))
        sd = np.sort(d.step(True))
        np.testing.assert_array_equal(sd, sh, err_msg=f"step {i}: spike sets differ")
        if i % 100 == 0:
            assert_state_equal(d.neurons(), h.neurons(), f"step {i}")
        if i in (299, 300, 601):  # twice in a row: the pending events are applied only once
            assert_state_equal(d.synapses(ell=False), h.synapses(ell=False), f"weights after step {i}")
    assert_state_equal(d.synapses(ell=False), h.synapses(ell=False), "weights")
    st = d.stats()
    # events applied by deliveries: only with deferred events
    assert (st["plastic_events"] > 0) == (model == "brunel_3pop_stdp_deferred"), st
    assert st["flush_rows"] > 0 and st["flush_synapses"] > 0


@pytest.mark.parametrize("delay", [2, 15])
def test_overlapped_step_equals_two_kernel_step(gpu, delay):
    """brunel_3pop_stdp at delay >= 2 runs overlapped steps (the update of step T beside the
    delivery of T + 1 and the passes of T - 1, one kernel per step boundary; the gate of
    receive() tested by the update); brunel_3pop_stdp_deferred runs the two-kernel step with
    deferred events. Same network, fixed point: identical spikes, neuron state, weights and
    event counts after runs of assorted lengths (graph blocks, single steps, step() calls and
    weight exports in between), and equal to the checker."""
    lay = three_pop_layout(2500)
    a = Snn("brunel_3pop_stdp", Layout(*lay), 1e-4, delay, seed=21, fixed_point=True)
    b = Snn("brunel_3pop_stdp_deferred", Layout(*lay), 1e-4, delay, seed=21, fixed_point=True)
    h = O.CpuSnn("brunel_3pop_stdp", lay, 1e-4, delay, seed=21, fixed_point=True)
    done = 0
    for n in (1, 2, 3, 37, 64, 5, 128, 1, 260):
        for net in (a, b):
            if n == 1:
                net.step()
            else:
                net.run(n)
        h.run(n)
        done += n
        assert_state_equal(a.neurons(), b.neurons(), f"neurons after {done} steps")
        if n in (3, 128, 260):
            assert_state_equal(a.synapses(ell=False), b.synapses(ell=False), f"weights after {done} steps")
            assert_state_equal(a.synapses(ell=False), h.synapses(ell=False), f"weights vs checker after {done} steps")
    assert_state_equal(a.neurons(), h.neurons(), "neurons vs checker")
    sa, sb = a.stats(), b.stats()
    assert sa["spikes"] == sb["spikes"] > 0 and sa["events"] == sb["events"], (sa, sb)
    assert sa["plastic_events"] == 0 and sb["plastic_events"] > 0, (sa, sb)
    a.close(), b.close(), h.close()


def test_run_graph_equals_step(gpu):
    """spice_hip_run (captured HIP graph) == repeated spice_hip_step."""
    lay = three_pop_layout(2000)
    a = Snn("brunel_3pop", Layout(*lay), 1e-4, 15, seed=9, fixed_point=True)
    b = Snn("brunel_3pop", Layout(*lay), 1e-4, 15, seed=9, fixed_point=True)
    for _ in range(100):
        a.step()
    b.run(100)
    assert a.clock() == b.clock()
    assert_state_equal(a.neurons(), b.neurons(), "run vs step")
    assert a.stats() == b.stats()


@pytest.mark.parametrize("model", ["brunel_3pop", "brunel_3pop_stdp", "brunel_with_plasticity"])
def test_short_and_unaligned_runs_replay_graphs(gpu, model):
    """Runs of any length, from any step, replay captured graphs (1-step graphs per clock
    phase, 4/8/16/32-step blocks) and equal the same number of spice_hip_step calls, fixed
    point: run(20) from step 0, then lengths that start and end off the 4-step phase, then runs
    of 100 and 2,500 steps (the latter: 79 graph launches, more than the 64 kept in flight)."""
    lay = three_pop_layout(2000) if "3pop" in model else brunel_ref_layout(2000)
    a = Snn(model, Layout(*lay), 1e-4, 15, seed=17, fixed_point=True)
    b = Snn(model, Layout(*lay), 1e-4, 15, seed=17, fixed_point=True)
    total = 0
    for n in (20, 1, 3, 5, 7, 2, 17, 33, 1, 100, 4, 2500, 3):
        b.run(n)
        for _ in range(n):
            a.step()
        total += n
        assert a.clock() == b.clock() and b.clock()[0] == total
        assert_state_equal(b.neurons(), a.neurons(), f"after run({n}), step {total}")
        assert np.array_equal(np.sort(a.step_spikes_of_last()), np.sort(b.step_spikes_of_last())), total
    assert_state_equal(b.synapses(ell=False), a.synapses(ell=False), "synapses")
    assert a.stats() == b.stats()


@pytest.mark.parametrize("model", ["brunel_3pop_stdp", "brunel_3pop_stdp_deferred"])
@pytest.mark.parametrize("launch", [(256, 128, 1024, 256, 1), (256, 128, 1024, 512, 2), (256, 128, 1024, 1024, 1)],
                         ids=["block256", "block512", "block1024"])
def test_run_profiled_equals_run(gpu, launch, model):
    """spice_hip_run_profiled (each kernel launched with events on its dispatch) steps the net
    exactly as spice_hip_run, at each delivery workgroup size, and reports positive kernel
    times: overlapped steps (brunel_3pop_stdp at delay 15) time their first delivery, last
    update and the overlapped kernels between; two-kernel steps (the deferred-event variant)
    their deliveries and updates."""
    lay = three_pop_layout(2000)
    a = Snn(model, Layout(*lay), 1e-4, 15, seed=9, fixed_point=True)
    b = Snn(model, Layout(*lay), 1e-4, 15, seed=9, fixed_point=True)
    a.set_launch(*launch)
    b.set_launch(*launch)
    a.run(40)
    ms = a.run_profiled(150)
    a.run(10)
    b.run(200)
    assert a.clock() == b.clock()
    assert_state_equal(a.neurons(), b.neurons(), "profiled vs run")
    assert_state_equal(a.synapses(ell=False), b.synapses(ell=False), "weights")
    assert ms["receive"] > 0 and ms["update"] > 0
    assert (ms["step"] > 0) == (model == "brunel_3pop_stdp"), ms


def test_cpp_drop_in_api(gpu):
    """The public C++ API (spice::hip::snn<Model>, include/spice/hip/snn.h) driven by a C++
    program that mirrors the reference's dSNN tests (spice_test/cuda/snn.d.cpp:72-157)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dsnn_test")
    assert os.path.exists(exe), "build first (__graft_entry__.build)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_plastic_delay_limit(gpu):
    """Plastic models keep a 256-step spike history, so delays above 255 are rejected."""
    with pytest.raises(ValueError):
        Snn("brunel_3pop_stdp", Layout(*three_pop_layout(500)), 1e-4, 256)
    Snn("brunel_3pop_stdp", Layout(*three_pop_layout(500)), 1e-4, 255).close()
    Snn("brunel_3pop", Layout(*three_pop_layout(500)), 1e-4, 1000).close()  # static models: any delay


def test_unknown_flags_are_rejected(gpu):
    """Only SPICE_FLAG_FIXED_POINT is a flag; other bits are an invalid argument."""
    with pytest.raises(ValueError):
        Snn("brunel_3pop", Layout(*three_pop_layout(500)), 1e-4, 15, flags=2)
    Snn("brunel_3pop", Layout(*three_pop_layout(500)), 1e-4, 15, flags=1).close()


def test_launch_shapes_do_not_change_results(gpu):
    """Delivery tile width, flush work-item size and flush grid changed between steps (fixed
    point, plastic model): state, weights and counters equal the default-shaped run."""
    lay = three_pop_layout(2500)
    a = Snn("brunel_3pop_stdp", Layout(*lay), 1e-4, 15, seed=21, fixed_point=True)
    b = Snn("brunel_3pop_stdp", Layout(*lay), 1e-4, 15, seed=21, fixed_point=True)
    shapes = [(64, 64, 1, 256, 1), (1024, 2048, 4096, 1024, 3), (320, 128, 37, 512, 16), (0, 0, 0, 0, 0)]
    for k, (tw, fs, fg, bs, rp) in enumerate(shapes):
        b.set_launch(tw, fs, fg, bs, rp)
        got = b.launch()
        if tw:
            assert got == {"tile_w": tw, "flush_slots": fs, "flush_grid": fg, "recv_block": bs, "recv_parts": rp}
        a.run(400)
        b.run(400)
        assert_state_equal(b.neurons(), a.neurons(), f"shape {k}")
    assert_state_equal(b.synapses(ell=False), a.synapses(ell=False), "weights")
    assert a.stats() == b.stats()
    with pytest.raises(ValueError):
        b.set_launch(100, 0, 0)  # not a multiple of 64


def test_autotune_keeps_the_network_in_place(gpu):
    """The autotuner runs its candidates on a copy: the tuned network has not advanced and
    its results equal an untuned twin's; a second call with the same shape hits the cache."""
    lay = three_pop_layout(3000)
    a = Snn("brunel_3pop_stdp", Layout(*lay), 1e-4, 15, seed=5, fixed_point=True)
    a.run(50)
    before = a.neurons()
    cfg, us = a.autotune(warmup=300, steps=50)
    assert us > 0 and a.launch() == cfg and a.clock()[0] == 50
    assert_state_equal(a.neurons(), before, "autotune must not step the network")
    b = Snn("brunel_3pop_stdp", Layout(*lay), 1e-4, 15, seed=5, fixed_point=True)
    b.run(50)
    b.run(1200)
    a.run(1200)
    assert_state_equal(a.neurons(), b.neurons(), "tuned vs untuned")
    assert_state_equal(a.synapses(ell=False), b.synapses(ell=False), "weights")
    c = Snn("brunel_3pop_stdp", Layout(*lay), 1e-4, 15, seed=5, fixed_point=True)
    cfg2, us2 = c.autotune(warmup=300, steps=50)
    assert cfg2 == cfg and us2 == us  # cached


@pytest.mark.parametrize("model", ["brunel_3pop_stdp", "brunel_3pop_stdp_deferred"])
@pytest.mark.parametrize("delay", [1, 15])
def test_trace_stdp_graph_replay_bit_exact(gpu, model, delay):
    """HIP-graph replay of plastic steps (the plasticity passes of each step inside the next
    kernel; with delay 15 brunel_3pop_stdp runs overlapped steps) equals the dense checker bit
    for bit over 1,500 steps, weights exported between the graph runs, with autotuned-style
    launch shapes (split delivery lists, small flush items)."""
    lay = three_pop_layout(2000)
    h = O.CpuSnn(model, lay, 1e-4, delay, seed=13, fixed_point=True)
    d = Snn(model, Layout(*lay), 1e-4, delay, seed=13, fixed_point=True)
    d.set_launch(256, 64, 64, 256, 3)
    for k in range(3):
        h.run(500)
        d.run(500)  # 500 = 15 graphs of 32 steps + eager steps
        assert_state_equal(d.neurons(), h.neurons(), f"after {500 * (k + 1)} steps")
        assert_state_equal(d.synapses(ell=False), h.synapses(ell=False), f"weights after {500 * (k + 1)} steps")
    st = d.stats()
    assert (st["plastic_events"] > 0) == (model == "brunel_3pop_stdp_deferred"), st
    assert st["flush_rows"] > 0 and st["flush_synapses"] > 0


@pytest.mark.parametrize("model", ["brunel_3pop", "brunel_3pop_stdp"])
@pytest.mark.parametrize("fixed", [False, True])
def test_lif_closed_form_on_gpu(gpu, model, fixed):
    """The headline model's neuron on the GPU meets the closed forms of
    tests/test_lif_closed_form.py (threshold, exact reset, refractory period, tau, 5/N and
    -25/N weights, delivery delay, the first STDP pre event)."""
    from tests.test_lif_closed_form import check_lif_closed_form

    def make(m, layout, dt, delay):
        return Snn(m, Layout(*layout), dt, delay, seed=3, fixed_point=fixed)

    for delay in (1, 15):
        check_lif_closed_form(make, model, delay).close()


@pytest.mark.parametrize("name,layout,seed", GRAPH_CASES[2:6], ids=[c[0] for c in GRAPH_CASES[2:6]])
def test_generate_graph_into_device_buffers(gpu, name, layout, seed):
    """spice_hip_generate_graph writes the same graph as the host build into caller-owned
    device memory (torch tensors here): offsets alone first, then offsets and targets; a
    column buffer that is too small is rejected with the graph's size reported."""
    import torch
    from spice_amd.snn import generate_graph
    rp_h, col_h, w_h = O.generate_csr(layout, seed)
    n = rp_h.size - 1
    rp = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
    nnz, w = generate_graph(layout, seed, rp.data_ptr())
    assert nnz == col_h.size and w == w_h
    np.testing.assert_array_equal(rp.cpu().numpy(), rp_h)
    col = torch.full((max(nnz, 1),), -1, dtype=torch.int32, device="cuda")
    rp.fill_(-1)
    assert generate_graph(layout, seed, rp.data_ptr(), col.data_ptr(), col.numel()) == (nnz, w)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(rp.cpu().numpy(), rp_h)
    np.testing.assert_array_equal(col.cpu().numpy()[:nnz], col_h)
    if nnz > 1:
        with pytest.raises(ValueError, match="column buffer"):
            generate_graph(layout, seed, rp.data_ptr(), col.data_ptr(), nnz - 1)


def test_probe_random_lines(gpu):
    """The random-line probe (spice_hip_probe_random_lines) times reads and read-modify-writes
    of 4-byte values at pseudo-random places: positive times, a write-back costs time, and
    invalid sizes are rejected before any device work."""
    from spice_amd.snn import probe_random_lines
    rd = probe_random_lines(1 << 28, 1 << 20, False, reps=5)
    rmw = probe_random_lines(1 << 28, 1 << 20, True, reps=5)
    assert rd > 0 and rmw > 0
    assert rmw > 0.5 * rd
    for args in ((32, 1 << 20, False), (1 << 28, 0, False), (1 << 40, 1, True)):
        with pytest.raises(ValueError):
            probe_random_lines(*args, reps=1)


@pytest.mark.parametrize("model", ["synth", "vogels_abbott"])
def test_large_network_update_blocks_bit_exact(gpu, model):
    """From 131,072 neurons the two-kernel step updates in 1024-thread workgroups (fewer of
    them reserve ring slots on the step's spike counter). Delay 1 (two kernels per step), N =
    150,000, fixed point, against the checker: spikes every single step and state bit for bit
    after single and graph-replayed steps (Vogels-Abbott: the whole population's first,
    synchronous spike within them)."""
    n = 150_000
    lay = ([n], [(0, 0, 0.0002)])
    h = O.CpuSnn(model, lay, 1e-4, 1, seed=51, fixed_point=True)
    d = Snn(model, Layout(*lay), 1e-4, 1, seed=51, fixed_point=True)
    for i in range(20):
        np.testing.assert_array_equal(np.sort(d.step(True)), np.sort(h.step(True)), err_msg=f"step {i}")
    d.run(280)  # Vogels-Abbott's background current brings every neuron to threshold at ~140 steps
    h.run(280)
    assert_state_equal(d.neurons(), h.neurons(), "after 300 steps")
    assert d.stats()["spikes"] > 0
    d.close(), h.close()


@pytest.mark.parametrize("model,tile_w,block", [("synth", 2048, 256), ("synth", 4096, 1024), ("synth", 8192, 512),
                                                  ("vogels_abbott", 8192, 256)])
def test_wide_tiles_bit_exact(gpu, model, tile_w, block):
    """Delivery tiles wider than 1,024 targets (a second instantiation of the delivery kernel,
    static models with two kernels per step): fixed point, against the checker, spikes every
    single step and state bit for bit after graph-replayed steps; plastic models and overlapped
    steps refuse them."""
    n = 20_000
    lay = ([n], [(0, 0, 0.01)])
    h = O.CpuSnn(model, lay, 1e-4, 1, seed=61, fixed_point=True)
    d = Snn(model, Layout(*lay), 1e-4, 1, seed=61, fixed_point=True)
    d.set_launch(tile_w, 0, 0, block, 2)
    assert d.launch()["tile_w"] == tile_w
    for i in range(10):
        np.testing.assert_array_equal(np.sort(d.step(True)), np.sort(h.step(True)), err_msg=f"step {i}")
    d.run(290)
    h.run(290)
    assert_state_equal(d.neurons(), h.neurons(), "after 300 steps")
    assert d.stats()["events"] > 0
    d.close(), h.close()
    for m, delay in (("brunel_3pop_stdp", 1), ("brunel_3pop", 15)):
        x = Snn(m, Layout(*three_pop_layout(800)), 1e-4, delay, seed=1)
        with pytest.raises(ValueError):
            x.set_launch(2048)
        x.close()


@pytest.mark.parametrize("model,n,p,launch", [
    ("vogels_abbott", 6000, 0.02, (320, 0, 0, 256, 1)),  # a partial second chunk, a ragged last tile
    ("synth", 5000, 0.05, (1024, 0, 0, 256, 1)),         # four update chunks per tile
    ("vogels_abbott", 4000, 0.05, (128, 0, 0, 512, 1)),  # tiles narrower than the workgroup
    ("synth", 3000, 0.05, (512, 0, 0, 1024, 1)),
    ("vogels_abbott", 4000, 0.05, (256, 0, 0, 256, 3)),  # several parts per tile: two kernels
    ("synth", 3000, 0.05, (2048, 0, 0, 256, 1))])        # more than four chunks: two kernels
def test_fused_step_bit_exact(gpu, model, n, p, launch):
    """Delay-1 static networks whose tiles all receive input, one delivery workgroup per tile
    of at most four workgroup widths, run one kernel per step: the workgroup that delivers a
    tile's inputs updates the tile's neurons. Fixed point, against the checker: spikes every
    single step, state bit for bit after graph-replayed and profiled steps; the profiled steps
    report the one step kernel only (the other shapes: two kernels)."""
    fused = launch[4] == 1 and launch[0] <= 4 * launch[3]
    lay = ([n], [(0, 0, p)])
    h = O.CpuSnn(model, lay, 1e-4, 1, seed=71, fixed_point=True)
    d = Snn(model, Layout(*lay), 1e-4, 1, seed=71, fixed_point=True)
    d.set_launch(*launch)
    for i in range(12):
        np.testing.assert_array_equal(np.sort(d.step(True)), np.sort(h.step(True)), err_msg=f"step {i}")
    d.run(250)
    h.run(250)
    assert_state_equal(d.neurons(), h.neurons(), "after 262 steps")
    ms = d.run_profiled(38)
    h.run(38)
    if fused:
        assert ms["receive"] == 0 and ms["update"] == 0 and ms["step"] > 0, ms
    else:
        assert ms["receive"] > 0 and ms["update"] > 0 and ms["step"] == 0, ms
    assert_state_equal(d.neurons(), h.neurons(), "after 300 steps")
    for i in range(3):
        np.testing.assert_array_equal(np.sort(d.step(True)), np.sort(h.step(True)), err_msg=f"step {300 + i}")
    assert d.stats()["events"] > 0 and d.stats()["spikes"] > 0
    d.close(), h.close()
