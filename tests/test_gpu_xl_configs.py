"""The benchmark configurations at their full size on one MI355X, against the CPU checker and
the host build of the generator:

  * 1e9 synapses, Brunel + E->E STDP (N = 100,000, delay 15): the device graph against the host
    generator (every row offset, sampled rows' targets, the expected synapse count); 40 steps
    in fp32 each checked from identical state against the checker (spikes equal, state within
    1e-5, every one of the 1e9 weights equal); 2,000 fixed-point steps replayed as graphs
    equal to 2,000 single steps, with plastic weights in [0, wmax] and fixed ones unchanged;
  * 1e10 synapses, plain Brunel (N = 316,228, col offsets beyond 2^31, 2^32 and 2^33): every
    row offset and the rows straddling those offsets against the host generator; 15 steps,
    fixed point, bit-exact against a checker resumed from the GPU's state, clock and spikes in
    flight (holding only the rows those spikes use);
  * 1e8 synapses, Vogels-Abbott (N = 70,710, delay 1): 10 s of biological time, per-population
    firing rate and ISI CV within 2 % of the checker, 4 seeds pooled;
  * the 512-net parameter sweep, fixed point: batched equal to each net alone, bit for bit;
  * a served stream that holds a 1e9-synapse STDP job beside smaller ones, fixed point: each
    job's spike stream and counters equal the same job run alone.
"""
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bench
from oracle import oracle as O
from spice_amd import Batch, Layout, Server, Snn
from spice_amd.raster import Raster
from spice_amd.serve import spike_digest
from tests.helpers import assert_state_close, assert_state_equal, three_pop_layout

pytestmark = pytest.mark.gpu

DT = 1e-4


def _say(msg):
    print(f"[xl] {msg}", flush=True)


def _expected_nnz(pops, conns):
    """Mean and standard deviation of the synapse count: per (row, connection) the degree is
    a rounded normal of mean range*p and variance range*p*(1-p)."""
    mean = sum(pops[s] * pops[d] * p for s, d, p in conns)
    var = sum(pops[s] * pops[d] * p * (1 - p) for s, d, p in conns)
    return mean, np.sqrt(var)


def _check_rows(net, lay, seed, r0, r1):
    rp_d, col_d = net.csr_rows(r0, r1)
    rp_h, col_h = O.generate_rows(lay, seed, r0, r1)
    np.testing.assert_array_equal(rp_d, rp_h, err_msg=f"row offsets [{r0}, {r1})")
    np.testing.assert_array_equal(col_d, col_h, err_msg=f"targets of rows [{r0}, {r1})")
    return rp_d


def test_brunel_stdp_1e9(gpu):
    model, seed, D = "brunel_3pop_stdp", 1, 15
    lay = three_pop_layout(100_000)
    n = sum(lay[0])
    npois, ne_end = lay[0][0], lay[0][0] + lay[0][1]

    _say("1e9: fp32 net, graph against the host generator")
    d = Snn(model, Layout(*lay), DT, D, seed=seed)
    rp, col = d.csr()
    nnz = int(rp[-1])
    mean, sd = _expected_nnz(*lay)
    assert abs(nnz - mean) < 6 * sd, (nnz, mean, sd)
    assert rp[0] == 0 and np.all(np.diff(rp) >= 0) and np.all(np.diff(rp) <= d.width())
    rp_h, _ = O.generate_rows(lay, seed, 0, n, cols=False)
    np.testing.assert_array_equal(rp, rp_h)
    rng = np.random.default_rng(0)
    for r in [0, npois - 1, npois, ne_end - 1, ne_end, n - 1] + list(rng.integers(0, n, 10)):
        _, c_h = O.generate_rows(lay, seed, int(r), int(r) + 1)
        np.testing.assert_array_equal(col[rp[r]:rp[r + 1]], c_h, err_msg=f"row {r}")

    _say("1e9: 40 fp32 steps, each from the checker's state")
    h = O.CpuSnn.from_csr(model, rp, col, d.width(), DT, D, seed=seed)
    v0 = rng.uniform(0.0, 0.02, n).astype(np.float32)  # LIF neurons fire from the first steps
    for net in (h, d):
        st = net.neurons()
        st["V"][npois:] = v0[npois:]
        net.set_neurons(st)
    spikes = 0
    for t in range(40):
        d.set_neurons(h.neurons())
        sh = np.sort(h.step(True))
        sd_ = np.sort(d.step(True))
        np.testing.assert_array_equal(sd_, sh, err_msg=f"step {t}")
        assert_state_close(d.neurons(), h.neurons(), 1e-5, 1e-9, f"step {t}")
        spikes += sh.size
    assert spikes > 5_000, spikes
    w_h = h.synapses(ell=False)["W"]
    h.close()
    w_d = d.synapses(ell=False)["W"]
    moved = np.count_nonzero((w_d != np.float32(1e-4)) & (w_d != np.float32(-5e-4)))
    assert np.array_equal(w_d.view(np.uint32), w_h.view(np.uint32)), "weights differ from the checker"
    assert moved > 1000, moved  # STDP events were applied in the window
    del w_h, w_d
    d.close()
    _say(f"1e9: {spikes} spikes in 40 steps, weights equal ({moved} off their initial value)")

    _say("1e9: 2,000 fixed-point steps, graph replay vs single steps")
    a = Snn(model, Layout(*lay), DT, D, seed=seed, fixed_point=True)
    b = Snn(model, Layout(*lay), DT, D, seed=seed, fixed_point=True)
    a.run(2000)
    for _ in range(2000):
        b.step()
    assert a.clock() == b.clock()
    assert_state_equal(a.neurons(), b.neurons(), "run vs step")
    st = a.stats()
    assert st == b.stats() and st["flush_synapses"] > 0, st
    wa = a.synapses(ell=False)["W"]
    a.close()
    wb = b.synapses(ell=False)["W"]
    b.close()
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), "run vs step weights"
    del wb
    changed = 0
    for r0 in range(0, n, 10_000):  # plastic: E -> E; bounds and fixed weights row block by row block
        r1 = min(n, r0 + 10_000)
        src = np.repeat(np.arange(r0, r1, dtype=np.int32), np.diff(rp[r0:r1 + 1]))
        c, w = col[rp[r0]:rp[r1]], wa[rp[r0]:rp[r1]]
        pl = (src >= npois) & (src < ne_end) & (c >= npois) & (c < ne_end)
        assert np.all((w[pl] >= 0) & (w[pl] <= np.float32(3e-4))), r0
        w0 = np.where(src < ne_end, np.float32(1e-4), np.float32(-5e-4))
        assert np.array_equal(w[~pl], w0[~pl]), r0
        changed += int(np.count_nonzero(w[pl] != np.float32(1e-4)))
    assert changed > 1_000_000, changed
    _say(f"1e9: {changed} plastic weights moved, all within [0, wmax]")


def test_brunel_1e10_past_32_bit_offsets(gpu):
    model, seed, D = "brunel_3pop", 3, 15
    lay = three_pop_layout(316_228)
    n = sum(lay[0])
    npois = lay[0][0]

    _say("1e10: fixed-point net, every row offset against the host generator")
    d = Snn(model, Layout(*lay), DT, D, seed=seed, fixed_point=True)
    rp, _ = d.csr_rows(0, n, cols=False)
    rp_h, _ = O.generate_rows(lay, seed, 0, n, cols=False)
    np.testing.assert_array_equal(rp, rp_h)
    nnz = int(rp[-1])
    mean, sd = _expected_nnz(*lay)
    assert nnz > 2 ** 33 and abs(nnz - mean) < 6 * sd, (nnz, mean, sd)
    for edge in (2 ** 31, 2 ** 32, 2 ** 33):
        r = int(np.searchsorted(rp, edge, side="right")) - 1  # rp[r] <= edge < rp[r + 1]
        assert rp[r] <= edge < rp[r + 1]
        _check_rows(d, lay, seed, r - 1, r + 2)
    _check_rows(d, lay, seed, n - 3, n)

    _say("1e10: 45 steps, then 15 steps against a resumed checker")
    rng = np.random.default_rng(1)
    d.run(30)
    st = d.neurons()
    forced = rng.choice(np.arange(npois, n), 3000, replace=False)  # excitatory and inhibitory sources
    st["V"][forced] = 0.0205
    d.set_neurons(st)
    flight = [np.sort(d.step(True)) for _ in range(D)]  # steps 30 .. 44: delivered at 45 .. 59
    assert np.all(np.isin(forced, flight[0]))
    rows = np.unique(np.concatenate(flight))
    lens = np.zeros(n, np.int64)
    parts = []
    for r in rows:
        rr, cc = d.csr_rows(int(r), int(r) + 1)
        lens[r] = rr[1] - rr[0]
        parts.append(cc)
    for r in rng.choice(rows, 8, replace=False):
        _check_rows(d, lay, seed, int(r), int(r) + 1)
    rp_red = np.concatenate(([0], np.cumsum(lens)))
    col_red = np.concatenate(parts)
    _say(f"1e10: {rows.size} rows in flight, {col_red.size} synapses delivered in the window")
    h = O.CpuSnn.from_csr(model, rp_red, col_red, d.width(), DT, D, seed=seed, fixed_point=True)
    del parts, col_red
    h.set_neurons(d.neurons())
    h.set_clock(*d.clock())
    h.set_in_flight(flight)
    ev0 = d.stats()["events"]
    for t in range(D):
        sh = np.sort(h.step(True))
        sd_ = np.sort(d.step(True))
        np.testing.assert_array_equal(sd_, sh, err_msg=f"step {45 + t}")
    assert_state_equal(d.neurons(), h.neurons(), "after 15 resumed steps")
    ev = d.stats()["events"] - ev0
    assert ev == int(sum(lens[np.asarray(s, np.int64)].sum() for s in flight)), ev
    ne_end = npois + lay[0][1]
    assert np.any(forced >= ne_end) and np.any(forced < ne_end)  # both weight signs were delivered
    h.close()
    d.close()
    _say(f"1e10: {ev} events in 15 steps, bit-exact")


def test_vogels_1e8_rates_and_isi_cv_over_10s(gpu):
    n = 70_710
    lay = ([n], [(0, 0, 0.02)])
    nexc = int(np.float32(0.8) * np.float32(n))  # vogels_abbott: sources below 0.8 N excite
    pops = [nexc, n - nexc]
    seeds, steps = range(1, 5), 100_000

    def cpu(seed):
        h = O.CpuSnn("vogels_abbott", lay, DT, 1, seed=seed)
        off, ids = h.run_record(steps, cap=60_000_000)
        h.close()
        r = Raster(n, pops)
        r.add(0, off, ids)
        return r.rates(DT), r.isi_cv()

    _say("1e8: 4 checker runs of 10 s (threads) and the GPU runs")
    rasters, lock = {}, threading.Lock()

    def sink(job, first, off, ids):
        with lock:
            r = rasters.setdefault(job, Raster(n, pops))
        r.add(first, off, ids)

    with ThreadPoolExecutor(4) as ex:
        fut = [ex.submit(cpu, s) for s in seeds]
        with Server(budget_bytes=16e9, chunk_steps=5000, on_spikes=sink) as sv:
            ids = [sv.submit("vogels_abbott", Layout(*lay), DT, steps, delay=1, seed=s) for s in seeds]
            recs = [sv.wait(i, timeout=600) for i in ids]
        assert all(r["status"] == "done" for r in recs), recs
        c = [f.result() for f in fut]
    g = [(rasters[i].rates(DT), rasters[i].isi_cv()) for i in ids]
    c_rate, c_cv = np.mean([x[0] for x in c], 0), np.mean([x[1] for x in c], 0)
    g_rate, g_cv = np.mean([x[0] for x in g], 0), np.mean([x[1] for x in g], 0)
    _say(f"1e8: rates cpu {c_rate} gpu {g_rate}; ISI CV cpu {c_cv} gpu {g_cv}")
    assert np.all(c_rate > 1.0), c_rate
    np.testing.assert_allclose(g_rate, c_rate, rtol=0.02, err_msg="population rates")
    np.testing.assert_allclose(g_cv, c_cv, rtol=0.02, err_msg="population ISI CV")


def test_sweep_512_nets_batched_equals_alone(gpu):
    nets = bench.sweep_nets(512, 7)
    steps = 1000
    groups = {}
    for k, (m, *_rest) in enumerate(nets):
        groups.setdefault(m, []).append(k)
    _say(f"sweep: {len(nets)} nets, batched, fixed point")
    got = {}
    for m, ks in groups.items():
        b = Batch(m, [(Layout(nets[k][1], nets[k][2]), nets[k][3], nets[k][4], nets[k][5]) for k in ks], DT,
                  fixed_point=True)
        b.run(steps)
        b.synchronize()
        for i, k in enumerate(ks):
            net = b.net(i)
            got[k] = (net.stats(), net.neurons(), net.clock())
        b.close()
    _say("sweep: every net alone")
    spikes = 0
    for k, (m, pops, conns, d, seed, par) in enumerate(nets):
        net = Snn(m, Layout(pops, conns), DT, d, seed=seed, fixed_point=True, params=par)
        net.run(steps)
        st, nrn, clk = got[k]
        assert net.stats() == st, (k, m, net.stats(), st)
        assert net.clock() == clk, k
        assert_state_equal(net.neurons(), nrn, f"net {k} ({m})")
        spikes += st["spikes"]
        net.close()
        if k % 64 == 63:
            _say(f"sweep: nets 0..{k} equal alone")
    assert spikes > 1_000_000, spikes
    _say(f"sweep: 512 nets equal, {spikes} spikes")


def test_served_stream_with_a_1e9_job(gpu):
    jobs = [
        ("brunel_3pop_stdp", three_pop_layout(100_000), 2000, 15, 11),
        ("vogels_abbott", ([30_000], [(0, 0, 0.02)]), 1500, 1, 12),
        ("brunel_3pop", three_pop_layout(20_000), 1700, 15, 13),
        ("brunel_3pop_stdp", three_pop_layout(5_000), 900, 15, 14),
        ("vogels_abbott", ([3_000], [(0, 0, 0.02)]), 2500, 1, 15),
        ("brunel_3pop", three_pop_layout(1_000), 600, 15, 16),
    ]
    _say("serve: a 1e9-synapse STDP job among smaller ones, fixed point")
    with Server(budget_bytes=250e9, chunk_steps=500) as sv:
        ids = []
        for k, (m, lay, steps, d, seed) in enumerate(jobs):
            ids.append(sv.submit(m, Layout(*lay), DT, steps, delay=d, seed=seed, fixed_point=True))
            if k == 0:
                time.sleep(0.5)  # the others arrive while the large job runs
        recs = [sv.wait(i, timeout=600) for i in ids]
        info = sv.info()
    assert info["finished"] == len(jobs) and info["reserved"] == 0
    assert recs[0]["synapses"] > 9.9e8
    for k, ((m, lay, steps, d, seed), rec) in enumerate(zip(jobs, recs)):
        assert rec["status"] == "done" and rec["done"] == steps, (k, rec)
        net = Snn(m, Layout(*lay), DT, d, seed=seed, fixed_point=True)
        ref = [np.sort(net.step(spikes=True)) for _ in range(steps)]
        st = net.stats()
        net.close()
        assert rec["digest"] == spike_digest(enumerate(ref)), (k, m)
        assert rec["spikes"] == st["spikes"] and rec["events"] == st["events"], (k, rec, st)
        assert rec["plastic_events"] == st["plastic_events"], k
        _say(f"serve: job {k} ({m}, {steps} steps) equals the job alone")
    _say(f"serve: {len(jobs)} jobs equal alone; the large job's latency {recs[0]['t_done'] - recs[0]['t_submit']:.2f}s")
