"""Serving loop on the GPU (spice_server_* through the C ABI): jobs that arrive while others
run are admitted, folded into the batched chunks and streamed back, and every job's result
is the same as the same network run alone -- spike for spike and counter for counter in
fixed-point mode -- whatever else shares its chunks."""
import time

import numpy as np
import pytest

from spice_amd import Layout, Server, Snn
from spice_amd.serve import spike_digest
from tests.helpers import three_pop_layout

pytestmark = pytest.mark.gpu

DT = 1e-4


def _jobs():
    """(model, layout, steps, delay, seed, params): every model family of the serving mix,
    uneven sizes, durations (not multiples of the chunk) and delays."""
    return [
        ("brunel_3pop", Layout(*three_pop_layout(1500)), 700, 15, 101, [12.0]),
        ("vogels_abbott", Layout([2000], [(0, 0, 0.05)]), 450, 1, 102, None),
        ("brunel_3pop_stdp", Layout(*three_pop_layout(1200)), 800, 15, 103, None),  # past 255-step flushes
        ("brunel_3pop", Layout(*three_pop_layout(800)), 333, 4, 104, [25.0]),
        ("vogels_abbott", Layout([1200], [(0, 0, 0.05)]), 901, 1, 105, None),
        ("brunel_3pop_stdp", Layout(*three_pop_layout(600)), 280, 5, 106, None),
    ]


def _alone(model, lay, steps, delay, seed, params):
    """The same job as a plain network stepped one step at a time: per-step spike sets."""
    net = Snn(model, lay, DT, delay, seed=seed, fixed_point=True, params=params)
    spikes = [np.sort(net.step(spikes=True)) for _ in range(steps)]
    st = net.stats()
    net.close()
    return spikes, st


def test_jobs_in_the_mix_equal_jobs_alone(gpu):
    jobs = _jobs()
    with Server(budget_bytes=8e9, chunk_steps=256, keep_spikes=True) as sv:
        ids = []
        for k, (model, lay, steps, delay, seed, params) in enumerate(jobs):
            ids.append(sv.submit(model, lay, DT, steps, delay=delay, seed=seed, params=params, fixed_point=True))
            if k % 2 == 1:
                time.sleep(0.05)  # later jobs arrive while earlier ones run
        recs = [sv.wait(i, timeout=120) for i in ids]
        streams = [sv.spikes(i) for i in ids]
        info = sv.info()
    assert info["reserved"] == 0 and info["resident"] == 0 and info["finished"] == len(jobs)
    assert info["peak_reserved"] <= 8e9
    for k, ((model, lay, steps, delay, seed, params), rec, stream) in enumerate(zip(jobs, recs, streams)):
        assert rec["status"] == "done", (k, rec)
        assert rec["done"] == steps and len(stream) == steps, (k, rec["done"], len(stream))
        ref, st = _alone(model, lay, steps, delay, seed, params)
        for t in range(steps):
            np.testing.assert_array_equal(np.sort(stream[t]), ref[t], err_msg=f"job {k} ({model}) step {t}")
        assert rec["spikes"] == st["spikes"] and rec["events"] == st["events"], (k, rec, st)
        assert rec["plastic_events"] == st["plastic_events"], (k, rec, st)
        assert rec["digest"] == spike_digest(enumerate(ref)), k
        assert rec["t_submit"] <= rec["t_admit"] <= rec["t_done"]


def test_each_job_alone_in_a_server_has_the_mix_digest(gpu):
    """Digest equality between a job served alone and the same job served in the mix."""
    jobs = _jobs()[:4]
    with Server(budget_bytes=8e9, chunk_steps=100) as sv:
        ids = [sv.submit(m, lay, DT, steps, delay=d, seed=s, params=p, fixed_point=True)
               for m, lay, steps, d, s, p in jobs]
        mixed = [sv.wait(i, timeout=120) for i in ids]
    for k, (m, lay, steps, d, s, p) in enumerate(jobs):
        with Server(budget_bytes=8e9, chunk_steps=64) as sv:
            alone = sv.wait(sv.submit(m, lay, DT, steps, delay=d, seed=s, params=p, fixed_point=True), timeout=120)
        assert alone["status"] == mixed[k]["status"] == "done"
        assert alone["digest"] == mixed[k]["digest"] and alone["spikes"] == mixed[k]["spikes"], k
        assert alone["events"] == mixed[k]["events"], k


def test_spike_jobs_filter_delivers_only_the_listed_jobs(gpu):
    """Server(spike_jobs=...): only the listed jobs' spikes reach on_spikes and keep_spikes
    (an id added before its submit counts); every job's spikes are still counted and digested
    by the library, equal to the job alone."""
    jobs = _jobs()[:3]
    listed = {2}
    seen = set()
    with Server(budget_bytes=8e9, chunk_steps=200, keep_spikes=True, spike_jobs=listed,
                on_spikes=lambda job, first, off, ids: seen.add(job)) as sv:
        ids = [sv.submit(m, lay, DT, steps, delay=d, seed=s, params=p, fixed_point=True)
               for m, lay, steps, d, s, p in jobs]
        recs = [sv.wait(i, timeout=120) for i in ids]
        streams = [sv.spikes(i) for i in ids]
    assert ids == [1, 2, 3] and seen == {2}
    assert len(streams[0]) == 0 and len(streams[2]) == 0 and len(streams[1]) == jobs[1][2]
    for k, ((m, lay, steps, d, s, p), rec) in enumerate(zip(jobs, recs)):
        ref, st = _alone(m, lay, steps, d, s, p)
        assert rec["status"] == "done" and rec["spikes"] == st["spikes"], (k, rec, st)
        assert rec["digest"] == spike_digest(enumerate(ref)), k
        if k == 1:
            for t in range(steps):
                np.testing.assert_array_equal(np.sort(streams[1][t]), ref[t], err_msg=f"step {t}")


def test_fp32_jobs_match_alone_statistically(gpu):
    """fp32 accumulation: sums are ordered differently in every launch shape, so served and
    plain runs agree in their statistics, not spike for spike."""
    lay = Layout(*three_pop_layout(3000))
    steps = 3000
    with Server(budget_bytes=8e9, chunk_steps=500) as sv:
        ids = [sv.submit("brunel_3pop", lay, DT, steps, delay=15, seed=7 + k) for k in range(4)]
        recs = [sv.wait(i, timeout=120) for i in ids]
    for k, rec in enumerate(recs):
        net = Snn("brunel_3pop", lay, DT, 15, seed=7 + k)
        net.run(steps)
        st = net.stats()
        net.close()
        assert rec["status"] == "done"
        assert abs(rec["spikes"] - st["spikes"]) <= 0.02 * st["spikes"] + 4 * np.sqrt(st["spikes"]), (k, rec, st)


def test_admission_budget_fifo_and_cancel(gpu):
    small = Layout(*three_pop_layout(1000))
    big = Layout(*three_pop_layout(4000))
    probe = Server(budget_bytes=1, chunk_steps=100)
    need = probe.estimate("brunel_3pop", small, DT, 100, delay=15)
    need_big = probe.estimate("brunel_3pop", big, DT, 100, delay=15)
    with pytest.raises(ValueError):
        probe.submit("brunel_3pop", small, DT, 100, delay=15)  # can never fit
    probe.close()
    assert need_big > need
    # room for one small job at a time: the second waits for the first (FIFO)
    with Server(budget_bytes=int(need * 1.5), chunk_steps=100) as sv:
        a = sv.submit("brunel_3pop", small, DT, 400, delay=15, seed=1)
        b = sv.submit("brunel_3pop", small, DT, 200, delay=15, seed=2)
        ra, rb = sv.wait(a, timeout=120), sv.wait(b, timeout=120)
        assert ra["status"] == rb["status"] == "done"
        assert rb["t_admit"] >= ra["t_done"] - 1e-6
        assert sv.info()["peak_reserved"] <= int(need * 1.5)
        # cancel: a long running job stops at a chunk boundary, a queued one never starts
        c = sv.submit("brunel_3pop", small, DT, 10_000_000, delay=15, seed=3)
        d = sv.submit("brunel_3pop", small, DT, 100, delay=15, seed=4)
        time.sleep(0.3)
        sv.cancel(d)
        sv.cancel(c)
        rc, rd = sv.wait(c, timeout=120), sv.wait(d, timeout=120)
        assert rc["status"] == "cancelled" and 0 < rc["done"] < 10_000_000 and rc["done"] % 100 == 0
        assert rd["status"] == "cancelled" and rd["done"] == 0
        sv.drain()
        assert sv.info()["reserved"] == 0


def test_invalid_jobs_are_rejected_at_submit(gpu):
    with Server(budget_bytes=1e9, chunk_steps=100) as sv:
        with pytest.raises(ValueError):
            sv.submit("brunel_3pop", Layout(*three_pop_layout(500)), DT, 0)
        with pytest.raises(ValueError):
            sv.submit("brunel_3pop", Layout(*three_pop_layout(500)), DT, 10, delay=0)
        with pytest.raises(ValueError):
            sv.submit("brunel_3pop_stdp", Layout(*three_pop_layout(500)), DT, 10, delay=256)  # plastic delay limit
        with pytest.raises(ValueError):
            sv.submit("brunel_3pop", Layout(*three_pop_layout(500)), DT, 10, flags=2)  # unknown flag bit
