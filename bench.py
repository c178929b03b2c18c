#!/usr/bin/env python3
"""Benchmark of the single-GPU simulation path on MI355X.

Workload (default): Brunel + E->E STDP at 1e9 synapses -- three populations (Poisson 20 %,
excitatory, inhibitory; N = 100,000), every population pair connected at p = 0.1, delay 15,
dt = 0.1 ms, additive pair STDP on excitatory->excitatory synapses (model brunel_3pop_stdp).
Synthetic by construction: the random graph is generated on the device from a fixed seed and
all state starts from the model's init.

One "step" is one simulation time step of the whole network (spike delivery incl. lazy
plasticity, neuron update, plasticity flush). The timed region replays captured HIP graphs
of steps with all state resident in HBM; value = wall seconds / simulated (bio) seconds.

Printed: one JSON line with the metric, synapse events/s, a roofline object for the
dominant kernel (algorithmic bytes per launch / mean kernel duration from HIP events attached
to each dispatch on the net's stream) and a cpu_baseline object (the single-threaded CPU checker, a restatement of
the reference's cpu::snn, timed on this host on a bounded number of steps of the same net).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "wall/bio ratio @1e9 syn; batched syn-events/s; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0


def three_pop(n, p):
    npois = n // 5
    ne = (n - npois) * 4 // 5
    return [npois, ne, n - npois - ne], [(s, d, p) for s in range(3) for d in range(3)]


WORKLOADS = {
    # name: (model, n, p, delay); brunel* models: three populations (Poisson / E / I), every
    # pair at p; vogels: one population at p
    "brunel_stdp_1e9": ("brunel_3pop_stdp", 100_000, 0.1, 15),
    # the CPU-runnable configuration: three-population Brunel at 1e6 synapses (N = 3,162)
    "brunel_1e6": ("brunel_3pop", 3_162, 0.1, 15),
    "brunel_1e9": ("brunel_3pop", 100_000, 0.1, 15),
    "vogels_1e8": ("vogels_abbott", 70_710, 0.02, 1),
    "brunel_1e10": ("brunel_3pop", 316_228, 0.1, 15),
    "vogels_1e9": ("vogels_abbott", 223_607, 0.02, 1),
    # the reference's own benchmark networks (samples/main.cpp:1265-1283, 1389-1402):
    # brunel / brunel+: pops {N/2, N/2}, 0->1 and 1->1 at p = 0.1, N = sqrt(S / 0.05), delay 15;
    # vogels: one population at p = 0.02, N = sqrt(S / 0.02), delay 8
    "ref_brunel_1e9": ("brunel", 141_421, 0.1, 15),
    "ref_vogels_1e9": ("vogels_abbott", 223_606, 0.02, 8),
    "ref_brunel_plus_6e8": ("brunel_with_plasticity", 109_544, 0.1, 15),
    # synth (samples/main.cpp:1285-1291: one population, N = sqrt(S / p)): the published
    # simulation row (p = 0.00156, delay 1) and the setup-time row (p = 0.05, delay 1)
    "ref_synth_1e9": ("synth", 800_640, 0.00156, 1),
    "ref_synth_setup_1e9": ("synth", 141_421, 0.05, 1),
}
# wall/bio the reference publishes for those networks on one GPU (results/spice.json:218-225,
# 122-129, 378-385, 26-33; GPU model not stated)
PUBLISHED = {"ref_brunel_1e9": 0.98772, "ref_vogels_1e9": 1.08669, "ref_brunel_plus_6e8": 21.8024,
             "ref_synth_1e9": 1.39947}
# setup time (graph generation + initial state, device synchronised) the reference publishes
# (results/spice.json:1274-1281, method samples/main.cpp:1337-1348)
PUBLISHED_SETUP = {"ref_synth_setup_1e9": 0.253356}


def workload_layout(name):
    model, n, p, delay = WORKLOADS[name]
    if name.startswith("ref_brunel"):
        return [n // 2, n // 2], [(0, 1, p), (1, 1, p)]
    if "brunel" in model:
        return three_pop(n, p)
    return [n], [(0, 0, p)]


def algorithmic_bytes(model, n, st, steps):
    """Algorithmic HBM bytes per launch of each kernel, from the device counters of the
    profiled steps (per-neuron arrays that stay L2-resident -- spike history, traces' current
    values -- are counted once per neuron in the update kernel, not per synapse):
      receive: per delivered spike its id (4) and row offsets (16); per delivered event the
               target id (4) and, for models with per-synapse state, the weight read (4); per
               plastic synapse changed, the write-back of its state (4 for neuron-trace
               synapses, 12 + 8 re-read for per-synapse-trace synapses);
      update : neuron state read + write, accumulator read + clear, plus for plastic models
               history word read + write (16) and row bookkeeping (9 lazy: synced, pending
               event, has-plastic flag; 1 eager: the flag), for neuron-trace synapses (eager)
               the traces read + written (16) and the step's flags (1) per neuron;
               4 per emitted spike;
      flush  : the plasticity work (counted with the kernel that runs it): 4 per
               index word read (lazy flush: the target id
               of every scanned slot of the flushed rows; eager: the target id of every plastic
               slot of the arriving rows and the (source, offset) pair of every by-target entry
               of the spiking targets) + per synapse brought up to date its state read (4
               neuron-trace, 12 per-synapse-trace) + per synapse written back its state write
               (eager passes write only synapses whose state the event changed; the lazy flush
               writes every one);
      step   : the overlapped step kernel (update of step T, delivery of T + 1, plasticity work
               of T - 1): one of each per launch.
    Target-neuron accumulation is counted as atomics, not bytes."""
    from spice_amd.models import MODELS
    _, nattrs, sattrs = MODELS[model]
    s_n = sum(np.dtype(t).itemsize for _, t in nattrs)
    s_s = sum(np.dtype(t).itemsize for _, t in sattrs)
    plastic = bool(sattrs)
    traced = model.startswith("brunel_3pop_stdp")
    acc_bytes = 4 * (1 if "brunel" in model else 2)  # receive-target attributes per neuron
    spk = st["spikes"] / steps
    ev = st["events"] / steps
    pev = st["plastic_events"] / steps
    wb = 4 if traced else (2 * s_s - 4)
    recv = spk * 20 + ev * (8 if plastic else 4) + pev * (wb if plastic else 0)
    upd = n * (2 * s_n + 2 * acc_bytes + ((17 + 17) if traced else (25 if plastic else 0))) + spk * 4
    fl = (st["flush_slots"] * 4 + (st["flush_synapses"] + st["flush_written"]) * s_s) / steps if plastic else 0.0
    # the plasticity work of step T - 1 runs inside the delivery kernel of step T (two-kernel
    # step) or inside the overlapped kernel of step T
    return {"receive": recv + fl, "update": upd, "step": recv + fl + upd}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs, scaled by the copy calibration), or (None, None)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)[workload]["kernels"][kernel]
        return int(t["fetch_bytes"] + t["write_bytes"]), "profiles/pmc_traffic.json"
    except (OSError, KeyError, ValueError):
        return None, None


def measured_atomics(workload, kernel):
    """Atomic operations per launch of `kernel` on `workload` from the committed rocprofv3 PMC
    summary: LDS atomic lane operations (SQ_INSTS_LDS_ATOMIC_BANDWIDTH, counted in 64-byte units
    of atomic data, x 16 for 4-byte operands) and L2 atomic requests (TCC_ATOMIC_sum), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)[workload]["kernels"][kernel]
        return {"lds_atomic_ops": int(t["lds_atomic_ops"]), "l2_atomic_requests": int(t["l2_atomic_requests"])}
    except (OSError, KeyError, ValueError):
        return None


def sweep_nets(count, seed):
    """Parameter sweep of the batched configuration: `count` independent nets, alternately
    Brunel (brunel_3pop, p = 0.1, delay 15, Poisson rate ~ U[5, 30] Hz) and Vogels-Abbott
    (p = 0.02, delay 1), synapse counts log-uniform in [1e4, 1e8], seed = base + index."""
    rng = np.random.default_rng(seed)
    nets = []
    for k in range(count):
        syn = 10 ** rng.uniform(4, 8)
        if k % 2 == 0:
            n = max(20, int(round(np.sqrt(syn / 0.1))))
            pops, conns = three_pop(n, 0.1)
            nets.append(("brunel_3pop", pops, conns, 15, 1000 + k, [float(rng.uniform(5, 30))]))
        else:
            n = max(20, int(round(np.sqrt(syn / 0.02))))
            nets.append(("vogels_abbott", [n], [(0, 0, 0.02)], 1, 1000 + k, None))
    return nets


def stat_pops(model, pops):
    """Populations for rate / ISI CV statistics: the layout's, except Vogels-Abbott's single
    layout population, which the model splits into excitatory (ids below 0.8 N) and inhibitory."""
    if model == "vogels_abbott" and len(pops) == 1:
        nexc = int(np.float32(0.8) * np.float32(pops[0]))
        return [nexc, pops[0] - nexc]
    return list(pops)


def _served_rasters(jobs, fixed, together):
    """Rasters (and job records) of `jobs` = [(model, pops, conns, delay, seed, params, steps)]
    streamed by the library's server: all jobs in one server (stepped together in batched
    chunks) or each job in a server of its own."""
    from spice_amd import Layout, Server
    from spice_amd.raster import Raster
    rasters, recs = {}, []
    groups = [list(range(len(jobs)))] if together else [[k] for k in range(len(jobs))]
    for g in groups:
        ids = {}

        def sink(job, first, off, ids_, _ids=ids):
            rasters[_ids[job]].add(first, off, ids_)

        with Server(budget_bytes=250e9, chunk_steps=1000, on_spikes=sink) as sv:
            for k in g:
                m, pops, conns, d, seed, par, steps = jobs[k]
                rasters[k] = Raster(sum(pops), stat_pops(m, pops))
                ids[sv.submit(m, Layout(pops, conns), 1e-4, steps, delay=d, seed=seed, params=par,
                              fixed_point=fixed)] = k
            for i in list(ids):
                recs.append((ids[i], sv.wait(i)))
    return rasters, dict(recs)


def compare_rasters(a, b, dt, min_spikes):
    """Per-net, per-population rate and mean ISI CV of rasters b against a (relative
    differences): the worst net overall and among nets with >= min_spikes spikes, and the
    share of compared nets whose every population is within 2 %."""
    rows = []
    for k in a:
        ra, rb = np.array(a[k].rates(dt)), np.array(b[k].rates(dt))
        ca, cb = np.array(a[k].isi_cv()), np.array(b[k].isi_cv())
        dr = np.abs(rb - ra) / np.maximum(ra, 1e-9)
        ok = np.isfinite(ca) & np.isfinite(cb)
        dc = np.abs(cb[ok] - ca[ok]) / np.maximum(ca[ok], 1e-9)
        rows.append((k, int(a[k].count.sum()), float(dr.max(initial=0.0)), float(dc.max(initial=0.0))))
    big = [r for r in rows if r[1] >= min_spikes]
    worst = max(rows, key=lambda r: max(r[2], r[3]), default=(None, 0, 0.0, 0.0))
    worst_big = max(big, key=lambda r: max(r[2], r[3]), default=(None, 0, 0.0, 0.0))
    return {"nets": len(rows), "max_rate_rel_diff": round(worst[2], 5), "max_isi_cv_rel_diff": round(worst[3], 5),
            "worst_net": worst[0], "worst_net_spikes": worst[1],
            f"nets_with_{min_spikes}_spikes": len(big),
            f"max_rate_rel_diff_{min_spikes}": round(worst_big[2], 5),
            f"max_isi_cv_rel_diff_{min_spikes}": round(worst_big[3], 5),
            "within_2pct_share": round(sum(1 for r in rows if r[2] <= 0.02 and r[3] <= 0.02) / max(1, len(rows)), 4),
            f"within_2pct_share_{min_spikes}": round(sum(1 for r in big if r[2] <= 0.02 and r[3] <= 0.02)
                                                     / max(1, len(big)), 4)}


# Sequential estimate of a net or job run alone: SEQ_WARM steps untimed, SEQ_WINDOW steps timed,
# the full run's time extrapolated linearly from the window (shorter runs run whole).
SEQ_WARM, SEQ_WINDOW = 1000, 500


def sequential_time(model, pops, conns, delay, seed, params, steps):
    """(set-up s, run s, extrapolated?) of one plain network run alone for `steps` steps."""
    from spice_amd import Layout, Snn
    t0 = time.perf_counter()
    net = Snn(model, Layout(pops, conns), 1e-4, delay, seed=seed, params=params)
    net.synchronize()
    t1 = time.perf_counter()
    if steps <= SEQ_WARM + SEQ_WINDOW:
        net.run(steps)
        net.synchronize()
        run, ext = time.perf_counter() - t1, False
    else:
        net.run(SEQ_WARM)
        net.synchronize()
        t2 = time.perf_counter()
        net.run(SEQ_WINDOW)
        net.synchronize()
        run, ext = (time.perf_counter() - t2) / SEQ_WINDOW * steps, True
    net.close()
    return t1 - t0, run, ext


def run_batched(count, steps, seq_stride):
    """Aggregate synapse events/s of the sweep stepped as batches (one batch per model, the
    two batches on their own streams) vs. the same nets stepped one at a time on the same GPU
    (every net alone, its run time extrapolated from a timed window: sequential_time; the
    events are the batched run's). Statistics of every seq_stride-th net: per net and population, firing rate and mean ISI CV with the nets stepped together
    in batched chunks (spikes streamed by the server) against each net stepped alone -- in fp32,
    where summation order differs and trajectories diverge, so only statistics agree, and in
    fixed point, where the spike streams must be identical."""
    from spice_amd import Batch, Layout
    nets = sweep_nets(count, 7)
    dt = 1e-4
    groups = {}
    for k, (m, pops, conns, d, seed, par) in enumerate(nets):
        groups.setdefault(m, []).append(k)
    t0 = time.perf_counter()
    batches = {m: Batch(m, [(Layout(nets[k][1], nets[k][2]), nets[k][3], nets[k][4], nets[k][5]) for k in ks], dt)
               for m, ks in groups.items()}
    for b in batches.values():
        b.synchronize()
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    for b in batches.values():
        b.run(steps)
    for b in batches.values():
        b.synchronize()
    wall = time.perf_counter() - t0
    events, spikes, syn = 0, {}, 0
    for m, ks in groups.items():
        for i, k in enumerate(ks):
            st = batches[m].net(i).stats()
            events += st["events"]
            spikes[k] = st["spikes"]
            syn += batches[m].net(i).num_edges()
    for b in batches.values():
        b.close()
    # every net alone (plain runs, one at a time)
    seq_wall, seq_setup = 0.0, 0.0
    for m, pops, conns, d, seed, par in nets:
        su, run, _ = sequential_time(m, pops, conns, d, seed, par, steps)
        seq_setup += su
        seq_wall += run
    sample = list(range(0, count, seq_stride))
    agg = events / wall
    seq = events / seq_wall
    print(f"[bench] batched {count} nets ({syn:.3g} synapses) x {steps} steps: {wall:.3f}s, "
          f"{agg:.3g} ev/s; sequential (all nets alone, extrapolated) {seq:.3g} ev/s", file=sys.stderr, flush=True)
    jobs = [nets[k] + (steps,) for k in sample]
    t1 = time.perf_counter()
    together, _ = _served_rasters(jobs, False, True)
    alone, _ = _served_rasters(jobs, False, False)
    stats_fp32 = compare_rasters(alone, together, dt, 10_000)
    _, fx_together = _served_rasters(jobs, True, True)
    fx_alone_r, fx_alone = _served_rasters(jobs, True, False)
    # noise floor: the same nets alone, fp32 against fixed-point accumulation -- only the
    # rounding of the sums differs, so this is how far chaos alone moves a net's statistics
    stats_noise = compare_rasters(fx_alone_r, alone, dt, 10_000)
    fx_same = all(fx_together[k]["digest"] == fx_alone[k]["digest"] and fx_together[k]["spikes"] == fx_alone[k]["spikes"]
                  for k in range(len(jobs)))
    print(f"[bench] sweep statistics ({len(jobs)} nets, {time.perf_counter() - t1:.1f}s): {stats_fp32}; "
          f"noise floor {stats_noise}; fixed point identical: {fx_same}", file=sys.stderr, flush=True)
    return {"nets": count, "synapses": syn, "steps": steps, "bio_s": steps * dt, "setup_s": round(setup, 3),
            "wall_s": round(wall, 4), "syn_events_per_s": round(agg, 1),
            "sequential_syn_events_per_s": round(seq, 1), "speedup_vs_sequential": round(agg / seq, 3),
            "sequential_method": f"all {count} nets alone (plain Snn, same seeds and parameters), run time "
                                 f"extrapolated from {SEQ_WINDOW} timed steps after {SEQ_WARM} untimed ones; "
                                 f"events = the batched run's",
            "sequential_run_s": round(seq_wall, 3), "sequential_setup_s": round(seq_setup, 3),
            "spikes_total": int(sum(spikes.values())),
            "statistics_sample": f"the same {len(sample)} nets, {steps} steps: stepped together in batched chunks "
                                 f"(server) vs each alone; per net and population",
            "statistics_fp32": stats_fp32,
            "statistics_noise_floor": stats_noise,
            "statistics_noise_floor_method": "the same nets alone, fp32 against fixed-point accumulation (rounding "
                                             "of the sums is all that differs)",
            "fixed_point_identical": fx_same}


def serve_stream(count, rate, seed):
    """Served job stream: Poisson arrivals at `rate` jobs/s; 40 % Brunel (brunel_3pop, p = 0.1,
    delay 15), 40 % Vogels-Abbott (p = 0.02, delay 1), 20 % Brunel+STDP (brunel_3pop_stdp,
    p = 0.1, delay 15); synapse counts log-uniform in [1e5, 2e9], biological duration uniform
    in [0.1, 10] s at dt = 0.1 ms, random seeds."""
    rng = np.random.default_rng(seed)
    jobs, t = [], 0.0
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        u = rng.random()
        syn = float(np.exp(rng.uniform(np.log(1e5), np.log(2e9))))
        steps = int(round(rng.uniform(0.1, 10.0) / 1e-4))
        job_seed = int(rng.integers(1, 2**31))
        if u < 0.8 and u >= 0.4:
            n = int(round(np.sqrt(syn / 0.02)))
            jobs.append((t, "vogels_abbott", [n], [(0, 0, 0.02)], 1, steps, job_seed))
        else:
            n = int(round(np.sqrt(syn / 0.1)))
            pops, conns = three_pop(n, 0.1)
            jobs.append((t, "brunel_3pop" if u < 0.4 else "brunel_3pop_stdp", pops, conns, 15, steps, job_seed))
    return jobs


def run_serving(count, rate, budget, chunk, seq_stride, seed):
    """Continuous batching of a Poisson job stream through the library's server (admission
    against the HBM budget, device-resident jobs folded into batched chunks, spikes streamed
    back per chunk) vs the same jobs run one at a time as plain networks: every job alone, its
    set-up timed and its run time extrapolated from a timed window (seq_window). Statistics:
    every seq_stride-th job served in the mix vs the same job served alone. Latency = submit ->
    last spikes delivered."""
    from spice_amd import Layout, Server
    from spice_amd.raster import Raster
    jobs = serve_stream(count, rate, seed)
    sample = list(range(0, count, seq_stride))
    chunks = {}  # job id -> the sampled job's spike chunks, turned into a raster after the run
    sampled_ids = set()

    def sink(job, first, off, ids):
        # runs on the server's worker thread between chunks: only queue the chunk
        chunks[job].append((first, off, ids))

    sv = Server(budget_bytes=budget, chunk_steps=chunk, on_spikes=sink, spike_jobs=sampled_ids)
    t0 = time.perf_counter()
    ids = []
    for k, (t, m, pops, conns, d, steps, js) in enumerate(jobs):
        lag = t - (time.perf_counter() - t0)
        if lag > 0:
            time.sleep(lag)
        lay = Layout(pops, conns)
        if k in sample:
            # registered before submit (ids count up from 1): the worker may deliver the first
            # chunk at once
            chunks[len(ids) + 1] = []
            sampled_ids.add(len(ids) + 1)
        ids.append(sv.submit(m, lay, 1e-4, steps, delay=d, seed=js))
        assert k not in sample or ids[-1] in chunks
    recs = [sv.wait(i) for i in ids]
    makespan = time.perf_counter() - t0
    info = sv.info()
    sv.close()
    failed = [r for r in recs if r["status"] != "done"]
    lat = np.array([r["t_done"] - r["t_submit"] for r in recs])
    qdelay = np.array([r["t_admit"] - r["t_submit"] for r in recs])  # waiting for admission (budget)
    events = sum(r["events"] for r in recs)
    syn = sum(r["synapses"] for r in recs)
    streamed = sum(r["spikes"] for r in recs)  # every job's spikes reached host memory (counted there)
    rasters = {}
    for k in sample:
        _, m, pops, conns, d, steps, js = jobs[k]
        r = rasters[ids[k]] = Raster(sum(pops), pops)
        for first, off, spk in chunks[ids[k]]:
            r.add(first, off, spk)
    print(f"[bench] served {count} jobs ({syn:.3g} synapses, {events:.3g} events) in {makespan:.1f}s: "
          f"p50 {np.percentile(lat, 50):.2f}s p99 {np.percentile(lat, 99):.2f}s, {len(failed)} not done",
          file=sys.stderr, flush=True)
    # sequential plain runs of every job (a sample of a log-uniform size mix is dominated by
    # whether it holds the few largest jobs): each job alone (sequential_time); the events are
    # the served run's
    seq_run, seq_setup, extrapolated = 0.0, 0.0, 0
    t1 = time.perf_counter()
    for _, m, pops, conns, d, steps, js in jobs:
        su, run, ext = sequential_time(m, pops, conns, d, js, None, steps)
        seq_setup += su
        seq_run += run
        extrapolated += int(ext)
    print(f"[bench] sequential estimate of {count} jobs alone: {seq_run:.1f}s run + {seq_setup:.1f}s set-up "
          f"({time.perf_counter() - t1:.1f}s to measure)", file=sys.stderr, flush=True)
    seq = events / seq_run
    seq_incl = events / (seq_run + seq_setup)
    served = events / makespan
    # statistics of the sampled jobs served in the mix vs the same jobs served alone
    stat = []
    for k in sample:
        _, m, pops, conns, d, steps, js = jobs[k]
        alone = Raster(sum(pops), pops)
        with Server(budget_bytes=budget, chunk_steps=chunk,
                    on_spikes=lambda job, first, off, ids_: alone.add(first, off, ids_)) as one:
            one.wait(one.submit(m, Layout(pops, conns), 1e-4, steps, delay=d, seed=js))
        mixed = rasters[ids[k]]
        ra, rm = float(alone.count.sum()), float(mixed.count.sum())
        cva, cvm = float(np.nanmean(alone.isi_cv())), float(np.nanmean(mixed.isi_cv()))
        stat.append((ra, rm, cva, cvm))
    return {
        "jobs": count, "arrival_rate_per_s": rate, "budget_bytes": budget, "chunk_steps": chunk,
        "mix": "40% brunel_3pop / 40% vogels_abbott / 20% brunel_3pop_stdp, synapses log-uniform 1e5-2e9, "
               "bio U[0.1, 10] s",
        "synapses_total": syn, "events_total": events, "spikes_streamed": streamed,
        "makespan_s": round(makespan, 3), "busy_s": round(info["busy_s"], 3),
        "worker_s": {k: round(info[k], 3) for k in ("launch_s", "admit_s", "wait_s", "collect_s")},
        "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
        "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
        "latency_max_s": round(float(lat.max()), 4),
        "admission_delay_p50_s": round(float(np.percentile(qdelay, 50)), 4),
        "admission_delay_p99_s": round(float(np.percentile(qdelay, 99)), 4),
        "admission_delay_max_s": round(float(qdelay.max()), 4),
        "jobs_waiting_over_1s_for_admission": int(np.count_nonzero(qdelay > 1.0)),
        "syn_events_per_s": round(served, 1),
        "sequential_syn_events_per_s": round(seq, 1),
        "sequential_syn_events_per_s_incl_setup": round(seq_incl, 1),
        "speedup_vs_sequential": round(served / seq, 3),
        "sequential_method": f"all {count} jobs alone (plain Snn, same seeds): set-up timed; run time of "
                             f"{extrapolated} jobs extrapolated from {SEQ_WINDOW} timed steps after {SEQ_WARM} untimed "
                             f"ones, the others run whole; events = the served run's",
        "sequential_run_s": round(seq_run, 3), "sequential_setup_s": round(seq_setup, 3),
        "speedup_vs_sequential_incl_setup": round(events / makespan / seq_incl, 3),
        "peak_reserved_bytes": info["peak_reserved"], "chunks": info["chunks"],
        "not_done": len(failed),
        "statistics_sample": f"{len(stat)} jobs served in the mix vs the same jobs served alone (fp32)",
        "rate_rel_diff_pooled": round(abs(sum(x[1] for x in stat) - sum(x[0] for x in stat)) / max(1.0, sum(x[0] for x in stat)), 5),
        "isi_cv_rel_diff_pooled": round(abs(np.nanmean([x[3] for x in stat]) - np.nanmean([x[2] for x in stat]))
                                        / max(1e-9, abs(np.nanmean([x[2] for x in stat]))), 5),
        "rate_rel_diff_max": round(max(abs(x[1] - x[0]) / max(1.0, x[0]) for x in stat), 5),
    }


def host_cpu():
    """Model name of the host CPU (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(model, lay, delay, seed, net, steps):
    """Times the CPU checker (single thread, the process pinned to one core while it runs) on
    `steps` steps of the same network: the graph is copied from the GPU, state starts from the
    model's init (as the GPU run did). Returns (s per step, init s, core)."""
    from oracle import oracle as O
    rp, col = net.csr()
    t0 = time.perf_counter()
    h = O.CpuSnn.from_csr(model, rp, col, net.width(), 1e-4, delay, seed=seed)
    t_init = time.perf_counter() - t0
    del rp, col
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        for _ in range(steps):
            h.step()
        dt_wall = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    h.close()
    return dt_wall / steps, t_init, core


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100_000,
                    help="timed steps (100,000 = 10 s bio: the configuration's and the reference's benchmark length)")
    ap.add_argument("--warmup", type=int, default=2_000)
    ap.add_argument("--workload", default="brunel_stdp_1e9", choices=sorted(WORKLOADS))
    ap.add_argument("--profile-steps", type=int, default=1_000)
    ap.add_argument("--cpu-steps", type=int, default=20, help="steps of the CPU baseline (0 = skip); ~0.55 s each at 1e9 synapses")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--flags", type=int, default=0, help="spice_flag bits (1: int64 fixed-point accumulation; see include/spice_hip.h)")
    ap.add_argument("--batch-nets", type=int, default=512, help="nets of the batched sweep (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=10_000, help="steps of the sweep (10,000 = 1 s bio)")
    ap.add_argument("--batch-seq-stride", type=int, default=8)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--launch", default=None,
                    help="tile_w,flush_slots,flush_grid,recv_block,recv_parts: fixed launch shapes instead of autotuning")
    ap.add_argument("--serve-jobs", type=int, default=200, help="jobs of the served stream (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=50.0,
                    help="Poisson arrival rate, jobs per second (50: the stream's reservations reach the budget)")
    ap.add_argument("--serve-budget", type=float, default=250e9, help="HBM admission budget, bytes")
    ap.add_argument("--serve-chunk", type=int, default=1000, help="steps per serving chunk")
    ap.add_argument("--serve-seq-stride", type=int, default=20, help="every n-th served job: statistics vs the job served alone")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the simulator runs one network per GPU on one GPU only (asked for --gpus {args.gpus}); "
              "sharding a network over GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    from spice_amd import Layout, Snn

    model, n, p, delay = WORKLOADS[args.workload]
    pops, conns = workload_layout(args.workload)
    dt = 1e-4

    t0 = time.perf_counter()
    net = Snn(model, Layout(pops, conns), dt, delay, seed=args.seed, flags=args.flags)
    net.synchronize()
    setup_s = time.perf_counter() - t0
    E = net.num_edges()
    print(f"[bench] {args.workload}: N={n} E={E} setup {setup_s:.3f}s", file=sys.stderr, flush=True)
    tune = None
    if args.launch:
        net.set_launch(*(int(x) for x in args.launch.split(",")))
    elif not args.no_autotune:
        t0 = time.perf_counter()
        cfg, us = net.autotune(warmup=500, steps=300)
        tune = {"launch": cfg, "predicted_us_per_step": round(us, 3), "tune_s": round(time.perf_counter() - t0, 3)}
        print(f"[bench] autotuned {cfg} ({us:.1f} us/step) in {tune['tune_s']}s", file=sys.stderr, flush=True)

    net.run(args.warmup)
    net.synchronize()

    # Roofline: per-kernel mean duration from HIP events attached to each dispatch on the net's
    # stream, over the profile steps right after the warmup (so that a profiler run of the
    # same command with fewer timed steps measures the same steps of the network)
    net.reset_stats()
    ms = net.run_profiled(args.profile_steps)
    pst = net.stats()
    net.synchronize()
    net.reset_stats()

    t0 = time.perf_counter()
    net.run(args.steps)
    net.synchronize()
    wall = time.perf_counter() - t0
    st = net.stats()
    bio = args.steps * dt
    print(f"[bench] timed {args.steps} steps in {wall:.3f}s, stats {st}", file=sys.stderr, flush=True)
    bytes_per = algorithmic_bytes(model, n, pst, args.profile_steps)
    kern = max(ms, key=lambda k: ms[k])
    # launches of each kernel in the profiled steps: two per step, or (overlapped steps) the
    # first delivery, one overlapped kernel per step boundary and the last update, or (fused
    # steps: no delivery kernel of its own) one step kernel per step
    fused = ms["step"] > 0 and ms["receive"] == 0
    overlapped = ms["step"] > 0 and not fused
    launches = ({"receive": 1, "update": 1, "step": args.profile_steps - 1} if overlapped
                else {"receive": args.profile_steps, "update": args.profile_steps,
                      "step": args.profile_steps if fused else 1})
    avg_ms = ms[kern] / launches[kern]
    achieved = bytes_per[kern] / (avg_ms * 1e-3) / 1e9
    kernel_names = {"receive": "k_receive_tiled", "update": "k_step_overlap" if overlapped else "k_update",
                    "step": "k_step_fused" if fused else "k_step_overlap"}
    traffic, traffic_src = measured_traffic(args.workload, kernel_names[kern])
    roofline = {
        "bound": "hbm", "kernel": kernel_names[kern], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
        "bytes_per_launch": round(bytes_per[kern]), "avg_launch_us": round(avg_ms * 1e3, 3),
        "kernel_us_per_step": {k: round(v / args.profile_steps * 1e3, 3) for k, v in ms.items()},
        # delivery accumulates every delivered event with one LDS atomic add per receive target
        # (the kernels that deliver: the delivery kernels, or the overlapped step kernels)
        "receive_lds_atomics_per_s": round(pst["events"] / ((ms["receive"] + ms["step"]) * 1e-3), 1),
    }
    # the same throughput from the PMC counters of a rocprofv3 run of this workload: atomic
    # operations counted per launch / this run's mean launch duration
    counted = measured_atomics(args.workload, kernel_names[kern])
    if counted is not None:
        roofline["atomics_counted"] = {
            "lds_atomic_ops_per_launch": counted["lds_atomic_ops"],
            "lds_atomic_ops_per_s": round(counted["lds_atomic_ops"] / (avg_ms * 1e-3), 1),
            "l2_atomic_requests_per_launch": counted["l2_atomic_requests"],
            "l2_atomic_requests_per_s": round(counted["l2_atomic_requests"] / (avg_ms * 1e-3), 1),
            "source": "profiles/pmc_traffic.json"}

    # The passes' traffic is random 64-byte lines, whose rate -- not the streaming peak -- bounds
    # the plastic step: the same kernel's line transfers per second next to the device's rate for
    # random 4-byte loads (one line in) and read-modify-writes (one line in, one out) in an array
    # the size of the plastic weights (~1.6 GB at 1e9 synapses)
    if model.startswith("brunel_3pop_stdp"):
        from spice_amd.snn import probe_random_lines
        acc = 1 << 21
        nbytes = 1_600_000_000
        t_rd = probe_random_lines(nbytes, acc, False)
        t_rmw = probe_random_lines(nbytes, acc, True)
        rl = {"probe_bytes": nbytes, "probe_accesses": acc,
              "probe_read_lines_per_s": round(acc / (t_rd * 1e-6), 1),
              "probe_rmw_line_transfers_per_s": round(2 * acc / (t_rmw * 1e-6), 1)}
        if traffic:
            rl["kernel_line_transfers_per_s"] = round(traffic / 64 / (avg_ms * 1e-3), 1)
            rl["kernel_frac_of_rmw_probe"] = round(rl["kernel_line_transfers_per_s"] / rl["probe_rmw_line_transfers_per_s"], 4)
        roofline["random_lines"] = rl

    cpu = None
    if args.cpu_steps > 0:
        s_per_step, t_init, core = cpu_baseline(model, (pops, conns), delay, args.seed, net, args.cpu_steps)
        cpu = {"value": round(s_per_step / dt, 3), "unit": "wall s / bio s", "cores": 1, "kind": "port",
               "cpu": host_cpu(),
               "sample": f"{args.cpu_steps} steps of the same {E:.3g}-synapse network (graph copied from the GPU, "
                         f"state from init; dense per-step plasticity as in cpu::snn), one thread pinned to core "
                         f"{core}, {s_per_step:.2f} s/step, wall/bio extrapolated linearly; init {t_init:.1f}s excluded"}

    net_launch = net.launch()
    net.close()
    batched = run_batched(args.batch_nets, args.batch_steps, args.batch_seq_stride) if args.batch_nets > 0 else None
    served = (run_serving(args.serve_jobs, args.serve_rate, args.serve_budget, args.serve_chunk,
                          args.serve_seq_stride, 11) if args.serve_jobs > 0 else None)

    out = {
        "metric": METRIC,
        "value": round(wall / bio, 5),
        "unit": "wall s / bio s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(wall / bio / PUBLISHED[args.workload], 5) if args.workload in PUBLISHED else None,
        "dtype": "f32",
        "data": "synthetic (device-generated random graph, fixed seed)",
        "config": {"workload": args.workload, "model": model, "neurons": n, "synapses": E,
                   "populations": pops, "p": p, "delay": delay, "dt": dt, "bio_s": bio},
        "syn_events_per_s": round(st["events"] / wall, 1),
        "spikes_per_step": round(st["spikes"] / args.steps, 2),
        "events_per_step": round(st["events"] / args.steps, 1),
        "setup_s": round(setup_s, 4),
        "setup_vs_published": (round(setup_s / PUBLISHED_SETUP[args.workload], 4)
                               if args.workload in PUBLISHED_SETUP else None),
        "launch": net_launch,
        "autotune": tune,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
