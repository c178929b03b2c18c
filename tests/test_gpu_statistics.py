"""Statistics over 10 s of biological time: per-population firing rates and ISI coefficients of
variation of the GPU (fp32 accumulation) within 2 % of the CPU checker (plain Brunel, Vogels-
Abbott, and Brunel with E->E STDP, whose dense checker takes ~2.5 minutes for its 8 seeds).

fp32 sums are ordered differently on the two sides, so the trajectories diverge after a few
hundred steps and only statistics can agree. For the Brunel network at 1e6 synapses (3,162
neurons, a bursting regime, ISI CV ~ 2.6) a single 10 s run's rate moves by ~2 % under a mere
change of rounding (checker fp32 vs checker fixed point: 2.1 % at seed 1), so both sides pool
16 seeds (spread then ~0.6 %); the Vogels-Abbott network pools 4. The Poisson population's
spikes are identical on both sides (same counter-based random stream)."""
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from spice_amd import Layout, Server
from spice_amd.raster import Raster
from tests.helpers import three_pop_layout

pytestmark = pytest.mark.gpu

DT = 1e-4
STEPS = 100_000  # 10 s
TOL = 0.02


def _cpu_stats(model, lay, delay, seed, steps):
    h = O.CpuSnn(model, lay, DT, delay, seed=seed)
    off, ids = h.run_record(steps)
    h.close()
    r = Raster(sum(lay[0]), lay[0])
    r.add(0, off, ids)
    return r.rates(DT), r.isi_cv()


def _gpu_stats(model, lay, delay, seeds, steps):
    pops = lay[0]
    rasters, lock = {}, threading.Lock()

    def sink(job, first, off, ids):
        with lock:
            r = rasters.setdefault(job, Raster(sum(pops), pops))
        r.add(first, off, ids)

    with Server(budget_bytes=16e9, chunk_steps=2000, on_spikes=sink) as sv:
        ids = [sv.submit(model, Layout(*lay), DT, steps, delay=delay, seed=s) for s in seeds]
        recs = [sv.wait(i, timeout=600) for i in ids]
    assert all(r["status"] == "done" for r in recs), recs
    return [(rasters[i].rates(DT), rasters[i].isi_cv()) for i in ids]


@pytest.mark.parametrize("model,lay,delay,seeds,steps", [
    ("brunel_3pop", three_pop_layout(3162), 15, range(1, 17), STEPS),
    ("vogels_abbott", ([10000], [(0, 0, 0.02)]), 1, range(1, 5), STEPS),
    # fp32 STDP (eager on the GPU, dense per step in the checker): weights move throughout
    pytest.param("brunel_3pop_stdp", three_pop_layout(3162), 15, range(1, 9), STEPS, marks=pytest.mark.timeout(900)),
])
def test_rates_and_isi_cv_match_checker(gpu, model, lay, delay, seeds, steps):
    with ThreadPoolExecutor(8) as ex:  # the checker releases the GIL while it runs
        futs = [ex.submit(_cpu_stats, model, lay, delay, s, steps) for s in seeds]
        cpu = []
        for k, f in enumerate(futs):
            cpu.append(f.result())
            print(f"{model}: checker seed {k + 1} of {len(futs)} done", flush=True)
    gpu_ = _gpu_stats(model, lay, delay, seeds, steps)
    c_rate, c_cv = np.mean([c[0] for c in cpu], 0), np.mean([c[1] for c in cpu], 0)
    g_rate, g_cv = np.mean([g[0] for g in gpu_], 0), np.mean([g[1] for g in gpu_], 0)
    print(f"{model}: rates cpu {c_rate} gpu {g_rate}; ISI CV cpu {c_cv} gpu {g_cv}")
    assert np.all(c_rate > 1.0), c_rate  # the network is active
    np.testing.assert_allclose(g_rate, c_rate, rtol=TOL, err_msg="population rates")
    np.testing.assert_allclose(g_cv, c_cv, rtol=TOL, err_msg="population ISI CV")
