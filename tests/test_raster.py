"""Spike raster: reference text format round trip and incremental statistics equal to a
brute-force computation over the whole raster (any chunking)."""
import numpy as np
import pytest

from spice_amd.raster import Raster, read_raster


def _random_raster(n, steps, p, seed):
    rng = np.random.default_rng(seed)
    return [rng.permutation(np.flatnonzero(rng.random(n) < p)).astype(np.int32) for _ in range(steps)]


def _brute(n, pops, steps_ids, dt):
    times = [[] for _ in range(n)]
    for t, ids in enumerate(steps_ids):
        for i in ids:
            times[int(i)].append(t)
    rates, cvs = [], []
    b = np.cumsum([0] + pops)
    for k in range(len(pops)):
        cnt = sum(len(times[i]) for i in range(b[k], b[k + 1]))
        rates.append(cnt / (pops[k] * len(steps_ids) * dt))
        cv = []
        for i in range(b[k], b[k + 1]):
            if len(times[i]) >= 3:
                isi = np.diff(times[i]).astype(float)
                cv.append(isi.std() / isi.mean())
        cvs.append(np.mean(cv) if cv else float("nan"))
    return rates, cvs


@pytest.mark.parametrize("chunk", [1, 7, 64, 1000])
def test_incremental_statistics_match_brute_force(chunk):
    n, pops = 300, [60, 200, 40]
    steps_ids = _random_raster(n, 500, 0.03, 5)
    r = Raster(n, pops)
    for c0 in range(0, len(steps_ids), chunk):
        part = steps_ids[c0:c0 + chunk]
        off = np.concatenate(([0], np.cumsum([len(x) for x in part]))).astype(np.uint64)
        r.add(c0, off, np.concatenate(part) if part else np.zeros(0, np.int32))
    rates, cvs = _brute(n, pops, steps_ids, 1e-4)
    np.testing.assert_allclose(r.rates(1e-4), rates, rtol=1e-12)
    np.testing.assert_allclose(r.isi_cv(), cvs, rtol=1e-9)


def test_text_format_round_trip_and_layout():
    r = Raster(10, keep_lines=True)
    r.add_step(0, [7, 2, 5])
    r.add_step(1, [])
    r.add_step(2, [9])
    txt = r.text()
    assert txt == "10\n2,5,7\n\n9\n"
    n, steps = read_raster(txt)
    assert n == 10 and [list(s) for s in steps] == [[2, 5, 7], [], [9]]


def test_regular_spiking_has_zero_cv_and_exact_rate():
    r = Raster(2)
    for t in range(1000):
        r.add_step(t, [0] if t % 10 == 0 else [])
    assert r.rates(1e-3) == [pytest.approx(50.0)]  # 100 spikes / (2 neurons * 1 s)
    assert r.isi_cv() == [pytest.approx(0.0)]


def test_rejects_out_of_order_and_out_of_range():
    r = Raster(5)
    r.add_step(0, [1])
    with pytest.raises(ValueError):
        r.add_step(2, [1])
    with pytest.raises(ValueError):
        r.add_step(1, [5])
