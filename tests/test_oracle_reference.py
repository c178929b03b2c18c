"""Pins the CPU checker to the reference.

* Golden fixtures (tests/golden/ref_*.npz, written by tests/golden/make_golden.py from the
  reference's own CPU build): the checker, run on the reference's graph with the reference's
  random stream (ref_rng), must emit the same spikes at every step and end in the same state
  (a few ulps apart: the reference is compiled with -ffast-math).
* Live reference build (oracle/_ref), when present: bit-exact against the IEEE variant of the
  reference build (same sources without -ffast-math) for all four reference models.
"""
import json
import os

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODEL = {"vogels": "vogels_abbott", "vogels_d8": "vogels_abbott", "brunel": "brunel",
         "brunel_plus": "brunel_with_plasticity", "synth": "synth"}


@pytest.mark.parametrize("name", sorted(MODEL))
def test_checker_reproduces_reference_golden_trajectory(name):
    g = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    model = MODEL[name]
    h = O.CpuSnn.from_ell(model, g["ell"], int(g["width"]), 1e-4, int(g["delay"]), seed=int(g["seed"]), ref_rng=True)
    off = 0
    for i, c in enumerate(g["spike_counts"]):
        want = g["spike_ids"][off:off + c]
        off += c
        np.testing.assert_array_equal(np.sort(h.step(True)), want, err_msg=f"{name} step {i}")
    n = h.neurons()
    for k, v in n.items():
        ref = g["n_" + k]
        if v.dtype.kind == "i":
            np.testing.assert_array_equal(v, ref)
        else:
            np.testing.assert_allclose(v, ref, rtol=1e-5, atol=1e-9)
    if model == "brunel_with_plasticity":
        s = h.synapses(ell=True)
        for k, v in s.items():
            np.testing.assert_allclose(v, g["s_" + k], rtol=1e-5, atol=1e-9)


def test_layouts_match_reference_golden():
    from spice_amd import Layout
    with open(os.path.join(GOLDEN, "ref_layouts.json")) as f:
        cases = json.load(f)
    for c in cases:
        l = Layout(c["pops"], [tuple(x) for x in c["conns"]])
        assert l.size() == c["size"] and l.max_degree() == c["max_degree"]
        assert [list(e) for e in l.connections()] == [list(e) for e in c["edges"]]


def test_degrees_match_reference_golden_distribution():
    from tests.golden.make_golden import DEGREE_LAYOUTS
    g = np.load(os.path.join(GOLDEN, "ref_degrees.npz"))
    for i, (pops, conns) in enumerate(DEGREE_LAYOUTS):
        ref = g[f"deg{i}"]
        ours = O.degrees((pops, conns), 100 + i)[: pops[0], 0]
        assert stats.ks_2samp(ref, ours).pvalue > 1e-3
        n, p = pops[1], conns[0][2]
        assert abs(ours.mean() - n * p) < 4 * np.sqrt(n * p * (1 - p) / pops[0])


LIVE = [
    ("vogels_abbott", ([3000], [(0, 0, 0.03)]), 1),
    ("vogels_abbott", ([3000], [(0, 0, 0.03)]), 8),
    ("brunel", ([1000, 1000], [(0, 1, 0.1), (1, 1, 0.1)]), 15),
    ("brunel_with_plasticity", ([600, 600], [(0, 1, 0.1), (1, 1, 0.1)]), 15),
    ("synth", ([2000], [(0, 0, 0.05)]), 1),
]


@pytest.mark.skipif(not O.ref_available(ieee=True), reason="reference CPU build not present")
@pytest.mark.parametrize("model,lay,delay", LIVE, ids=[f"{m}-d{d}" for m, _, d in LIVE])
def test_checker_bit_exact_vs_reference_ieee_build(model, lay, delay):
    r = O.RefSnn(model, lay, 1e-4, delay, ieee=True)
    ell, w = r.adj()
    h = O.CpuSnn.from_ell(model, ell, w, 1e-4, delay, seed=r.seed, ref_rng=True)
    for i in range(600):
        np.testing.assert_array_equal(np.sort(h.step(True)), np.sort(r.step()), err_msg=f"step {i}")
    for k, v in r.neurons().items():
        np.testing.assert_array_equal(h.neurons()[k], v, err_msg=k)
    for k, v in r.synapses().items():
        np.testing.assert_array_equal(h.synapses(ell=True)[k], v, err_msg=k)
