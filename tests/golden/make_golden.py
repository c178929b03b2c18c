#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own CPU build
(oracle/_ref/libspice_ref.so, compiled from the reference's sources with its own flags by
`make -C oracle ref`). Run in a fresh process: the reference seeds the random stream of its
n-th constructed network with 1337 + n, which the fixtures record.

Fixtures (inputs and outputs only; no reference source):
  ref_<model>.npz : the reference's ELL graph, width, delay, seed, per-step spike lists of
                    `steps` steps (concatenated ids + per-step counts) and the final neuron
                    (and synapse) state.
  ref_layouts.json: layout flattening / ELL width of a few layouts.
  ref_degrees.npz : row degrees produced by the reference generator for three layouts.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

CASES = [
    # name, model, layout, delay, steps
    ("vogels", "vogels_abbott", ([600], [(0, 0, 0.1)]), 1, 300),
    ("vogels_d8", "vogels_abbott", ([600], [(0, 0, 0.1)]), 8, 300),
    ("brunel", "brunel", ([300, 300], [(0, 1, 0.1), (1, 1, 0.1)]), 15, 300),
    ("brunel_plus", "brunel_with_plasticity", ([200, 200], [(0, 1, 0.1), (1, 1, 0.1)]), 15, 300),
    ("synth", "synth", ([500], [(0, 0, 0.05)]), 1, 300),
]

LAYOUTS = [
    ([1], [(0, 0, 0.0)]),
    ([100], [(0, 0, 1.0)]),
    ([100], [(0, 0, 0.25)]),
    ([10, 20, 30], [(1, 0, 0.5), (2, 1, 0.25), (1, 2, 0.125)]),
    ([632, 2024, 506], [(s, d, 0.1) for s in range(3) for d in range(3)]),
    ([56568, 14142], [(0, 0, 0.02), (0, 1, 0.02), (1, 0, 0.02), (1, 1, 0.02)]),
]

DEGREE_LAYOUTS = [([4000, 1000], [(0, 1, 0.1)]), ([4000, 400], [(0, 1, 0.5)]), ([4000, 3000], [(0, 1, 0.02)])]


def main():
    assert O.ref_available(), "build the reference CPU path first: make -C oracle ref"
    for name, model, lay, delay, steps in CASES:
        r = O.RefSnn(model, lay, 1e-4, delay)
        ell, w = r.adj()
        ids, counts = [], []
        for _ in range(steps):
            s = np.sort(r.step())
            ids.append(s)
            counts.append(s.size)
        out = {"ell": ell, "width": w, "delay": delay, "seed": r.seed, "steps": steps,
               "spike_ids": np.concatenate(ids).astype(np.int32), "spike_counts": np.asarray(counts, np.int32)}
        for k, v in r.neurons().items():
            out["n_" + k] = v
        for k, v in r.synapses().items():
            out["s_" + k] = v
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **out)
        print(name, "spikes", int(np.sum(counts)))
    lays = []
    for pops, conns in LAYOUTS:
        e, n, md = O.ref_layout(pops, conns)
        lays.append({"pops": pops, "conns": conns, "edges": e, "size": n, "max_degree": md})
    with open(os.path.join(HERE, "ref_layouts.json"), "w") as f:
        json.dump(lays, f, indent=1)
    deg = {}
    for i, (pops, conns) in enumerate(DEGREE_LAYOUTS):
        ell, w = O.ref_generate(pops, conns)
        deg[f"deg{i}"] = (ell.reshape(-1, w)[: pops[0]] != np.iinfo(np.int32).max).sum(axis=1).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_degrees.npz"), **deg)


if __name__ == "__main__":
    main()
