"""The three-population Brunel neuron (models brunel_3pop and brunel_3pop_stdp) pinned by
closed forms that the mathematics of the configuration fixes, independently of how the model
header is written:

  * LIF with tau = 20 ms: between spikes and inputs V_k = V_0 (1 - dt / tau)^k;
  * threshold 20 mV: V = 20.5 mV spikes, V = 19.5 mV does not;
  * reset to exactly 10 mV, then 20 refractory steps (Twait = 20 - k) during which V holds
    and inputs are dropped; the decay resumes at the 20th step;
  * weights 5/N V from excitatory sources and -25/N V from inhibitory ones (the 1e-4 and
    -5e-4 "standard weights" times 5e4/N), received D steps after the spike;
  * brunel_3pop_stdp: W starts at 1e-4 (excitatory source) / -5e-4 (inhibitory), receive()
    adds W * 5e4/N, the pre event of an arrival subtracts A- wmax y_post with
    A- = 0.0105, wmax = 3e-4 and y_post = (1 - dt / tau)^t for a target that spiked t steps
    earlier.

The checks run on the CPU checker (oracle/liboracle.so, which compiles the same model headers
as the GPU library). test_mutated_checker_fails_closed_form builds copies of the checker from
model headers with one constant changed and shows that every such change fails them; the GPU
library meets the same closed forms in tests/test_gpu_parity.py.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DT = 1e-4
N = 1000
# brunel_3pop populations of N = 1000: Poisson [0, 200), excitatory [200, 840), inhibitory
# [840, 1000). Layout populations isolate the probe neurons:
#   a = 200 (E source), b = 201 (E target of a), c = 202 (E target of a, refractory when a's
#   spike arrives), q = 203 (sub-threshold probe), e = 204 (target of i), r = 205 (spikes),
#   i = 840 (I source)
POPS = [200, 1, 1, 1, 1, 1, 1, 634, 1, 159]
CONNS = [(1, 2, 1.0), (1, 3, 1.0), (8, 5, 1.0)]
A, B, C, Q, E, R, I = 200, 201, 202, 203, 204, 205, 840


def check_lif_closed_form(make, model, delay, steps=260):
    """make(model, layout, dt, delay) -> a net with step(True), neurons(), set_neurons(),
    csr() and, for plastic models, synapses(ell=False)."""
    h = make(model, (POPS, CONNS), DT, delay)
    rp, col = h.csr()
    src = np.repeat(np.arange(N), np.diff(rp))
    assert sorted(zip(src.tolist(), col.tolist())) == [(A, B), (A, C), (I, E)]
    st = h.neurons()
    assert np.all(st["V"] == 0) and np.all(st["Twait"] == 0)
    V = st["V"].copy()
    V[[A, C, I]] = 0.0205
    V[R] = 0.0201
    V[Q] = 0.0195
    h.set_neurons({"V": V, "Twait": st["Twait"]})

    a = 1.0 - DT / 0.020                 # LIF decay per step (tau = 20 ms)
    wex, win = 5.0 / N, -25.0 / N        # volts
    reset = np.float32(0.01)
    for t in range(steps):
        s = h.step(True)
        lif = set(int(x) for x in s if x >= 200)
        assert lif == ({A, C, R, I} if t == 0 else set()), (model, delay, t, sorted(lif))
        st = h.neurons()
        V, Tw = st["V"].astype(np.float64), st["Twait"]
        # spiked at step 0: reset to exactly 10 mV, 20 refractory steps holding V (inputs to c
        # are dropped), then the decay from 10 mV
        for x in (A, C, R, I):
            assert Tw[x] == 20 - t, (model, t, x, Tw[x])
            if t < 20:
                assert st["V"][x] == reset, (model, t, x, st["V"][x])
            else:
                np.testing.assert_allclose(V[x], 0.01 * a ** (t - 19), rtol=3e-5, err_msg=f"{model} t={t} x={x}")
        # below threshold: decays from 19.5 mV from the first step on
        np.testing.assert_allclose(V[Q], 0.0195 * a ** (t + 1), rtol=3e-5, err_msg=f"{model} t={t} q")
        # a's spike (emitted at step 0) reaches b at step `delay`, i's reaches e
        if t < delay:
            assert V[B] == 0.0 and V[E] == 0.0, (model, t)
        else:
            np.testing.assert_allclose(V[B], wex * a ** (t - delay + 1), rtol=3e-5, err_msg=f"{model} t={t} b")
            np.testing.assert_allclose(V[E], win * a ** (t - delay + 1), rtol=3e-5, err_msg=f"{model} t={t} e")
    if model == "brunel_3pop_stdp":
        w = h.synapses(ell=False)["W"].astype(np.float64)
        k = {(int(s_), int(d_)): n_ for n_, (s_, d_) in enumerate(zip(src, col))}
        # a -> b: b never spiked, y_b = 0: the pre event leaves W at 1e-4
        assert np.float32(w[k[(A, B)]]) == np.float32(1e-4)
        # a -> c: c spiked at step 0, y_c at step `delay` = a^delay
        np.testing.assert_allclose(w[k[(A, C)]], 1e-4 - 0.0105 * 3e-4 * a ** delay, rtol=1e-6)
        # i -> e is not plastic
        assert np.float32(w[k[(I, E)]]) == np.float32(-5e-4)
    return h


def _checker(model, layout, dt, delay):
    from oracle import oracle as O
    return O.CpuSnn(model, layout, dt, delay, seed=3)


@pytest.mark.parametrize("model", ["brunel_3pop", "brunel_3pop_stdp"])
@pytest.mark.parametrize("delay", [1, 3, 15])
def test_checker_lif_closed_form(model, delay):
    check_lif_closed_form(_checker, model, delay)


# (header, original text, mutated text, models whose checks must fail)
MUTATIONS = [
    ("brunel_3pop.h", "float const Vreset = 0.01f;", "float const Vreset = 0.0f;", ("brunel_3pop", "brunel_3pop_stdp")),
    ("brunel_3pop.h", "float const Vthres = 0.02f;", "float const Vthres = 0.015f;", ("brunel_3pop", "brunel_3pop_stdp")),
    ("brunel_3pop.h", "float const TmemInv = 1.0f / 0.02f;", "float const TmemInv = 1.0f / 0.01f;",
     ("brunel_3pop", "brunel_3pop_stdp")),
    ("brunel_3pop.h", "int_ const Tref = 20;", "int_ const Tref = 10;", ("brunel_3pop", "brunel_3pop_stdp")),
    ("brunel_3pop.h", "0.0001f * 50'000 / info.num_neurons", "0.0001f * 20'000 / info.num_neurons", ("brunel_3pop",)),
    ("brunel_3pop.h", "-0.0005f * 50'000 / info.num_neurons", "-0.0005f * 20'000 / info.num_neurons", ("brunel_3pop",)),
    ("brunel_3pop_stdp.h", "get<W>( syn ) * 50'000", "get<W>( syn ) * 20'000", ("brunel_3pop_stdp",)),
    ("brunel_3pop_stdp.h", "? 0.0001f : -0.0005f", "? 0.0001f : -0.0002f", ("brunel_3pop_stdp",)),
    ("brunel_3pop_stdp.h", "float const Aminus = 0.0105f * 0.0003f;", "float const Aminus = 0.01f * 0.0003f;",
     ("brunel_3pop_stdp",)),
]

_RUN = (
    "import sys; sys.path[:0] = [{root!r}, {tests!r}]\n"
    "import test_lif_closed_form as t\n"
    "for d in (1, 3):\n"
    "    t.check_lif_closed_form(t._checker, {model!r}, d)\n"
)


def _build_checker(dst, include):
    out = os.path.join(dst, "liboracle.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-w", "-I" + include,
                    "-shared", "-o", out, os.path.join(ROOT, "oracle", "oracle_capi.cpp")], check=True)
    return out


def _copy_include(dst):
    import shutil
    inc = os.path.join(dst, "include")
    shutil.copytree(os.path.join(ROOT, "include"), inc)
    return inc


def _run_checks(lib, model):
    env = dict(os.environ, SPICE_ORACLE_LIB=lib)
    code = _RUN.format(root=ROOT, tests=os.path.join(ROOT, "tests"), model=model)
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)


def test_mutated_checker_fails_closed_form(tmp_path):
    """Each changed constant of the headline model makes the closed-form checks fail on a
    checker built from the changed header; the unchanged copy, built the same way, passes."""
    jobs = [("control", None, None, None, ("brunel_3pop", "brunel_3pop_stdp"))] + \
           [(f"m{k}",) + m for k, m in enumerate(MUTATIONS)]

    def one(job):
        name, header, old, new, models = job
        d = tmp_path / name
        d.mkdir()
        inc = _copy_include(str(d))
        if header is not None:
            path = os.path.join(inc, "spice", "models", header)
            with open(path) as f:
                text = f.read()
            assert text.count(old) == 1, (header, old)
            with open(path, "w") as f:
                f.write(text.replace(old, new))
        lib = _build_checker(str(d), inc)
        return name, header, new, [(m, _run_checks(lib, m)) for m in models]

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        results = list(ex.map(one, jobs))
    for name, header, new, runs in results:
        for model, r in runs:
            if header is None:
                assert r.returncode == 0, (model, r.stdout[-2000:], r.stderr[-2000:])
            else:
                assert r.returncode != 0 and "AssertionError" in r.stderr, \
                    f"{header}: '{new}' passed the closed-form checks of {model}"
