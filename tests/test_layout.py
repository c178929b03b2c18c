"""layout: the reference's known answers (spice_test/util/layout.cpp:9-47,
spice_test/util/adj_list.cpp:35-59) through the product's C ABI (host logic, no GPU) and
against the reference's own layout class where its CPU build is present."""
import numpy as np
import pytest

from oracle import oracle as O
from spice_amd import Layout


def test_single_population_p0():
    l = Layout(1, 0.0)
    assert l.size() == 1 and l.max_degree() == 0
    assert l.connections() == [(0, 1, 0, 1, 0.0)]


def test_no_connections():
    l = Layout([1], [])
    assert l.size() == 1 and l.connections() == [] and l.max_degree() == 0


def test_from_np():
    l = Layout(100, 0.25)
    assert l.size() == 100
    assert l.connections() == [(0, 100, 0, 100, 0.25)]


def test_from_pop_connect_sorted():
    l = Layout([10, 20, 30], [(1, 0, 0.5), (2, 1, 0.25), (1, 2, 0.125)])
    assert l.size() == 60
    assert l.connections() == [(10, 30, 0, 10, 0.5), (10, 30, 30, 60, 0.125), (30, 60, 10, 30, 0.25)]


def test_full_connectivity_width():
    # layout(100, 1) -> ELL width 128 (spice_test/util/adj_list.cpp:35-46)
    assert Layout(100, 1.0).max_degree() == 128
    assert Layout(100, 0.0).max_degree() == 0


@pytest.mark.parametrize("pops,conns", [
    ([], []),
    ([0], [(0, 0, 0.1)]),
    ([10], [(1, 0, 0.1)]),
    ([10], [(0, 0, 1.5)]),
    ([10], [(0, 0, -0.1)]),
    ([10, 10], [(0, 1, 0.1), (0, 1, 0.2)]),
])
def test_invalid_layouts_raise(pops, conns):
    with pytest.raises(ValueError):
        Layout(pops, conns)


def test_matches_checker_restatement():
    for pops, conns in [([10, 20, 30], [(0, 0, 0.5), (0, 1, 0.1), (0, 2, 0.5), (1, 0, 1.0), (1, 2, 0.5)]),
                        ([632, 2024, 506], [(s, d, 0.1) for s in range(3) for d in range(3)])]:
        l = Layout(pops, conns)
        e, n, md = O.layout_flatten(pops, conns)
        assert l.connections() == e and l.size() == n and l.max_degree() == md


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU build not present")
def test_matches_reference_layout_class():
    rng = np.random.default_rng(5)
    for _ in range(50):
        npop = int(rng.integers(1, 5))
        pops = [int(x) for x in rng.integers(1, 5000, npop)]
        pairs = {(int(rng.integers(0, npop)), int(rng.integers(0, npop))) for _ in range(int(rng.integers(0, 6)))}
        conns = [(s, d, float(np.float32(rng.uniform(0, 1)))) for s, d in sorted(pairs)]
        l = Layout(pops, conns)
        e, n, md = O.ref_layout(pops, conns)
        assert l.size() == n and l.max_degree() == md
        assert l.connections() == e
