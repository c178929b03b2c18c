"""Random-graph generator (host build of the device generator): the reference's structural
known answers (spice_test/util/adj_list.cpp:9-149, spice_test/cuda/algorithm.d.cpp:13-56)
and distribution checks against closed forms and against the reference's own generator
(spice/util/adj_list.cpp:49-91) where its CPU build is present."""
import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O

INT_MAX = np.iinfo(np.int32).max


def rows(rp, col):
    return [col[rp[i]:rp[i + 1]] for i in range(rp.size - 1)]


def test_full_connectivity_is_identity():
    rp, col, w = O.generate_csr(([100], [(0, 0, 1.0)]), 3)
    assert w == 128
    for r in rows(rp, col):
        np.testing.assert_array_equal(r, np.arange(100))


def test_zero_probability_is_empty():
    rp, col, w = O.generate_csr(([100], [(0, 0, 0.0)]), 3)
    assert w == 0 and col.size == 0 and np.all(rp == 0)


def test_three_population_ranges():
    lay = ([10, 20, 30], [(0, 0, 0.5), (0, 1, 0.1), (0, 2, 0.5), (1, 0, 1.0), (1, 2, 0.5)])
    in_band_a = in_band_b = 0
    for seed in range(1, 41):
        rp, col, w = O.generate_csr(lay, seed)
        rs = rows(rp, col)
        for i in range(10):
            r = rs[i]
            assert np.all((r >= 0) & (r < 60)) and np.all(np.diff(r) > 0)
            in_band_a += 12 <= r.size <= 32
        for i in range(10, 30):
            r = rs[i]
            assert 10 <= r.size <= 40
            assert np.all((r < 10) | (r >= 30)) and np.all(np.diff(r) > 0)
            # B -> A at p = 1 is complete
            np.testing.assert_array_equal(r[:10], np.arange(10))
            in_band_b += 15 <= r.size <= 35
        for i in range(30, 60):
            assert rs[i].size == 0
    # the reference's rng-dependent bands hold for (almost) every row
    assert in_band_a >= 0.95 * 400 and in_band_b >= 0.95 * 800


def test_sorted_in_range_and_clipped_to_width():
    lay = ([3000, 1000], [(0, 0, 0.3), (0, 1, 0.7), (1, 0, 0.05)])
    rp, col, w = O.generate_csr(lay, 11)
    lens = np.diff(rp)
    assert lens.max() <= w
    for i, r in enumerate(rows(rp, col)):
        assert np.all(np.diff(r) > 0)
        assert r.size == 0 or (r[0] >= 0 and r[-1] < 4000)
        if i >= 3000:
            assert r.size == 0 or r[-1] < 3000


def test_deterministic_per_seed():
    lay = ([2000], [(0, 0, 0.1)])
    a = O.generate_csr(lay, 5)
    b = O.generate_csr(lay, 5)
    c = O.generate_csr(lay, 6)
    np.testing.assert_array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])


def test_degree_mean_and_variance_match_binomial():
    n, p, rows_ = 1000, 0.1, 20000
    d = O.degrees(([rows_, n], [(0, 1, p)]), 9)[:rows_, 0].astype(np.float64)
    mean, var = n * p, n * p * (1 - p)
    assert abs(d.mean() - mean) < 4 * np.sqrt(var / rows_)
    assert abs(d.var() / var - 1) < 0.05
    # rounding of the normal approximation: integer degrees, never outside [0, range]
    assert d.min() >= 0 and d.max() <= n


def test_target_ids_uniform_marginal():
    n = 2000
    rp, col, _ = O.generate_csr(([n], [(0, 0, 0.05)]), 13)
    hist = np.bincount(col // 40, minlength=n // 40)
    chi2, pval = stats.chisquare(hist)
    assert pval > 1e-4, (chi2, pval)


def test_gaps_match_uniform_subset():
    """For a uniform random k-subset of [0, R), the first element has mean (R - k)/(k + 1)."""
    n = 1000
    rp, col, _ = O.generate_csr(([20000, n], [(0, 1, 0.05)]), 17)
    rs = [r for r in rows(rp, col) if r.size]
    first = np.array([r[0] - 20000 for r in rs], np.float64)
    k = np.array([r.size for r in rs], np.float64)
    expected = np.mean((n - k) / (k + 1))
    assert abs(first.mean() / expected - 1) < 0.05


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU build not present")
@pytest.mark.parametrize("n,p", [(1000, 0.1), (400, 0.5), (3000, 0.02)])
def test_degree_distribution_matches_reference_generator(n, p):
    """Same distribution as the reference generator (two-sample KS on row degrees)."""
    rows_ = 6000
    lay = ([rows_, n], [(0, 1, p)])
    ell, w = O.ref_generate(*lay)
    ref_deg = (ell.reshape(-1, w)[:rows_] != INT_MAX).sum(axis=1)
    ours = O.degrees(lay, 21)[:rows_, 0]
    ks = stats.ks_2samp(ref_deg, ours)
    assert ks.pvalue > 1e-3, ks
    assert abs(ref_deg.mean() - ours.mean()) < 0.5 + 4 * np.sqrt(n * p * (1 - p) * 2 / rows_)


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU build not present")
def test_target_distribution_matches_reference_generator():
    lay = ([4000, 1000], [(0, 1, 0.1)])
    ell, w = O.ref_generate(*lay)
    e = ell.reshape(-1, w)[:4000]
    ref_ids = e[e != INT_MAX] - 4000
    rp, col, _ = O.generate_csr(lay, 23)
    ours = col - 4000
    h1 = np.bincount(ref_ids // 50, minlength=20)
    h2 = np.bincount(ours // 50, minlength=20)
    chi2, pval, _, _ = stats.chi2_contingency(np.vstack([h1, h2]))
    assert pval > 1e-3, (chi2, pval)


def test_rows_only_generation_equals_the_whole_graph():
    """The host build of a row range (used where the whole graph is too large for the host,
    e.g. 1e10 synapses) equals the same rows of the whole graph: absolute offsets and targets."""
    lay = ([200, 640, 160], [(s, d, 0.1) for s in range(3) for d in range(3)])
    rp, col, _ = O.generate_csr(lay, 5)
    for r0, r1 in [(0, 1000), (0, 0), (17, 18), (333, 777), (999, 1000), (1000, 1000)]:
        a, b = O.generate_rows(lay, 5, r0, r1)
        np.testing.assert_array_equal(a, rp[r0:r1 + 1])
        np.testing.assert_array_equal(b, col[rp[r0]:rp[r1]])
    a, b = O.generate_rows(lay, 5, 10, 20, cols=False)
    assert b is None and np.array_equal(a, rp[10:21])
    with pytest.raises(ValueError):
        O.generate_rows(lay, 5, 5, 1001)
