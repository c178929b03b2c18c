"""Known-answer and accuracy tests of the shared deterministic primitives (host build):
Philox4x32-10, the deterministic log / sqrt / cos(2 pi u), the fixed-point exponential draw
and the Kahan clock. These pin the generator's building blocks independently of the code
that uses them."""
import numpy as np
import pytest

from oracle import oracle as O


# Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11; kat_vectors).
PHILOX_KAT = [
    ([0x00000000] * 4, [0x00000000] * 2, [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,expected", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expected):
    assert O.philox(ctr, key) == expected


def _philox_numpy(ctr, key):
    """Independent vectorised Philox4x32-10 from the paper's round function (uint64 math)."""
    c = [np.asarray(x, np.uint64) for x in ctr]
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    W0, W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
    mask = np.uint64(0xffffffff)
    for r in range(10):
        if r:
            k0, k1 = (k0 + W0) & mask, (k1 + W1) & mask
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & mask, p1 & mask, ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & mask,
             p0 & mask]
    return [int(x) for x in c]


def test_philox_random_counters_match_independent_implementation():
    rng = np.random.default_rng(7)
    for _ in range(200):
        ctr = [int(x) for x in rng.integers(0, 2**32, 4)]
        key = [int(x) for x in rng.integers(0, 2**32, 2)]
        assert O.philox(ctr, key) == _philox_numpy(ctr, key)


def _ulps(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


def test_det_log_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(1e-9, 1.0, 20000), rng.uniform(1.0, 2**32, 20000),
                        2.0 ** rng.uniform(-300, 300, 20000), np.arange(1, 5000, dtype=np.float64),
                        [np.sqrt(0.5), np.sqrt(2.0), 1.0 - 2**-53, 1.0 + 2**-52]])
    y = O.det_log(x)
    ref = np.log(x)
    err = np.abs(y - ref)
    # within 2 ulp of |log x| (absolute error 2e-16 near x = 1 where log x -> 0)
    assert np.all(err <= np.maximum(2 * np.spacing(np.abs(ref)), 2.3e-16)), float(np.max(_ulps(y, ref)))
    assert O.det_log(np.array([1.0]))[0] == 0.0


def test_det_sqrt_accuracy():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(0, 50, 20000), 2.0 ** rng.uniform(-200, 200, 20000), [0.0, 1.0, 4.0, 2.0]])
    y = O.det_sqrt(x)
    pos = x > 0
    assert np.all(_ulps(y[pos], np.sqrt(x[pos])) <= 1)
    assert O.det_sqrt(np.array([0.0]))[0] == 0.0
    np.testing.assert_array_equal(O.det_sqrt(np.array([4.0, 9.0, 16.0])), [2.0, 3.0, 4.0])


def test_det_cos2pi_accuracy():
    """Against 40-digit mpmath values (numpy's cos(2*pi*u) itself carries ~1 ulp of argument error)."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    u = np.concatenate([np.linspace(0, 1, 4001), np.random.default_rng(3).uniform(0, 1, 6000)])
    y = O.det_cos2pi(u)
    exact = np.array([float(mpmath.cos(2 * mpmath.pi * mpmath.mpf(float(x)))) for x in u])
    assert np.max(np.abs(y - exact)) <= 2.3e-16
    np.testing.assert_array_equal(O.det_cos2pi(np.array([0.0, 0.5, 1.0])), [1.0, -1.0, 1.0])


def test_exp_fix_is_floor_of_scaled_exponential():
    """floor(-ln((x+1)/2^32) * 2^24) for 32-bit x; compared with numpy's log in float64 (the
    two may differ by one unit only where the scaled value is within 1e-6 of an integer)."""
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.integers(0, 2**32, 200000, dtype=np.uint64), [0, 1, 2**31, 2**32 - 2, 2**32 - 1]])
    x = x.astype(np.uint32)
    got = O.exp_fix(x).astype(np.int64)
    e = -np.log((x.astype(np.float64) + 1.0) / 2**32) * 2**24
    want = np.floor(e).astype(np.int64)
    diff = got - want
    assert np.all(np.abs(diff) <= 1)
    assert np.all((diff == 0) | (np.abs(e - np.round(e)) < 1e-6))
    assert got[-1] == 0  # U = 1 -> E = 0
    # mean of Exp(1) in 2^24 units
    assert abs(got[:200000].mean() / 2**24 - 1.0) < 0.01


def test_kahan_clock_known_answer():
    """Reference KAT (spice_test/util/numeric.cpp:16-35): 1000 steps of 0.001 sum to exactly 1
    with Kahan compensation, and every returned increment is within 1 % of dt."""
    d, s = O.kahan(0.001, 1000)
    assert np.float32(s[-1]) == np.float32(1.0)
    assert np.float32(d.sum(dtype=np.float32)) == np.float32(1.0) or abs(float(d.sum()) - 1) < 1e-6
    assert np.all(np.abs(d - np.float32(0.001)) <= 0.001 / 100)
    # against a float32 re-computation with numpy
    c = np.float32(0)
    tot = np.float32(0)
    for i in range(1000):
        y = np.float32(np.float32(0.001) - c)
        t = np.float32(tot + y)
        c = np.float32(np.float32(t - tot) - y)
        tot = t
        assert d[i] == y and s[i] == t
