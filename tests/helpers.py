"""Shared helpers for the parity tests (layouts of the benchmark configurations, state
comparison with explicit tolerances)."""
import numpy as np

# Reference test scenario (spice_test/cuda/snn.d.cpp:104-157): vogels, layout{4000, 0.02}.
VOGELS_4000 = ([4000], [(0, 0, 0.02)])


def brunel_ref_layout(n, p=0.1):
    """The reference samples' Brunel topology: {N/2, N/2}, {0->1, 1->1} (samples/main.cpp:1392-1396)."""
    return ([n // 2, n // 2], [(0, 1, p), (1, 1, p)])


def three_pop_layout(n, p=0.1):
    """Benchmark topology: Poisson 20 % / E 80 % x 0.8 / I, every population pair at p."""
    npois = n // 5
    ne = (n - npois) * 4 // 5
    ni = n - npois - ne
    pops = [npois, ne, ni]
    return (pops, [(s, d, p) for s in range(3) for d in range(3)])


def assert_state_close(a, b, rtol, atol, what=""):
    """|a - b| <= atol + rtol * (|b| + max|b|) element-wise for every float attribute, i.e.
    within rtol relative to the value and to the attribute's magnitude over the population
    (summing the same inputs in a different order moves a sum near zero by a few ulps of its
    terms, which a purely element-wise relative bound would reject); integers exactly."""
    for k in a:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, (what, k, x.shape, y.shape)
        if x.dtype.kind in "iu":
            np.testing.assert_array_equal(x, y, err_msg=f"{what} {k}")
        else:
            scale = float(np.max(np.abs(y))) if y.size else 0.0
            np.testing.assert_allclose(x, y, rtol=rtol, atol=atol + rtol * scale, err_msg=f"{what} {k}")


def assert_state_equal(a, b, what=""):
    for k in a:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, (what, k)
        np.testing.assert_array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                                      y.view(np.uint32) if y.dtype == np.float32 else y, err_msg=f"{what} {k}")
