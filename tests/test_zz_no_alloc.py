"""spice_hip_generate_graph_noalloc -- the generator entry point that allocates no device
memory (its scratch is a fixed region of the library's own device memory, shared by every call
on a GPU) -- against spice_hip_generate_graph, which allocates its scratch per call: the same
offsets, targets and ELL width bit for bit on two calls in a row, the second with a smaller
graph (so the shared region holds the first call's data when the second runs); and inputs
beyond the region's limits are rejected before any device work."""
import numpy as np
import pytest

from oracle import oracle as O
from spice_amd.snn import generate_graph
from tests.helpers import three_pop_layout

pytestmark = pytest.mark.gpu


def _generate(layout, seed, noalloc):
    import torch
    n = sum(layout[0])
    rp = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
    nnz, w = generate_graph(layout, seed, rp.data_ptr(), noalloc=noalloc)  # offsets only
    rp_only = rp.cpu().numpy()
    col = torch.full((max(nnz, 1),), -1, dtype=torch.int32, device="cuda")
    rp.fill_(-1)
    assert generate_graph(layout, seed, rp.data_ptr(), col.data_ptr(), col.numel(), noalloc=noalloc) == (nnz, w)
    torch.cuda.synchronize()
    out = rp.cpu().numpy(), col.cpu().numpy()[:nnz], w
    np.testing.assert_array_equal(rp_only, out[0])
    return out


def test_noalloc_equals_allocating_entry_point_twice_larger_then_smaller(gpu):
    # 1e8-synapse three-population graph (N = 31,623: 31 scan tiles), then a 2e4-neuron,
    # one-population one (fewer connections and tiles than the first call left behind)
    cases = [(three_pop_layout(31_623), 5), (([20_000], [(0, 0, 0.02)]), 9)]
    for k, (lay, seed) in enumerate(cases):
        print(f"[no-alloc] call {k + 1}: N={sum(lay[0])}, {len(lay[1])} connections", flush=True)
        rp_b, col_b, w_b = _generate(lay, seed, True)
        rp_a, col_a, w_a = _generate(lay, seed, False)
        assert w_b == w_a
        np.testing.assert_array_equal(rp_b, rp_a, err_msg=f"call {k + 1}: row offsets")
        np.testing.assert_array_equal(col_b, col_a, err_msg=f"call {k + 1}: targets")
        if k == 1:  # the smaller graph also against the host build of the generator
            rp_h, col_h, w_h = O.generate_csr(lay, seed)
            assert w_b == w_h
            np.testing.assert_array_equal(rp_b, rp_h)
            np.testing.assert_array_equal(col_b, col_h)


def test_noalloc_rejects_inputs_beyond_its_scratch(gpu):
    import torch
    # 33 populations, all pairs: 1089 flattened connections (the region holds 1024)
    lay = ([8] * 33, [(s, d, 0.5) for s in range(33) for d in range(33)])
    rp = torch.zeros(8 * 33 + 1, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError, match="too many connections"):
        generate_graph(lay, 1, rp.data_ptr(), noalloc=True)
    assert generate_graph(lay, 1, rp.data_ptr())[0] > 0  # the allocating entry point takes it
    # one neuron more than the region's scan sums cover; the offset buffer is sized for the
    # whole graph, although the call is rejected before it writes anything
    n = 130_023_425
    rp = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError, match="too many neurons"):
        generate_graph(([n], [(0, 0, 0.0)]), 1, rp.data_ptr(), noalloc=True)
    torch.cuda.synchronize()
    assert int(rp.abs().sum().item()) == 0
