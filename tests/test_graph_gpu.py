"""Plain calls captured into a hipGraph (torch.cuda.graph) and replayed: the
multisplit (prescan, single-pass scan whose status words the prescan clears,
postscan, bucket starts), the radix sort (its passes) and the allocation-free
multisplit.  Between replays the input keys are rewritten by eager work, and
every replay must equal the oracle on the keys it saw; the library records no
memset and no allocation into the graph.
"""
import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W
from gpu_util import banner, filled_u32, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _replays(dev, n, B, run, check, seed):
    """Capture run() once, then replay it 4 times with new keys each time."""
    import torch
    kin = to_dev(W.uniform_keys(n, seed=seed), dev)
    run(kin)                                 # eager call first (allocations, lookups)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g):
            run(kin)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    for rep in range(4):
        keys = W.uniform_keys(n, seed=seed + 1 + rep)
        if rep == 2:
            keys = np.sort(keys)             # one replay on sorted keys
        kin.copy_(to_dev(keys, dev))         # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check(keys, rep)


@pytest.mark.parametrize("B,kv,n", [(256, True, 1 << 20), (7, False, 100_003), (1024, True, 300_001)])
def test_multisplit_graph_replays(dev, B, kv, n):
    banner(f"graph multisplit B={B} kv={kv} n={n}")
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(n, ctx)
    vin = to_dev(np.arange(n, dtype=np.uint32), dev) if kv else None
    kout = filled_u32(n, dev)
    vout = filled_u32(n, dev) if kv else None
    offs = filled_u32(B, dev, 0xFFFFFFFF)
    f = gms.range_bucket(B)

    def run(kin):
        gms.multisplit(kin, kout, n, ctx, f, d_value_in=vin, d_value_out=vout, d_offsets=offs)

    def check(keys, rep):
        wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0,
                                       np.arange(n, dtype=np.uint32) if kv else None)
        np.testing.assert_array_equal(from_dev(kout)[:n], wk, err_msg=f"replay {rep} keys")
        np.testing.assert_array_equal(from_dev(offs), wo, err_msg=f"replay {rep} offsets")
        if kv:
            np.testing.assert_array_equal(from_dev(vout)[:n], wv, err_msg=f"replay {rep} values")

    _replays(dev, n, B, run, check, seed=B + n)
    gms.multisplit_release_memory(ctx)


def test_sort_graph_replays(dev):
    n = 500_001
    banner(f"graph sort n={n}")
    ctx = gms.ms_sort_context()
    gms.multisplit_sort_8bit_allocate_pairs(n, ctx)
    vin = to_dev(np.arange(n, dtype=np.uint32), dev)
    kout, vout = filled_u32(n, dev), filled_u32(n, dev)

    def run(kin):
        gms.multisplit_sort_8bit_pairs(kin, kout, vin, vout, n, ctx)

    def check(keys, rep):
        wk, wv = oracle.sort(keys, 8, np.arange(n, dtype=np.uint32))
        np.testing.assert_array_equal(from_dev(kout)[:n], wk, err_msg=f"replay {rep} keys")
        np.testing.assert_array_equal(from_dev(vout)[:n], wv, err_msg=f"replay {rep} values")

    _replays(dev, n, 256, run, check, seed=9)
    gms.multisplit_sort_release_memory(ctx)


def test_noalloc_graph_replays(dev):
    import torch
    n, B = 2_000_003, 64
    banner(f"graph noalloc B={B} n={n}")
    vin = to_dev(np.arange(n, dtype=np.uint32), dev)
    kout, vout = filled_u32(n, dev), filled_u32(n, dev)
    offs = filled_u32(B, dev, 0xFFFFFFFF)

    def run(kin):
        gms.multisplit_noalloc(kin, kout, n, B, gms.range_bucket(B), d_value_in=vin,
                               d_value_out=vout, d_offsets=offs)

    def check(keys, rep):
        wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0, np.arange(n, dtype=np.uint32))
        np.testing.assert_array_equal(from_dev(kout)[:n], wk, err_msg=f"replay {rep} keys")
        np.testing.assert_array_equal(from_dev(vout)[:n], wv, err_msg=f"replay {rep} values")
        np.testing.assert_array_equal(from_dev(offs), wo, err_msg=f"replay {rep} offsets")

    _replays(dev, n, B, run, check, seed=21)
    torch.cuda.synchronize()
