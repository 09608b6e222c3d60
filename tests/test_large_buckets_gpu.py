"""GPU parity of multisplits into more than 256 buckets (the reduced-bit-sort
path: bucket ids, LSD passes over the id bits, gather), bit-exact against the
oracle for every identifier kind, key-only and key-value, with bucket starts.
"""
import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W
from gpu_util import banner, filled_u32, from_dev, to_dev

pytestmark = pytest.mark.gpu

CASES = [  # (B, functor, oracle kind, oracle param, key transform)
    (257, gms.range_bucket(257), oracle.RANGE, 0, None),
    (1000, gms.delta_func(4294968), oracle.DELTA, 4294968, None),
    (4096, gms.range_bucket(4096), oracle.RANGE, 0, None),
    (70_000, gms.identity_bucket(), oracle.IDENTITY, 0, lambda k: k % np.uint32(70_100)),
    (1 << 16, gms.test_bucket(16), oracle.LOWBITS, 16, None),
    (1 << 20, gms.range_bucket(1 << 20), oracle.RANGE, 0, None),
]


def _run(dev, keys, B, f, kv):
    n = keys.size
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(n, ctx)
    kin = to_dev(keys, dev) if n else filled_u32(0, dev)
    kout = filled_u32(n, dev)
    offs = np.zeros(B, dtype=np.uint32)
    if kv:
        vin = to_dev(np.arange(n, dtype=np.uint32), dev) if n else filled_u32(0, dev)
        vout = filled_u32(n, dev)
        gms.multisplit_key_value(kin, vin, kout, vout, n, ctx, f, offs)
        res = from_dev(kout)[:n], from_dev(vout)[:n], offs
    else:
        gms.multisplit_key_only(kin, kout, n, ctx, f, offs)
        res = from_dev(kout)[:n], None, offs
    gms.multisplit_release_memory(ctx)
    return res


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("kv", [False, True])
def test_more_than_256_buckets(dev, case, kv):
    B, f, kind, param, tr = CASES[case]
    for n in (0, 1, 6145, 300_007):
        banner(f"B={B} kind={kind} kv={kv} n={n}")
        keys = W.uniform_keys(n, seed=B + n)
        if tr is not None:
            keys = tr(keys).astype(np.uint32)
        got = _run(dev, keys, B, f, kv)
        wk, wv, wo = oracle.multisplit(keys, B, kind, param,
                                       np.arange(n, dtype=np.uint32) if kv else None)
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[2], wo)
        if kv:
            np.testing.assert_array_equal(got[1], wv)


def test_digit_identifier_wider_than_8_bits(dev):
    # (key >> 7) & (2^12 - 1) into 4096 buckets
    n = 200_003
    keys = W.uniform_keys(n, seed=12)
    got = _run(dev, keys, 4096, gms.digit_bucket(7, 12), True)
    wk, wv, wo = oracle.multisplit(keys, 4096, oracle.DIGIT, 7, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[1], wv)
    np.testing.assert_array_equal(got[2], wo)


@pytest.mark.slow
def test_more_than_256_buckets_2e25(dev):
    n, B = 1 << 25, 65536
    banner(f"B={B} n={n}")
    keys = W.uniform_keys(n, seed=65536)
    got = _run(dev, keys, B, gms.range_bucket(B), True)
    wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[1], wv)
    np.testing.assert_array_equal(got[2], wo)


@pytest.mark.slow
@pytest.mark.parametrize("B,n", [(1024, (1 << 24) - 1), (1000, (1 << 24) + 9215), (1024, 1 << 24)])
def test_1024_buckets_around_long_tile_switch(dev, B, n):
    """1024-bucket key-value inputs switch to the long tiles at 2^24 keys: just
    below, at, and past the switch with a ragged last tile, against the oracle."""
    banner(f"B={B} n={n} tile={gms.tile_keys(B, True, n)}")
    keys = W.uniform_keys(n, seed=B + n)
    got = _run(dev, keys, B, gms.range_bucket(B), True)
    wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[1], wv)
    np.testing.assert_array_equal(got[2], wo)
