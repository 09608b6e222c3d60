"""Golden fixtures (tests/golden/golden.npz, written by make_golden.py):
the oracle must reproduce them on the CPU and the HIP library on the GPU.
The fixtures pin results that an edit of the oracle could otherwise move
together with the library."""
import hashlib
import os

import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")
MS_CASES = ["ms_range8_ko", "ms_range256_kv", "ms_range100_kv", "ms_zipf64_kv", "ms_delta32_kv",
            "ms_identity16_ko", "ms_lowbits2_kv", "ms_digit128_kv"]


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _functor(B, kind, param):
    if kind == oracle.RANGE:
        return gms.range_bucket(B)
    if kind == oracle.DELTA:
        return gms.delta_func(param)
    if kind == oracle.IDENTITY:
        return gms.identity_bucket()
    if kind == oracle.LOWBITS:
        return gms.test_bucket(param)
    return {7: gms.ms_bucket_sort_7bit, 8: gms.ms_bucket_sort_8bit}[B.bit_length() - 1](param)


@pytest.mark.parametrize("name", MS_CASES)
def test_oracle_reproduces_multisplit_fixture(g, name):
    n, B, kind, param, kv = (int(x) for x in g[f"{name}__meta"])
    keys, perm = g[f"{name}__keys"], g[f"{name}__perm"].astype(np.uint32)
    wk, wv, wo = oracle.multisplit(keys, B, kind, param, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(wv, perm)
    np.testing.assert_array_equal(wk, keys[perm])
    np.testing.assert_array_equal(wo, g[f"{name}__offsets"])


def test_oracle_reproduces_other_fixtures(g):
    for bits in (7, 8):
        keys = g[f"sort{bits}__keys"]
        wk, wv = oracle.sort(keys, bits, np.arange(keys.size, dtype=np.uint32))
        np.testing.assert_array_equal(wv, g[f"sort{bits}__perm"])
        np.testing.assert_array_equal(wk, keys[g[f"sort{bits}__perm"]])
    np.testing.assert_array_equal(oracle.histogram(g["hist256__keys"], 256, oracle.DIGIT, 0),
                                  g["hist256__hist"])
    inv = float(np.float32(32) / np.float32(256))
    np.testing.assert_array_equal(oracle.histogram_even_f32(g["histeven32__samples"], 32, 0.0, inv),
                                  g["histeven32__hist"])
    c, wk, _ = oracle.compaction(g["compact__keys"], oracle.LOWBITS, 1)
    assert c == int(g["compact__count"][0])
    np.testing.assert_array_equal(wk, g["compact__out_keys"])
    n, B, seed = (int(x) for x in g["big8__meta"])
    wk, _, wo = oracle.multisplit(W.uniform_keys(n, seed), B, oracle.RANGE, 0, None)
    assert hashlib.sha256(wk.tobytes()).digest() == g["big8__sha256_out_keys"].tobytes()
    np.testing.assert_array_equal(wo, g["big8__offsets"])


@pytest.mark.gpu
def test_library_reproduces_fixtures(dev, g):
    import torch
    from gpu_util import filled_u32, from_dev, to_dev
    for name in MS_CASES:
        n, B, kind, param, kv = (int(x) for x in g[f"{name}__meta"])
        keys, perm = g[f"{name}__keys"], g[f"{name}__perm"].astype(np.uint32)
        ctx = gms.multisplit_context(B)
        gms.multisplit_allocate_key_value(n, ctx)
        kout = filled_u32(n, dev)
        offs = np.zeros(B, np.uint32)
        if kv:
            vout = filled_u32(n, dev)
            gms.multisplit_key_value(to_dev(keys, dev), to_dev(np.arange(n, dtype=np.uint32), dev),
                                     kout, vout, n, ctx, _functor(B, kind, param), offs)
            np.testing.assert_array_equal(from_dev(vout)[:n], perm, err_msg=name)
        else:
            gms.multisplit_key_only(to_dev(keys, dev), kout, n, ctx, _functor(B, kind, param), offs)
        np.testing.assert_array_equal(from_dev(kout)[:n], keys[perm], err_msg=name)
        np.testing.assert_array_equal(offs, g[f"{name}__offsets"], err_msg=name)
        gms.multisplit_release_memory(ctx)
    for bits in (7, 8):
        keys = g[f"sort{bits}__keys"]
        n = keys.size
        ctx = gms.ms_sort_context()
        (gms.multisplit_sort_7bit_allocate_pairs if bits == 7 else gms.multisplit_sort_8bit_allocate_pairs)(n, ctx)
        kout, vout = filled_u32(n, dev), filled_u32(n, dev)
        (gms.multisplit_sort_7bit_pairs if bits == 7 else gms.multisplit_sort_8bit_pairs)(
            to_dev(keys, dev), kout, to_dev(np.arange(n, dtype=np.uint32), dev), vout, n, ctx)
        np.testing.assert_array_equal(from_dev(vout)[:n], g[f"sort{bits}__perm"])
        gms.multisplit_sort_release_memory(ctx)
    h = torch.zeros(256, dtype=torch.int32, device=dev)
    gms.histogram(to_dev(g["hist256__keys"], dev), g["hist256__keys"].size, h, 256, gms.ms_bucket_sort_8bit(0))
    np.testing.assert_array_equal(from_dev(h), g["hist256__hist"])
    x = g["histeven32__samples"]
    he = torch.zeros(32, dtype=torch.int32, device=dev)
    gms.multisplit2_histogram_even(torch.from_numpy(x).to(dev), x.size, he, 0.0, 256.0,
                                   float(np.float32(32) / np.float32(256)), 32)
    np.testing.assert_array_equal(from_dev(he).view(np.int32), g["histeven32__hist"])
    keys = g["compact__keys"]
    ctx = gms.compaction_context()
    gms.compaction_allocate_key_only(keys.size, ctx)
    kout = filled_u32(keys.size, dev)
    c = gms.compaction_key_only(to_dev(keys, dev), kout, keys.size, ctx, gms.test_bucket(1))
    assert c == int(g["compact__count"][0])
    np.testing.assert_array_equal(from_dev(kout)[:keys.size], g["compact__out_keys"])
    gms.compaction_release_memory(ctx)
    # the 2^20-key, 8-bucket key-only configuration
    n, B, seed = (int(x) for x in g["big8__meta"])
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_only(n, ctx)
    kout = filled_u32(n, dev)
    offs = np.zeros(B, np.uint32)
    gms.multisplit_key_only(to_dev(W.uniform_keys(n, seed), dev), kout, n, ctx, gms.range_bucket(B), offs)
    assert hashlib.sha256(from_dev(kout)[:n].tobytes()).digest() == g["big8__sha256_out_keys"].tobytes()
    np.testing.assert_array_equal(offs, g["big8__offsets"])
    gms.multisplit_release_memory(ctx)
