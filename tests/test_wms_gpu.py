"""GPU parity of the wave-level multisplit variants (WMS: one subproblem per
wave, ballot ranking; DMS: the same with the keys scattered from registers,
no LDS reorder; the reference's warp-wide path, src/kernels/wms/),
bit-exact against the oracle and against the block-wide variant on the same
inputs, for every bucket count up to 64, key-only and key-value."""
import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W
from gpu_util import banner, filled_u32, from_dev, to_dev

pytestmark = pytest.mark.gpu


def run_variant(dev, keys, B, f, kv, variant):
    n = keys.size
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(max(n, 1), ctx)
    gms.set_variant(ctx, variant)
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


WAVE_VARIANTS = [gms.VARIANT_WMS, gms.VARIANT_DMS]


@pytest.mark.parametrize("variant", WAVE_VARIANTS)
@pytest.mark.parametrize("B", [2, 3, 4, 8, 16, 17, 32, 33, 64])
@pytest.mark.parametrize("kv", [False, True])
def test_wms_matches_oracle(dev, B, kv, variant):
    for n in (0, 1, 1023, 1025, 4096 * 3 + 1000, 1_000_003):
        banner(f"variant {variant} B={B} kv={kv} n={n}")
        keys = W.uniform_keys(n, seed=B * 31 + n)
        got = run_variant(dev, keys, B, gms.range_bucket(B), kv, variant)
        wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0,
                                       np.arange(n, dtype=np.uint32) if kv else None)
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[2], wo)
        if kv:
            np.testing.assert_array_equal(got[1], wv)


@pytest.mark.parametrize("variant", WAVE_VARIANTS)
@pytest.mark.parametrize("dist", ["zipf", "onehot90"])
@pytest.mark.parametrize("B", [8, 64])
def test_wms_equals_bms_on_skewed_keys(dev, dist, B, variant):
    n = (1 << 22) + 17
    keys = W.keys_for(dist, n, B, seed=B)
    a = run_variant(dev, keys, B, gms.range_bucket(B), True, variant)
    b = run_variant(dev, keys, B, gms.range_bucket(B), True, gms.VARIANT_BMS)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("variant", WAVE_VARIANTS)
def test_wms_other_identifiers(dev, variant):
    n = 777_777
    for B, f, kind, param, keys in [
            (64, gms.delta_func(1 << 26), oracle.DELTA, 1 << 26, W.uniform_keys(n, seed=1)),
            (16, gms.ms_bucket_sort_4bit(28), oracle.DIGIT, 28, W.uniform_keys(n, seed=2)),
            (50, gms.identity_bucket(), oracle.IDENTITY, 0, W.uniform_keys(n, seed=3) % np.uint32(60)),
            (2, gms.test_bucket(1), oracle.LOWBITS, 1, W.uniform_keys(n, seed=4))]:
        keys = keys.astype(np.uint32)
        got = run_variant(dev, keys, B, f, True, variant)
        wk, wv, wo = oracle.multisplit(keys, B, kind, param, np.arange(n, dtype=np.uint32))
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[1], wv)
        np.testing.assert_array_equal(got[2], wo)


def test_wms_sizes_and_rejection(dev):
    assert gms.subproblem_size_wms_key_only(32) == (1024, 4096)
    assert gms.subproblem_size_wms_key_value(64) == (1024, 4096)
    assert gms.subproblem_size_wms_key_only(65) == (0, 0)
    ctx = gms.multisplit_context(256)
    with pytest.raises(gms.GmsError):
        gms.set_variant(ctx, gms.VARIANT_WMS)
    with pytest.raises(gms.GmsError):
        gms.set_variant(ctx, gms.VARIANT_DMS)
    with pytest.raises(gms.GmsError):
        gms.set_variant(gms.multisplit_context(8), 4)


@pytest.mark.slow
@pytest.mark.parametrize("variant", WAVE_VARIANTS)
def test_wms_2e27_kv_32(dev, variant):
    n, B = 1 << 27, 32
    banner(f"variant {variant} B={B} n={n}")
    keys = W.uniform_keys(n, seed=27)
    got = run_variant(dev, keys, B, gms.range_bucket(B), True, variant)
    wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[1], wv)
    np.testing.assert_array_equal(got[2], wo)
