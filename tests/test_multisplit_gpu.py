"""Parity of the HIP multisplit with the CPU oracle, through the C ABI.

Mirrors the reference's driver checks (src/main/main_multisplit.cu:172-192,
:239-267: run the GPU path, run cpu_multisplit_*_general, compare element by
element) and closes its blind spots: values are the original index (so
stability is observable), n is ragged around tile boundaries, skewed bucket
distributions are exercised, and bucket offsets are compared.  At the full
sizes (2^25, 2^27) results are checked through properties that fix a stable
multisplit completely: values are a permutation of 0..n-1, keys equal the
input gathered by values, bucket ids are non-decreasing, values increase
inside each bucket, and offsets are the exclusive scan of the bucket counts.
"""
import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W
from gpu_util import banner, empty_u32, filled_u32, from_dev, to_dev

pytestmark = pytest.mark.gpu


def run_ms(dev, keys, B, functor, kv=True, capacity=None, values=None, want_offsets=True):
    n = keys.size
    ctx = gms.multisplit_context(B)
    (gms.multisplit_allocate_key_value if kv else gms.multisplit_allocate_key_only)(
        capacity if capacity is not None else n, ctx)
    kin = to_dev(keys, dev)
    kout = filled_u32(n, dev)
    offs = np.zeros(B, dtype=np.uint32) if want_offsets else None
    if kv:
        vals = np.arange(n, dtype=np.uint32) if values is None else values
        vin = to_dev(vals, dev)
        vout = filled_u32(n, dev)
        gms.multisplit_key_value(kin, vin, kout, vout, n, ctx, functor, offs)
        res = from_dev(kout)[:n], from_dev(vout)[:n], offs
    else:
        gms.multisplit_key_only(kin, kout, n, ctx, functor, offs)
        res = from_dev(kout)[:n], None, offs
    gms.multisplit_release_memory(ctx)
    return res


def oracle_ms(keys, B, kind, param=0, kv=True, values=None):
    vals = (np.arange(keys.size, dtype=np.uint32) if values is None else values) if kv else None
    return oracle.multisplit(keys, B, kind, param, vals)


def check_equal(got, want, kv):
    np.testing.assert_array_equal(got[0], want[0])
    if kv:
        np.testing.assert_array_equal(got[1], want[1])
    if got[2] is not None:
        np.testing.assert_array_equal(got[2], want[2])


def tile_edge_sizes(B, kv):
    """Sizes around the tile boundaries of this bucket count's own tile
    (gms.tile_keys): empty, tiny, one wave, tile-1 / tile / tile+1, two
    tiles +-1, and a ragged multi-tile input."""
    tile = gms.tile_keys(B, kv, 1 << 20) if B > 1 else 6144
    return sorted({0, 1, 2, 63, 64, 65, 1000, tile - 1, tile, tile + 1, 2 * tile - 1, 2 * tile,
                   2 * tile + 1, 3 * tile + 17, 100_003})


@pytest.mark.parametrize("kv", [False, True])
@pytest.mark.parametrize("B", [1, 2, 3, 4, 5, 8, 16, 17, 32, 33, 64, 100, 128, 129, 255, 256])
def test_range_identifier_all_bucket_counts_ragged(dev, B, kv):
    for n in tile_edge_sizes(B, kv):
        banner(f"range B={B} kv={kv} n={n}")
        keys = W.uniform_keys(n, seed=B * 131 + n)
        got = run_ms(dev, keys, B, gms.range_bucket(B), kv)
        check_equal(got, oracle_ms(keys, B, oracle.RANGE, kv=kv), kv)


@pytest.mark.parametrize("kv", [False, True])
@pytest.mark.parametrize("dist", ["uniform", "zipf", "onehot90"])
@pytest.mark.parametrize("B", [2, 8, 32, 64, 256])
def test_sweep_distributions_2e20(dev, dist, B, kv):
    n = 1 << 20
    banner(f"sweep {dist} B={B} kv={kv} n={n}")
    keys = W.keys_for(dist, n, B, seed=7 * B + len(dist))
    got = run_ms(dev, keys, B, gms.range_bucket(B), kv)
    check_equal(got, oracle_ms(keys, B, oracle.RANGE, kv=kv), kv)


def test_config0_keyonly_2e20_8_buckets(dev):
    # 2^20 uint32 keys uniform over [0, 2^32), 8 buckets by range identifier,
    # key-only, bit-exact against the CPU path.
    keys = W.uniform_keys(1 << 20, seed=20)
    got = run_ms(dev, keys, 8, gms.range_bucket(8), kv=False)
    check_equal(got, oracle_ms(keys, 8, oracle.RANGE, kv=False), False)


@pytest.mark.parametrize("kv", [False, True])
def test_delta_identifier_like_reference_driver(dev, kv):
    # main_multisplit.cu:98-108: B = 32, delta = ceil(n / B), keys in [b*delta, (b+1)*delta)
    for n in (1 << 16, (1 << 20) + 7):
        B = 32
        delta = (n + B - 1) // B
        keys = W.delta_keys(n, B, delta, seed=n)
        banner(f"delta n={n} kv={kv}")
        got = run_ms(dev, keys, B, gms.delta_func(delta), kv)
        check_equal(got, oracle_ms(keys, B, oracle.DELTA, delta, kv=kv), kv)


@pytest.mark.parametrize("B", [2, 7, 64, 256])
def test_delta_identifier_odd_deltas_and_clamping(dev, B):
    keys = W.uniform_keys(50_001, seed=B)
    for delta in (1, 3, 1000, 0x01000001, 0xFFFFFFFF):
        banner(f"delta B={B} delta={delta}")
        got = run_ms(dev, keys, B, gms.delta_func(delta), True)
        check_equal(got, oracle_ms(keys, B, oracle.DELTA, delta), True)


@pytest.mark.parametrize("B", [2, 16, 256])
def test_identity_identifier(dev, B):
    # main_wms.cu drives identity_bucket with keys = shuffled i % B
    n = 40_000
    keys = (W.permutation_keys(n, seed=B) % B).astype(np.uint32)
    got = run_ms(dev, keys, B, gms.identity_bucket(), True)
    check_equal(got, oracle_ms(keys, B, oracle.IDENTITY), True)
    # ids >= B are clamped into bucket B-1 (undefined in the reference)
    keys2 = W.uniform_keys(n, seed=B + 1) % np.uint32(2 * B)
    got = run_ms(dev, keys2, B, gms.identity_bucket(), True)
    check_equal(got, oracle_ms(keys2, B, oracle.IDENTITY), True)


@pytest.mark.parametrize("bits,shift", [(8, 0), (8, 24), (7, 4), (7, 25), (4, 28), (2, 30), (6, 13)])
def test_digit_identifier(dev, bits, shift):
    keys = W.uniform_keys(70_000, seed=bits * 32 + shift)
    B = 1 << bits
    got = run_ms(dev, keys, B, gms._digit_functor(bits)(shift), True)
    check_equal(got, oracle_ms(keys, B, oracle.DIGIT, shift), True)


def test_lowbits_identifier(dev):
    keys = W.permutation_keys(33_333, seed=5)
    for flag, B in ((1, 2), (3, 8), (2, 8)):
        got = run_ms(dev, keys, B, gms.test_bucket(flag), False)
        check_equal(got, oracle_ms(keys, B, oracle.LOWBITS, flag, kv=False), False)


def test_random_payload_values(dev):
    keys = W.uniform_keys(123_457, seed=9)
    vals = W.uniform_keys(123_457, seed=10)
    got = run_ms(dev, keys, 64, gms.range_bucket(64), True, values=vals)
    check_equal(got, oracle_ms(keys, 64, oracle.RANGE, values=vals), True)


def test_call_smaller_than_capacity_and_reuse(dev):
    B = 256
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(1 << 18, ctx)
    for n in (1 << 18, 5000, 1, 0, 77_777):
        banner(f"reuse n={n}")
        keys = W.uniform_keys(n, seed=n + 3)
        vals = np.arange(n, dtype=np.uint32)
        kin, vin = to_dev(keys, dev), to_dev(vals, dev)
        kout, vout = filled_u32(n, dev), filled_u32(n, dev)
        offs = np.full(B, 7, dtype=np.uint32)
        gms.multisplit_key_value(kin, vin, kout, vout, n, ctx, gms.range_bucket(B), offs)
        wk, wv, wo = oracle_ms(keys, B, oracle.RANGE)
        np.testing.assert_array_equal(from_dev(kout)[:n], wk)
        np.testing.assert_array_equal(from_dev(vout)[:n], wv)
        np.testing.assert_array_equal(offs, wo)
    gms.multisplit_release_memory(ctx)


def test_unaligned_pointers_take_scalar_path(dev):
    import torch
    n, B = 20_000, 128
    keys = W.uniform_keys(n + 3, seed=12)
    vals = np.arange(n + 3, dtype=np.uint32)
    kin_full, vin_full = to_dev(keys, dev), to_dev(vals, dev)
    kout_full = torch.zeros(n + 3, dtype=torch.int32, device=dev)
    vout_full = torch.zeros(n + 3, dtype=torch.int32, device=dev)
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(n, ctx)
    gms.multisplit_key_value(kin_full[1:], vin_full[3:], kout_full[2:], vout_full[1:], n, ctx,
                             gms.range_bucket(B))
    wk, wv, _ = oracle.multisplit(keys[1:n + 1], B, oracle.RANGE, 0, vals[3:n + 3])
    np.testing.assert_array_equal(from_dev(kout_full)[2:n + 2], wk)
    np.testing.assert_array_equal(from_dev(vout_full)[1:n + 1], wv)
    gms.multisplit_release_memory(ctx)


def test_device_offsets_and_determinism(dev):
    import torch
    n, B = 300_000, 256
    keys = W.keys_for("zipf", n, B, seed=4)
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(n, ctx)
    kin, vin = to_dev(keys, dev), to_dev(np.arange(n, dtype=np.uint32), dev)
    outs = []
    for rep in range(3):
        kout, vout = filled_u32(n, dev), filled_u32(n, dev)
        d_off = torch.zeros(B, dtype=torch.int32, device=dev)
        gms.multisplit(kin, kout, n, ctx, gms.range_bucket(B), vin, vout, d_off)
        torch.cuda.synchronize()
        outs.append((from_dev(kout), from_dev(vout), from_dev(d_off)))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            np.testing.assert_array_equal(a, b)
    wk, wv, wo = oracle_ms(keys, B, oracle.RANGE)
    np.testing.assert_array_equal(outs[0][0], wk)
    np.testing.assert_array_equal(outs[0][1], wv)
    np.testing.assert_array_equal(outs[0][2], wo)
    gms.multisplit_release_memory(ctx)


def test_prescan_matrix_stage(dev):
    import torch
    for B, n, kv in ((256, 100_000, True), (8, 50_001, True), (3, 9000, True),
                     (256, (1 << 26) + 5, False), (64, (1 << 26) + 3, True),
                     (1000, 70_001, True), (1024, 300_000, False)):
        ctx = gms.multisplit_context(B)
        gms.multisplit_allocate_key_value(n, ctx)
        keys = W.uniform_keys(n, seed=B)
        tile = gms.tile_keys(B, kv, n)
        assert tile >= gms.subproblem_size_bms_key_value(B)
        nt = (n + tile - 1) // tile
        mat = torch.zeros(B * nt, dtype=torch.int32, device=dev)
        gms.prescan_matrix(to_dev(keys, dev), n, ctx, gms.range_bucket(B), mat, key_value=kv)
        np.testing.assert_array_equal(from_dev(mat), oracle.tile_histograms(keys, B, tile))
        gms.multisplit_release_memory(ctx)


@pytest.mark.parametrize("m", [1, 5, 4095, 4096, 4097, 1 << 20, (1 << 22) + 3, 2048 * 4096, 2048 * 4096 + 1])
def test_exclusive_scan_stage(dev, m):
    import torch
    ctx = gms.multisplit_context(256)
    gms.multisplit_allocate_key_value(m * 32 + 4096, ctx)   # matrix (and scan scratch) >= m
    x = (W.uniform_keys(m, seed=m) % 1000).astype(np.uint32)
    d_in = to_dev(x, dev)
    d_out = torch.zeros(m, dtype=torch.int32, device=dev)
    gms.exclusive_scan(d_in, d_out, m, ctx)
    csum = np.cumsum(x.astype(np.uint64))
    want = (np.concatenate([np.zeros(1, np.uint64), csum[:-1]]) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    np.testing.assert_array_equal(from_dev(d_out), want)
    gms.exclusive_scan(d_in, d_in, m, ctx)                    # in place
    np.testing.assert_array_equal(from_dev(d_in), want)
    gms.multisplit_release_memory(ctx)


def test_error_paths(dev):
    import torch
    ctx = gms.multisplit_context(16)
    k = torch.zeros(100, dtype=torch.int32, device=dev)
    o = torch.zeros(100, dtype=torch.int32, device=dev)
    with pytest.raises(gms.GmsError) as e:
        gms.multisplit_key_only(k, o, 100, ctx, gms.range_bucket(16))
    assert e.value.status == 4                                   # not allocated
    gms.multisplit_allocate_key_only(50, ctx)
    with pytest.raises(gms.GmsError) as e:
        gms.multisplit_key_only(k, o, 60, ctx, gms.range_bucket(16))
    assert e.value.status == 3                                   # capacity
    with pytest.raises(gms.GmsError) as e:
        gms.multisplit_key_only(k, k, 50, ctx, gms.range_bucket(16))
    assert e.value.status == 7                                   # in/out alias
    with pytest.raises(gms.GmsError) as e:
        gms.multisplit_key_only(k[:50], k[25:75], 50, ctx, gms.range_bucket(16))
    assert e.value.status == 7
    with pytest.raises(gms.GmsError) as e:
        gms.multisplit_key_only(k, o, 50, ctx, gms.ms_bucket_sort_2bit(0).__class__(40))
    assert e.value.status == 1                                   # shift >= 32
    gms.multisplit_release_memory(ctx)


_KEY_CACHE = {}


def _device_skewed_keys(dev, n, dist, B, seed):
    """Keys of the skewed sweep distributions generated on the GPU (2^27 keys
    in well under a second), for a power-of-two B under the range identifier:
    bucket ids first, then key = id << (32 - log2 B) | uniform low bits.
    zipf: ids by inverse CDF of P(b) ~ 1/(b+1)^1.1; onehot90: 90 % of keys in
    one random bucket, the rest uniform, in random positions."""
    import torch
    assert B & (B - 1) == 0 and B >= 2
    lb = B.bit_length() - 1
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    if dist == "zipf":
        w = 1.0 / torch.arange(1, B + 1, dtype=torch.float64, device=dev) ** 1.1
        cdf = torch.cumsum(w / w.sum(), 0)
        u = torch.rand(n, dtype=torch.float64, device=dev, generator=g)
        ids = torch.searchsorted(cdf, u, right=True).clamp_(max=B - 1)
    elif dist == "onehot90":
        hot = int(torch.randint(0, B, (1,), device=dev, generator=g).item())
        ids = torch.randint(0, B, (n,), dtype=torch.int64, device=dev, generator=g)
        ids[torch.rand(n, device=dev, generator=g) < 0.9] = hot
    else:
        raise ValueError(dist)
    low = torch.randint(0, 1 << (32 - lb), (n,), dtype=torch.int64, device=dev, generator=g)
    k = (ids.to(torch.int64) << (32 - lb)) | low
    return (k - ((k >> 31) << 32)).to(torch.int32)   # the same 32 bits as int32


def _full_size_keys(dev, n, dist, B):
    import torch
    if dist == "uniform":
        g = torch.Generator(device=dev)
        g.manual_seed(n)
        return torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=dev, generator=g)
    if n > (1 << 25):
        return _device_skewed_keys(dev, n, dist, B, seed=n + B)
    key = (n, dist, B)
    if key not in _KEY_CACHE:
        _KEY_CACHE.clear()
        _KEY_CACHE[key] = W.keys_for(dist, n, B, seed=3)
    return to_dev(_KEY_CACHE[key], dev)


def _check_properties(dev, n, dist, B, kv):
    """Complete property check of one full-size multisplit on the GPU: output
    keys are a stable permutation of the input grouped by bucket, values
    travel with their keys, offsets are the exclusive scan of the counts."""
    import torch
    banner(f"full-size {dist} n={n} B={B} kv={kv}")
    kin = _full_size_keys(dev, n, dist, B)
    vin = torch.arange(n, dtype=torch.int32, device=dev)
    kout = torch.empty_like(kin)
    vout = torch.empty_like(vin)
    d_off = torch.zeros(B, dtype=torch.int32, device=dev)
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(n, ctx)
    if kv:
        gms.multisplit(kin, kout, n, ctx, gms.range_bucket(B), vin, vout, d_off)
    else:
        gms.multisplit(kin, kout, n, ctx, gms.range_bucket(B), d_offsets=d_off)
    torch.cuda.synchronize()
    gms.multisplit_release_memory(ctx)
    ku = kin.to(torch.int64) & 0xFFFFFFFF
    bid_in = (ku * B) >> 32
    ko = kout.to(torch.int64) & 0xFFFFFFFF
    bid_out = (ko * B) >> 32
    # bucket ids non-decreasing
    assert bool((bid_out[1:] >= bid_out[:-1]).all())
    # offsets = exclusive scan of bucket counts
    cnt = torch.bincount(bid_in, minlength=B)
    want = torch.cumsum(cnt, 0) - cnt
    assert bool((d_off.to(torch.int64) == want).all())
    if kv:
        vo = vout.to(torch.int64)
        # Range check first: the values index device tensors below, and an
        # out-of-range index would fault instead of failing the test.
        assert bool(((vo >= 0) & (vo < n)).all()), "output values outside [0, n)"
        seen = torch.zeros(n, dtype=torch.int32, device=dev)
        seen.index_add_(0, vo, torch.ones(n, dtype=torch.int32, device=dev))
        assert bool((seen == 1).all())                      # a permutation of 0..n-1
        assert bool((kout == kin[vo]).all())                 # keys travel with their values
        same = bid_out[1:] == bid_out[:-1]
        assert bool((vo[1:][same] > vo[:-1][same]).all())   # stable inside each bucket
    else:
        # key-only: inside each bucket the keys keep input order; compare with
        # the stable order computed by torch (a sort by bucket id, stable)
        order = torch.sort(bid_in, stable=True).indices
        assert bool((kout == kin[order]).all())


@pytest.mark.slow
@pytest.mark.parametrize("n,dist", [(1 << 27, "uniform"), (1 << 25, "uniform"), (1 << 25, "zipf"),
                                    (1 << 25, "onehot90")])
def test_full_size_256_kv_properties(dev, n, dist):
    """2^25 / 2^27 keys, 256 buckets, key-value."""
    _check_properties(dev, n, dist, 256, True)


@pytest.mark.slow
@pytest.mark.parametrize("dist", ["uniform", "zipf", "onehot90"])
@pytest.mark.parametrize("B", [2, 8, 32, 64, 256])
@pytest.mark.parametrize("kv", [False, True])
def test_plain_sweep_2e25(dev, dist, B, kv):
    """The plain-multisplit sweep at 2^25 keys: every bucket count and key
    distribution, key-only and key-value."""
    _check_properties(dev, 1 << 25, dist, B, kv)


@pytest.mark.slow
@pytest.mark.parametrize("B", [2, 8, 32, 64, 256])
@pytest.mark.parametrize("kv", [False, True])
def test_plain_sweep_2e27_uniform(dev, B, kv):
    _check_properties(dev, 1 << 27, "uniform", B, kv)


@pytest.mark.slow
@pytest.mark.parametrize("dist", ["zipf", "onehot90"])
@pytest.mark.parametrize("B", [2, 8, 32, 64, 256])
def test_plain_sweep_2e27_skewed(dev, dist, B):
    """The plain-multisplit sweep at 2^27 keys with Zipf s=1.1 and one-hot
    90 % bucket ids, key-only and key-value, by the complete properties."""
    for kv in (False, True):
        _check_properties(dev, 1 << 27, dist, B, kv)


@pytest.mark.slow
@pytest.mark.parametrize("n,dist", [(1 << 27, "uniform"), (1 << 25, "zipf"), (1 << 25, "onehot90")])
@pytest.mark.parametrize("B", [512, 1000, 1024])
@pytest.mark.parametrize("kv", [False, True])
def test_one_pass_up_to_1024_buckets_properties(dev, n, dist, B, kv):
    """257..1024 buckets run as one pass with three LDS ranking passes; full
    sizes (long tiles from 2^26 keys) and skewed keys, by the same properties."""
    _check_properties(dev, n, dist, B, kv)


@pytest.mark.slow
@pytest.mark.parametrize("B,kv,extra", [(128, False, 12345), (64, True, 777), (256, False, 1),
                                        (256, True, 6143), (1024, True, 9215), (512, True, 9217)])
def test_long_tiles_against_oracle(dev, B, kv, extra):
    """Inputs of 2^26 keys and more take the long-tile shape of some classes;
    ragged sizes there, bit-exact against the oracle."""
    n = (1 << 26) + extra
    banner(f"long tiles B={B} kv={kv} n={n} tile={gms.tile_keys(B, kv, n)}")
    keys = W.uniform_keys(n, seed=B + extra)
    got = run_ms(dev, keys, B, gms.range_bucket(B), kv)
    check_equal(got, oracle_ms(keys, B, oracle.RANGE), kv)


@pytest.mark.slow
def test_2e25_kv_256_against_oracle(dev):
    n, B = 1 << 25, 256
    keys = W.uniform_keys(n, seed=25)
    got = run_ms(dev, keys, B, gms.range_bucket(B), True)
    check_equal(got, oracle_ms(keys, B, oracle.RANGE), True)


@pytest.mark.slow
@pytest.mark.parametrize("B,kv", [(256, False), (16, True)])
def test_past_2e31_keys(dev, B, kv):
    """More than 2^31 keys (8.6 GB per array): positions past the signed
    32-bit range.  The GPU runs only the multisplit (inputs come from torch's
    generators); results are checked on the host with numpy, chunk by chunk:
    bucket order, offsets = scan of counts, the multiset of keys per bucket
    (key-only), and for key-value (value = index) a permutation whose keys
    travel with their values, stable inside each bucket."""
    import torch
    n = (1 << 31) + 12345
    banner(f"past 2^31: n={n} B={B} kv={kv}")
    g = torch.Generator(device=dev)
    g.manual_seed(31)
    kin = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=dev, generator=g)
    kout = torch.empty_like(kin)
    vin = torch.arange(n, dtype=torch.int64, device=dev).to(torch.int32) if kv else None
    vout = torch.empty_like(kin) if kv else None
    d_off = torch.zeros(B, dtype=torch.int32, device=dev)
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(n, ctx)
    gms.multisplit(kin, kout, n, ctx, gms.range_bucket(B), vin, vout, d_off)
    torch.cuda.synchronize()
    gms.multisplit_release_memory(ctx)
    hk = from_dev(kin)
    del kin, vin
    ho = from_dev(kout)
    del kout
    hv = from_dev(vout) if kv else None
    del vout
    offs = from_dev(d_off)
    CH = 1 << 26

    def bid(a):
        return ((a.astype(np.uint64) * np.uint64(B)) >> np.uint64(32)).astype(np.int64)

    def key_sums(b, k):
        # per-bucket sums of the low and high 16 bits (exact in float64)
        return np.stack([np.bincount(b, weights=(k & np.uint32(0xFFFF)).astype(np.float64), minlength=B),
                         np.bincount(b, weights=(k >> np.uint32(16)).astype(np.float64), minlength=B)])

    cnt = np.zeros(B, np.int64)
    ksum_in = np.zeros((2, B))
    for c in range(0, n, CH):
        b = bid(hk[c:c + CH])
        cnt += np.bincount(b, minlength=B)
        if not kv:
            ksum_in += key_sums(b, hk[c:c + CH])
    want_off = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    np.testing.assert_array_equal(offs.astype(np.int64), want_off)
    prev_b = -1
    ksum_out = np.zeros((2, B))
    for c in range(0, n, CH):
        b = bid(ho[c:c + CH])
        assert b[0] >= prev_b and bool((b[1:] >= b[:-1]).all()), f"bucket order broken near {c}"
        prev_b = int(b[-1])
        if not kv:
            ksum_out += key_sums(b, ho[c:c + CH])
    if not kv:
        np.testing.assert_array_equal(ksum_out, ksum_in)       # same keys in every bucket
        return
    seen = np.zeros(n, dtype=bool)
    prev_v, prev_b = -1, -1
    for c in range(0, n, CH):
        v = hv[c:c + CH].astype(np.int64)
        assert bool(((v >= 0) & (v < n)).all()), "output values outside [0, n)"
        seen[v] = True
        np.testing.assert_array_equal(hk[v], ho[c:c + CH])     # keys travel with values
        b = bid(ho[c:c + CH])
        bb = np.concatenate([[prev_b], b])
        vv = np.concatenate([[prev_v], v])
        same = bb[1:] == bb[:-1]
        assert bool((vv[1:][same] > vv[:-1][same]).all()), f"unstable near {c}"
        prev_v, prev_b = int(v[-1]), int(b[-1])
    assert bool(seen.all())                                     # a permutation of 0..n-1
