"""GPU parity of the multisplit radix sort, histogram and compaction (C ABI).

Sort: the reference checks its multisplit sort against CUB's radix sort
(src/main/main_sort.cu:162-229, main_sort_pairs.cu); here the check is the
CPU oracle (LSD over the same digit schedule) and numpy's stable sort, with
values = original index so that instability shows.  Histogram: the
reference compares against CUB's HistogramEven (src/main/main_histogram.cu:
185-186) and CUB's HistogramRange (main_histogram.cu:224-310); here numpy
bincount / the oracle's float formula / the oracle's edge scan.  Compaction: the
reference compares with its CPU multisplit (src/main/main_compaction.cu:
102-115, 167-181); here the oracle's stable partition and its count.
"""
import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W
from gpu_util import banner, filled_u32, from_dev, to_dev
from hist_cases import range_level_sets, range_samples

pytestmark = pytest.mark.gpu


def dev_in(a, dev):
    """Input buffer; an empty input still gets a (never read) one-element buffer."""
    return to_dev(a, dev) if a.size else filled_u32(0, dev)

SORT_FUNCS = {
    4: (gms.multisplit_sort_4bit_allocate, gms.multisplit_sort_4bit_allocate_pairs,
        gms.multisplit_sort_4bit, gms.multisplit_sort_4bit_pairs),
    5: (gms.multisplit_sort_5bit_allocate, gms.multisplit_sort_5bit_allocate_pairs,
        gms.multisplit_sort_5bit, gms.multisplit_sort_5bit_pairs),
    6: (gms.multisplit_sort_6bit_allocate, gms.multisplit_sort_6bit_allocate_pairs,
        gms.multisplit_sort_6bit, gms.multisplit_sort_6bit_pairs),
    7: (gms.multisplit_sort_7bit_allocate, gms.multisplit_sort_7bit_allocate_pairs,
        gms.multisplit_sort_7bit, gms.multisplit_sort_7bit_pairs),
    8: (gms.multisplit_sort_8bit_allocate, gms.multisplit_sort_8bit_allocate_pairs,
        gms.multisplit_sort_8bit, gms.multisplit_sort_8bit_pairs),
}


def gpu_sort(dev, keys, bits, kv):
    n = keys.size
    alloc, alloc_pairs, sort, sort_pairs = SORT_FUNCS[bits]
    ctx = gms.ms_sort_context()
    (alloc_pairs if kv else alloc)(n, ctx)
    kin = dev_in(keys, dev)
    kout = filled_u32(n, dev)
    if kv:
        vin = dev_in(np.arange(n, dtype=np.uint32), dev)
        vout = filled_u32(n, dev)
        sort_pairs(kin, kout, vin, vout, n, ctx)
        res = from_dev(kout)[:n], from_dev(vout)[:n]
    else:
        sort(kin, kout, n, ctx)
        res = from_dev(kout)[:n], None
    # the input is never written
    np.testing.assert_array_equal(from_dev(kin)[:n], keys)
    gms.multisplit_sort_release_memory(ctx)
    return res


@pytest.mark.parametrize("bits", [4, 5, 6, 7, 8])
@pytest.mark.parametrize("kv", [False, True])
def test_sort_all_digit_widths(dev, bits, kv):
    for n in (0, 1, 7, 6143, 6145, 100_003):
        banner(f"sort bits={bits} kv={kv} n={n}")
        keys = W.uniform_keys(n, seed=bits * 1000 + n)
        if n > 10:
            keys[::9] = keys[3]               # duplicates expose instability
        got = gpu_sort(dev, keys, bits, kv)
        wk, wv = oracle.sort(keys, bits, np.arange(n, dtype=np.uint32) if kv else None)
        np.testing.assert_array_equal(got[0], wk)
        if kv:
            np.testing.assert_array_equal(got[1], wv)


@pytest.mark.slow
@pytest.mark.parametrize("bits", [7, 8])
@pytest.mark.parametrize("kv", [False, True])
def test_sort_2e25_like_reference_config(dev, bits, kv):
    # 2^25 uint32 keys uniform, key-only and key-value (value = original index)
    n = 1 << 25
    keys = W.uniform_keys(n, seed=bits + 25)
    got = gpu_sort(dev, keys, bits, kv)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(got[0], keys[order])
    if kv:
        np.testing.assert_array_equal(got[1], order.astype(np.uint32))


def test_sort_reference_driver_keys(dev):
    # main_sort.cu feeds rand() keys (31 bits) and pairs with value = key
    n = 1 << 20
    keys = (W.uniform_keys(n, seed=31) >> np.uint32(1)).astype(np.uint32)
    got = gpu_sort(dev, keys, 7, True)
    np.testing.assert_array_equal(got[0], np.sort(keys))


@pytest.mark.parametrize("num_bins", [2, 32, 64, 100, 256])
def test_histogram_u32_bins(dev, num_bins):
    import torch
    for n in (0, 1, 4095, 123_457):
        keys = W.uniform_keys(n, seed=num_bins + n)
        h = torch.zeros(num_bins, dtype=torch.int32, device=dev)
        gms.histogram(dev_in(keys, dev), n, h, num_bins, gms.range_bucket(num_bins))
        np.testing.assert_array_equal(from_dev(h), oracle.histogram(keys, num_bins, oracle.RANGE))


def test_histogram_unaligned_and_tails(dev):
    import torch
    for n, off in ((4096 * 3 + 17, 1), (4096 * 512 * 2 + 5, 3), (100, 2), (4096 * 600, 0)):
        keys = W.uniform_keys(n, seed=n + off)
        buf = to_dev(np.concatenate([np.zeros(off, np.uint32), keys]), dev)
        h = torch.zeros(256, dtype=torch.int32, device=dev)
        gms.histogram(buf[off:], n, h, 256, gms.ms_bucket_sort_8bit(8))
        np.testing.assert_array_equal(from_dev(h), oracle.histogram(keys, 256, oracle.DIGIT, 8))
        x = (W.rng(n).random(n, dtype=np.float32) * np.float32(300.0) - np.float32(20.0)).astype(np.float32)
        xb = torch.from_numpy(np.concatenate([np.zeros(off, np.float32), x])).to(dev)
        he = torch.zeros(64, dtype=torch.int32, device=dev)
        inv = float(np.float32(64) / np.float32(256))
        gms.multisplit2_histogram_even(xb[off:], n, he, 0.0, 256.0, inv, 64)
        np.testing.assert_array_equal(from_dev(he).view(np.int32), oracle.histogram_even_f32(x, 64, 0.0, inv))


def skewed_keys(kind, n, seed):
    """Keys whose bins coincide across whole waves: sorted (long runs), one
    bucket holding 90 % of shuffled keys, and every key equal."""
    g = W.rng(seed)
    k = g.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    if kind == "sorted":
        return np.sort(k)
    if kind == "dominant":
        k[g.random(n) < 0.9] = np.uint32(0x9ABCDEF0)
        return k
    return np.full(n, 0x12345678, np.uint32)


@pytest.mark.parametrize("kind", ["sorted", "dominant", "equal"])
def test_skewed_keys_histogram_and_multisplit(dev, kind):
    import torch
    # Lanes of a wave that share a bin are counted together (histogram copies,
    # ballot-combined prescan counts); the counts must stay exact.
    for n in (4096 * 37 + 11, (1 << 22) + 3):
        keys = skewed_keys(kind, n, seed=n)
        d = to_dev(keys, dev)
        for bins in (16, 256):
            h = torch.zeros(bins, dtype=torch.int32, device=dev)
            gms.histogram(d, n, h, bins, gms.range_bucket(bins))
            np.testing.assert_array_equal(from_dev(h), oracle.histogram(keys, bins, oracle.RANGE))
        for B in (16, 256):
            vals = np.arange(n, dtype=np.uint32)
            kout, vout = torch.empty_like(d), torch.empty_like(d)
            offs = np.zeros(B, np.uint32)
            ctx = gms.multisplit_context(B)
            gms.multisplit_allocate_key_value(n, ctx)
            gms.multisplit_key_value(d, to_dev(vals, dev), kout, vout, n, ctx, gms.range_bucket(B),
                                     offs)
            gms.multisplit_release_memory(ctx)
            wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0, vals)
            np.testing.assert_array_equal(from_dev(kout), wk)
            np.testing.assert_array_equal(from_dev(vout), wv)
            np.testing.assert_array_equal(offs, wo)


def test_histogram_range_sorted_samples(dev):
    import torch
    # Sorted samples put whole waves into one bin (the kernel's histogram
    # copies keep that from serialising); counts must stay exact.
    g = W.rng(48)
    n = (1 << 22) + 9
    x = np.sort((g.random(n, dtype=np.float32) * np.float32(1024.0)).astype(np.float32))
    lv = np.sort((g.random(257, dtype=np.float32) * np.float32(1024.0)).astype(np.float32))
    lv[0], lv[-1] = 0.0, 1024.0
    h = torch.zeros(256, dtype=torch.int32, device=dev)
    gms.multisplit2_histogram_range(torch.from_numpy(x).to(dev), n, h, torch.from_numpy(lv).to(dev),
                                    257)
    np.testing.assert_array_equal(from_dev(h).view(np.int32), oracle.histogram_range(x, lv))


def test_histogram_256_low_byte_accumulates(dev):
    import torch
    # 256-bin histogram of the low 8 bits, added twice into the same array
    n = (1 << 22) + 5
    keys = W.uniform_keys(n, seed=256)
    h = torch.zeros(256, dtype=torch.int32, device=dev)
    d = to_dev(keys, dev)
    gms.histogram(d, n, h, 256, gms.ms_bucket_sort_8bit(0))
    gms.histogram(d, n, h, 256, gms.ms_bucket_sort_8bit(0))
    want = 2 * np.bincount(keys & 0xFF, minlength=256)
    np.testing.assert_array_equal(from_dev(h).astype(np.int64), want)


def test_histogram_even_float_like_reference_driver(dev):
    import torch
    # main_histogram.cu:40-120: float samples in [0, 256), 32 bins, inv_delta = 32/256
    n = 1 << 20
    g = W.rng(7)
    x = (g.random(n, dtype=np.float32) * np.float32(256.0)).astype(np.float32)
    x[:4] = np.array([0.0, 255.99998, 256.0, -0.5], dtype=np.float32)
    lower, upper, bins = 0.0, 256.0, 32
    inv = float(np.float32(1.0) / np.float32((upper - lower) / bins))
    h = torch.zeros(bins, dtype=torch.int32, device=dev)
    gms.multisplit2_histogram_even(torch.from_numpy(x).to(dev), n, h, lower, upper, inv, bins)
    np.testing.assert_array_equal(from_dev(h).view(np.int32),
                                  oracle.histogram_even_f32(x, bins, lower, inv))


@pytest.mark.parametrize("num_levels", [2, 3, 33, 65, 129, 257])
def test_histogram_range_float(dev, num_levels):
    import torch
    # main_histogram.cu:224-310 drives the range histograms with float samples
    # and edges; here uneven ascending edges over [-10, 300), samples spilling
    # out of range on both sides, samples exactly on edges, and NaN.
    g = W.rng(num_levels)
    levels = np.sort(g.random(num_levels, dtype=np.float32) * np.float32(310) - np.float32(10))
    levels[0], levels[-1] = np.float32(-10), np.float32(300)
    for n, off in ((0, 0), (1, 0), (4095, 1), (4096 * 3 + 7, 0), (1_000_003, 2)):
        banner(f"range f32 nl={num_levels} n={n} off={off}")
        x = (g.random(n, dtype=np.float32) * np.float32(340) - np.float32(25)).astype(np.float32)
        if n > 10:
            x[:num_levels if num_levels < n else n] = levels[:n]
            x[-1], x[-2] = np.float32("nan"), np.float32(300)
        xb = torch.from_numpy(np.concatenate([np.zeros(off, np.float32), x])).to(dev)
        h = torch.zeros(num_levels - 1, dtype=torch.int32, device=dev)
        gms.multisplit2_histogram_range(xb[off:], n, h, torch.from_numpy(levels).to(dev), num_levels)
        np.testing.assert_array_equal(from_dev(h).view(np.int32), oracle.histogram_range(x, levels))


@pytest.mark.parametrize("num_levels", [2, 17, 256, 257])
def test_histogram_range_uint(dev, num_levels):
    import torch
    g = W.rng(1000 + num_levels)
    levels = np.unique(g.integers(0, 1 << 32, num_levels * 2, dtype=np.uint64)).astype(np.uint32)
    levels = np.sort(levels[:num_levels])
    if num_levels == 257:
        levels[0], levels[-1] = 0, 0xFFFFFFFF
    for n in (0, 5, 123_457, (1 << 22) + 3):
        banner(f"range u32 nl={num_levels} n={n}")
        keys = W.uniform_keys(n, seed=n + num_levels)
        h = torch.zeros(num_levels - 1, dtype=torch.int32, device=dev)
        gms.multisplit2_histogram_range(dev_in(keys, dev), n, h, to_dev(levels, dev), num_levels)
        np.testing.assert_array_equal(from_dev(h).view(np.int32), oracle.histogram_range(keys, levels))


@pytest.mark.parametrize("case", range(13))
def test_histogram_range_edge_sets(dev, case):
    import torch
    sets, isets = range_level_sets()
    lv = (sets + isets)[case]
    for n_rand in (3 * 4096 + 123, (1 << 22) + 1):
        x = range_samples(lv, n_rand, seed=case)
        banner(f"range edge set {case} ({lv.dtype}, {lv.size} levels) n={x.size}")
        h = torch.zeros(lv.size - 1, dtype=torch.int32, device=dev)
        if lv.dtype == np.float32:
            gms.multisplit2_histogram_range(torch.from_numpy(x).to(dev), x.size, h,
                                            torch.from_numpy(lv).to(dev), lv.size)
        else:
            gms.multisplit2_histogram_range(to_dev(x, dev), x.size, h, to_dev(lv, dev), lv.size)
        np.testing.assert_array_equal(from_dev(h).view(np.int32), oracle.histogram_range(x, lv))


def test_histogram_range_rejects_bad_levels(dev):
    import torch
    x = torch.zeros(16, dtype=torch.float32, device=dev)
    h = torch.zeros(300, dtype=torch.int32, device=dev)
    lv = torch.zeros(300, dtype=torch.float32, device=dev)
    for nl in (0, 1, 258):
        with pytest.raises(gms.GmsError):
            gms.multisplit2_histogram_range(x, 16, h, lv, nl)


@pytest.mark.parametrize("selectivity", [0.5, 0.05])
@pytest.mark.parametrize("kv", [False, True])
def test_compaction_selectivity(dev, selectivity, kv):
    # predicate key < p * 2^32: bucket 0 = kept; built as a 2-bucket delta identifier
    n = (1 << 20) + 3
    keys = W.uniform_keys(n, seed=int(selectivity * 100) + kv)
    thr = int(selectivity * (1 << 32))
    f = gms.delta_func(thr)                       # key / thr: 0 below thr, clamped 1 above
    ctx = gms.compaction_context()
    (gms.compaction_allocate_key_value if kv else gms.compaction_allocate_key_only)(n, ctx)
    kin = to_dev(keys, dev)
    kout = filled_u32(n, dev)
    vals = np.arange(n, dtype=np.uint32)
    if kv:
        vout = filled_u32(n, dev)
        cnt = gms.compaction_key_value(kin, to_dev(vals, dev), kout, vout, n, ctx, f)
    else:
        cnt = gms.compaction_key_only(kin, kout, n, ctx, f)
    gms.compaction_release_memory(ctx)
    keep = keys < np.uint32(thr) if thr < (1 << 32) else np.ones(n, bool)
    order = np.concatenate([np.nonzero(keep)[0], np.nonzero(~keep)[0]])
    assert cnt == int(keep.sum())
    np.testing.assert_array_equal(from_dev(kout)[:n], keys[order])
    if kv:
        np.testing.assert_array_equal(from_dev(vout)[:n], vals[order])


def test_compaction_like_reference_driver_and_device_count(dev):
    import torch
    # main_compaction.cu: keys = permutation of 0..n-1, bucket = key & 1
    for n in (0, 1, 6145, 1 << 18):
        keys = W.permutation_keys(n, seed=n + 1)
        ctx = gms.compaction_context()
        gms.compaction_allocate_key_only(n, ctx)
        kin = dev_in(keys, dev)
        kout = filled_u32(n, dev)
        cnt = gms.compaction_key_only(kin, kout, n, ctx, gms.test_bucket(1))
        c2, wk, _ = oracle.compaction(keys, oracle.LOWBITS, 1)
        assert cnt == c2
        np.testing.assert_array_equal(from_dev(kout)[:n], wk)
        d_cnt = torch.full((1,), -1, dtype=torch.int32, device=dev)
        kout2 = filled_u32(n, dev)
        gms.compaction_device_count(kin, kout2, n, ctx, gms.test_bucket(1), d_cnt)
        torch.cuda.synchronize()
        assert int(d_cnt.item()) == c2
        np.testing.assert_array_equal(from_dev(kout2)[:n], wk)
        gms.compaction_release_memory(ctx)


@pytest.mark.slow
def test_histogram_256_low_byte_2e27(dev):
    """The histogram cell of the sweep at its stated size: 256 bins of the
    low 8 bits of 2^27 uniform keys, against the oracle's counting loop."""
    import torch
    n = 1 << 27
    banner(f"histogram 256 bins n={n}")
    g = torch.Generator(device=dev)
    g.manual_seed(227)
    kin = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=dev, generator=g)
    h = torch.zeros(256, dtype=torch.int32, device=dev)
    gms.histogram(kin, n, h, 256, gms.ms_bucket_sort_8bit(0))
    keys = from_dev(kin)
    np.testing.assert_array_equal(from_dev(h).astype(np.int64),
                                  oracle.histogram(keys, 256, oracle.DIGIT, 0).astype(np.int64))
    assert int(from_dev(h).astype(np.int64).sum()) == n


@pytest.mark.slow
@pytest.mark.parametrize("selectivity", [0.5, 0.05])
def test_compaction_2e27(dev, selectivity):
    """2-bucket compaction of 2^27 keys at the sweep's selectivities, key-only
    and key-value: the output is the stable partition (kept keys first, in
    input order, then the rest in input order) and the count is the
    partition point."""
    import torch
    n = 1 << 27
    thr = int(selectivity * (1 << 32))
    g = torch.Generator(device=dev)
    g.manual_seed(27 + int(selectivity * 100))
    kin = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=dev, generator=g)
    keep = (kin.to(torch.int64) & 0xFFFFFFFF) < thr
    order = torch.cat([torch.nonzero(keep).flatten(), torch.nonzero(~keep).flatten()])
    want_cnt = int(keep.sum().item())
    del keep
    for kv in (False, True):
        banner(f"compaction n={n} selectivity={selectivity} kv={kv}")
        ctx = gms.compaction_context()
        (gms.compaction_allocate_key_value if kv else gms.compaction_allocate_key_only)(n, ctx)
        kout = torch.empty_like(kin)
        if kv:
            vin = torch.arange(n, dtype=torch.int32, device=dev)
            vout = torch.empty_like(vin)
            cnt = gms.compaction_key_value(kin, vin, kout, vout, n, ctx, gms.delta_func(thr))
        else:
            cnt = gms.compaction_key_only(kin, kout, n, ctx, gms.delta_func(thr))
        torch.cuda.synchronize()
        gms.compaction_release_memory(ctx)
        assert cnt == want_cnt
        assert bool((kout == kin[order]).all())
        if kv:
            assert bool((vout.to(torch.int64) == order).all())
            del vin, vout
        del kout
