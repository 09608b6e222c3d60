"""GPU parity of multisplit (up to 2^20 buckets), radix sort and compaction
with 64-bit keys and/or 64-bit values (the reference's KeyT/ValueT templates,
src/api/multisplit.cuh:81-313, multisplit_sort.cuh:366-527,
multisplit_sort_pairs.cuh:67-228, multisplit2_compaction.cuh:525-573), bit-exact
against the oracle's 64-bit restatement (itself pinned against Python
integers and numpy's stable argsort) and numpy's stable sort."""
import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpu_util import banner

pytestmark = pytest.mark.gpu


def to_dev_any(a, dev):
    import torch
    a = np.ascontiguousarray(a)
    view = {np.dtype(np.uint64): np.int64, np.dtype(np.uint32): np.int32}.get(a.dtype)
    return torch.from_numpy(a.view(view) if view else a).to(dev)


def from_dev_any(t, dtype):
    return t.cpu().numpy().view(dtype)


def out_like(a, dev):
    import torch
    dt = torch.int64 if a.dtype.itemsize == 8 else torch.int32
    return torch.full((max(a.size, 1),), 0x5A, dtype=dt, device=dev)


def keys_of(n, seed, kbytes):
    g = np.random.default_rng(seed)
    if kbytes == 8:
        k = g.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
        k[::5] &= np.uint64(0xFF)
        return k
    return g.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)


def vals_of(n, vbytes):
    if not vbytes:
        return None
    return (np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)).astype(
        np.uint64 if vbytes == 8 else np.uint32)


class _Ident64(gms._Bucket):
    def __init__(self, kind, param, width):
        self.kind, self.param, self.width = kind, param, width


IDENTS = [  # (B, kind, param, width)
    (256, gms.BUCKET_RANGE, 0, 0), (2, gms.BUCKET_RANGE, 0, 0), (3, gms.BUCKET_RANGE, 0, 0),
    (100, gms.BUCKET_DELTA, (1 << 57) + 5, 0), (32, gms.BUCKET_DIGIT, 40, 5),
    (256, gms.BUCKET_DIGIT, 56, 8), (17, gms.BUCKET_IDENTITY, 0, 0), (8, gms.BUCKET_LOWBITS, 3, 0),
]


def run_wide(dev, keys, vals, B, f, off=0):
    n = keys.size
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(max(n, 1), ctx)
    pad_k = np.concatenate([np.zeros(off, keys.dtype), keys])
    kin = to_dev_any(pad_k, dev)[off:] if n else out_like(keys, dev)
    kout = out_like(keys, dev)
    offs = np.zeros(B, np.uint32)
    if vals is None:
        gms.multisplit_key_only(kin, kout, n, ctx, f, offs)
        res = from_dev_any(kout, keys.dtype)[:n], None, offs
    else:
        pad_v = np.concatenate([np.zeros(off, vals.dtype), vals])
        vin = to_dev_any(pad_v, dev)[off:] if n else out_like(vals, dev)
        vout = out_like(vals, dev)
        gms.multisplit_key_value(kin, vin, kout, vout, n, ctx, f, offs)
        res = from_dev_any(kout, keys.dtype)[:n], from_dev_any(vout, vals.dtype)[:n], offs
    gms.multisplit_release_memory(ctx)
    return res


@pytest.mark.parametrize("case", range(len(IDENTS)))
@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 4), (8, 8), (4, 8)])
def test_multisplit_wide_matches_oracle(dev, case, kbytes, vbytes):
    B, kind, param, width = IDENTS[case]
    if kbytes == 4:
        param = {gms.BUCKET_DELTA: 1000, gms.BUCKET_DIGIT: 32 - max(width, 1)}.get(kind, param)
    f = _Ident64(kind, param, width)
    for n in (0, 1, 3073, 123_457, (1 << 20) + 7):
        banner(f"wide k{kbytes} v{vbytes} B={B} kind={kind} n={n}")
        keys = keys_of(n, seed=n + case, kbytes=kbytes)
        vals = vals_of(n, vbytes)
        got = run_wide(dev, keys, vals, B, f)
        wk, wv, wo = oracle.multisplit_wide(keys, B, kind, param, vals, width)
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[2], wo)
        if vbytes:
            np.testing.assert_array_equal(got[1], wv)


@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 8), (4, 8)])
def test_multisplit_wide_8_byte_aligned_inputs(dev, kbytes, vbytes):
    # inputs one element past a 16-byte boundary take the unvectorised loads
    n = 50_001
    keys, vals = keys_of(n, 5, kbytes), vals_of(n, vbytes)
    got = run_wide(dev, keys, vals, 64, gms.range_bucket(64), off=1)
    wk, wv, wo = oracle.multisplit_wide(keys, 64, oracle.RANGE, 0, vals)
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[2], wo)
    if vbytes:
        np.testing.assert_array_equal(got[1], wv)


def test_multisplit_wide_device_offsets_async(dev):
    import torch
    n, B = 300_000, 200
    keys = keys_of(n, 11, 8)
    vals = vals_of(n, 8)
    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(n, ctx)
    kout, vout = out_like(keys, dev), out_like(vals, dev)
    d_off = torch.zeros(B, dtype=torch.int32, device=dev)
    gms.multisplit(to_dev_any(keys, dev), kout, n, ctx, gms.delta_func(((1 << 64) - 1) // 150),
                   to_dev_any(vals, dev), vout, d_off)
    torch.cuda.synchronize()
    wk, wv, wo = oracle.multisplit_wide(keys, B, oracle.DELTA, ((1 << 64) - 1) // 150, vals)
    np.testing.assert_array_equal(from_dev_any(kout, np.uint64), wk)
    np.testing.assert_array_equal(from_dev_any(vout, np.uint64), wv)
    np.testing.assert_array_equal(d_off.cpu().numpy().view(np.uint32), wo)


@pytest.mark.parametrize("bits", [8, 6, 4])
@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 8), (8, 4), (4, 8)])
def test_sort_wide_matches_numpy(dev, bits, kbytes, vbytes):
    n = (1 << 20) + 3
    banner(f"sort wide bits={bits} k{kbytes} v{vbytes}")
    keys = keys_of(n, bits, kbytes)
    keys[::13] = keys[0]  # duplicates expose instability
    vals = None if not vbytes else np.arange(n, dtype=np.uint64 if vbytes == 8 else np.uint32)
    ctx = gms.ms_sort_context()
    alloc, alloc_pairs, sort, sort_pairs = {
        8: (gms.multisplit_sort_8bit_allocate, gms.multisplit_sort_8bit_allocate_pairs,
            gms.multisplit_sort_8bit, gms.multisplit_sort_8bit_pairs),
        6: (gms.multisplit_sort_6bit_allocate, gms.multisplit_sort_6bit_allocate_pairs,
            gms.multisplit_sort_6bit, gms.multisplit_sort_6bit_pairs),
        4: (gms.multisplit_sort_4bit_allocate, gms.multisplit_sort_4bit_allocate_pairs,
            gms.multisplit_sort_4bit, gms.multisplit_sort_4bit_pairs)}[bits]
    kin, kout = to_dev_any(keys, dev), out_like(keys, dev)
    if vals is None:
        alloc(n, ctx, key_bytes=kbytes)
        sort(kin, kout, n, ctx)
    else:
        alloc_pairs(n, ctx, key_bytes=kbytes, value_bytes=vbytes)
        vout = out_like(vals, dev)
        sort_pairs(kin, kout, to_dev_any(vals, dev), vout, n, ctx)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(from_dev_any(kout, keys.dtype)[:n], keys[order])
    np.testing.assert_array_equal(from_dev_any(kin, keys.dtype), keys)  # input not written
    if vals is not None:
        np.testing.assert_array_equal(from_dev_any(vout, vals.dtype)[:n], vals[order])
    gms.multisplit_sort_release_memory(ctx)


def test_wide_rejects_what_it_cannot_do(dev):
    import torch
    k = torch.zeros(16, dtype=torch.int64, device=dev)
    ctx2 = gms.multisplit_context(256)
    gms.multisplit_allocate_key_only(16, ctx2)
    with pytest.raises(gms.GmsError):  # overlapping input and output
        gms.multisplit_key_only(k, k, 16, ctx2, gms.range_bucket(256))
    with pytest.raises(gms.GmsError):  # digit shift past 64 bits
        gms.multisplit_key_only(k, torch.zeros_like(k), 16, ctx2, gms.ms_bucket_sort_8bit(64))
    s = gms.ms_sort_context()
    gms.multisplit_sort_8bit_allocate(16, s)  # 32-bit temporaries
    with pytest.raises(gms.GmsError):
        gms.multisplit_sort_8bit(k, torch.zeros_like(k), 16, s)


@pytest.mark.slow
def test_multisplit_wide_kv_256_2e26(dev):
    n, B = 1 << 26, 256
    banner(f"wide kv 64/64 B={B} n={n}")
    keys = keys_of(n, 26, 8)
    vals = np.arange(n, dtype=np.uint64)
    got = run_wide(dev, keys, vals, B, gms.range_bucket(B))
    # size-independent checks: bucket ids non-decreasing, starts match the
    # counts, every value a permutation index, keys follow their values
    ids = (keys[got[1].astype(np.int64)] >> np.uint64(56)).astype(np.int64)
    np.testing.assert_array_equal(keys[got[1].astype(np.int64)], got[0])
    assert np.all(np.diff(ids) >= 0)
    assert np.array_equal(np.sort(got[1]), vals)
    np.testing.assert_array_equal(got[2], np.concatenate([[0], np.cumsum(np.bincount(ids, minlength=B))[:-1]]))
    # stability: within every bucket the original indices increase
    same = np.diff(ids) == 0
    assert np.all(np.diff(got[1].astype(np.int64))[same] > 0)


@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 4), (8, 8), (4, 8)])
def test_multisplit_wide_more_than_256_buckets(dev, kbytes, vbytes):
    for B, kind, param, width in [(257, gms.BUCKET_RANGE, 0, 0), (5000, gms.BUCKET_DIGIT, 30, 12),
                                  (1 << 20, gms.BUCKET_RANGE, 0, 0),
                                  (70000, gms.BUCKET_DELTA,
                                   ((1 << 64) - 1) // 65000 if kbytes == 8 else 60000, 0)]:
        if kbytes == 4 and kind == gms.BUCKET_DIGIT:
            param = 8
        for n in (0, 1, 300_007):
            banner(f"wide large k{kbytes} v{vbytes} B={B} n={n}")
            keys = keys_of(n, seed=B + n, kbytes=kbytes)
            vals = vals_of(n, vbytes)
            got = run_wide(dev, keys, vals, B, _Ident64(kind, param, width))
            wk, wv, wo = oracle.multisplit_wide(keys, B, kind, param, vals, width)
            np.testing.assert_array_equal(got[0], wk)
            np.testing.assert_array_equal(got[2], wo)
            if vbytes:
                np.testing.assert_array_equal(got[1], wv)


@pytest.mark.parametrize("case", range(len(IDENTS)))
def test_histogram_wide(dev, case):
    # 64-bit keys through the functor histogram; the counts are the
    # differences of the oracle's bucket starts, and the histogram accumulates.
    import torch
    B, kind, param, width = IDENTS[case]
    f = _Ident64(kind, param, width)
    for n, off in ((0, 0), (1, 0), (3073, 1), (1_000_003, 0), (1_000_003, 1), ((1 << 24) + 5, 0)):
        banner(f"histogram wide B={B} kind={kind} n={n} offset={off}")
        keys = keys_of(n, seed=n + 7 * case, kbytes=8)
        hist = torch.full((B,), 3, dtype=torch.int32, device=dev)
        # off = 1: an 8- but not 16-byte-aligned key array
        kin = to_dev_any(np.concatenate([np.zeros(off, np.uint64), keys]), dev)[off:]
        gms.histogram(kin if n else out_like(keys, dev), n, hist, B, f)
        _, _, wo = oracle.multisplit_wide(keys, B, kind, param, None, width)
        want = np.diff(np.append(wo.astype(np.int64), n))
        np.testing.assert_array_equal(hist.cpu().numpy().astype(np.int64) - 3, want)


@pytest.mark.slow
def test_histogram_wide_full_size(dev):
    # 2^28 + 5 uint64 keys (2 GiB) from torch's generator: the 256-bin range
    # histogram is the count of each top byte, counted here by torch.
    import torch
    n = (1 << 28) + 5
    banner(f"histogram wide full size n={n}")
    g = torch.Generator(device=dev)
    g.manual_seed(28)
    k = torch.randint(-(1 << 62), 1 << 62, (n,), dtype=torch.int64, device=dev, generator=g) * 2
    k[::7] = -1  # 2^64 - 1: the last bin
    hist = torch.zeros(256, dtype=torch.int32, device=dev)
    gms.histogram(k, n, hist, 256, gms.range_bucket(256))
    want = torch.bincount((k >> 56) & 0xFF, minlength=256)
    assert int(hist.to(torch.int64).sum()) == n
    assert bool((hist.to(torch.int64) == want).all())


@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 8), (4, 8)])
def test_compaction_wide(dev, kbytes, vbytes):
    import torch
    n = 1_000_003
    keys, vals = keys_of(n, 21, kbytes), vals_of(n, vbytes)
    ctx = gms.compaction_context()
    gms.compaction_allocate_key_value(n, ctx)
    kin, kout = to_dev_any(keys, dev), out_like(keys, dev)
    f = gms.test_bucket(1)
    if vals is None:
        cnt = gms.compaction_key_only(kin, kout, n, ctx, f)
    else:
        vout = out_like(vals, dev)
        cnt = gms.compaction_key_value(kin, to_dev_any(vals, dev), kout, vout, n, ctx, f)
    wk, wv, wo = oracle.multisplit_wide(keys, 2, oracle.LOWBITS, 1, vals)
    assert cnt == wo[1] == int(np.sum((keys & 1) == 0))
    np.testing.assert_array_equal(from_dev_any(kout, keys.dtype)[:n], wk)
    if vals is not None:
        np.testing.assert_array_equal(from_dev_any(vout, vals.dtype)[:n], wv)
    d_cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    gms.compaction_device_count(kin, kout, n, ctx, f, d_cnt)
    torch.cuda.synchronize()
    assert int(d_cnt.item()) == cnt
