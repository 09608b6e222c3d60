"""GPU parity of the batched multisplit (gms_batch_*): every instance of a
batch must equal the oracle run on that instance alone, under every tuning,
through a captured hipGraph replayed several times, and after the autotuner
has chosen a tuning.  At the batched-workload shape (4096 key-value
instances, sizes log-uniform in [2^8, 2^22], 64 buckets) every instance is
compared with the plain multisplit of the same instance.
"""
import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W
from gpu_util import banner, filled_u32, from_dev, to_dev

pytestmark = pytest.mark.gpu

# (n, B, functor factory, oracle kind, oracle param, kv)
MIXED = [
    (0, 8, lambda: gms.range_bucket(8), oracle.RANGE, 0, True),
    (1, 2, lambda: gms.range_bucket(2), oracle.RANGE, 0, False),
    (255, 1, lambda: gms.range_bucket(1), oracle.RANGE, 0, True),
    (6143, 5, lambda: gms.range_bucket(5), oracle.RANGE, 0, True),
    (6144, 64, lambda: gms.range_bucket(64), oracle.RANGE, 0, True),
    (6145, 64, lambda: gms.range_bucket(64), oracle.RANGE, 0, False),
    (100_003, 256, lambda: gms.range_bucket(256), oracle.RANGE, 0, True),
    (70_000, 32, lambda: gms.delta_func((1 << 32) // 32 + 1), oracle.DELTA, (1 << 32) // 32 + 1, True),
    (50_000, 7, lambda: gms.delta_func(1000), oracle.DELTA, 1000, True),
    (65_536, 256, lambda: gms.ms_bucket_sort_8bit(16), oracle.DIGIT, 16, False),
    (33_333, 128, lambda: gms.ms_bucket_sort_7bit(25), oracle.DIGIT, 25, True),
    (20_000, 16, lambda: gms.identity_bucket(), oracle.IDENTITY, 0, True),
    (12_345, 2, lambda: gms.test_bucket(1), oracle.LOWBITS, 1, False),
    (300_001, 100, lambda: gms.range_bucket(100), oracle.RANGE, 0, True),
]


def _make(dev, cases, seed):
    items, want = [], []
    for i, (n, B, ff, kind, param, kv) in enumerate(cases):
        keys = W.uniform_keys(n, seed=seed + i)
        if kind == oracle.IDENTITY:
            keys = keys % np.uint32(B + 5)
        if kind == oracle.DELTA and param == 1000:
            keys = keys % np.uint32(9000)
        vals = (np.arange(n, dtype=np.uint32) * np.uint32(3) + np.uint32(i)).astype(np.uint32)
        it = gms.BatchItem(to_dev(keys, dev) if n else filled_u32(0, dev), filled_u32(n, dev), n, B,
                           ff(), d_value_in=(to_dev(vals, dev) if n else filled_u32(0, dev)) if kv else None,
                           d_value_out=filled_u32(n, dev) if kv else None,
                           d_offsets=filled_u32(B, dev, 0xFFFFFFFF))
        items.append(it)
        want.append(oracle.multisplit(keys, B, kind, param, vals if kv else None))
    return items, want


def _check(items, want, label):
    for i, (it, (wk, wv, wo)) in enumerate(zip(items, want)):
        n = it.n
        np.testing.assert_array_equal(from_dev(it.d_key_out)[:n], wk, err_msg=f"{label} inst {i} keys")
        np.testing.assert_array_equal(from_dev(it.d_offsets), wo, err_msg=f"{label} inst {i} offsets")
        if it.d_value_in is not None:
            np.testing.assert_array_equal(from_dev(it.d_value_out)[:n], wv,
                                          err_msg=f"{label} inst {i} values")


def _clear(items):
    for it in items:
        it.d_key_out.fill_(0x5A5A5A5A)
        it.d_offsets.fill_(-1)
        if it.d_value_out is not None:
            it.d_value_out.fill_(0x5A5A5A5A)


@pytest.mark.parametrize("tuning", [(8, 0), (8, 1), (16, 0), (16, 1), (24, 0), (24, 1), (32, 0), (32, 1)])
def test_batch_mixed_matches_oracle(dev, tuning):
    import torch
    banner(f"batch mixed tuning={tuning}")
    items, want = _make(dev, MIXED, seed=17)
    b = gms.Batch(items)
    b.tuning = {"items_per_thread": tuning[0], "persistent": tuning[1]}
    b.run()
    torch.cuda.synchronize()
    _check(items, want, f"run {tuning}")
    # run again on the same plan: same result
    _clear(items)
    b.run()
    torch.cuda.synchronize()
    _check(items, want, f"rerun {tuning}")


def test_batch_graph_replays_match_eager(dev):
    import torch
    banner("batch graph replay")
    items, want = _make(dev, MIXED, seed=99)
    b = gms.Batch(items)
    b.capture()
    for rep in range(4):
        _clear(items)                      # eager work between replays
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        _check(items, want, f"replay {rep}")


def test_batch_autotune_keeps_results(dev):
    import torch
    banner("batch autotune")
    items, want = _make(dev, MIXED, seed=5)
    b = gms.Batch(items)
    t = b.autotune(reps=2)
    assert t["items_per_thread"] in (8, 16, 24, 32) and t["persistent"] in (0, 1)
    assert t["ms_per_run"] > 0
    assert b.tuning == {"items_per_thread": t["items_per_thread"], "persistent": t["persistent"]}
    _clear(items)
    b.run()
    torch.cuda.synchronize()
    _check(items, want, "autotuned")
    st = b.stats()
    assert st["keys"] == sum(c[0] for c in MIXED)
    groups = len({(max(1, int(np.ceil(np.log2(c[1])))), c[5]) for c in MIXED})
    assert st["launches"] == t["persistent"] + 3 + 2 * groups


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_unaligned_instances(dev, persistent):
    """Instances whose keys/values start off a 16-byte boundary take their
    own group (scalar loads) next to aligned ones."""
    import torch
    banner(f"batch unaligned persistent={persistent}")
    items, want = [], []
    for i, (n, off) in enumerate([(40_000, 1), (6144 * 3, 0), (25_000, 3), (7, 2), (6145, 0)]):
        keys = W.uniform_keys(n, seed=300 + i)
        vals = np.arange(n, dtype=np.uint32)
        kbuf = to_dev(np.concatenate([np.zeros(off, np.uint32), keys]), dev)
        vbuf = to_dev(np.concatenate([np.zeros(off, np.uint32), vals]), dev)
        kout = filled_u32(n + off, dev)
        vout = filled_u32(n + off, dev)
        it = gms.BatchItem(kbuf[off:], kout[off:], n, 64, gms.range_bucket(64), d_value_in=vbuf[off:],
                           d_value_out=vout[off:], d_offsets=filled_u32(64, dev, 0xFFFFFFFF))
        items.append(it)
        want.append(oracle.multisplit(keys, 64, oracle.RANGE, 0, vals))
    b = gms.Batch(items)
    b.tuning = {"items_per_thread": 24, "persistent": persistent}
    b.run()
    torch.cuda.synchronize()
    _check(items, want, "unaligned")


def test_batch_validation(dev):
    a = filled_u32(64, dev)
    o = filled_u32(64, dev)
    with pytest.raises(gms.GmsError):          # 300 buckets
        gms.Batch([gms.BatchItem(a, o, 64, 300, gms.range_bucket(300))])
    x = filled_u32(200, dev)
    with pytest.raises(gms.GmsError):          # outputs overlap
        gms.Batch([gms.BatchItem(a, x[:64], 64, 4), gms.BatchItem(a, x[32:96], 64, 4)])
    with pytest.raises(gms.GmsError):          # output overlaps another item's input
        gms.Batch([gms.BatchItem(x[:64], o, 64, 4), gms.BatchItem(a, x[32:96], 64, 4)])
    gms.Batch([gms.BatchItem(a, x[:64], 64, 4), gms.BatchItem(a, x[64:128], 64, 4)])  # shared input ok


# (n, digit_bits, kv) of the radix sorts planned beside multisplits
SORTS = [(0, 8, True), (1, 4, False), (100_003, 8, True), (70_001, 5, False), (6144, 6, True),
         (300_000, 7, True), (1 << 20, 4, True), (50_000, 8, False)]


def _make_sorts(dev, seed):
    sorts, want, inputs = [], [], []
    for i, (n, bits, kv) in enumerate(SORTS):
        keys = W.uniform_keys(n, seed=seed + i)
        if i == 3:
            keys = keys & np.uint32(0xFFF)       # few distinct keys: stability across passes
        vals = (np.arange(n, dtype=np.uint32) * np.uint32(5) + np.uint32(i)).astype(np.uint32)
        kin = to_dev(keys, dev) if n else filled_u32(0, dev)
        vin = (to_dev(vals, dev) if n else filled_u32(0, dev)) if kv else None
        s = gms.BatchSortItem(kin, filled_u32(n, dev), n, bits, d_value_in=vin,
                              d_value_out=filled_u32(n, dev) if kv else None)
        sorts.append(s)
        inputs.append((keys, vals if kv else None))
        want.append(oracle.sort(keys, bits, vals if kv else None))
    return sorts, want, inputs


def _check_sorts(sorts, want, inputs, label):
    for i, (s, (wk, wv), (keys, vals)) in enumerate(zip(sorts, want, inputs)):
        n = s.n
        np.testing.assert_array_equal(from_dev(s.d_key_in)[:n], keys, err_msg=f"{label} sort {i} input")
        np.testing.assert_array_equal(from_dev(s.d_key_out)[:n], wk, err_msg=f"{label} sort {i} keys")
        if vals is not None:
            np.testing.assert_array_equal(from_dev(s.d_value_in)[:n], vals,
                                          err_msg=f"{label} sort {i} value input")
            np.testing.assert_array_equal(from_dev(s.d_value_out)[:n], wv,
                                          err_msg=f"{label} sort {i} values")


def _clear_sorts(sorts):
    for s in sorts:
        s.d_key_out.fill_(0x5A5A5A5A)
        if s.d_value_out is not None:
            s.d_value_out.fill_(0x5A5A5A5A)


@pytest.mark.parametrize("tuning", [(8, 0), (16, 1), (24, 0), (24, 1), (32, 0)])
def test_batch_sorts_and_multisplits_match_oracle(dev, tuning):
    """Sorts as passes of the batch beside multisplits: every sort equals the
    oracle's sort with the same digit schedule, every multisplit the oracle's
    multisplit, inputs unchanged; a second run on the same plan agrees."""
    import torch
    banner(f"batch sorts tuning={tuning}")
    items, want = _make(dev, MIXED, seed=23)
    sorts, swant, inputs = _make_sorts(dev, seed=41)
    b = gms.Batch(items, sorts=sorts)
    b.tuning = {"items_per_thread": tuning[0], "persistent": tuning[1]}
    b.run()
    torch.cuda.synchronize()
    _check(items, want, f"run {tuning}")
    _check_sorts(sorts, swant, inputs, f"run {tuning}")
    _clear(items)
    _clear_sorts(sorts)
    b.run()
    torch.cuda.synchronize()
    _check(items, want, f"rerun {tuning}")
    _check_sorts(sorts, swant, inputs, f"rerun {tuning}")
    passes = sum(len(oracle.sort_schedule(bits)) * n for n, bits, _ in SORTS)
    assert b.stats()["keys"] == passes + sum(c[0] for c in MIXED)


def test_batch_sorts_only_graph_replays_and_autotune(dev):
    """A batch of sorts only: graph replays with eager work between them, and
    the autotuner's choice, each equal the oracle."""
    import torch
    banner("batch sorts only: graph + autotune")
    sorts, swant, inputs = _make_sorts(dev, seed=3)
    b = gms.Batch([], sorts=sorts)
    b.capture()
    for rep in range(3):
        _clear_sorts(sorts)
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        _check_sorts(sorts, swant, inputs, f"replay {rep}")
    t = b.autotune(reps=1)
    assert t["ms_per_run"] > 0
    _clear_sorts(sorts)
    b.run()
    torch.cuda.synchronize()
    _check_sorts(sorts, swant, inputs, "autotuned")


def test_batch_sort_validation(dev):
    a = filled_u32(64, dev)
    o = filled_u32(64, dev)
    with pytest.raises(gms.GmsError):          # 3-bit digits
        gms.Batch([], sorts=[gms.BatchSortItem(a, o, 64, 3)])
    with pytest.raises(gms.GmsError):          # sort output overlaps a multisplit's input
        gms.Batch([gms.BatchItem(a, o, 64, 4)], sorts=[gms.BatchSortItem(o, a, 64, 8)])


@pytest.mark.slow
def test_batch_4096_loguniform_equals_plain_calls(dev):
    """The batched workload: 4096 key-value multisplits, sizes round(2^U),
    U ~ U[8, 22], 64 buckets by range, uniform keys.  Each instance equals
    the plain multisplit of that instance (itself checked against the oracle
    elsewhere), and a sample of instances equals the oracle directly."""
    import torch
    sizes = W.log_uniform_sizes(4096, 8, 22, seed=4096)
    total = int(sizes.sum())
    banner(f"batch 4096 instances, {total} keys")
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    kin = torch.randint(-(1 << 31), (1 << 31) - 1, (total,), dtype=torch.int32, device=dev, generator=g)
    vin = torch.arange(total, dtype=torch.int32, device=dev)
    kout = torch.empty_like(kin)
    vout = torch.empty_like(kin)
    offs = torch.empty(4096 * 64, dtype=torch.int32, device=dev)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    items = [gms.BatchItem(kin[s:s + n], kout[s:s + n], int(n), 64, gms.range_bucket(64),
                           d_value_in=vin[s:s + n], d_value_out=vout[s:s + n],
                           d_offsets=offs[64 * i:64 * i + 64])
             for i, (s, n) in enumerate(zip(starts, sizes))]
    b = gms.Batch(items)
    b.autotune(reps=1)
    b.capture()
    b.replay()
    torch.cuda.synchronize()
    ctx = gms.multisplit_context(64)
    gms.multisplit_allocate_key_value(int(sizes.max()), ctx)
    pk = torch.empty(int(sizes.max()), dtype=torch.int32, device=dev)
    pv = torch.empty_like(pk)
    po = torch.empty(64, dtype=torch.int32, device=dev)
    bad = 0
    for i, (s, n) in enumerate(zip(starts, sizes)):
        n = int(n)
        gms.multisplit(kin[s:s + n], pk, n, ctx, gms.range_bucket(64), d_value_in=vin[s:s + n],
                       d_value_out=pv, d_offsets=po)
        ok = (torch.equal(pk[:n], kout[s:s + n]) and torch.equal(pv[:n], vout[s:s + n])
              and torch.equal(po, offs[64 * i:64 * i + 64]))
        bad += 0 if ok else 1
    assert bad == 0, f"{bad} instances differ from the plain call"
    for i in (0, 1, 2047, 4095):
        s, n = int(starts[i]), int(sizes[i])
        keys = from_dev(kin[s:s + n])
        wk, wv, wo = oracle.multisplit(keys, 64, oracle.RANGE, 0, np.arange(s, s + n, dtype=np.uint32))
        np.testing.assert_array_equal(from_dev(kout[s:s + n]), wk)
        np.testing.assert_array_equal(from_dev(vout[s:s + n]), wv)
    gms.multisplit_release_memory(ctx)


# ---------------------------------------------------------------------------
# 64-bit elements in the batch: wide multisplits of every shape and 64-bit
# sorts beside 32-bit items; every instance against the oracle's 64-bit
# multisplit / sort.
# ---------------------------------------------------------------------------
def _u64_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to(dev)


def _from_dev64(t):
    return t.cpu().numpy().view(np.uint64)


# (n, B, key bytes, value bytes (0 = key-only), kind, param, width)
WIDE = [(0, 8, 8, 8, oracle.RANGE, 0, 0), (1, 2, 8, 0, oracle.RANGE, 0, 0),
        (300_001, 256, 8, 8, oracle.RANGE, 0, 0), (70_000, 64, 8, 4, oracle.RANGE, 0, 0),
        (123_457, 17, 4, 8, oracle.RANGE, 0, 0), (1 << 20, 16, 8, 0, oracle.DIGIT, 40, 4),
        (50_000, 100, 8, 8, oracle.DELTA, (1 << 64) // 100 + 7, 0),
        (4096 * 3, 128, 8, 4, oracle.IDENTITY, 0, 0)]
# (n, digit bits, key bytes, value bytes (0 = keys only))
WIDE_SORTS = [(200_000, 8, 8, 8), (65_537, 6, 8, 0), (30_000, 5, 8, 4), (100_000, 8, 4, 8)]


def _rng64(n, seed):
    return np.random.default_rng(seed).integers(0, 2**64, size=n, dtype=np.uint64)


def _make_wide(dev, seed):
    import torch
    items, want = [], []
    for i, (n, B, kb, vb, kind, param, width) in enumerate(WIDE):
        keys = _rng64(n, seed + i) if kb == 8 else W.uniform_keys(n, seed=seed + i)
        if kind == oracle.IDENTITY:
            keys = keys % np.uint64(B + 9)             # some ids past B-1: clamped
        vals = None if vb == 0 else (_rng64(n, seed + 50 + i) if vb == 8
                                     else np.arange(n, dtype=np.uint32))
        kdt = torch.int64 if kb == 8 else torch.int32
        kin = (_u64_dev(keys, dev) if kb == 8 else to_dev(keys, dev)) if n else \
            torch.zeros(1, dtype=kdt, device=dev)
        kout = torch.full((max(n, 1),), 0x5A, dtype=kdt, device=dev)
        vin = vout = None
        if vb:
            vdt = torch.int64 if vb == 8 else torch.int32
            vin = (_u64_dev(vals, dev) if vb == 8 else to_dev(vals, dev)) if n else \
                torch.zeros(1, dtype=vdt, device=dev)
            vout = torch.full((max(n, 1),), 0x5A, dtype=vdt, device=dev)
        f = gms.ms_bucket_sort_4bit(param) if kind == oracle.DIGIT else \
            gms.delta_func(param) if kind == oracle.DELTA else \
            gms.identity_bucket() if kind == oracle.IDENTITY else gms.range_bucket(B)
        items.append(gms.BatchWideItem(kin, kout, n, B, f, d_value_in=vin, d_value_out=vout,
                                       d_offsets=filled_u32(B, dev, 0xFFFFFFFF)))
        want.append(oracle.multisplit_wide(keys, B, kind, param, vals, width))
    return items, want


def _check_wide(items, want, label):
    for i, (it, (wk, wv, wo)) in enumerate(zip(items, want)):
        n = it.n
        ko = _from_dev64(it.d_key_out) if it.key_bytes == 8 else from_dev(it.d_key_out)
        np.testing.assert_array_equal(ko[:n], wk, err_msg=f"{label} wide {i} keys")
        np.testing.assert_array_equal(from_dev(it.d_offsets), wo, err_msg=f"{label} wide {i} offsets")
        if it.d_value_out is not None:
            vo = _from_dev64(it.d_value_out) if it.value_bytes == 8 else from_dev(it.d_value_out)
            np.testing.assert_array_equal(vo[:n], wv, err_msg=f"{label} wide {i} values")


def _make_wide_sorts(dev, seed):
    import torch
    sorts, want = [], []
    for i, (n, bits, kb, vb) in enumerate(WIDE_SORTS):
        keys = _rng64(n, seed + i) if kb == 8 else W.uniform_keys(n, seed=seed + i)
        if i == 1:
            keys &= np.uint64(0xFFFF0000000000FF)          # equal digits in most passes
        vals = None if vb == 0 else (_rng64(n, seed + 40 + i) if vb == 8
                                     else np.arange(n, dtype=np.uint32))
        kin = _u64_dev(keys, dev) if kb == 8 else to_dev(keys, dev)
        kout = torch.full_like(kin, 0x5A)
        vin = vout = None
        if vb:
            vin = _u64_dev(vals, dev) if vb == 8 else to_dev(vals, dev)
            vout = torch.full_like(vin, 0x5A)
        sorts.append(gms.BatchSortItem(kin, kout, n, bits, d_value_in=vin, d_value_out=vout))
        if kb == 8:
            want.append(oracle.sort_wide(keys, bits, vals))
        else:   # 32-bit keys, 64-bit values: the stable order applied to the values
            wk, perm = oracle.sort(keys, bits, np.arange(n, dtype=np.uint32))
            want.append((wk, vals[perm]))
    return sorts, want


def _check_wide_sorts(sorts, want, label):
    for i, (s, (wk, wv)) in enumerate(zip(sorts, want)):
        n = s.n
        ko = _from_dev64(s.d_key_out) if s.key_bytes == 8 else from_dev(s.d_key_out)
        np.testing.assert_array_equal(ko[:n], wk, err_msg=f"{label} sort {i} keys")
        if s.d_value_out is not None:
            vo = _from_dev64(s.d_value_out) if s.value_bytes == 8 else from_dev(s.d_value_out)
            np.testing.assert_array_equal(vo[:n], wv, err_msg=f"{label} sort {i} values")


@pytest.mark.parametrize("tuning", [(24, 0), (16, 1)])
def test_batch_wide_items_and_sorts_match_oracle(dev, tuning):
    """64-bit multisplits and sorts beside 32-bit multisplits and sorts in one
    batch: every instance equals the oracle; again after a rerun, through
    graph replays with eager work between them, and after autotuning."""
    import torch
    banner(f"batch wide tuning={tuning}")
    items, want = _make(dev, MIXED[:8], seed=61)
    wide, wwant = _make_wide(dev, seed=71)
    wsorts, wswant = _make_wide_sorts(dev, seed=81)
    sorts, swant, inputs = _make_sorts(dev, seed=91)
    b = gms.Batch(items, sorts=sorts + wsorts, wide=wide)
    b.tuning = {"items_per_thread": tuning[0], "persistent": tuning[1]}
    b.run()
    torch.cuda.synchronize()
    _check(items, want, "run")
    _check_wide(wide, wwant, "run")
    _check_wide_sorts(wsorts, wswant, "run")
    _check_sorts(sorts, swant, inputs, "run")
    b.capture()
    for rep in range(2):
        _clear(items)
        _clear_sorts(sorts + wsorts)
        for it in wide:
            it.d_key_out.fill_(0x5A)
            if it.d_value_out is not None:
                it.d_value_out.fill_(0x5A)
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        _check(items, want, f"replay {rep}")
        _check_wide(wide, wwant, f"replay {rep}")
        _check_wide_sorts(wsorts, wswant, f"replay {rep}")
    b.autotune(reps=1)
    b.run()
    torch.cuda.synchronize()
    _check_wide(wide, wwant, "autotuned")
    _check_wide_sorts(wsorts, wswant, "autotuned")


def test_batch_wide_validation(dev):
    import torch
    a = torch.zeros(64, dtype=torch.int64, device=dev)
    o = torch.zeros(64, dtype=torch.int64, device=dev)
    with pytest.raises(gms.GmsError):          # 257 buckets
        gms.Batch([], wide=[gms.BatchWideItem(a, o, 64, 257)])
    with pytest.raises(gms.GmsError):          # output overlaps the input
        gms.Batch([], wide=[gms.BatchWideItem(a[:32], a[16:48], 32, 8)])
    with pytest.raises(gms.GmsError):          # 3-bit digits
        gms.Batch([], sorts=[gms.BatchSortItem(a, o, 64, 3)])
