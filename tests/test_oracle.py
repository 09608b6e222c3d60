"""Pins the CPU oracle against independent computations (CPU only).

Each check uses a different algorithm from the oracle's count/scan/scatter
(numpy stable argsort, bincount, Python big-integer arithmetic, brute-force
closed forms), chosen so that a dropped term, a wrong index, an off-by-one in
the scan or an unstable scatter fails at least one of them.
"""
import numpy as np
import pytest

import oracle
from gpumultisplit_amd import workloads as W

U32 = 1 << 32


def py_bucket(kind, B, param, k):
    """Bucket identifiers in Python big-integer arithmetic."""
    if kind == oracle.RANGE:
        b = (k * B) >> 32
    elif kind == oracle.DELTA:
        b = k // param
    elif kind == oracle.DIGIT:
        b = (k >> param) % B
    elif kind == oracle.IDENTITY:
        b = k
    elif kind == oracle.LOWBITS:
        b = k % (1 << param)
    else:
        raise ValueError
    return min(b, B - 1)


EDGE_KEYS = np.array([0, 1, 2, 3, 255, 256, 65535, 65536, 0x7FFFFFFF, 0x80000000,
                      0x80000001, 0xFFFFFFFE, 0xFFFFFFFF, 0x12345678, 0xDEADBEEF], dtype=np.uint32)


@pytest.mark.parametrize("kind,B,param", [
    (oracle.RANGE, 2, 0), (oracle.RANGE, 3, 0), (oracle.RANGE, 8, 0), (oracle.RANGE, 100, 0),
    (oracle.RANGE, 256, 0), (oracle.DELTA, 32, (1 << 25) // 32), (oracle.DELTA, 7, 3),
    (oracle.DELTA, 256, 0x1000001), (oracle.DIGIT, 256, 0), (oracle.DIGIT, 256, 24), (oracle.DIGIT, 128, 25),
    (oracle.DIGIT, 16, 4), (oracle.IDENTITY, 64, 0), (oracle.LOWBITS, 2, 1), (oracle.LOWBITS, 8, 3),
])
def test_bucket_ids_match_python_integers(kind, B, param):
    keys = np.concatenate([EDGE_KEYS, W.uniform_keys(2000, seed=B * 7 + kind)])
    if kind == oracle.DELTA:
        d = param
        extra = [m * d + o for m in range(0, 300) for o in (-1, 0, 1) if 0 <= m * d + o < U32]
        keys = np.concatenate([keys, np.array(extra, dtype=np.uint64).astype(np.uint32)])
    got = oracle.bucket_ids(keys, B, kind, param)
    want = np.array([py_bucket(kind, B, param, int(k)) for k in keys], dtype=np.uint32)
    np.testing.assert_array_equal(got, want)


def brute_force_positions(ids, B):
    """pos(i) = sum_{c < b_i} count(c) + #{j < i : b_j = b_i}, by direct counting."""
    n = len(ids)
    pos = []
    for i in range(n):
        below = sum(1 for j in range(n) if ids[j] < ids[i])
        earlier_same = sum(1 for j in range(i) if ids[j] == ids[i])
        pos.append(below + earlier_same)
    return pos


@pytest.mark.parametrize("B", [1, 2, 3, 5, 8, 33, 256])
@pytest.mark.parametrize("n", [0, 1, 2, 7, 64, 97])
def test_multisplit_closed_form_brute_force(B, n):
    keys = W.uniform_keys(n, seed=1000 + n * 3 + B)
    vals = np.arange(n, dtype=np.uint32)
    kout, vout, offs = oracle.multisplit(keys, B, oracle.RANGE, 0, vals)
    ids = [py_bucket(oracle.RANGE, B, 0, int(k)) for k in keys]
    pos = brute_force_positions(ids, B)
    want_k = np.zeros(n, dtype=np.uint32)
    want_v = np.zeros(n, dtype=np.uint32)
    for i, p in enumerate(pos):
        want_k[p] = keys[i]
        want_v[p] = i
    np.testing.assert_array_equal(kout, want_k)
    np.testing.assert_array_equal(vout, want_v)
    want_offs = [sum(1 for b in ids if b < c) for c in range(B)]
    np.testing.assert_array_equal(offs, np.array(want_offs, dtype=np.uint32))


@pytest.mark.parametrize("dist", ["uniform", "zipf", "onehot90"])
@pytest.mark.parametrize("B", [2, 8, 32, 64, 256])
def test_multisplit_equals_numpy_stable_argsort(dist, B):
    n = 50_000
    keys = W.keys_for(dist, n, B, seed=B + len(dist))
    vals = np.arange(n, dtype=np.uint32)
    kout, vout, offs = oracle.multisplit(keys, B, oracle.RANGE, 0, vals)
    ids = ((keys.astype(np.uint64) * B) >> 32).astype(np.int64)
    order = np.argsort(ids, kind="stable")
    np.testing.assert_array_equal(kout, keys[order])
    np.testing.assert_array_equal(vout, vals[order])
    cnt = np.bincount(ids, minlength=B)
    np.testing.assert_array_equal(offs, np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.uint32))


def test_multisplit_special_cases():
    keys = W.uniform_keys(5000, seed=3)
    # one bucket: identity
    k1, _, o1 = oracle.multisplit(keys, 1)
    np.testing.assert_array_equal(k1, keys)
    assert o1.tolist() == [0]
    # already bucket-ordered input: identity
    srt = np.sort(keys)
    k2, _, _ = oracle.multisplit(srt, 256)
    np.testing.assert_array_equal(k2, srt)
    # all keys in the top bucket
    top = np.full(100, 0xFFFFFFFF, dtype=np.uint32)
    k3, _, o3 = oracle.multisplit(top, 16)
    np.testing.assert_array_equal(k3, top)
    assert o3.tolist() == [0] * 16


def test_delta_functor_like_reference_driver():
    # main_multisplit.cu:104-107: delta = ceil(n / B), keys in [b*delta, (b+1)*delta)
    n, B = 1 << 14, 32
    delta = (n + B - 1) // B
    keys = W.delta_keys(n, B, delta, seed=11)
    vals = np.arange(n, dtype=np.uint32)
    kout, vout, _ = oracle.multisplit(keys, B, oracle.DELTA, delta, vals)
    order = np.argsort(keys // delta, kind="stable")
    np.testing.assert_array_equal(kout, keys[order])
    np.testing.assert_array_equal(vout, vals[order])


def test_out_of_range_ids_are_clamped_to_last_bucket():
    keys = np.array([5, 100, 2, 7, 300, 1], dtype=np.uint32)
    kout, _, offs = oracle.multisplit(keys, 8, oracle.IDENTITY, 0)
    # ids 5,7(clamp),2,7,7(clamp),1 -> [1],[2],[5],[100,7,300]
    assert kout.tolist() == [1, 2, 5, 100, 7, 300]
    assert offs.tolist() == [0, 0, 1, 2, 2, 2, 3, 3]


@pytest.mark.parametrize("bits", [4, 5, 6, 7, 8])
def test_sort_schedule_covers_all_32_bits(bits):
    sched = oracle.sort_schedule(bits)
    covered = []
    for s, b in sched:
        covered.extend(range(s, s + b))
    assert covered == list(range(32))
    want_passes = {8: 4, 7: 5, 6: 6, 5: 7, 4: 8}[bits]
    assert len(sched) == want_passes


@pytest.mark.parametrize("bits", [4, 5, 6, 7, 8])
def test_sort_matches_numpy(bits):
    n = 20_000
    keys = W.uniform_keys(n, seed=bits)
    keys[::7] = keys[3]                       # many duplicates to expose instability
    vals = np.arange(n, dtype=np.uint32)
    kout, vout = oracle.sort(keys, bits, vals)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(kout, keys[order])
    np.testing.assert_array_equal(vout, vals[order])
    k_only, none = oracle.sort(keys, bits)
    assert none is None
    np.testing.assert_array_equal(k_only, np.sort(keys))


def test_histogram_matches_bincount():
    keys = W.uniform_keys(100_000, seed=5)
    h = oracle.histogram(keys, 256, oracle.DIGIT, 0)
    np.testing.assert_array_equal(h, np.bincount(keys & 0xFF, minlength=256).astype(np.uint32))
    # closed form: keys 0..n-1 with bin = key & 0xFF gives n/256 per bin
    h2 = oracle.histogram(np.arange(256 * 40, dtype=np.uint32), 256, oracle.DIGIT, 0)
    assert h2.tolist() == [40] * 256


def test_histogram_even_matches_numpy_float32():
    g = W.rng(9)
    x = (g.random(50_000, dtype=np.float32) * np.float32(256.0)).astype(np.float32)
    x[:5] = np.array([0.0, 255.99998, 256.0, -1.0, np.nan], dtype=np.float32)
    lower, upper, nb = np.float32(0.0), np.float32(256.0), 32
    inv = np.float32(1.0) / ((upper - lower) / np.float32(nb))
    h = oracle.histogram_even_f32(x, nb, float(lower), float(inv))
    with np.errstate(invalid="ignore"):
        s = (x - lower).astype(np.float32) * inv
        ok = (x - lower >= 0) & (s < nb)
    want = np.bincount(s[ok].astype(np.uint32), minlength=nb)
    np.testing.assert_array_equal(h, want.astype(np.int32))


@pytest.mark.parametrize("sel", [0.5, 0.05])
def test_compaction_matches_stable_partition(sel):
    n = 30_001
    keys = W.uniform_keys(n, seed=int(sel * 100))
    vals = np.arange(n, dtype=np.uint32)
    thr = int(sel * U32)
    # predicate key < thr  <=>  bucket 0 of the range identifier with B=2 only at sel=0.5;
    # use LOWBITS(1) like the reference driver, and check against a Python partition.
    c, kout, vout = oracle.compaction(keys, oracle.LOWBITS, 1, vals)
    zero = [i for i in range(n) if keys[i] & 1 == 0]
    one = [i for i in range(n) if keys[i] & 1 == 1]
    assert c == len(zero)
    np.testing.assert_array_equal(kout, keys[zero + one])
    np.testing.assert_array_equal(vout, vals[zero + one])
    assert thr >= 0


def test_tile_histograms_match_reshape_bincount():
    n, B, tile = 10_000, 64, 1024
    keys = W.uniform_keys(n, seed=2)
    H = oracle.tile_histograms(keys, B, tile)
    nt = (n + tile - 1) // tile
    ids = (keys.astype(np.uint64) * B) >> 32
    for t in range(nt):
        c = np.bincount(ids[t * tile:(t + 1) * tile].astype(np.int64), minlength=B)
        np.testing.assert_array_equal(H.reshape(B, nt)[:, t], c)


def _unstable_multisplit(keys, B):
    """A deliberately broken multisplit (reverses order inside buckets)."""
    ids = ((keys.astype(np.uint64) * B) >> 32).astype(np.int64)
    order = np.lexsort((-np.arange(keys.size), ids))
    return keys[order]


def test_checks_catch_a_planted_stability_bug():
    keys = W.uniform_keys(4096, seed=17)
    good, _, _ = oracle.multisplit(keys, 8)
    bad = _unstable_multisplit(keys, 8)
    ids = ((keys.astype(np.uint64) * 8) >> 32).astype(np.int64)
    ref = keys[np.argsort(ids, kind="stable")]
    assert np.array_equal(good, ref)
    assert not np.array_equal(bad, ref)


def _np_histogram_range(x, levels):
    b = np.searchsorted(levels, x, side="right") - 1
    ok = (b >= 0) & (b < levels.size - 1)
    if x.dtype == np.float32:
        ok &= ~np.isnan(x)
    return np.bincount(b[ok], minlength=levels.size - 1).astype(np.int32)


@pytest.mark.parametrize("nbins", [1, 2, 7, 32, 100, 256])
def test_oracle_histogram_range_against_searchsorted(nbins):
    g = W.rng(nbins)
    # float: random sorted edges in [-10, 300), samples around them incl. edges, outside, NaN
    lv = np.sort(g.uniform(-10, 300, size=nbins + 1).astype(np.float32))
    x = g.uniform(-20, 320, size=20000).astype(np.float32)
    x[:nbins + 1] = lv                               # exactly on every edge
    x[-3:] = np.array([np.nan, lv[0] - 1, lv[-1]], np.float32)
    np.testing.assert_array_equal(oracle.histogram_range(x, lv), _np_histogram_range(x, lv))
    # uint32, with repeated edges (empty bins)
    lu = np.sort(g.integers(0, 1 << 32, size=nbins + 1, dtype=np.uint64)).astype(np.uint32)
    if nbins > 3:
        lu[2] = lu[1]
    k = W.uniform_keys(20000, seed=nbins + 1)
    k[:nbins + 1] = lu
    np.testing.assert_array_equal(oracle.histogram_range(k, lu), _np_histogram_range(k, lu))


# --------------------------------------------------------------- 64-bit keys
def _py_bucket64(kind, B, param, width, k):
    if kind == oracle.RANGE:
        b = (k * B) >> 64
    elif kind == oracle.DELTA:
        b = k // param
    elif kind == oracle.DIGIT:
        b = (k >> param) & (((1 << width) - 1) if width else B - 1)
    elif kind == oracle.IDENTITY:
        b = k
    else:
        b = k & ((1 << param) - 1)
    return min(b, B - 1)


WIDE_KINDS = [  # (kind, B, param, width)
    (oracle.RANGE, 256, 0, 0), (oracle.RANGE, 3, 0, 0), (oracle.RANGE, 2, 0, 0),
    (oracle.DELTA, 100, (1 << 57) + 12345, 0), (oracle.DELTA, 256, 3, 0),
    (oracle.DIGIT, 256, 56, 8), (oracle.DIGIT, 16, 31, 4), (oracle.DIGIT, 64, 0, 0),
    (oracle.IDENTITY, 200, 0, 0), (oracle.LOWBITS, 32, 5, 0), (oracle.LOWBITS, 256, 40, 0),
]


def _wide_keys(n, seed):
    g = np.random.default_rng(seed)
    k = g.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
    edge = np.array([0, 1, (1 << 64) - 1, (1 << 63), (1 << 63) - 1, (1 << 32), (1 << 32) - 1,
                     199, 200, 201], dtype=np.uint64)
    return np.concatenate([edge, k, k[:50] & np.uint64(0xFF)]).astype(np.uint64)


@pytest.mark.parametrize("case", range(len(WIDE_KINDS)))
def test_oracle_bucket_ids64_against_python_integers(case):
    kind, B, param, width = WIDE_KINDS[case]
    keys = _wide_keys(500, case)
    if kind == oracle.RANGE:  # every range boundary j * 2^64 / B and its neighbours
        bnd = [(j << 64) // B for j in range(B)] + [-(-(j << 64) // B) for j in range(1, B)]
        extra = [v + d for v in bnd for d in (-1, 0, 1) if 0 <= v + d < (1 << 64)]
        keys = np.concatenate([keys, np.array(extra, dtype=np.uint64)])
    if kind == oracle.DELTA:
        extra = [j * param + d for j in range(1, 8) for d in (-1, 0, 1) if j * param + d < (1 << 64)]
        keys = np.concatenate([keys, np.array(extra, dtype=np.uint64)])
    got = oracle.bucket_ids64(keys, B, kind, param, width)
    want = [_py_bucket64(kind, B, param, width, int(k)) for k in keys]
    np.testing.assert_array_equal(got, np.array(want, dtype=np.uint32))


@pytest.mark.parametrize("case", range(len(WIDE_KINDS)))
@pytest.mark.parametrize("vbits", [0, 32, 64])
def test_oracle_multisplit_wide_against_stable_argsort(case, vbits):
    kind, B, param, width = WIDE_KINDS[case]
    keys = _wide_keys(3000, 100 + case)
    ids = np.array([_py_bucket64(kind, B, param, width, int(k)) for k in keys], dtype=np.int64)
    order = np.argsort(ids, kind="stable")
    vals = None
    if vbits:
        vt = np.uint64 if vbits == 64 else np.uint32
        vals = (np.arange(keys.size, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)).astype(vt)
    ko, vo, offs = oracle.multisplit_wide(keys, B, kind, param, vals, width)
    np.testing.assert_array_equal(ko, keys[order])
    np.testing.assert_array_equal(offs, np.concatenate([[0], np.cumsum(np.bincount(ids, minlength=B))[:-1]]))
    if vbits:
        np.testing.assert_array_equal(vo, vals[order])


def test_oracle_multisplit_wide_u32_keys_match_32bit_oracle():
    # 32-bit keys with 64-bit values: the bucket ids equal the 32-bit oracle's
    keys = W.uniform_keys(5000, seed=9)
    vals = np.arange(5000, dtype=np.uint64) << np.uint64(20)
    ko, vo, offs = oracle.multisplit_wide(keys, 100, oracle.RANGE, 0, vals)
    wk, wv, wo = oracle.multisplit(keys, 100, oracle.RANGE, 0, np.arange(5000, dtype=np.uint32))
    np.testing.assert_array_equal(ko, wk)
    np.testing.assert_array_equal(vo, vals[wv])
    np.testing.assert_array_equal(offs, wo)


@pytest.mark.parametrize("digit_bits", [8, 7, 5])
@pytest.mark.parametrize("vbits", [0, 32, 64])
def test_oracle_sort_wide_against_numpy(digit_bits, vbits):
    keys = _wide_keys(4000, digit_bits)
    keys[::7] = keys[0]  # duplicates expose instability
    order = np.argsort(keys, kind="stable")
    vals = None if not vbits else np.arange(keys.size, dtype=np.uint64 if vbits == 64 else np.uint32)
    ko, vo = oracle.sort_wide(keys, digit_bits, vals)
    np.testing.assert_array_equal(ko, keys[order])
    if vbits:
        np.testing.assert_array_equal(vo, vals[order])


def test_standard_library_baselines_against_numpy():
    """The CPU baselines bench.py times (oracle/baseline.cpp) compute what the
    rows compute: a stable sort equals numpy's stable argsort order, the
    counting loop equals bincount, the stable partition equals boolean
    selection in input order."""
    g = np.random.default_rng(5)
    n = 50_003
    keys = g.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    keys[::7] = keys[3]                                    # ties
    vals = np.arange(n, dtype=np.uint32)
    ko, vo = oracle.baseline_stable_sort(keys, vals)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(ko, keys[order])
    np.testing.assert_array_equal(vo, vals[order])
    ko, vo = oracle.baseline_stable_sort(keys)
    np.testing.assert_array_equal(ko, np.sort(keys))
    assert vo is None
    for shift in (0, 8, 24):
        np.testing.assert_array_equal(oracle.baseline_histogram(keys, 256, shift),
                                      np.bincount((keys >> np.uint32(shift)) & 0xFF, minlength=256))
    for thr in (0, 1 << 31, int(0.05 * (1 << 32)), 1 << 32):
        keep = keys.astype(np.uint64) < thr
        c, ko, vo = oracle.baseline_stable_partition(keys, thr, vals)
        assert c == int(keep.sum())
        np.testing.assert_array_equal(ko, np.concatenate([keys[keep], keys[~keep]]))
        np.testing.assert_array_equal(vo, np.concatenate([vals[keep], vals[~keep]]))
