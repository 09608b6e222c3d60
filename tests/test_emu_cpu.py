"""CPU parity of the library's HIP kernel sources, built for the host emulation
(tests/emu): the same prescan/scan/postscan/sort/histogram code runs with one
host thread per GPU thread and is compared with the oracle at small sizes.

This pins the kernels' index arithmetic and stability on the CPU; the GPU
tests pin the compiled gfx950 code.  The AddressSanitizer build of the same
sources is `make -C tests/emu emu_check_asan && tests/emu/emu_check_asan`.
"""
import ctypes
import fcntl
import os
import subprocess

import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W
from hist_cases import range_level_sets, range_samples

EMU_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")
EMU_LIB = os.path.join(EMU_DIR, "libgpumultisplit_emu.so")


@pytest.fixture(scope="module")
def emu():
    # one build at a time when pytest-xdist workers start together
    with open(os.path.join(EMU_DIR, ".build.lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-C", EMU_DIR, "libgpumultisplit_emu.so"], check=True)
    L = ctypes.CDLL(EMU_LIB)
    for name, res, args in gms._SIGS:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def p(a):
    return a.ctypes.data if a is not None else None


def emu_multisplit(L, keys, B, functor, kv):
    n = keys.size
    h = ctypes.c_void_p()
    assert L.gms_multisplit_create(B, ctypes.byref(h)) == 0
    try:
        assert L.gms_multisplit_allocate(h, n, int(kv)) == 0
        vals = np.arange(n, dtype=np.uint32)
        kout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        vout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        offs = np.zeros(B, np.uint32)
        if kv:
            st = L.gms_multisplit_key_value(h, p(keys), p(vals), p(kout), p(vout), n,
                                            functor.desc(), p(offs), None)
        else:
            st = L.gms_multisplit_key_only(h, p(keys), p(kout), n, functor.desc(), p(offs), None)
        assert st == 0
        return kout[:n], (vout[:n] if kv else None), offs
    finally:
        L.gms_multisplit_destroy(h)


@pytest.mark.parametrize("B", [1, 2, 3, 8, 17, 64, 128, 256])
@pytest.mark.parametrize("kv", [False, True])
def test_emulated_multisplit_matches_oracle(emu, B, kv):
    tile = gms.subproblem_size_bms_key_value(256)
    for n in (0, 1, 65, tile, tile + 1):
        keys = W.uniform_keys(n, seed=B * 7 + n)
        got = emu_multisplit(emu, keys, B, gms.range_bucket(B), kv)
        wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0,
                                       np.arange(n, dtype=np.uint32) if kv else None)
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[2], wo)
        if kv:
            np.testing.assert_array_equal(got[1], wv)


@pytest.mark.parametrize("B,kv,extra", [(2, True, 0), (8, False, 5), (64, True, 777),
                                        (256, True, 6143), (256, False, 1)])
def test_emulated_many_tiles_ragged_last(emu, B, kv, extra):
    """19 full tiles and a ragged (or absent) last tile, on zipf keys where
    whole tiles fall into one bucket."""
    tile = gms.tile_keys(B, kv, 0)
    n = 19 * tile + extra
    keys = W.keys_for("zipf", n, B, seed=B + extra) if B >= 8 else W.uniform_keys(n, seed=B)
    got = emu_multisplit(emu, keys, B, gms.range_bucket(B), kv)
    wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0,
                                   np.arange(n, dtype=np.uint32) if kv else None)
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[2], wo)
    if kv:
        np.testing.assert_array_equal(got[1], wv)


def test_emulated_skewed_and_delta(emu):
    tile = gms.subproblem_size_bms_key_value(256)
    n = 2 * tile + 3
    keys = W.keys_for("onehot90", n, 256, seed=5)
    got = emu_multisplit(emu, keys, 256, gms.range_bucket(256), True)
    wk, wv, wo = oracle.multisplit(keys, 256, oracle.RANGE, 0, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[1], wv)
    delta = (n + 31) // 32
    keys = W.delta_keys(n, 32, delta, seed=6)
    got = emu_multisplit(emu, keys, 32, gms.delta_func(delta), True)
    wk, wv, _ = oracle.multisplit(keys, 32, oracle.DELTA, delta, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[1], wv)


@pytest.mark.parametrize("bits", [7, 8])
def test_emulated_sort(emu, bits):
    n = 5000
    keys = W.uniform_keys(n, seed=bits)
    keys[::5] = keys[1]
    vals = np.arange(n, dtype=np.uint32)
    kout, vout = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    h = ctypes.c_void_p()
    assert emu.gms_sort_create(ctypes.byref(h)) == 0
    assert emu.gms_sort_allocate(h, n, bits, 1) == 0
    assert emu.gms_sort(h, p(keys), p(kout), p(vals), p(vout), n, bits, None) == 0
    emu.gms_sort_destroy(h)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(kout, keys[order])
    np.testing.assert_array_equal(vout, vals[order])


# ---------------------------------------------------------------------------
# Batched multisplit on the emulation: a mixed batch (every bucket class,
# every identifier kind, key-only and key-value, empty and ragged instances)
# under every tuning, each instance against the oracle run alone.
# ---------------------------------------------------------------------------
BATCH_CASES = [  # (n, B, functor, oracle kind, oracle param, kv)
    (0, 8, gms.range_bucket(8), oracle.RANGE, 0, True),
    (1, 2, gms.range_bucket(2), oracle.RANGE, 0, False),
    (100, 1, gms.range_bucket(1), oracle.RANGE, 0, True),
    (2047, 5, gms.range_bucket(5), oracle.RANGE, 0, True),
    (4097, 64, gms.delta_func(1 << 26), oracle.DELTA, 1 << 26, True),
    (6145, 256, gms.ms_bucket_sort_8bit(8), oracle.DIGIT, 8, False),
    (3000, 128, gms.identity_bucket(), oracle.IDENTITY, 0, True),
    (5000, 16, gms.test_bucket(3), oracle.LOWBITS, 3, False),
    (9000, 256, gms.range_bucket(256), oracle.RANGE, 0, True),
    (777, 32, gms.ms_bucket_sort_5bit(27), oracle.DIGIT, 27, True),
]


def _emu_batch(L, cases, tuning, seed):
    keep, items, want = [], [], []
    for i, (n, B, f, kind, param, kv) in enumerate(cases):
        keys = W.uniform_keys(n, seed=seed + i)
        if kind == oracle.IDENTITY:
            keys = keys % np.uint32(B + 3)        # some ids past B-1: clamped
        vals = np.arange(n, dtype=np.uint32) + np.uint32(1000 * i)
        kout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        vout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32) if kv else None
        offs = np.full(B, 0xFFFFFFFF, np.uint32)
        keep += [keys, vals, kout, vout, offs]
        items.append(gms.gms_batch_item(p(keys) if n else None, p(vals) if kv else None,
                                        p(kout), p(vout) if kv else None, n, B, f.desc(), p(offs)))
        wk, wv, wo = oracle.multisplit(keys, B, kind, param, vals if kv else None)
        want.append((wk, wv, wo))
    h = ctypes.c_void_p()
    assert L.gms_batch_create(ctypes.byref(h)) == 0
    try:
        assert L.gms_batch_set_tuning(h, gms.gms_batch_tuning(*tuning)) == 0
        arr = (gms.gms_batch_item * len(items))(*items)
        assert L.gms_batch_plan(h, arr, len(items), None) == 0
        assert L.gms_batch_run(h, None) == 0
    finally:
        L.gms_batch_destroy(h)
    for i, (n, B, *_rest) in enumerate(cases):
        kv = cases[i][5]
        kout, vout, offs = keep[5 * i + 2], keep[5 * i + 3], keep[5 * i + 4]
        wk, wv, wo = want[i]
        np.testing.assert_array_equal(kout[:n], wk, err_msg=f"instance {i} keys")
        np.testing.assert_array_equal(offs, wo, err_msg=f"instance {i} offsets")
        if kv:
            np.testing.assert_array_equal(vout[:n], wv, err_msg=f"instance {i} values")


@pytest.mark.parametrize("tuning", [(8, 0), (8, 1), (16, 1), (24, 0), (32, 1)])
def test_emulated_batch_matches_oracle(emu, tuning):
    _emu_batch(emu, BATCH_CASES, tuning, seed=sum(tuning))


def test_emulated_batch_rejects_overlapping_outputs(emu):
    a = np.zeros(64, np.uint32)
    b = np.zeros(64, np.uint32)
    it1 = gms.gms_batch_item(p(a), None, p(b), None, 32, 4, gms.range_bucket(4).desc(), None)
    it2 = gms.gms_batch_item(p(a), None, p(b) + 64, None, 32, 4, gms.range_bucket(4).desc(), None)
    h = ctypes.c_void_p()
    assert emu.gms_batch_create(ctypes.byref(h)) == 0
    arr = (gms.gms_batch_item * 2)(it1, it2)
    assert emu.gms_batch_plan(h, arr, 2, None) == 7           # GMS_ERR_ALIAS
    it2.d_key_out = p(b) + 128                                  # adjacent: fine
    arr = (gms.gms_batch_item * 2)(it1, it2)
    assert emu.gms_batch_plan(h, arr, 2, None) == 0
    bad = gms.gms_batch_item(p(a), None, p(b), None, 32, 300, gms.range_bucket(4).desc(), None)
    assert emu.gms_batch_plan(h, (gms.gms_batch_item * 1)(bad), 1, None) == 2
    emu.gms_batch_destroy(h)


# (n, digit_bits, kv) of the radix sorts planned beside multisplits
SORT_CASES = [(0, 8, True), (1, 4, False), (3000, 8, True), (4099, 5, False), (777, 6, True),
              (6145, 7, True), (2048, 4, True)]


@pytest.mark.parametrize("with_items,tuning", [(True, (24, 0)), (True, (8, 1)), (False, (16, 0))])
def test_emulated_batch_with_sorts(emu, with_items, tuning):
    """Radix sorts as passes of the batch (step 0: the multisplits and every
    sort's first pass, step p: the p-th passes) equal the oracle's sort, the
    multisplits beside them equal the oracle, and no input is written."""
    keep, sorts, want = [], [], []
    for i, (n, bits, kv) in enumerate(SORT_CASES):
        keys = W.uniform_keys(n, seed=70 + i)
        if i == 3:
            keys = keys & np.uint32(0x3FF)          # many equal digits: stability
        vals = np.arange(n, dtype=np.uint32) + np.uint32(7 * i)
        kout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        vout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        keep.append((keys, keys.copy(), vals, kout, vout))
        sorts.append(gms.gms_batch_sort_item(p(keys) if n else None, p(vals) if kv else None,
                                             p(kout), p(vout) if kv else None, n, bits))
        want.append(oracle.sort(keys, bits, vals if kv else None))
    cases = BATCH_CASES[:6] if with_items else []
    ikeep, items, iwant = [], [], []
    for i, (n, B, f, kind, param, kv) in enumerate(cases):
        keys = W.uniform_keys(n, seed=90 + i)
        vals = np.arange(n, dtype=np.uint32)
        kout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        vout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        offs = np.full(B, 0xFFFFFFFF, np.uint32)
        ikeep.append((keys, vals, kout, vout, offs))
        items.append(gms.gms_batch_item(p(keys) if n else None, p(vals) if kv else None, p(kout),
                                        p(vout) if kv else None, n, B, f.desc(), p(offs)))
        iwant.append(oracle.multisplit(keys, B, kind, param, vals if kv else None))
    h = ctypes.c_void_p()
    assert emu.gms_batch_create(ctypes.byref(h)) == 0
    try:
        assert emu.gms_batch_set_tuning(h, gms.gms_batch_tuning(*tuning)) == 0
        iarr = (gms.gms_batch_item * max(len(items), 1))(*items)
        sarr = (gms.gms_batch_sort_item * len(sorts))(*sorts)
        assert emu.gms_batch_plan_mixed(h, iarr, len(items), sarr, len(sorts), None) == 0
        assert emu.gms_batch_run(h, None) == 0
        k, m = ctypes.c_uint64(), ctypes.c_uint64()
        w, lc = ctypes.c_uint32(), ctypes.c_uint32()
        assert emu.gms_batch_stats(h, ctypes.byref(k), ctypes.byref(w), ctypes.byref(m),
                                   ctypes.byref(lc)) == 0
        passes = sum(len(oracle.sort_schedule(b)) * n for n, b, _ in SORT_CASES)
        assert k.value == passes + sum(c[0] for c in cases)
    finally:
        emu.gms_batch_destroy(h)
    for i, ((n, bits, kv), (keys, keys0, vals, kout, vout), (wk, wv)) in enumerate(
            zip(SORT_CASES, keep, want)):
        np.testing.assert_array_equal(keys, keys0, err_msg=f"sort {i} input written")
        np.testing.assert_array_equal(kout[:n], wk, err_msg=f"sort {i} keys")
        if kv:
            np.testing.assert_array_equal(vout[:n], wv, err_msg=f"sort {i} values")
    for i, ((keys, vals, kout, vout, offs), (wk, wv, wo)) in enumerate(zip(ikeep, iwant)):
        n = keys.size
        np.testing.assert_array_equal(kout[:n], wk, err_msg=f"item {i} keys")
        np.testing.assert_array_equal(offs, wo, err_msg=f"item {i} offsets")
        if cases[i][5]:
            np.testing.assert_array_equal(vout[:n], wv, err_msg=f"item {i} values")


def test_emulated_batch_sort_validation(emu):
    a = np.zeros(64, np.uint32)
    b = np.zeros(64, np.uint32)
    h = ctypes.c_void_p()
    assert emu.gms_batch_create(ctypes.byref(h)) == 0
    try:
        def plan(*sorts):
            arr = (gms.gms_batch_sort_item * len(sorts))(*sorts)
            return emu.gms_batch_plan_mixed(h, None, 0, arr, len(sorts), None)
        assert plan(gms.gms_batch_sort_item(p(a), None, p(b), None, 64, 3)) == 1   # 3-bit digits
        assert plan(gms.gms_batch_sort_item(p(a), p(a), p(b), None, 64, 8)) == 1   # value out missing
        assert plan(gms.gms_batch_sort_item(p(a), None, p(a) + 128, None, 64, 8)) == 7  # in/out overlap
        assert plan(gms.gms_batch_sort_item(p(a), None, p(b), None, 32, 8),
                    gms.gms_batch_sort_item(p(a), None, p(b) + 64, None, 32, 8)) == 7  # outputs overlap
        assert plan(gms.gms_batch_sort_item(p(a), None, p(b), None, 64, 8)) == 0
    finally:
        emu.gms_batch_destroy(h)


# ---------------------------------------------------------------------------
# Request server on the emulation: mixed multisplit and sort requests, each
# against the oracle; a budget that holds only some sorts at once (the rest
# wait and are admitted as others finish); a request that can never fit.
# ---------------------------------------------------------------------------
def _emu_server(L, budget, max_inflight=2, max_step_keys=0):
    cfg = gms.gms_server_config(budget, max_inflight, max_step_keys, 8)
    h = ctypes.c_void_p()
    assert L.gms_server_create(ctypes.byref(cfg), ctypes.byref(h)) == 0
    return h


def _server_requests(seed, count):
    g = W.rng(seed)
    reqs = []
    for i in range(count):
        n = int(g.integers(0, 9000))
        keys = W.uniform_keys(n, seed=seed * 100 + i)
        kv = bool(i % 2)
        vals = np.arange(n, dtype=np.uint32)
        kout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        vout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        if i % 3 == 2:
            bits = (8, 7, 4)[i % 3 if i % 9 else 0]
            req = gms.gms_request(1, p(keys), p(vals) if kv else None, p(kout), p(vout) if kv else None,
                                  n, 0, gms.gms_bucket_desc(0, 0, 0), bits, None)
            wk, wv = oracle.sort(keys, bits, vals if kv else None)
            want = (wk, wv, None)
            offs = None
        else:
            B = (2, 16, 64, 256)[i % 4]
            offs = np.full(B, 0xFFFFFFFF, np.uint32)
            req = gms.gms_request(0, p(keys), p(vals) if kv else None, p(kout), p(vout) if kv else None,
                                  n, B, gms.range_bucket(B).desc(), 0, p(offs))
            want = oracle.multisplit(keys, B, oracle.RANGE, 0, vals if kv else None)
        reqs.append((req, (keys, vals, kout, vout, offs), want, n, kv))
    return reqs


def _check_server_results(reqs):
    for i, (_req, (keys, vals, kout, vout, offs), (wk, wv, wo), n, kv) in enumerate(reqs):
        np.testing.assert_array_equal(kout[:n], wk, err_msg=f"request {i} keys")
        if kv:
            np.testing.assert_array_equal(vout[:n], wv, err_msg=f"request {i} values")
        if offs is not None:
            np.testing.assert_array_equal(offs, wo, err_msg=f"request {i} offsets")


@pytest.mark.parametrize("max_inflight", [1, 2])
def test_emulated_server_mixed_requests(emu, max_inflight):
    h = _emu_server(emu, 64 << 20, max_inflight)
    try:
        reqs = _server_requests(seed=max_inflight, count=18)
        tickets = []
        for req, *_ in reqs:
            t = ctypes.c_uint64()
            assert emu.gms_server_submit(h, ctypes.byref(req), ctypes.byref(t)) == 0
            tickets.append(t.value)
        for t in tickets:
            lat = ctypes.c_double()
            assert emu.gms_server_wait(h, t, ctypes.byref(lat)) == 0
            assert lat.value > 0
        st = gms.gms_server_stats()
        assert emu.gms_server_get_stats(h, ctypes.byref(st)) == 0
        assert st.completed == len(reqs) and st.submitted == len(reqs)
        assert st.active == 0 and st.pending == 0
        assert st.steps >= 4          # sorts take one step per digit
        # every pass of every request went through one step
        passes = sum(len(oracle.sort_schedule(r[0].digit_bits)) if r[0].op == 1 else 1
                     for r in reqs)
        assert st.step_items == passes
        assert st.keys_processed == sum(
            r[3] * (len(oracle.sort_schedule(r[0].digit_bits)) if r[0].op == 1 else 1) for r in reqs)
    finally:
        assert emu.gms_server_destroy(h) == 0
    _check_server_results(reqs)


def test_emulated_server_budget_admission(emu):
    # 1 MiB budget: a 60000-key key-value sort reserves ~0.5 MiB + its step
    # shares, so only one such sort is admitted at a time; a 200000-key sort
    # can never fit and is rejected at submit.
    h = _emu_server(emu, 1 << 20, 2, 1 << 16)
    try:
        n = 60000
        bufs, tickets = [], []
        for i in range(3):
            keys = W.uniform_keys(n, seed=70 + i)
            vals = np.arange(n, dtype=np.uint32)
            kout = np.zeros(n, np.uint32)
            vout = np.zeros(n, np.uint32)
            bufs.append((keys, vals, kout, vout))
            req = gms.gms_request(1, p(keys), p(vals), p(kout), p(vout), n, 0,
                                  gms.gms_bucket_desc(0, 0, 0), 8, None)
            t = ctypes.c_uint64()
            assert emu.gms_server_submit(h, ctypes.byref(req), ctypes.byref(t)) == 0
            tickets.append(t.value)
        big = np.zeros(600000, np.uint32)
        t = ctypes.c_uint64()
        req = gms.gms_request(1, p(big), p(big), p(big) + 4 * 300000, None, 300000, 0,
                              gms.gms_bucket_desc(0, 0, 0), 8, None)
        assert emu.gms_server_submit(h, ctypes.byref(req), ctypes.byref(t)) == 1  # value in, no value out
        req = gms.gms_request(1, p(big), None, p(big) + 4 * 300000, None, 300000, 0,
                              gms.gms_bucket_desc(0, 0, 0), 8, None)
        assert emu.gms_server_submit(h, ctypes.byref(req), ctypes.byref(t)) == 3  # never fits
        for t in tickets:
            assert emu.gms_server_wait(h, t, None) == 0
        st = gms.gms_server_stats()
        emu.gms_server_get_stats(h, ctypes.byref(st))
        assert st.peak_scratch_bytes <= 1 << 20
        assert st.rejected == 1 and st.completed == 3
    finally:
        emu.gms_server_destroy(h)
    for keys, vals, kout, vout in bufs:
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(kout, keys[order])
        np.testing.assert_array_equal(vout, vals[order])


def test_emulated_server_fragmented_arena_fails_request_instead_of_hanging(emu):
    """The budget check counts reserved bytes, not contiguous ones: with the
    arena's free space cut into 4 KiB holes (the emulation build's hold hook
    plays the sorts' buffers), a request whose step tables need more than a
    hole is admitted but can never be placed.  With nothing in flight it must
    fail with GMS_ERR_CAPACITY (not wait forever); once the space is whole
    again the same request succeeds."""
    emu.gms_emu_server_hold.restype = ctypes.c_int
    emu.gms_emu_server_hold.argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t)]
    emu.gms_emu_server_unhold.restype = ctypes.c_int
    emu.gms_emu_server_unhold.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t]
    h = _emu_server(emu, 1 << 20, 2)
    piece = 4096
    held = []
    try:
        off = ctypes.c_size_t()
        while emu.gms_emu_server_hold(h, piece, ctypes.byref(off)) == 0:
            held.append(off.value)
        assert len(held) == (1 << 20) // piece
        for o in held[::2]:                              # 4 KiB holes
            assert emu.gms_emu_server_unhold(h, o, piece) == 0
        held = held[1::2]
        n, B = 50000, 256
        keys = W.uniform_keys(n, seed=404)
        vals = np.arange(n, dtype=np.uint32)
        kout = np.zeros(n, np.uint32)
        vout = np.zeros(n, np.uint32)
        req = gms.gms_request(0, p(keys), p(vals), p(kout), p(vout), n, B,
                              gms.range_bucket(B).desc(), 0, None)
        t = ctypes.c_uint64()
        assert emu.gms_server_submit(h, ctypes.byref(req), ctypes.byref(t)) == 0
        assert emu.gms_server_wait(h, t.value, None) == 3        # GMS_ERR_CAPACITY
        for o in held:
            assert emu.gms_emu_server_unhold(h, o, piece) == 0
        held = []
        assert emu.gms_server_submit(h, ctypes.byref(req), ctypes.byref(t)) == 0
        assert emu.gms_server_wait(h, t.value, None) == 0
        wk, wv, _ = oracle.multisplit(keys, B, oracle.RANGE, 0, vals)
        np.testing.assert_array_equal(kout, wk)
        np.testing.assert_array_equal(vout, wv)
        st = gms.gms_server_stats()
        emu.gms_server_get_stats(h, ctypes.byref(st))
        assert st.completed == 2 and st.active == 0
    finally:
        for o in held:
            emu.gms_emu_server_unhold(h, o, piece)
        emu.gms_server_destroy(h)


@pytest.mark.parametrize("B,kind,param,width", [(257, oracle.RANGE, 0, 0), (1000, oracle.DELTA, 4294968, 0),
                                                (512, oracle.DIGIT, 5, 9), (1024, oracle.RANGE, 0, 0),
                                                (1025, oracle.RANGE, 0, 0),
                                                (4096, oracle.DIGIT, 7, 12), (70000, oracle.IDENTITY, 0, 0),
                                                (1 << 20, oracle.RANGE, 0, 0)])
def test_emulated_more_than_256_buckets(emu, B, kind, param, width):
    """Up to 1024 buckets: one pass with three LDS ranking passes; beyond:
    LSD passes over 8-bit digits of the id, bucket starts read off the output.
    Sizes span several tiles and a ragged last one."""
    n = 2 * 6144 + 7001
    keys = W.uniform_keys(n, seed=B)
    if kind == oracle.IDENTITY:
        keys = keys % np.uint32(B + 100)
    desc = gms.gms_bucket_desc(kind, param, width)
    for kv in (False, True):
        vals = np.arange(n, dtype=np.uint32)
        kout, vout = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        offs = np.zeros(B, np.uint32)
        h = ctypes.c_void_p()
        assert emu.gms_multisplit_create(B, ctypes.byref(h)) == 0
        assert emu.gms_multisplit_allocate(h, n, int(kv)) == 0
        if kv:
            st = emu.gms_multisplit_key_value(h, p(keys), p(vals), p(kout), p(vout), n, desc, p(offs), None)
        else:
            st = emu.gms_multisplit_key_only(h, p(keys), p(kout), n, desc, p(offs), None)
        emu.gms_multisplit_destroy(h)
        assert st == 0
        okind = oracle.DIGIT if kind == oracle.DIGIT else kind
        wk, wv, wo = oracle.multisplit(keys, B, okind, param, vals)
        np.testing.assert_array_equal(kout, wk)
        np.testing.assert_array_equal(offs, wo)
        if kv:
            np.testing.assert_array_equal(vout, wv)


def test_emulated_histogram_range_edge_sets(emu):
    # Custom-edge histograms through the cell-table kernel: spread edges,
    # 200 edges inside one cell, duplicate edges, a range whose width
    # overflows to inf, ranges a few ulps wide and denormal, integer ranges
    # of every width; samples include every edge, its neighbours, NaN, +-inf.
    sets, isets = range_level_sets()
    for lv in sets:
        x = range_samples(lv, 3 * 4096 + 123, seed=5)
        h = np.zeros(lv.size - 1, np.int32)
        assert emu.gms_histogram_range_f32(p(x), x.size, p(h), lv.size, p(lv), None) == 0
        want = oracle.histogram_range(x, lv)
        assert want.sum() > lv.size // 2
        np.testing.assert_array_equal(h, want, err_msg=str(lv[:8]))
    for lv in isets:
        x = range_samples(lv, 3 * 4096 + 123, seed=6)
        h = np.zeros(lv.size - 1, np.int32)
        assert emu.gms_histogram_range_u32(p(x), x.size, p(h), lv.size, p(lv), None) == 0
        want = oracle.histogram_range(x, lv)
        assert want.sum() > lv.size // 2
        np.testing.assert_array_equal(h, want, err_msg=str(lv[:8]))


# ------------------------------------------------- 64-bit keys and values
WIDE_CASES = [  # (B, kind, param, width)
    (256, gms.BUCKET_RANGE, 0, 0), (2, gms.BUCKET_RANGE, 0, 0), (100, gms.BUCKET_DELTA, (1 << 57) + 5, 0),
    (64, gms.BUCKET_DIGIT, 58, 6), (17, gms.BUCKET_IDENTITY, 0, 0), (8, gms.BUCKET_LOWBITS, 3, 0),
    (1, gms.BUCKET_RANGE, 0, 0),
]


def _wide_keys(n, seed, kbytes):
    g = np.random.default_rng(seed)
    if kbytes == 8:
        k = g.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
        k[::5] &= np.uint64(0xFF)  # small keys for identity / low-bit identifiers
        return k
    return g.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)


def emu_multisplit_wide(L, keys, vals, B, kind, param, width):
    n = keys.size
    kb = keys.dtype.itemsize
    vb = 0 if vals is None else vals.dtype.itemsize
    h = ctypes.c_void_p()
    assert L.gms_multisplit_create(B, ctypes.byref(h)) == 0
    try:
        assert L.gms_multisplit_allocate(h, n, 1) == 0
        kout = np.full(max(n, 1), 0x5A, keys.dtype)
        vout = None if vals is None else np.full(max(n, 1), 0x5A, vals.dtype)
        offs = np.zeros(B, np.uint32)
        st = L.gms_multisplit_wide(h, p(keys), p(vals), p(kout), p(vout), n, kb, vb,
                                   gms.gms_bucket_desc64(kind, width, param), None, p(offs), None)
        assert st == 0, st
        return kout[:n], (None if vals is None else vout[:n]), offs
    finally:
        L.gms_multisplit_destroy(h)


@pytest.mark.parametrize("case", range(len(WIDE_CASES)))
@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 4), (8, 8), (4, 8)])
def test_emulated_multisplit_wide_matches_oracle(emu, case, kbytes, vbytes):
    B, kind, param, width = WIDE_CASES[case]
    if kbytes == 4 and param > 0xFFFFFFFF:
        param = 1000
    if kbytes == 4 and kind == gms.BUCKET_DIGIT:
        param = 26
    tile = 256 * (16 if kbytes + vbytes <= 8 else 12)
    for n in (0, 1, 77, tile, 2 * tile + 5):
        keys = _wide_keys(n, seed=n + case, kbytes=kbytes)
        vals = None
        if vbytes:
            vals = (np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)).astype(
                np.uint64 if vbytes == 8 else np.uint32)
        got = emu_multisplit_wide(emu, keys, vals, B, kind, param, width)
        wk, wv, wo = oracle.multisplit_wide(keys, B, kind, param, vals, width)
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[2], wo)
        if vbytes:
            np.testing.assert_array_equal(got[1], wv)


def wide_counts(keys, B, kind, param, width):
    """Per-bucket counts from the oracle's bucket starts (64- or 32-bit keys)."""
    _, _, wo = oracle.multisplit_wide(keys, B, kind, param, None, width)
    return np.diff(np.append(wo.astype(np.int64), keys.size)).astype(np.uint32)


@pytest.mark.parametrize("case", range(len(WIDE_CASES)))
@pytest.mark.parametrize("kbytes", [8, 4])
def test_emulated_histogram_wide(emu, case, kbytes):
    B, kind, param, width = WIDE_CASES[case]
    if kbytes == 4 and param > 0xFFFFFFFF:
        param = 1000
    if kbytes == 4 and kind == gms.BUCKET_DIGIT:
        param = 26
    d = gms.gms_bucket_desc64(kind, width, param)
    for n in (0, 1, 77, 4 * 1024 + 5, 3 * 4 * 1024 * 8 + 11):
        keys = _wide_keys(n, seed=n + case, kbytes=kbytes)
        h = np.full(B, 3, np.uint32)  # the histogram accumulates
        assert emu.gms_histogram_wide(p(keys), n, kbytes, p(h), B, d, None) == 0
        np.testing.assert_array_equal(h - 3, wide_counts(keys, B, kind, param, width))
    invalid_arg, unsupported_buckets = 1, 2
    h, k = np.zeros(4, np.uint32), np.zeros(4, np.uint64)
    assert emu.gms_histogram_wide(p(k), 4, 2, p(h), 4, d, None) == invalid_arg
    assert emu.gms_histogram_wide(p(k), 4, 8, p(h), 257, gms.gms_bucket_desc64(0, 0, 0),
                                  None) == unsupported_buckets


@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 8), (8, 4), (4, 8)])
def test_emulated_sort_wide(emu, kbytes, vbytes):
    n = 2 * 256 * 12 + 9
    keys = _wide_keys(n, seed=3, kbytes=kbytes)
    keys[::11] = keys[0]
    vals = None if not vbytes else np.arange(n, dtype=np.uint64 if vbytes == 8 else np.uint32)
    h = ctypes.c_void_p()
    assert emu.gms_sort_create(ctypes.byref(h)) == 0
    try:
        assert emu.gms_sort_allocate_wide(h, n, 8, kbytes, vbytes) == 0
        kout = np.zeros(n, keys.dtype)
        vout = None if vals is None else np.zeros(n, vals.dtype)
        assert emu.gms_sort_wide(h, p(keys), p(kout), p(vals), p(vout), n, kbytes, vbytes, None) == 0
    finally:
        emu.gms_sort_destroy(h)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(kout, keys[order])
    if vbytes:
        np.testing.assert_array_equal(vout, vals[order])


# ------------------------------------------------- wave-level variant (WMS)
def emu_multisplit_variant(L, keys, B, functor, kv, variant):
    n = keys.size
    h = ctypes.c_void_p()
    assert L.gms_multisplit_create(B, ctypes.byref(h)) == 0
    try:
        assert L.gms_multisplit_allocate(h, n, int(kv)) == 0
        assert L.gms_multisplit_set_variant(h, variant) == 0
        vals = np.arange(n, dtype=np.uint32)
        kout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        vout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        offs = np.zeros(B, np.uint32)
        if kv:
            st = L.gms_multisplit_key_value(h, p(keys), p(vals), p(kout), p(vout), n,
                                            functor.desc(), p(offs), None)
        else:
            st = L.gms_multisplit_key_only(h, p(keys), p(kout), n, functor.desc(), p(offs), None)
        assert st == 0
        return kout[:n], (vout[:n] if kv else None), offs
    finally:
        L.gms_multisplit_destroy(h)


@pytest.mark.parametrize("B", [2, 3, 8, 17, 32, 64])
@pytest.mark.parametrize("kv", [False, True])
def test_emulated_wms_matches_oracle(emu, B, kv):
    sub, blk = 1024, 4096
    for n in (0, 1, 63, 65, sub, sub + 1, blk + 3 * 64 + 5, 2 * blk + 1):
        keys = W.uniform_keys(n, seed=B * 13 + n)
        got = emu_multisplit_variant(emu, keys, B, gms.range_bucket(B), kv, gms.VARIANT_WMS)
        wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0,
                                       np.arange(n, dtype=np.uint32) if kv else None)
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[2], wo)
        if kv:
            np.testing.assert_array_equal(got[1], wv)


def test_emulated_wms_other_identifiers_and_skew(emu):
    n = 3 * 4096 + 77
    cases = [(64, gms.delta_func(1 << 26), oracle.DELTA, 1 << 26, W.uniform_keys(n, seed=1)),
             (16, gms.ms_bucket_sort_4bit(8), oracle.DIGIT, 8, W.uniform_keys(n, seed=2)),
             (40, gms.identity_bucket(), oracle.IDENTITY, 0,
              W.uniform_keys(n, seed=3) % np.uint32(45)),
             (32, gms.test_bucket(5), oracle.LOWBITS, 5, W.uniform_keys(n, seed=4)),
             (64, gms.range_bucket(64), oracle.RANGE, 0, W.keys_for("onehot90", n, 64, seed=5))]
    for B, f, kind, param, keys in cases:
        keys = keys.astype(np.uint32)
        got = emu_multisplit_variant(emu, keys, B, f, True, gms.VARIANT_WMS)
        wk, wv, wo = oracle.multisplit(keys, B, kind, param, np.arange(n, dtype=np.uint32))
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[1], wv)
        np.testing.assert_array_equal(got[2], wo)


def test_wms_variant_validation(emu):
    h = ctypes.c_void_p()
    assert emu.gms_multisplit_create(65, ctypes.byref(h)) == 0
    assert emu.gms_multisplit_set_variant(h, gms.VARIANT_WMS) == 2  # more than 64 buckets
    assert emu.gms_multisplit_set_variant(h, 7) == 1
    assert emu.gms_multisplit_set_variant(h, gms.VARIANT_BMS) == 0
    emu.gms_multisplit_destroy(h)
    blk = ctypes.c_uint32()
    assert emu.gms_subproblem_size_wms(64, 1, ctypes.byref(blk)) == 1024 and blk.value == 4096
    assert emu.gms_subproblem_size_wms(65, 0, ctypes.byref(blk)) == 0
    assert emu.gms_subproblem_size_wms(1, 0, ctypes.byref(blk)) == 0


@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 8), (4, 8)])
def test_emulated_wide_more_than_256_buckets(emu, kbytes, vbytes):
    for B, kind, param, width in [(1000, gms.BUCKET_RANGE, 0, 0), (4096, gms.BUCKET_DIGIT, 20, 12),
                                  (70000, gms.BUCKET_DELTA, ((1 << 64) - 1) // 65000 if kbytes == 8 else 60000, 0)]:
        if kbytes == 4 and kind == gms.BUCKET_DIGIT:
            param = 10
        for n in (0, 1, 2 * 3072 + 17):
            keys = _wide_keys(n, seed=n + B, kbytes=kbytes)
            vals = None
            if vbytes:
                vals = (np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)).astype(
                    np.uint64 if vbytes == 8 else np.uint32)
            got = emu_multisplit_wide(emu, keys, vals, B, kind, param, width)
            wk, wv, wo = oracle.multisplit_wide(keys, B, kind, param, vals, width)
            np.testing.assert_array_equal(got[0], wk)
            np.testing.assert_array_equal(got[2], wo)
            if vbytes:
                np.testing.assert_array_equal(got[1], wv)


@pytest.mark.parametrize("kbytes,vbytes", [(8, 0), (8, 8), (4, 8)])
def test_emulated_compaction_wide(emu, kbytes, vbytes):
    n = 2 * 4096 + 33
    keys = _wide_keys(n, seed=9, kbytes=kbytes)
    vals = None if not vbytes else np.arange(n, dtype=np.uint64 if vbytes == 8 else np.uint32)
    h = ctypes.c_void_p()
    assert emu.gms_compaction_create(ctypes.byref(h)) == 0
    try:
        assert emu.gms_compaction_allocate(h, n, 1) == 0
        kout = np.zeros(n, keys.dtype)
        vout = None if vals is None else np.zeros(n, vals.dtype)
        cnt = ctypes.c_uint32()
        d = gms.gms_bucket_desc64(gms.BUCKET_LOWBITS, 0, 1)
        assert emu.gms_compaction_wide(h, p(keys), p(vals), p(kout), p(vout), n, kbytes, vbytes, d,
                                       ctypes.addressof(cnt), None, None) == 0
    finally:
        emu.gms_compaction_destroy(h)
    wk, wv, wo = oracle.multisplit_wide(keys, 2, oracle.LOWBITS, 1, vals)
    np.testing.assert_array_equal(kout, wk)
    assert cnt.value == wo[1] == int(np.sum((keys & 1) == 0))
    if vbytes:
        np.testing.assert_array_equal(vout, wv)


# ---------------------------------------------------------------------------
# gms_multisplit_noalloc on the emulation: equal to gms_multisplit on the same
# input, on two calls in a row (the second smaller, so stale words of the
# first call's scratch sit past the second call's matrix), with no allocation
# or free while it runs; over-capacity inputs rejected before any launch.
# ---------------------------------------------------------------------------
def test_emulated_noalloc_equals_context_call(emu):
    emu.gms_emu_alloc_calls.restype = ctypes.c_uint64
    for B, kv, sizes in ((256, True, (70_001, 9000)), (5, False, (20_000, 3)),
                         (1024, True, (30_000, 6145))):
        h = ctypes.c_void_p()
        assert emu.gms_multisplit_create(B, ctypes.byref(h)) == 0
        assert emu.gms_multisplit_allocate(h, max(sizes), int(kv)) == 0
        f = gms.range_bucket(B).desc()
        try:
            for n in sizes:
                keys = W.keys_for("zipf", n, B, seed=n) if B >= 8 else W.uniform_keys(n, seed=n)
                vals = np.arange(n, dtype=np.uint32) if kv else None
                outs = []
                for noalloc in (False, True):
                    kout = np.full(n, 0x5A5A5A5A, np.uint32)
                    vout = np.full(n, 0x5A5A5A5A, np.uint32) if kv else None
                    offs = np.full(B, 0xFFFFFFFF, np.uint32)
                    before = emu.gms_emu_alloc_calls()
                    if noalloc:
                        st = emu.gms_multisplit_noalloc(p(keys), p(vals), p(kout), p(vout), n, B, f,
                                                        p(offs), None)
                        assert emu.gms_emu_alloc_calls() == before, "no-alloc call allocated"
                    else:
                        st = emu.gms_multisplit(h, p(keys), p(vals), p(kout), p(vout), n, f,
                                                p(offs), None)
                    assert st == 0
                    outs.append((kout, vout, offs))
                (k0, v0, o0), (k1, v1, o1) = outs
                np.testing.assert_array_equal(k1, k0, err_msg=f"B={B} n={n} keys")
                np.testing.assert_array_equal(o1, o0, err_msg=f"B={B} n={n} offsets")
                if kv:
                    np.testing.assert_array_equal(v1, v0, err_msg=f"B={B} n={n} values")
                wk, _, wo = oracle.multisplit(keys, B, oracle.RANGE, 0, vals)
                np.testing.assert_array_equal(k1, wk)
                np.testing.assert_array_equal(o1, wo)
        finally:
            emu.gms_multisplit_destroy(h)
    # beyond the 32 MiB array: 2^31 keys in 1024 buckets; over 1024 buckets
    a = np.zeros(4, np.uint32)
    b = np.zeros(4, np.uint32)
    big = (1 << 31)
    assert emu.gms_multisplit_noalloc(p(a), None, p(b), None, big, 1024,
                                      gms.range_bucket(1024).desc(), None, None) == 3
    assert emu.gms_multisplit_noalloc(p(a), None, p(b), None, 4, 1025,
                                      gms.range_bucket(1025).desc(), None, None) == 2


# ---------------------------------------------------------------------------
# Wave-level variants on the emulation: WMS (LDS reorder) and DMS (direct
# scatter from registers) against the oracle, ragged last subproblem included.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("variant", [2, 3])
@pytest.mark.parametrize("B,kv", [(2, True), (5, False), (32, True), (64, False)])
def test_emulated_wave_variants(emu, variant, B, kv):
    for n in (1, 1000, 4096 + 77):
        keys = W.keys_for("zipf", n, B, seed=n + B) if B >= 8 else W.uniform_keys(n, seed=n)
        vals = np.arange(n, dtype=np.uint32)
        kout = np.full(n, 0x5A5A5A5A, np.uint32)
        vout = np.full(n, 0x5A5A5A5A, np.uint32)
        offs = np.zeros(B, np.uint32)
        h = ctypes.c_void_p()
        assert emu.gms_multisplit_create(B, ctypes.byref(h)) == 0
        try:
            assert emu.gms_multisplit_allocate(h, n, int(kv)) == 0
            assert emu.gms_multisplit_set_variant(h, variant) == 0
            assert emu.gms_multisplit(h, p(keys), p(vals) if kv else None, p(kout),
                                      p(vout) if kv else None, n, gms.range_bucket(B).desc(),
                                      p(offs), None) == 0
        finally:
            emu.gms_multisplit_destroy(h)
        wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0, vals if kv else None)
        np.testing.assert_array_equal(kout, wk, err_msg=f"variant {variant} B={B} n={n}")
        np.testing.assert_array_equal(offs, wo)
        if kv:
            np.testing.assert_array_equal(vout, wv)


# ---------------------------------------------------------------------------
# Batched 64-bit elements on the emulation: wide multisplits of every shape
# and 64-bit sorts beside 32-bit items, each against the oracle; a 32-bit
# shape given as a wide item runs as a plain one.
# ---------------------------------------------------------------------------
def _u64(n, seed):
    return np.random.default_rng(seed).integers(0, 2**64, size=n, dtype=np.uint64)


# (n, B, key bytes, value bytes (0 = key-only), kind, param)
BATCH_WIDE_CASES = [(0, 8, 8, 8, oracle.RANGE, 0), (1, 2, 8, 0, oracle.RANGE, 0),
              (3001, 256, 8, 8, oracle.RANGE, 0), (4097, 64, 8, 4, oracle.RANGE, 0),
              (2500, 17, 4, 8, oracle.RANGE, 0), (5000, 16, 8, 0, oracle.DIGIT, 40),
              (777, 32, 4, 4, oracle.DELTA, (1 << 32) // 32 + 1),
              (3333, 100, 8, 8, oracle.DELTA, (1 << 64) // 100 + 7)]
# (n, digit bits, key bytes, value bytes (0 = keys only))
BATCH_WIDE_SORTS = [(3000, 8, 8, 8), (2049, 6, 8, 0), (1500, 5, 8, 4), (2222, 8, 4, 8), (0, 8, 8, 8)]


def _wide_desc(B, kind, param):
    width = 4 if kind == oracle.DIGIT else 0
    return gms.gms_bucket_desc64(kind, width, param), width


@pytest.mark.parametrize("with_plain", [False, True])
def test_emulated_batch_wide_items_and_sorts(emu, with_plain):
    keep, witems, wwant = [], [], []
    for i, (n, B, kb, vb, kind, param) in enumerate(BATCH_WIDE_CASES):
        keys = _u64(n, 10 + i) if kb == 8 else W.uniform_keys(n, seed=10 + i)
        vals = (None if vb == 0 else (_u64(n, 50 + i) if vb == 8
                                      else np.arange(n, dtype=np.uint32)))
        kout = np.full(max(n, 1), 0x5A, np.uint64 if kb == 8 else np.uint32)
        vout = None if vb == 0 else np.full(max(n, 1), 0x5A, np.uint64 if vb == 8 else np.uint32)
        offs = np.full(B, 0xFFFFFFFF, np.uint32)
        desc, width = _wide_desc(B, kind, param)
        keep.append((keys, vals, kout, vout, offs))
        witems.append(gms.gms_batch_wide_item(p(keys) if n else None, p(vals), p(kout), p(vout), n, B,
                                              kb, vb, desc, p(offs)))
        wwant.append(oracle.multisplit_wide(keys, B, kind, param, vals, width))
    skeep, sitems, swant = [], [], []
    for i, (n, bits, kb, vb) in enumerate(BATCH_WIDE_SORTS):
        keys = _u64(n, 80 + i) if kb == 8 else W.uniform_keys(n, seed=80 + i)
        if i == 1:
            keys &= np.uint64(0xFFFF0000000000FF)   # equal digits in most passes
        vals = (None if vb == 0 else (_u64(n, 90 + i) if vb == 8
                                      else np.arange(n, dtype=np.uint32)))
        kout = np.full(max(n, 1), 0x5A, keys.dtype)
        vout = None if vb == 0 else np.full(max(n, 1), 0x5A, vals.dtype)
        skeep.append((keys, keys.copy(), vals, kout, vout))
        sitems.append(gms.gms_batch_sort_item(p(keys) if n else None, p(vals), p(kout), p(vout), n,
                                              bits, kb, vb if vb else 4))
        if kb == 8:
            swant.append(oracle.sort_wide(keys, bits, vals))
        else:   # 32-bit keys, 64-bit values: the stable order applied to the values
            wk, perm = oracle.sort(keys, bits, np.arange(n, dtype=np.uint32))
            swant.append((wk, vals[perm]))
    pkeep, pitems, pwant = [], [], []
    if with_plain:
        for i, (n, B, f, kind, param, kv) in enumerate(BATCH_CASES[3:7]):
            keys = W.uniform_keys(n, seed=120 + i)
            vals = np.arange(n, dtype=np.uint32)
            kout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
            vout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
            offs = np.full(B, 0xFFFFFFFF, np.uint32)
            pkeep.append((keys, vals, kout, vout, offs, kv))
            pitems.append(gms.gms_batch_item(p(keys), p(vals) if kv else None, p(kout),
                                             p(vout) if kv else None, n, B, f.desc(), p(offs)))
            pwant.append(oracle.multisplit(keys, B, kind, param, vals if kv else None))
    h = ctypes.c_void_p()
    assert emu.gms_batch_create(ctypes.byref(h)) == 0
    try:
        parr = (gms.gms_batch_item * max(len(pitems), 1))(*pitems)
        warr = (gms.gms_batch_wide_item * len(witems))(*witems)
        sarr = (gms.gms_batch_sort_item * len(sitems))(*sitems)
        assert emu.gms_batch_plan_wide(h, parr, len(pitems), warr, len(witems), sarr, len(sitems),
                                       None) == 0
        assert emu.gms_batch_run(h, None) == 0
    finally:
        emu.gms_batch_destroy(h)
    for i, ((keys, vals, kout, vout, offs), (wk, wv, wo)) in enumerate(zip(keep, wwant)):
        n = keys.size
        np.testing.assert_array_equal(kout[:n], wk, err_msg=f"wide item {i} keys")
        np.testing.assert_array_equal(offs, wo, err_msg=f"wide item {i} offsets")
        if vals is not None:
            np.testing.assert_array_equal(vout[:n], wv, err_msg=f"wide item {i} values")
    for i, ((keys, keys0, vals, kout, vout), (wk, wv)) in enumerate(zip(skeep, swant)):
        n = keys.size
        np.testing.assert_array_equal(keys, keys0, err_msg=f"wide sort {i} input written")
        np.testing.assert_array_equal(kout[:n], wk, err_msg=f"wide sort {i} keys")
        if vals is not None:
            np.testing.assert_array_equal(vout[:n], wv, err_msg=f"wide sort {i} values")
    for i, ((keys, vals, kout, vout, offs, kv), (wk, wv, wo)) in enumerate(zip(pkeep, pwant)):
        n = keys.size
        np.testing.assert_array_equal(kout[:n], wk, err_msg=f"plain item {i} keys")
        np.testing.assert_array_equal(offs, wo, err_msg=f"plain item {i} offsets")
        if kv:
            np.testing.assert_array_equal(vout[:n], wv, err_msg=f"plain item {i} values")


def test_emulated_batch_wide_validation(emu):
    a = np.zeros(64, np.uint64)
    b = np.zeros(64, np.uint64)
    h = ctypes.c_void_p()
    assert emu.gms_batch_create(ctypes.byref(h)) == 0
    rng = gms.gms_bucket_desc64(0, 0, 0)

    def plan(*items):
        arr = (gms.gms_batch_wide_item * len(items))(*items)
        return emu.gms_batch_plan_wide(h, None, 0, arr, len(items), None, 0, None)
    try:
        W_ = gms.gms_batch_wide_item
        assert plan(W_(p(a), None, p(b), None, 64, 257, 8, 0, rng, None)) == 2      # buckets
        assert plan(W_(p(a), None, p(b), None, 64, 8, 2, 0, rng, None)) == 1        # key bytes
        assert plan(W_(p(a) + 4, None, p(b), None, 32, 8, 8, 0, rng, None)) == 1    # alignment
        assert plan(W_(p(a), p(a), p(b), None, 64, 8, 8, 8, rng, None)) == 1        # value out
        assert plan(W_(p(a), None, p(a) + 256, None, 64, 8, 8, 0, rng, None)) == 7  # in/out overlap
        assert plan(W_(p(a), None, p(b), None, 32, 8, 8, 0, gms.gms_bucket_desc64(2, 4, 64),
                       None)) == 1                                                   # digit shift
        assert plan(W_(p(a), None, p(b), None, 64, 8, 8, 0, rng, None)) == 0
    finally:
        emu.gms_batch_destroy(h)


# ---------------------------------------------------------------------------
# 64-bit requests on the emulated server: multisplits and sorts of every
# element shape, interleaved with 32-bit ones, each against the oracle.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("max_inflight", [1, 2])
def test_emulated_server_wide_requests(emu, max_inflight):
    g = W.rng(700 + max_inflight)
    reqs = []
    shapes = [(8, 0), (8, 4), (8, 8), (4, 8), (4, 4), (4, 0)]
    for i in range(16):
        kb, vb = shapes[i % len(shapes)]
        n = int(g.integers(0, 6000))
        keys = _u64(n, 900 + i) if kb == 8 else W.uniform_keys(n, seed=900 + i)
        vals = None if vb == 0 else (_u64(n, 950 + i) if vb == 8 else np.arange(n, dtype=np.uint32))
        kout = np.full(max(n, 1), 0x5A, keys.dtype)
        vout = None if vb == 0 else np.full(max(n, 1), 0x5A, vals.dtype)
        if i % 2:
            bits = (8, 5, 6)[i % 3]
            req = gms.gms_request(1, p(keys) if n else None, p(vals), p(kout), p(vout), n, 0,
                                  gms.gms_bucket_desc(0, 0, 0), bits, None, kb, vb or 4,
                                  gms.gms_bucket_desc64(0, 0, 0))
            if kb == 8:
                wk, wv = oracle.sort_wide(keys, bits, vals)
            else:
                wk, perm = oracle.sort(keys, bits, np.arange(n, dtype=np.uint32))
                wv = None if vals is None else vals[perm]
            want = (wk, wv, None)
            offs = None
        else:
            B = (2, 16, 64, 256, 100)[i % 5]
            offs = np.full(B, 0xFFFFFFFF, np.uint32)
            desc32 = gms.range_bucket(B).desc() if kb == 4 and vb != 8 else gms.gms_bucket_desc(0, 0, 0)
            req = gms.gms_request(0, p(keys) if n else None, p(vals), p(kout), p(vout), n, B, desc32,
                                  0, p(offs), kb, vb or 4, gms.range_bucket(B).desc64())
            want = oracle.multisplit_wide(keys, B, oracle.RANGE, 0, vals, 0)
        reqs.append((req, (keys, vals, kout, vout, offs), want, n))
    h = _emu_server(emu, 64 << 20, max_inflight)
    try:
        tickets = []
        for req, *_ in reqs:
            t = ctypes.c_uint64()
            assert emu.gms_server_submit(h, ctypes.byref(req), ctypes.byref(t)) == 0
            tickets.append(t.value)
        for t in tickets:
            assert emu.gms_server_wait(h, t, None) == 0
    finally:
        assert emu.gms_server_destroy(h) == 0
    for i, (_req, (keys, vals, kout, vout, offs), (wk, wv, wo), n) in enumerate(reqs):
        np.testing.assert_array_equal(kout[:n], wk, err_msg=f"request {i} keys")
        if vals is not None:
            np.testing.assert_array_equal(vout[:n], wv, err_msg=f"request {i} values")
        if offs is not None:
            np.testing.assert_array_equal(offs, wo, err_msg=f"request {i} offsets")
