"""GPU parity of the continuous-batching request server (gms_server_*):
every request's result in the mix equals the same request run alone through
the plain entry points (themselves checked against the oracle), and small
requests equal the oracle directly; requests submitted from several threads
and under a budget that admits only a few at a time all complete.
"""
import threading

import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import served as S
from gpu_util import banner, from_dev

pytestmark = pytest.mark.gpu


def _alone(dev, rec):
    """The request run alone through the plain API."""
    import torch
    s = rec["spec"]
    n = s["n"]
    kout = torch.empty_like(rec["kin"])
    vout = torch.empty_like(rec["kin"]) if s["kv"] else None
    offs = None
    if s["op"] == "sort":
        ctx = gms.ms_sort_context()
        if s["kv"]:
            gms.multisplit_sort_8bit_allocate_pairs(n, ctx)
            gms.multisplit_sort_8bit_pairs(rec["kin"], kout, rec["vin"], vout, n, ctx)
        else:
            gms.multisplit_sort_8bit_allocate(n, ctx)
            gms.multisplit_sort_8bit(rec["kin"], kout, n, ctx)
        torch.cuda.synchronize()
        gms.multisplit_sort_release_memory(ctx)
    else:
        ctx = gms.multisplit_context(s["B"])
        gms.multisplit_allocate_key_value(n, ctx)
        offs = torch.empty(s["B"], dtype=torch.int32, device=dev)
        gms.multisplit(rec["kin"], kout, n, ctx, gms.range_bucket(s["B"]), d_value_in=rec["vin"],
                       d_value_out=vout, d_offsets=offs)
        torch.cuda.synchronize()
        gms.multisplit_release_memory(ctx)
    return kout, vout, offs


def _same_as_alone(dev, rec, i):
    import torch
    kout, vout, offs = _alone(dev, rec)
    n = rec["spec"]["n"]
    assert torch.equal(kout[:n], rec["kout"][:n]), f"request {i} keys differ from the call alone"
    if vout is not None:
        assert torch.equal(vout[:n], rec["vout"][:n]), f"request {i} values differ"
    if offs is not None:
        assert torch.equal(offs, rec["offs"]), f"request {i} offsets differ"


def _oracle_check(rec, i):
    s = rec["spec"]
    keys = from_dev(rec["kin"])
    vals = np.arange(s["n"], dtype=np.uint32) if s["kv"] else None
    if s["op"] == "sort":
        wk, wv = oracle.sort(keys, 8, vals)
        wo = None
    else:
        wk, wv, wo = oracle.multisplit(keys, s["B"], oracle.RANGE, 0, vals)
    np.testing.assert_array_equal(from_dev(rec["kout"])[:s["n"]], wk, err_msg=f"request {i}")
    if s["kv"]:
        np.testing.assert_array_equal(from_dev(rec["vout"])[:s["n"]], wv, err_msg=f"request {i}")
    if wo is not None:
        np.testing.assert_array_equal(from_dev(rec["offs"]), wo, err_msg=f"request {i}")


@pytest.mark.parametrize("max_inflight", [1, 2])
def test_server_mixed_burst_matches_alone_and_oracle(dev, max_inflight):
    banner(f"server burst inflight={max_inflight}")
    specs = S.request_specs(60, seed=max_inflight, lo_log2=0, hi_log2=18)
    recs = S.make_requests(gms, __import__("torch"), specs, dev, seed=3)
    with gms.Server(1 << 30, max_inflight_steps=max_inflight) as srv:
        tickets = [srv.submit(r["req"]) for r in recs]
        lats = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert all(x > 0 for x in lats)
    assert st["completed"] == len(recs) and st["active"] == 0 and st["pending"] == 0
    assert st["steps"] < sum(4 if s["op"] == "sort" else 1 for s in specs)  # batched, not 1:1
    for i, r in enumerate(recs):
        _same_as_alone(dev, r, i)
        if r["spec"]["n"] <= 20000:
            _oracle_check(r, i)


def test_server_open_loop_and_threads(dev):
    import torch
    banner("server open loop + 2 submitter threads")
    specs = S.request_specs(80, seed=11, lo_log2=8, hi_log2=20)
    recs = S.make_requests(gms, torch, specs, dev, seed=12)
    with gms.Server(4 << 30) as srv:
        # first half: Poisson arrivals from this thread
        half = len(recs) // 2
        lats, _ = S.run_open_loop(srv, recs[:half], S.poisson_arrivals(half, 2000.0, seed=1))
        assert len(lats) == half
        # second half: two threads submitting concurrently
        out = {}

        def worker(part, key):
            ts = [srv.submit(r["req"]) for r in part]
            out[key] = [srv.wait(t) for t in ts]

        th = [threading.Thread(target=worker, args=(recs[half::2], 0)),
              threading.Thread(target=worker, args=(recs[half + 1::2], 1))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert len(out[0]) + len(out[1]) == len(recs) - half
    for i, r in enumerate(recs):
        _same_as_alone(dev, r, i)


def test_server_budget_admits_in_turn(dev):
    import torch
    banner("server small budget")
    # 2^22-key key-value sorts reserve ~35 MiB of scratch each; a 96 MiB
    # budget admits two at a time, the rest wait for a finished one.
    specs = [{"op": "sort", "n": 1 << 22, "B": 256, "kv": True, "dist": "uniform"} for _ in range(5)]
    specs += [{"op": "multisplit", "n": 5000, "B": 16, "kv": False, "dist": "zipf"}]
    recs = S.make_requests(gms, torch, specs, dev, seed=21)
    with gms.Server(96 << 20) as srv:
        tickets = [srv.submit(r["req"]) for r in recs]
        for t in tickets:
            srv.wait(t)
        st = srv.stats()
        big = torch.empty(1 << 26, dtype=torch.int32, device=dev)
        with pytest.raises(gms.GmsError):        # 2^25-key sort cannot fit 96 MiB
            srv.submit(gms.sort_request(big[:1 << 25], big[1 << 25:], 1 << 25))
    assert st["peak_scratch_bytes"] <= 96 << 20
    for i, r in enumerate(recs):
        _same_as_alone(dev, r, i)


@pytest.mark.slow
def test_server_served_workload_matches_alone(dev):
    """The served workload's request mix (sizes log-uniform [2^10, 2^24])
    at a Poisson arrival rate, 64 GiB scratch budget; every request equals
    the call alone."""
    import torch
    specs = S.request_specs(300, seed=2024)
    banner(f"served workload, {sum(s['n'] for s in specs)} keys")
    recs = S.make_requests(gms, torch, specs, dev, seed=2025)
    with gms.Server(64 << 30) as srv:
        lats, wall = S.run_open_loop(srv, recs, S.poisson_arrivals(len(recs), 3000.0, seed=5))
    print(f"p50 {S.percentile(lats, 50):.1f} us p99 {S.percentile(lats, 99):.1f} us wall {wall:.3f} s")
    for i, r in enumerate(recs):
        _same_as_alone(dev, r, i)


def test_server_wide_requests_match_oracle(dev):
    """64-bit requests of every element shape -- multisplits and sorts --
    submitted together with 32-bit ones; each result equals the oracle's
    64-bit multisplit or sort (32-bit keys carrying 64-bit values: the oracle's
    stable order applied to the values)."""
    import torch
    banner("server wide requests")
    g = np.random.default_rng(77)
    shapes = [(8, 0), (8, 4), (8, 8), (4, 8), (4, 4), (4, 0)]
    recs = []

    def dev_arr(a):
        t = torch.from_numpy(np.ascontiguousarray(a).view(np.int64 if a.dtype == np.uint64 else np.int32))
        return t.to(dev)

    for i in range(24):
        kb, vb = shapes[i % len(shapes)]
        n = int(2 ** g.uniform(0, 20))
        keys = (g.integers(0, 2**64, size=n, dtype=np.uint64) if kb == 8
                else g.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32))
        vals = None if vb == 0 else (g.integers(0, 2**64, size=n, dtype=np.uint64) if vb == 8
                                     else np.arange(n, dtype=np.uint32))
        kin = dev_arr(keys)
        kout = torch.full_like(kin, 0x5A)
        vin = None if vals is None else dev_arr(vals)
        vout = None if vin is None else torch.full_like(vin, 0x5A)
        if i % 2:
            bits = (8, 6, 5)[i % 3]
            req = gms.sort_request(kin, kout, n, bits, d_value_in=vin, d_value_out=vout)
            if kb == 8:
                wk, wv = oracle.sort_wide(keys, bits, vals)
            else:
                wk, perm = oracle.sort(keys, bits, np.arange(n, dtype=np.uint32))
                wv = None if vals is None else vals[perm]
            wo, offs = None, None
        else:
            B = (2, 16, 64, 256)[i % 4]
            offs = torch.full((B,), -1, dtype=torch.int32, device=dev)
            req = gms.multisplit_request(kin, kout, n, B, gms.range_bucket(B), d_value_in=vin,
                                         d_value_out=vout, d_offsets=offs)
            wk, wv, wo = oracle.multisplit_wide(keys, B, oracle.RANGE, 0, vals, 0)
        # the request holds raw pointers: every buffer stays referenced here
        # until the request has completed
        recs.append((req, kout, vout, offs, wk, wv, wo, n, kb, vb, kin, vin))
    with gms.Server(1 << 30, max_inflight_steps=2) as srv:
        tickets = [srv.submit(r[0]) for r in recs]
        for t in tickets:
            srv.wait(t)
    for i, (_req, kout, vout, offs, wk, wv, wo, n, kb, vb, _kin, _vin) in enumerate(recs):
        ko = kout.cpu().numpy().view(np.uint64 if kb == 8 else np.uint32)
        np.testing.assert_array_equal(ko[:n], wk, err_msg=f"request {i} keys")
        if vout is not None:
            vo = vout.cpu().numpy().view(np.uint64 if vb == 8 else np.uint32)
            np.testing.assert_array_equal(vo[:n], wv, err_msg=f"request {i} values")
        if offs is not None:
            np.testing.assert_array_equal(from_dev(offs), wo, err_msg=f"request {i} offsets")
