"""Served workload: a stream of mixed multisplit and radix-sort requests fed
to the continuous-batching server (``gpumultisplit_amd.Server``).

Request mix: each request is a multisplit with B in {2, 16, 64, 256} buckets
(range identifier) or an 8-bit-digit radix sort, key-only or key-value
(value = original index), with log-uniform sizes and uniform or Zipf(s=1.1)
keys.  Keys are generated on the device, so inputs are resident in HBM
before the first request arrives.  ``run_open_loop`` submits requests at
given arrival times (open loop: arrivals do not wait for completions) and
collects the server's submit-to-completion latencies; ``run_steady`` replays
the request set at a Poisson rate for a stated duration, rotating the
requests over several copies of their output buffers so that no two
requests in flight write the same memory.
"""
from __future__ import annotations

import time

import numpy as np

from . import workloads as W


def request_specs(count: int, seed: int, lo_log2: float = 10, hi_log2: float = 24):
    g = W.rng(seed)
    sizes = W.log_uniform_sizes(count, lo_log2, hi_log2, seed + 1)
    specs = []
    for i in range(count):
        op = "sort" if g.random() < 0.5 else "multisplit"
        specs.append({
            "op": op,
            "n": int(sizes[i]),
            "B": 256 if op == "sort" else int(g.choice([2, 16, 64, 256])),
            "kv": bool(g.random() < 0.5),
            "dist": "zipf" if g.random() < 0.5 else "uniform",
        })
    return specs


def device_keys(torch, n: int, dist: str, num_buckets: int, gen, dev):
    """n uint32 keys (as int32) on the device: uniform, or range-identifier
    buckets ~ Zipf(1.1) over num_buckets with uniform bits below them."""
    if dist == "uniform" or num_buckets < 2:
        return torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=dev, generator=gen)
    lb = W.ceil_log2(num_buckets)
    w = 1.0 / np.power(np.arange(1, num_buckets + 1, dtype=np.float64), 1.1)
    cdf = torch.tensor(np.cumsum(w / w.sum()), dtype=torch.float64, device=dev)
    u = torch.rand(n, dtype=torch.float64, device=dev, generator=gen)
    b = torch.clamp(torch.searchsorted(cdf, u), max=num_buckets - 1).to(torch.int64)
    low = torch.randint(0, 1 << (32 - lb), (n,), dtype=torch.int64, device=dev, generator=gen)
    k = (b << (32 - lb)) | low
    return (k - (k >= (1 << 31)).to(torch.int64) * (1 << 32)).to(torch.int32)


def make_requests(gms, torch, specs, dev, seed: int):
    """Device buffers and a gms_request for every spec (buffers kept in the
    returned records so they stay alive)."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    recs = []
    for s in specs:
        n = s["n"]
        kin = device_keys(torch, n, s["dist"], s["B"], gen, dev)
        kout = torch.empty_like(kin)
        vin = torch.arange(n, dtype=torch.int32, device=dev) if s["kv"] else None
        vout = torch.empty_like(kin) if s["kv"] else None
        offs = torch.empty(s["B"], dtype=torch.int32, device=dev) if s["op"] == "multisplit" else None
        if s["op"] == "sort":
            req = gms.sort_request(kin, kout, n, 8, d_value_in=vin, d_value_out=vout)
        else:
            req = gms.multisplit_request(kin, kout, n, s["B"], gms.range_bucket(s["B"]),
                                         d_value_in=vin, d_value_out=vout, d_offsets=offs)
        recs.append({"spec": s, "req": req, "kin": kin, "kout": kout, "vin": vin, "vout": vout,
                     "offs": offs})
    torch.cuda.synchronize()
    return recs


def run_open_loop(server, recs, arrivals_s):
    """Submits recs[i] at time arrivals_s[i] after the start (busy-wait
    pacing), then waits for all.  Returns (latencies_us, wall_s)."""
    t0 = time.perf_counter()
    tickets = []
    for rec, ta in zip(recs, arrivals_s):
        target = t0 + ta
        while time.perf_counter() < target:
            pass
        tickets.append(server.submit(rec["req"]))
    lats = [server.wait(t) for t in tickets]
    return np.array(lats), time.perf_counter() - t0


def output_copies(gms, torch, recs, copies: int):
    """`copies` sets of the requests: the first is recs itself, the others
    read the same inputs and write their own output buffers."""
    sets = [recs]
    for _ in range(copies - 1):
        out = []
        for r in recs:
            s = r["spec"]
            kout = torch.empty_like(r["kin"])
            vout = torch.empty_like(r["kin"]) if s["kv"] else None
            offs = torch.empty_like(r["offs"]) if r["offs"] is not None else None
            if s["op"] == "sort":
                req = gms.sort_request(r["kin"], kout, s["n"], 8, d_value_in=r["vin"],
                                       d_value_out=vout)
            else:
                req = gms.multisplit_request(r["kin"], kout, s["n"], s["B"], gms.range_bucket(s["B"]),
                                             d_value_in=r["vin"], d_value_out=vout, d_offsets=offs)
            out.append({"spec": s, "req": req, "kin": r["kin"], "kout": kout, "vin": r["vin"],
                        "vout": vout, "offs": offs})
        sets.append(out)
    torch.cuda.synchronize()
    return sets


def run_steady(server, rec_sets, rate_per_s: float, duration_s: float, seed: int):
    """Open-loop Poisson arrivals at rate_per_s for duration_s (at least one
    pass over the request set): request k is rec_sets[(k // R) % C][k % R].
    Before a request's buffers are used again, the previous request on them
    must have completed; the number of times that had to be waited for is
    returned (0 when the copies keep uses apart).  Returns a dict with the
    latencies (us), the wall time from the first submit to the last
    completion, the submitted count and total keys."""
    R, C = len(rec_sets[0]), len(rec_sets)
    count = max(R, int(round(rate_per_s * duration_s)))
    arr = poisson_arrivals(count, rate_per_s, seed)
    last = {}       # (copy, index) -> ticket of the last request on those buffers
    lat_of = {}     # ticket -> latency, for tickets already waited for
    reuse_waits = 0
    tickets = []
    keys = 0
    t0 = time.perf_counter()
    for k in range(count):
        i, c = k % R, (k // R) % C
        prev = last.get((c, i))
        if prev is not None and prev not in lat_of:
            done, lat = server.poll(prev)
            if not done:
                reuse_waits += 1
            lat_of[prev] = server.wait(prev)
        target = t0 + arr[k]
        while time.perf_counter() < target:
            pass
        t = server.submit(rec_sets[c][i]["req"])
        last[(c, i)] = t
        tickets.append(t)
        keys += rec_sets[c][i]["spec"]["n"]
    for t in tickets:
        if t not in lat_of:
            lat_of[t] = server.wait(t)
    wall = time.perf_counter() - t0
    return {"lat_us": np.array([lat_of[t] for t in tickets]), "wall_s": wall,
            "requests": count, "keys": keys, "reuse_waits": reuse_waits,
            "arrival_span_s": float(arr[-1])}


def poisson_arrivals(count: int, rate_per_s: float, seed: int):
    gaps = W.rng(seed).exponential(1.0 / rate_per_s, size=count)
    return np.cumsum(gaps) - gaps[0]


def percentile(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else 0.0
