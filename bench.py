#!/usr/bin/env python3
"""Benchmark of the plain multisplit hot path on one MI355X.

Default workload: stable key-value multisplit of 2^27 uint32 keys (uniform over
[0, 2^32)) with uint32 values (value = original index) into 256 buckets by the
range identifier floor(key * 256 / 2^32).  A step is one full multisplit call
(prescan -> scan -> postscan) with inputs already resident in HBM.

Prints ONE JSON line: whole-job throughput in Gkeys/s, the postscan kernel's
achieved HBM bandwidth against the 8 TB/s peak (algorithmic bytes: 16 B/key,
key+value read and written once; the whole call moves 20 B/key), and the CPU
oracle timed on the host as a reported baseline.

Options other than the driver's (--gpus/--steps/--warmup) are for
measurement runs: --workload picks another configuration, --no-cpu skips the
CPU baseline, --no-stage-events times without per-stage HIP events.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", default="ms_kv_256",
                   help="ms_{kv,ko}_<B>[_<dist>][_n<log2n>] | sort{7,8}_{kv,ko} | "
                        "hist256 | histrange256 | compact{50,5} | copy | wide_{kv,ko}_256 | wide_sort8_kv | wide_hist256 | "
                        "variant_sweep")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-stage-events", action="store_true")
    p.add_argument("--cpu-reps", type=int, default=5)
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run one plain copy of n u32 before warm-up (HBM counter calibration)")
    p.add_argument("--no-batched", action="store_true", help="skip the batched-form measurement")
    p.add_argument("--batch-reps", type=int, default=5)
    p.add_argument("--no-served", action="store_true", help="skip the served-form measurement")
    p.add_argument("--served-requests", type=int, default=400)
    return p.parse_args()


def bench_served(args, torch, gms, dev):
    """Served form: a mixed request stream (multisplit B in {2,16,64,256} or
    8-bit radix sort, key-only or key-value, sizes log-uniform [2^10, 2^24],
    uniform or Zipf 1.1 keys) through the continuous-batching server with a
    64 GiB scratch budget.  Capacity is measured with all requests arriving
    at once (median of three runs after one untimed full-size run, whose time
    is reported as first_run_ms); then Poisson arrivals at 30/60/90 % of it,
    three runs each, reporting throughput and p50/p99 latency of the run
    with the median p99; then the steady state: the request set replayed at
    each rate of a ladder (30 % .. 130 % of capacity) for at least 0.5 s of
    arrivals, and the highest rate whose achieved rate is at least 98 % of
    the offered one.  The same requests as back-to-back plain calls give the
    comparison throughput."""
    import numpy as np
    from gpumultisplit_amd import served as S
    specs = S.request_specs(args.served_requests, seed=77)
    total = sum(s["n"] for s in specs)
    print(f"[bench] served: {len(specs)} requests, {total} keys", file=sys.stderr, flush=True)
    recs = S.make_requests(gms, torch, specs, dev, seed=78)
    out = {"workload": f"{len(specs)} requests: multisplit B in {{2,16,64,256}} or 8-bit sort, "
                       "KO/KV, sizes log-uniform [2^10,2^24], uniform/Zipf 1.1 keys",
           "requests": len(specs), "total_keys": total,
           "sorts": sum(1 for s in specs if s["op"] == "sort"), "scratch_budget_gib": 64}
    with gms.Server(64 << 30, max_inflight_steps=2) as srv:
        # the first full-size run after start-up (device pages and caches
        # cold), then the median of three steady-state runs
        _, wall_first = S.run_open_loop(srv, recs, np.zeros(len(recs)))
        runs = sorted((S.run_open_loop(srv, recs, np.zeros(len(recs))) for _ in range(3)),
                      key=lambda r: r[1])
        lat0, wall0 = runs[1]
        cap = len(recs) / wall0
        out["capacity"] = {"req_per_s": round(cap, 1), "gkeys_per_s": round(total / wall0 / 1e9, 3),
                           "p50_us": round(S.percentile(lat0, 50), 1),
                           "p99_us": round(S.percentile(lat0, 99), 1),
                           "runs_ms": [round(r[1] * 1e3, 3) for r in runs],
                           "first_run_ms": round(wall_first * 1e3, 3)}
        rates = []
        for i, frac in enumerate((0.3, 0.6, 0.9)):
            # three runs of the same arrival times; the one with the median p99
            # is reported, with every run's p99 beside it
            arr = S.poisson_arrivals(len(recs), frac * cap, seed=100 + i)
            runs = sorted((S.run_open_loop(srv, recs, arr) for _ in range(3)),
                          key=lambda r: S.percentile(r[0], 99))
            lat, wall = runs[1]
            rates.append({"offered_req_per_s": round(frac * cap, 1),
                          "achieved_req_per_s": round(len(recs) / wall, 1),
                          "gkeys_per_s": round(total / wall / 1e9, 3),
                          "p50_us": round(S.percentile(lat, 50), 1),
                          "p99_us": round(S.percentile(lat, 99), 1),
                          "p99_us_runs": [round(S.percentile(r[0], 99), 1) for r in runs]})
        out["poisson"] = rates
        # Steady state: the request set replayed at each Poisson rate for at
        # least 0.5 s of arrivals, over three copies of the output buffers.
        sets = S.output_copies(gms, torch, recs, 3)
        ladder = []
        for j, frac in enumerate((0.3, 0.6, 0.8, 0.9, 1.0, 1.1, 1.2, 1.3)):
            r = S.run_steady(srv, sets, frac * cap, 0.5, seed=300 + j)
            ladder.append({"offered_req_per_s": round(frac * cap, 1),
                           "achieved_req_per_s": round(r["requests"] / r["wall_s"], 1),
                           "gkeys_per_s": round(r["keys"] / r["wall_s"] / 1e9, 3),
                           "p50_us": round(S.percentile(r["lat_us"], 50), 1),
                           "p99_us": round(S.percentile(r["lat_us"], 99), 1),
                           "requests": r["requests"], "arrivals_s": round(r["arrival_span_s"], 3),
                           "wall_s": round(r["wall_s"], 3), "buffer_reuse_waits": r["reuse_waits"]})
        del sets
        out["steady"] = {"min_arrival_s": 0.5, "output_copies": 3, "ladder": ladder}
        st = srv.stats()
        out["steps"] = st["steps"]
        out["passes_per_step"] = round(st["step_items"] / max(st["steps"], 1), 2)
        out["peak_scratch_gib"] = round(st["peak_scratch_bytes"] / (1 << 30), 3)

    # the same requests as back-to-back plain calls on one stream
    nmax = max(s["n"] for s in specs)
    sort_ctx = gms.ms_sort_context()
    gms.multisplit_sort_8bit_allocate_pairs(nmax, sort_ctx)
    ms_ctx = {}
    for B in (2, 16, 64, 256):
        ms_ctx[B] = gms.multisplit_context(B)
        gms.multisplit_allocate_key_value(nmax, ms_ctx[B])
    L = gms.lib()
    stream = torch.cuda.current_stream(dev).cuda_stream

    def plain_all():
        for r in recs:
            s = r["spec"]
            vin = r["vin"].data_ptr() if s["kv"] else None
            vout = r["vout"].data_ptr() if s["kv"] else None
            if s["op"] == "sort":
                st_ = L.gms_sort(sort_ctx.handle, r["kin"].data_ptr(), r["kout"].data_ptr(), vin,
                                 vout, s["n"], 8, stream)
            else:
                st_ = L.gms_multisplit(ms_ctx[s["B"]].handle, r["kin"].data_ptr(), vin,
                                       r["kout"].data_ptr(), vout, s["n"],
                                       gms.range_bucket(s["B"]).desc(), r["offs"].data_ptr(),
                                       stream)
            if st_ != 0:
                raise gms.GmsError(st_, "plain call")

    plain_all()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_all()
    torch.cuda.synchronize()
    wall_plain = time.perf_counter() - t0
    out["plain_back_to_back"] = {"gkeys_per_s": round(total / wall_plain / 1e9, 3),
                                 "ms": round(wall_plain * 1e3, 3)}
    out["capacity_speedup_vs_plain"] = round(wall_plain / wall0, 3)
    plain_gkeys = total / wall_plain / 1e9
    ok = [x for x in out["steady"]["ladder"]
          if x["achieved_req_per_s"] >= 0.98 * x["offered_req_per_s"]]
    if ok:
        best = max(ok, key=lambda x: x["offered_req_per_s"])
        out["steady"]["max_sustained"] = dict(best, speedup_vs_plain=round(best["gkeys_per_s"] / plain_gkeys, 3))
    gms.multisplit_sort_release_memory(sort_ctx)
    for c in ms_ctx.values():
        gms.multisplit_release_memory(c)
    return out


def _check_status(st, gms, what):
    if st != 0:
        raise gms.GmsError(st, what)


def bench_batched(args, torch, gms, W, dev):
    """Batched form: 4096 independent key-value multisplits, sizes round(2^U),
    U ~ U[8, 22], 64 buckets by range, uniform keys, all in HBM.  Times the
    autotuned batch replayed from a hipGraph against the same 4096 instances
    as back-to-back plain calls on one stream."""
    import numpy as np
    count, B = 4096, 64
    sizes = W.log_uniform_sizes(count, 8, 22, seed=4096)
    total = int(sizes.sum())
    print(f"[bench] batched: {count} instances, {total} keys", file=sys.stderr, flush=True)
    # instances packed into one allocation, each starting on a 16-byte boundary
    # (as separately allocated buffers would)
    starts = np.zeros(count, dtype=np.int64)
    for i in range(1, count):
        starts[i] = (starts[i - 1] + sizes[i - 1] + 3) // 4 * 4
    span = int(starts[-1] + sizes[-1])
    g = torch.Generator(device=dev)
    g.manual_seed(4096)
    kin = torch.randint(-(1 << 31), 1 << 31, (span,), dtype=torch.int32, device=dev, generator=g)
    vin = torch.arange(span, dtype=torch.int32, device=dev)
    kout = torch.empty_like(kin)
    vout = torch.empty_like(kin)
    f = gms.range_bucket(B)
    items = [gms.BatchItem(kin[s:s + n], kout[s:s + n], int(n), B, f, d_value_in=vin[s:s + n],
                           d_value_out=vout[s:s + n]) for s, n in zip(starts, sizes)]
    batch = gms.Batch(items)
    tuned = batch.autotune(reps=2)
    batch.capture()
    batch.replay()
    torch.cuda.synchronize()
    reps = args.batch_reps
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.replay()
    torch.cuda.synchronize()
    ms_batch = (time.perf_counter() - t0) * 1e3 / reps

    ctx = gms.multisplit_context(B)
    gms.multisplit_allocate_key_value(int(sizes.max()), ctx)
    L = gms.lib()
    stream = torch.cuda.current_stream(dev).cuda_stream
    desc = f.desc()
    calls = [(kin[s:s + n].data_ptr(), vin[s:s + n].data_ptr(), kout[s:s + n].data_ptr(),
              vout[s:s + n].data_ptr(), int(n)) for s, n in zip(starts, sizes)]

    def plain_all():
        for kp, vp, kop, vop, n in calls:
            st = L.gms_multisplit(ctx.handle, kp, vp, kop, vop, n, desc, None, stream)
            if st != 0:
                raise gms.GmsError(st, "gms_multisplit")

    plain_all()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_all()
    torch.cuda.synchronize()
    ms_plain = (time.perf_counter() - t0) * 1e3 / reps
    gms.multisplit_release_memory(ctx)
    st = batch.stats()
    del batch, items

    # Batched radix sorts: 1024 key-value 8-bit-digit sorts of uneven sizes as
    # one batch (4 steps), against the same sorts as back-to-back gms_sort calls.
    scount = 1024
    ssizes = W.log_uniform_sizes(scount, 8, 20, seed=1024)
    sstarts = np.zeros(scount, dtype=np.int64)
    for i in range(1, scount):
        sstarts[i] = (sstarts[i - 1] + ssizes[i - 1] + 3) // 4 * 4
    stotal = int(ssizes.sum())
    print(f"[bench] batched sorts: {scount} sorts, {stotal} keys", file=sys.stderr, flush=True)
    sorts = [gms.BatchSortItem(kin[s:s + n], kout[s:s + n], int(n), 8, d_value_in=vin[s:s + n],
                               d_value_out=vout[s:s + n]) for s, n in zip(sstarts, ssizes)]
    sbatch = gms.Batch([], sorts=sorts)
    stuned = sbatch.autotune(reps=1)
    sbatch.capture()
    sbatch.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sbatch.replay()
    torch.cuda.synchronize()
    ms_sbatch = (time.perf_counter() - t0) * 1e3 / reps
    sctx = gms.ms_sort_context()
    _check_status(L.gms_sort_allocate(sctx.handle, int(ssizes.max()), 8, 1), gms, "gms_sort_allocate")
    scalls = [(kin[s:s + n].data_ptr(), kout[s:s + n].data_ptr(), vin[s:s + n].data_ptr(),
               vout[s:s + n].data_ptr(), int(n)) for s, n in zip(sstarts, ssizes)]

    def plain_sorts():
        for kp, kop, vp, vop, n in scalls:
            _check_status(L.gms_sort(sctx.handle, kp, kop, vp, vop, n, 8, stream), gms, "gms_sort")

    plain_sorts()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_sorts()
    torch.cuda.synchronize()
    ms_splain = (time.perf_counter() - t0) * 1e3 / reps
    sst = sbatch.stats()
    del sbatch, sorts
    batched_sorts = {
        "workload": f"{scount} key-value radix sorts (8-bit digits, 4 passes), sizes round(2^U) "
                    "U~U[8,20]",
        "total_keys": stotal, "sorts": scount,
        "ms_per_batch": round(ms_sbatch, 4), "gkeys_per_s": round(stotal / ms_sbatch / 1e6, 3),
        "plain_back_to_back_ms": round(ms_splain, 4),
        "plain_back_to_back_gkeys_per_s": round(stotal / ms_splain / 1e6, 3),
        "speedup_vs_plain": round(ms_splain / ms_sbatch, 3),
        "hbm_fraction": round(stotal * 80 / (ms_sbatch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "launches_per_batch": sst["launches"],
        "tuning": {"items_per_thread": stuned["items_per_thread"], "persistent": stuned["persistent"]},
    }
    return {
        "workload": f"{count} key-value multisplits, sizes round(2^U) U~U[8,22], {B} buckets (range), "
                    "16-byte-aligned instance buffers",
        "total_keys": total, "instances": count,
        "ms_per_batch": round(ms_batch, 4), "gkeys_per_s": round(total / ms_batch / 1e6, 3),
        "plain_back_to_back_ms": round(ms_plain, 4),
        "plain_back_to_back_gkeys_per_s": round(total / ms_plain / 1e6, 3),
        "speedup_vs_plain": round(ms_plain / ms_batch, 3),
        "hbm_fraction": round(total * 20 / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "launches_per_batch": st["launches"], "work_items": st["work_items"],
        "tuning": {"items_per_thread": tuned["items_per_thread"], "persistent": tuned["persistent"]},
        "hipgraph": True, "reps": reps,
        "sorts": batched_sorts,
    }


def bench_aux(args, torch, gms, W, dev):
    """The other hot-path rows, one JSON line each (measurement runs):
    sort{7,8}_{kv,ko}: radix sort of 2^25 uint32 keys (value = index);
    hist256: 256-bin histogram of the low 8 bits of 2^27 keys;
    histrange256: 256 custom bins (257 sorted random edges in [0, 1024)) of
      2^27 float samples uniform in [0, 1024), as the reference's range driver;
    compact{50,5}: 2-bucket compaction of 2^27 keys, predicate key < p*2^32;
    copy: the streaming copy kernel over 2^27 u32 (the measured copy ceiling);
    wide_kv_256 / wide_ko_256: 256-bucket range multisplit of 2^27 uint64 keys
      with (or without) uint64 values;
    wide_sort8_kv: 8-bit-digit radix sort of 2^25 uint64 keys with uint64
      values (8 passes);
    wide_hist256: 256-bin range histogram (top 8 bits) of 2^27 uint64 keys."""
    import numpy as np
    w = args.workload
    stream = torch.cuda.current_stream(dev).cuda_stream
    L = gms.lib()
    g = torch.Generator(device=dev)
    g.manual_seed(12345)
    if w.startswith("sort"):
        bits, kv = int(w[4]), w.endswith("kv")
        n = 1 << 25
        passes = {8: 4, 7: 5}[bits]
        bpk = passes * (20 if kv else 12)
        metric = f"{'key-value' if kv else 'key-only'} radix sort ({bits}-bit digits), {n} uint32 keys"
    elif w == "hist256":
        n, bpk, metric = 1 << 27, 4, "256-bin histogram of the low 8 bits, 134217728 uint32 keys"
    elif w == "histrange256":
        n, bpk, metric = 1 << 27, 4, "256-bin custom-edge histogram, 134217728 float samples"
    elif w.startswith("compact"):
        pct = int(w[7:])
        n, bpk = 1 << 27, 12
        metric = f"2-bucket compaction ({pct}% kept), 134217728 uint32 keys"
    elif w == "copy":
        n, bpk, metric = 1 << 27, 8, "device copy of 134217728 uint32"
    elif w in ("wide_kv_256", "wide_ko_256"):
        kv = w == "wide_kv_256"
        n, bpk = 1 << 27, (40 if kv else 24)
        metric = f"256-bucket multisplit, {n} uint64 keys{' + uint64 values' if kv else ''}"
    elif w == "wide_sort8_kv":
        n, bpk = 1 << 25, 8 * 40
        metric = f"key-value radix sort (8-bit digits), {n} uint64 keys + uint64 values"
    elif w == "wide_hist256":
        n, bpk, metric = 1 << 27, 8, "256-bin range histogram, 134217728 uint64 keys"
    else:
        raise SystemExit(f"unknown workload {w!r}")
    wide = w.startswith("wide")
    kdt = torch.int64 if wide else torch.int32
    kin = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=kdt, device=dev, generator=g)
    if wide:
        kin = kin * (1 << 32) + torch.randint(0, 1 << 32, (n,), dtype=kdt, device=dev, generator=g)
    kout = torch.empty_like(kin)
    if w.startswith("sort"):
        vin = torch.arange(n, dtype=torch.int32, device=dev) if kv else None
        vout = torch.empty_like(kin) if kv else None
        ctx = gms.ms_sort_context()
        L.gms_sort_allocate(ctx.handle, n, bits, int(kv))
        vp, vop = (vin.data_ptr(), vout.data_ptr()) if kv else (None, None)

        def step():
            st = L.gms_sort(ctx.handle, kin.data_ptr(), kout.data_ptr(), vp, vop, n, bits, stream)
            if st != 0:
                raise gms.GmsError(st, "gms_sort")
    elif w == "hist256":
        h = torch.zeros(256, dtype=torch.int32, device=dev)
        desc = gms.ms_bucket_sort_8bit(0).desc()

        def step():
            st = L.gms_histogram_u32(kin.data_ptr(), n, h.data_ptr(), 256, desc, stream)
            if st != 0:
                raise gms.GmsError(st, "gms_histogram_u32")
    elif w == "histrange256":
        h = torch.zeros(256, dtype=torch.int32, device=dev)
        kin = torch.rand(n, device=dev, generator=g) * 1024.0
        lv = torch.sort(torch.rand(257, device=dev, generator=g) * 1024.0).values
        lv[0], lv[-1] = 0.0, 1024.0

        def step():
            st = L.gms_histogram_range_f32(kin.data_ptr(), n, h.data_ptr(), 257, lv.data_ptr(),
                                           stream)
            if st != 0:
                raise gms.GmsError(st, "gms_histogram_range_f32")
    elif w.startswith("compact"):
        ctx = gms.compaction_context()
        gms.compaction_allocate_key_only(n, ctx)
        desc = gms.delta_func(int(pct / 100 * (1 << 32))).desc()
        cnt = torch.zeros(1, dtype=torch.int32, device=dev)

        def step():
            st = L.gms_compaction(ctx.handle, kin.data_ptr(), None, kout.data_ptr(), None, n, desc,
                                  None, cnt.data_ptr(), stream)
            if st != 0:
                raise gms.GmsError(st, "gms_compaction")
    elif w in ("wide_kv_256", "wide_ko_256"):
        ctx = gms.multisplit_context(256)
        gms.multisplit_allocate_key_value(n, ctx)
        vin = torch.arange(n, dtype=torch.int64, device=dev) if kv else None
        vout = torch.empty_like(kin) if kv else None
        desc = gms.range_bucket(256).desc64()
        vp, vop = (vin.data_ptr(), vout.data_ptr()) if kv else (None, None)

        def step():
            st = L.gms_multisplit_wide(ctx.handle, kin.data_ptr(), vp, kout.data_ptr(), vop, n, 8,
                                       8 if kv else 0, desc, None, None, stream)
            if st != 0:
                raise gms.GmsError(st, "gms_multisplit_wide")
    elif w == "wide_sort8_kv":
        ctx = gms.ms_sort_context()
        gms.multisplit_sort_8bit_allocate_pairs(n, ctx, key_bytes=8, value_bytes=8)
        vin = torch.arange(n, dtype=torch.int64, device=dev)
        vout = torch.empty_like(vin)

        def step():
            st = L.gms_sort_wide(ctx.handle, kin.data_ptr(), kout.data_ptr(), vin.data_ptr(),
                                 vout.data_ptr(), n, 8, 8, stream)
            if st != 0:
                raise gms.GmsError(st, "gms_sort_wide")
    elif w == "wide_hist256":
        h = torch.zeros(256, dtype=torch.int32, device=dev)
        desc = gms.range_bucket(256).desc64()

        def step():
            st = L.gms_histogram_wide(kin.data_ptr(), n, 8, h.data_ptr(), 256, desc, stream)
            if st != 0:
                raise gms.GmsError(st, "gms_histogram_wide")
    else:
        def step():
            gms.copy_u32(kin, kout, n)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # GPU time of the calls alone: HIP events on the stream the library
    # launches on, around a second pass of the same K calls
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    gpu_ms = e0.elapsed_time(e1) / args.steps
    achieved = n * bpk / (gpu_ms * 1e-3) / 1e9
    out = {"metric": metric, "value": round(n / ms / 1e6, 3), "unit": "Gkeys/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "u64" if wide else "u32",
           "data": "synthetic (uniform keys)", "config": {"workload": w, "n": n},
           "call_hbm_fraction": round(n * bpk / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "algorithmic_bytes_per_key": bpk,
           "roofline": {"kernel": "all kernels of one call", "bound": "hbm",
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "algorithmic_bytes": n * bpk, "avg_launch_us": round(gpu_ms * 1e3, 2)},
           "cpu_baseline": None if args.no_cpu else cpu_baseline_aux(w, n, args.cpu_reps)}
    print(json.dumps(out), flush=True)


def cpu_baseline_aux(w, n, reps):
    """The host-side baseline of an auxiliary row, single-threaded, median of
    `reps` runs on uniform keys from the host generator (the same
    distribution the GPU row uses): std::stable_sort for the sorts, a counting
    loop for the histogram, std::stable_partition for compaction (the
    reference drivers' CPU comparisons, oracle/baseline.cpp); the oracle's own
    restatements for the custom-edge histogram and the 64-bit rows."""
    import numpy as np
    import oracle
    from gpumultisplit_amd import workloads as W
    sample = n
    if w.startswith("sort"):
        keys = W.uniform_keys(n, seed=1)
        vals = np.arange(n, dtype=np.uint32) if w.endswith("kv") else None
        run = lambda: oracle.baseline_stable_sort(keys, vals)
        what = "std::stable_sort of (key, value) pairs" if vals is not None else "std::stable_sort of the keys"
    elif w == "hist256":
        keys = W.uniform_keys(n, seed=1)
        run, what = (lambda: oracle.baseline_histogram(keys, 256, 0)), "counting loop over key & 0xFF"
    elif w.startswith("compact"):
        keys = W.uniform_keys(n, seed=1)
        thr = int(int(w[7:]) / 100 * (1 << 32))
        run, what = (lambda: oracle.baseline_stable_partition(keys, thr)), "std::stable_partition of the keys"
    elif w == "histrange256":
        sample = 1 << 24
        g = W.rng(1)
        x = (g.random(sample, dtype=np.float32) * np.float32(1024.0)).astype(np.float32)
        lv = np.sort((g.random(257, dtype=np.float32) * np.float32(1024.0)).astype(np.float32))
        lv[0], lv[-1] = 0.0, 1024.0
        run, what = (lambda: oracle.histogram_range(x, lv)), "the oracle's edge scan per sample"
    elif w in ("wide_kv_256", "wide_ko_256", "wide_hist256", "wide_sort8_kv"):
        sample = 1 << 24 if w != "wide_sort8_kv" else 1 << 22
        g = W.rng(1)
        keys = g.integers(0, 1 << 63, sample, dtype=np.uint64) * np.uint64(2) + g.integers(0, 2, sample, dtype=np.uint64)
        vals = np.arange(sample, dtype=np.uint64)
        if w == "wide_hist256":
            run = lambda: np.bincount(oracle.bucket_ids64(keys, 256, oracle.RANGE), minlength=256)
            what = "the oracle's 64-bit range identifier + counting"
        elif w == "wide_sort8_kv":
            run, what = (lambda: oracle.sort_wide(keys, 8, vals)), "the oracle's LSD sort of 64-bit pairs"
        else:
            v = vals if w == "wide_kv_256" else None
            run, what = (lambda: oracle.multisplit_wide(keys, 256, oracle.RANGE, 0, v)), \
                "the oracle's 64-bit multisplit"
    else:
        return None
    ts = []
    for _ in range(reps):
        c0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - c0)
    med = statistics.median(ts)
    return {"value": round(sample / med / 1e9, 4), "unit": "Gkeys/s", "cores": 1, "kind": "port",
            "sample": f"{reps} runs (median) of {what} over {sample} keys on the host",
            "cpu_model": host_cpu_model()}


def bench_variants(args, torch, gms, dev):
    """Block-wide (BMS) against wave-level (WMS) and direct wave-level (DMS)
    multisplit: one JSON line per (buckets, key-only/key-value, n) with each
    variant's time per call, the measurement the automatic variant choice is
    set from.  Uniform keys, range identifier, values = index; --steps timed
    calls after --warmup."""
    L = gms.lib()
    stream = torch.cuda.current_stream(dev).cuda_stream
    g = torch.Generator(device=dev)
    g.manual_seed(4321)
    nmax = 1 << 27
    kin = torch.randint(-(1 << 31), 1 << 31, (nmax,), dtype=torch.int32, device=dev, generator=g)
    vin = torch.arange(nmax, dtype=torch.int32, device=dev)
    kout, vout = torch.empty_like(kin), torch.empty_like(kin)
    for B in (2, 4, 8, 16, 32, 64):
        ctx = gms.multisplit_context(B)
        gms.multisplit_allocate_key_value(nmax, ctx)
        desc = gms.range_bucket(B).desc()
        for kv in (False, True):
            for log2n in (16, 20, 24, 27):
                n = 1 << log2n
                res = {}
                for name, var in (("bms", gms.VARIANT_BMS), ("wms", gms.VARIANT_WMS),
                                  ("dms", gms.VARIANT_DMS)):
                    gms.set_variant(ctx, var)

                    def step():
                        st = L.gms_multisplit(ctx.handle, kin.data_ptr(),
                                              vin.data_ptr() if kv else None, kout.data_ptr(),
                                              vout.data_ptr() if kv else None, n, desc, None,
                                              stream)
                        if st != 0:
                            raise gms.GmsError(st, "gms_multisplit")
                    for _ in range(args.warmup):
                        step()
                    torch.cuda.synchronize()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record()
                    for _ in range(args.steps):
                        step()
                    e1.record()
                    torch.cuda.synchronize()
                    res[name] = e0.elapsed_time(e1) / args.steps
                print(json.dumps({"buckets": B, "key_value": kv, "n": n,
                                  "bms_ms": round(res["bms"], 5), "wms_ms": round(res["wms"], 5),
                                  "dms_ms": round(res["dms"], 5),
                                  "wms_over_bms": round(res["wms"] / res["bms"], 3),
                                  "dms_over_bms": round(res["dms"] / res["bms"], 3)}), flush=True)
        gms.set_variant(ctx, gms.VARIANT_AUTO)
        gms.multisplit_release_memory(ctx)


def host_cpu_model():
    """Model name of the host CPU (the cores the CPU baselines run on)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def traffic_from_profiles(workload):
    """HBM bytes per postscan launch from committed rocprofv3 PMC summaries."""
    path = os.path.join(PROFILES, f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def parse_workload(w):
    parts = w.split("_")
    cfg = {"kind": parts[0]}
    if parts[0] == "ms":
        cfg["kv"] = parts[1] == "kv"
        cfg["B"] = int(parts[2])
        cfg["dist"] = "uniform"
        cfg["log2n"] = 27
        for extra in parts[3:]:
            if extra.startswith("n"):
                cfg["log2n"] = int(extra[1:])
            else:
                cfg["dist"] = extra
    return cfg


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the multi-GPU sharding design is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gpumultisplit_amd as gms
    from gpumultisplit_amd import workloads as W

    gms.load_library()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev).cuda_stream
    cfg = parse_workload(args.workload)
    if args.workload == "variant_sweep":
        bench_variants(args, torch, gms, dev)
        return
    if cfg["kind"] != "ms":
        bench_aux(args, torch, gms, W, dev)
        return

    n = 1 << cfg["log2n"]
    B = cfg["B"]
    kv = cfg["kv"]
    print(f"[bench] workload={args.workload} n={n} B={B} kv={kv} dist={cfg['dist']}",
          file=sys.stderr, flush=True)
    if cfg["dist"] == "uniform":
        g = torch.Generator(device=dev)
        g.manual_seed(12345)
        kin = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=dev, generator=g)
    else:
        kin = torch.from_numpy(W.keys_for(cfg["dist"], n, B, seed=12345).view(np.int32)).to(dev)
    vin = torch.arange(n, dtype=torch.int32, device=dev) if kv else None
    kout = torch.empty_like(kin)
    vout = torch.empty_like(kin) if kv else None
    ctx = gms.multisplit_context(B)
    (gms.multisplit_allocate_key_value if kv else gms.multisplit_allocate_key_only)(n, ctx)
    f = gms.range_bucket(B)
    desc = f.desc()
    L = gms.lib()
    h, kp, vp = ctx.handle, kin.data_ptr(), (vin.data_ptr() if kv else None)
    kop, vop = kout.data_ptr(), (vout.data_ptr() if kv else None)

    def step():
        st = L.gms_multisplit(h, kp, vp, kop, vop, n, desc, None, stream)
        if st != 0:
            raise gms.GmsError(st, "gms_multisplit")

    if args.calibrate_copy:
        gms.copy_u32(kin, kout, n)
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # The timed region: K calls with nothing else recorded on the stream.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step", file=sys.stderr, flush=True)
    # Stage times for the roofline: a second pass of the same K calls with
    # HIP events recorded on the library's stream around each stage.
    stage_events = not args.no_stage_events
    pre_ms = scan_ms = post_ms = calls = 0
    if stage_events:
        gms.set_profiling(ctx, True)
        gms.read_profile(ctx)  # reset
        for _ in range(args.steps):
            step()
        pre_ms, scan_ms, post_ms, calls = gms.read_profile(ctx)
        gms.set_profiling(ctx, False)

    gkeys = n / (ms_per_step * 1e-3) / 1e9
    bpk_call = 20 if kv else 12
    bpk_post = 16 if kv else 8
    roofline = None
    stages = None
    if stage_events and calls:
        post_avg = post_ms / calls
        achieved = n * bpk_post / (post_avg * 1e-3) / 1e9
        tr = traffic_from_profiles(args.workload)
        roofline = {"kernel": "ms_postscan_kernel", "bound": "hbm", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": tr, "algorithmic_bytes": n * bpk_post,
                    "avg_launch_us": round(post_avg * 1e3, 2)}
        stages = {"prescan_us": round(pre_ms / calls * 1e3, 2), "scan_us": round(scan_ms / calls * 1e3, 2),
                  "postscan_us": round(post_avg * 1e3, 2)}

    cpu = None
    if not args.no_cpu:
        import oracle
        print("[bench] cpu baseline (oracle, 1 thread)...", file=sys.stderr, flush=True)
        keys = W.uniform_keys(n, seed=12345)
        vals = np.arange(n, dtype=np.uint32) if kv else None
        ts = []
        for _ in range(args.cpu_reps):
            c0 = time.perf_counter()
            oracle.multisplit(keys, B, oracle.RANGE, 0, vals)
            ts.append(time.perf_counter() - c0)
        med = statistics.median(ts)
        cpu = {"value": round(n / med / 1e9, 4), "unit": "Gkeys/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_reps} runs (median) of the full workload ({n} keys) with the "
                         f"single-threaded C oracle on the host, {os.cpu_count()} host threads present",
               "cpu_model": host_cpu_model()}

    out = {
        "metric": f"{'key-value' if kv else 'key-only'} multisplit throughput, {n} uint32 keys, "
                  f"{B} buckets (range identifier)",
        "value": round(gkeys, 3),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": f"synthetic ({cfg['dist']} keys, values = original index)",
        "config": {"workload": args.workload, "n": n, "buckets": B, "key_value": kv,
                   "bucket_identifier": "range", "tile_keys": gms.lib().gms_tile_keys(B, int(kv), n)},
        "call_hbm_fraction": round(gkeys * bpk_call / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "stages": stages,
        "cpu_baseline": cpu,
    }
    gms.multisplit_release_memory(ctx)
    del kin, vin, kout, vout
    if not args.no_batched:
        out["batched"] = bench_batched(args, torch, gms, W, dev)
    if not args.no_served:
        out["served"] = bench_served(args, torch, gms, dev)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
