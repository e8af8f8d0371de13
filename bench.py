#!/usr/bin/env python3
"""Two-stage ordinal retrieval benchmark on one MI355X.

Workload (the reference's two-stage path at realistic shape): 1,000,000 docs x
1024-d from the Gaussian mixture (4096 centroids, seed 2), sign-bitmap
XOR-popcount candidate generation to top-1000, then RankQuant b=2 asymmetric
rerank to k=10.  Index and queries are resident in HBM before timing.

* batched: one call scores a batch of 1024 queries (the headline `value`)
* plain:   one query per call (latency p50/p99 and QPS)
* cpu_baseline: the oracle (C restatement of the reference CPU path, same
  algorithm and tiling) on the host cores for a bounded query sample; the
  same sample's GPU top-10 ids are compared with it.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_INT32_PEAK_OPS = 256 * 64 * 2.4e9  # CUs x lanes/clk x clock (estimate)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-docs", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--m", type=int, default=1000)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--plain-queries", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve", action="store_true",
                    help="serving run instead: 20M x 1024 index, open-loop Poisson load at --rates")
    ap.add_argument("--serve-docs", type=int, default=20_000_000)
    ap.add_argument("--rates", type=str, default="1000,10000,50000", help="offered queries/s, comma separated")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--serve-max-batch", type=int, default=1024)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                          "(documents shard by id range with one all-gather of per-shard top-m keys)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import ordvec_amd as ov
    from ordvec_amd import _native as N
    from ordvec_amd.synthetic import MixtureTorch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.serve:
        return serve_main(args, dev)
    d, n, B, m, k = args.dim, args.n_docs, args.batch, args.m, args.k

    # ---- build the resident index ------------------------------------------------
    t0 = time.time()
    gen = MixtureTorch(d, seed=2, device=dev)
    sb = ov.SignBitmap(d)
    rq = ov.RankQuant(d, 2)
    chunk = 131072
    docs_keep = torch.empty((n, d), dtype=torch.float32, device=dev)
    for s0 in range(0, n, chunk):
        c = gen.docs(min(chunk, n - s0))
        docs_keep[s0:s0 + c.shape[0]] = c
        sb.add(c)
        rq.add(c)
    torch.cuda.synchronize()
    n_batches = args.warmup + args.steps
    qs_all, src = gen.queries(docs_keep, n_batches * B + args.plain_queries)
    del docs_keep
    torch.cuda.synchronize()
    log(f"index built: {n} docs x {d} in {time.time() - t0:.1f}s")

    batches = [qs_all[i * B:(i + 1) * B] for i in range(n_batches)]
    plain = qs_all[n_batches * B:]

    # ---- batched (headline) ----------------------------------------------------------
    for i in range(args.warmup):
        ov.two_stage_search(sb, rq, batches[i], m, k)
    torch.cuda.synchronize()
    N.kernel_timing_enable(True)
    t0 = time.perf_counter()
    outs = []
    for i in range(args.warmup, n_batches):
        outs.append(ov.two_stage_search(sb, rq, batches[i], m, k))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fam = {f: N.kernel_timing_read(f) for f in N.KERNEL_FAMILIES}
    N.kernel_timing_enable(False)
    elapsed = t1 - t0
    qps = args.steps * B / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    log(f"batched: {qps:,.0f} QPS, {ms_per_step:.3f} ms per {B}-query batch")
    for f, (cnt, ms) in fam.items():
        if cnt:
            log(f"  kernel family {f}: {cnt} launches, avg {ms / cnt * 1e3:.1f} us")

    scan_launches, scan_ms = fam["sign_scan"]
    scan_avg_s = scan_ms / 1e3 / max(1, scan_launches)
    alg_bytes = n * d // 8 + B * d // 8          # sign corpus + query bitmaps per launch
    scan_gbs = alg_bytes / scan_avg_s / 1e9
    int_ops = 2.0 * B * n * (d // 32)           # xor + popcount per 32-bit word and pair
    valu_frac = int_ops / scan_avg_s / VALU_INT32_PEAK_OPS
    step_bytes = n * d // 8 + B * (m * d * 2 // 8 + d * 4 + 12 * k)
    step_gbs = step_bytes / (elapsed / args.steps) / 1e9

    # ---- plain (one query per call) ------------------------------------------------
    lat = []
    P = plain.shape[0]
    for i in range(min(20, P)):
        ov.two_stage_search(sb, rq, plain[i:i + 1], m, k)
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    plain_ids = []
    for i in range(P):
        a = time.perf_counter()
        s_, i_ = ov.two_stage_search(sb, rq, plain[i:i + 1], m, k)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - a)
        plain_ids.append(i_)
    tp1 = time.perf_counter()
    lat_ms = np.array(lat) * 1e3
    plain_qps = P / (tp1 - tp0)
    log(f"plain: {plain_qps:,.0f} QPS, p50 {np.percentile(lat_ms, 50):.3f} ms, p99 {np.percentile(lat_ms, 99):.3f} ms")

    # consistency: a query alone and inside a batch gives the same result
    bi = torch.cat([p for p in plain_ids[:8]]).cpu().numpy()
    bs_, bi_b = ov.two_stage_search(sb, rq, plain[:8], m, k)
    same_alone_batched = bool(np.array_equal(bi, bi_b.cpu().numpy()))

    # ---- CPU baseline (oracle on host cores) + full-size parity sample -----------------
    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        import oracle as O

        O.build()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        codes = rq.codes()
        words = sb.words()
        qhost = batches[args.warmup].cpu().numpy()
        gpu_s, gpu_i = outs[0][0].cpu().numpy(), outs[0][1].cpu().numpy()

        def run(nq):
            q = qhost[:nq]
            a = time.perf_counter()
            cand = O.sign_top_m(words, O.sign_encode(q), m, threads)
            offs = np.arange(nq + 1, dtype=np.uint64) * cand.shape[1]
            s_, i_ = O.subset_batched(codes, d, 2, q, offs, cand.reshape(-1), k, threads)
            return time.perf_counter() - a, s_, i_

        nq = threads
        dt, _, _ = run(nq)
        while dt < args.cpu_seconds / 4 and nq < B:
            nq = min(B, nq * 4)
            dt, cs, ci = run(nq)
        if nq == threads:
            dt, cs, ci = run(nq)
        cpu = {"value": round(nq / dt, 2), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"{nq} queries of the timed batch, same 1M x 1024 index, two-stage m={m} k={k}, "
                         f"{dt:.2f} s wall on {threads} host threads"}
        parity = {"queries": nq, "ids_identical": bool(np.array_equal(ci, gpu_i[:nq])),
                  "scores_bit_identical": bool(np.array_equal(cs.view(np.uint32), gpu_s[:nq].view(np.uint32)))}
        log(f"cpu baseline: {cpu['value']:,.1f} QPS on {threads} threads; parity {parity}")

    out = {
        "metric": "QPS + HBM GB/s (% of 8 TB/s), 1M x 1024-d two-stage k=10, batched vs plain; p99 ms",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 popcount + f32",
        "data": "synthetic (Gaussian mixture, 4096 centroids, seed 2)",
        "config": {"workload": f"two-stage sign top-{m} -> RankQuant b=2 rerank k={k}", "n_docs": n, "dim": d,
                   "batch": B, "m": m, "k": k, "parallelism": "single GPU"},
        "step_hbm_gbs": round(step_gbs, 1),
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "plain": {"qps": round(plain_qps, 1), "p50_ms": round(float(np.percentile(lat_ms, 50)), 4),
                  "p99_ms": round(float(np.percentile(lat_ms, 99)), 4), "queries": P,
                  "batched_over_plain": round(qps / plain_qps, 2)},
        "roofline": {"kernel": "k_sign_scan (stage-1 XOR-popcount scan)", "bound": "hbm",
                     "achieved": round(scan_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(scan_gbs / HBM_PEAK_GBS, 5), "traffic": None,
                     "algorithmic_bytes_per_launch": alg_bytes, "avg_launch_us": round(scan_avg_s * 1e6, 2)},
        "valu_roofline": {"kernel": "k_sign_scan", "int_ops_per_launch": int_ops,
                          "achieved_tops": round(int_ops / scan_avg_s / 1e12, 3),
                          "peak_tops": round(VALU_INT32_PEAK_OPS / 1e12, 2), "frac": round(valu_frac, 4)},
        "kernel_avg_us": {f: round(ms / c * 1e3, 2) for f, (c, ms) in fam.items() if c},
        "alone_equals_batched": same_alone_batched,
        "cpu_baseline": cpu,
        "parity_full_size": parity,
    }
    print(json.dumps(out), flush=True)


def serve_main(args, dev):
    """Continuous-batching server under open-loop Poisson load: 20M x 1024
    mixture (seed 4) resident in HBM, requests of 1-64 queries, k uniform in
    [10, 100], m = 1000; latency from scheduled arrival to result on the host."""
    import numpy as np
    import torch

    import ordvec_amd as ov
    from ordvec_amd.synthetic import MixtureTorch

    d, n, m = args.dim, args.serve_docs, args.m
    t0 = time.time()
    gen = MixtureTorch(d, seed=4, device=dev)
    sb = ov.SignBitmap(d)
    rq = ov.RankQuant(d, 2)
    chunk = 262144
    pool = []
    per_chunk = max(1, 32768 * chunk // n)
    for s0 in range(0, n, chunk):
        c = gen.docs(min(chunk, n - s0))
        sb.add(c)
        rq.add(c)
        q, _ = gen.queries(c, per_chunk)
        pool.append(q.cpu().numpy())
        if (s0 // chunk) % 16 == 0:
            torch.cuda.synchronize()
            log(f"  built {s0 + c.shape[0]:,} docs ({time.time() - t0:.0f}s)")
        del c
    torch.cuda.synchronize()
    pool = np.ascontiguousarray(np.concatenate(pool))
    log(f"index built: {n:,} docs x {d} in {time.time() - t0:.1f}s; query pool {pool.shape[0]}")

    runs = []
    with ov.Server(sb, rq, m=m, max_k=100, slots=args.slots, max_batch_queries=args.serve_max_batch,
                   memory_budget_bytes=16 << 30) as srv:
        cap = srv.info()["max_batch_queries"]
        # Mixed batch vs alone: every request's rows must match a solo call.
        rng = np.random.default_rng(11)
        reqs = []
        for _ in range(24):
            nq = int(rng.integers(1, 65))
            st = int(rng.integers(0, pool.shape[0] - nq))
            reqs.append((pool[st:st + nq], int(rng.integers(10, 101))))
        tickets = [srv.submit(q, k) for q, k in reqs]
        mixed = [srv.wait(t) for t in tickets]
        same = all(np.array_equal(r[1], ov.two_stage_search(sb, rq, q, m, k)[1]) and
                   np.array_equal(r[0].view(np.uint32), ov.two_stage_search(sb, rq, q, m, k)[0].view(np.uint32))
                   for (q, k), r in zip(reqs, mixed))
        log(f"server: {args.slots} slots, batch cap {cap}; mixed == alone: {same}")
        srv.open_loop(pool, 2000, 1.0, seed=99)  # warm-up
        for rate in [float(r) for r in args.rates.split(",")]:
            rep = srv.open_loop(pool, rate, args.serve_seconds, nq_range=(1, 64), k_range=(10, 100), seed=int(rate))
            rep = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in rep.items()}
            log(f"offered {rate:,.0f} q/s: achieved {rep['achieved_qps']:,.0f} q/s ({rep['achieved_rps']:,.0f} req/s), "
                f"p50 {rep['p50_ms']:.2f} ms, p99 {rep['p99_ms']:.2f} ms, mean batch {rep['mean_batch_queries']:.1f}")
            runs.append(rep)
    ten_k = next((r for r in runs if abs(r["offered_qps"] - 10000) < 1), None)
    out = {
        "metric": "serving p99 latency (ms) at 10k offered queries/s, 20M x 1024-d two-stage",
        "value": ten_k["p99_ms"] if ten_k else None,
        "unit": "ms",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "u64 popcount + f32",
        "data": "synthetic (Gaussian mixture, 4096 centroids, seed 4)",
        "config": {"workload": "open-loop Poisson serving, requests of 1-64 queries, k in [10,100], m=1000",
                   "n_docs": n, "dim": d, "slots": args.slots, "batch_cap": cap,
                   "seconds_per_rate": args.serve_seconds},
        "mixed_equals_alone": bool(same),
        "runs": runs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
