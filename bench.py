"""Benchmark: batched Goldilocks NTT on one MI355X.

Workload (the plain batched-NTT configuration): 256 columns x 2^20 u64
Goldilocks elements, column-major with stride 2^20, uniform draws, resident
in HBM.  One step = four out-of-place calls over the whole matrix: forward and
inverse, each with natural input (n2b) and bit-reversed input (b2n).
`value` = transformed elements per second over the K timed steps.

roofline: HBM bound.  Algorithmic bytes per transform call = 16 B x elements
(each element read once and written once).  `achieved` uses the average
duration of one call measured with HIP events on the stream the library
launches on (torch's current stream) over the timed region.

cpu_baseline: the oracle's multithreaded restatement of boojum's CPU NTT
(oracle/gl_oracle.c), timed on a bounded sample of the same workload.

Prints ONE JSON line.  `--gpus` above 1 is refused: the measurement is one
GPU's share (columns are independent, so a batch shards by column range with
no communication).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "era-boojum-cuda_amd"))

HBM_PEAK_GBS = 8000.0
CPU_SHARE = 16  # host cores that come with one GPU of the node
METRIC = "Goldilocks NTT/LDE elems/s & %HBM peak; ragged-batched vs plain; p50/p99 latency"


def uniform_field(count, gen):
    from boojum_cuda_amd.workloads import uniform_field as uf
    return uf(count, gen)


def bench_lde(log_n, cols, log_blowup, reps, warmup):
    """Coset LDE: bit-reversed evaluations -> coefficients -> 2^k cosets,
    replayed from a captured hipGraph.  Algorithmic bytes: 8 * (1 + B) * n * cols."""
    import torch
    from boojum_cuda_amd import lde
    n = 1 << log_n
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    evals = uniform_field(cols * n, g)
    coeffs = torch.empty_like(evals)
    out = torch.empty(cols * (n << log_blowup), dtype=torch.int64, device="cuda")
    graph = lde.LdeGraph(evals, coeffs, out, log_n, cols, n, n, n << log_blowup, log_blowup, True)
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        graph.launch()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
    for r in range(reps):
        ev[2 * r].record(stream)
        graph.launch()
        ev[2 * r + 1].record(stream)
    torch.cuda.synchronize()
    ms = sum(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps)) / reps
    graph.destroy()
    blowup = 1 << log_blowup
    algo = 8.0 * (1 + blowup) * n * cols
    del evals, coeffs, out
    return {"workload": f"coset LDE {cols} x 2^{log_n}, blowup {blowup}, bit-reversed in/out, hipGraph",
            "ms_per_call": ms, "lde_elems_per_s": blowup * n * cols / (ms * 1e-3),
            "roofline": {"bound": "hbm", "achieved": algo / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_call": algo}}


def bench_ragged(count, log_lo, log_hi, reps, warmup, plain=True):
    """Ragged batch: the persistent work-queue kernel over every job (one
    cooperative launch, workgroups waiting for each other between dependency
    levels) vs the plain path (separate launch sequences
    per job, on the same stream).  Both must give
    the same bits.  Algorithmic bytes: 16 n (NTT/iNTT), 8 (1 + B) n (LDE)."""
    import torch
    from boojum_cuda_amd import ragged, workloads
    specs = workloads.ragged_specs(count, log_lo, log_hi, workloads.RAGGED_SHAPE["seed"])
    jobs = workloads.ragged_jobs(specs, workloads.RAGGED_SHAPE["seed"])
    algo, out_elems, in_elems = 0.0, 0, 0
    for log_n, op, lb in specs:
        n = 1 << log_n
        algo += 8.0 * (n + (n << lb))
        out_elems += n << lb
        in_elems += n
    plan = ragged.RaggedPlan(jobs)
    info = plan.info()
    stream = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        for r in range(reps):
            ev[2 * r].record(stream)
            fn()
            ev[2 * r + 1].record(stream)
        torch.cuda.synchronize()
        return sorted(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps))[reps // 2]

    ms_ragged = timed(plan.run)
    plan.wait_status()  # raises if a wait between levels gave up (outputs invalid)
    ragged_out = [j.outputs.clone() for j in jobs]
    ms_plain = timed(lambda: ragged.run_plain(jobs)) if plain else float("nan")
    same = all(torch.equal(a, j.outputs) for a, j in zip(ragged_out, jobs)) if plain else None
    plan.destroy()
    del ragged_out, jobs
    torch.cuda.empty_cache()
    gbs = algo / (ms_ragged * 1e-3) / 1e9
    return {"workload": f"ragged batch: {count} single-column jobs, log_n ~ U[{log_lo},{log_hi}], "
                        f"op ~ U{{NTT, iNTT, LDE x2, LDE x4}}, one persistent launch of the queue kernel",
            "ms_ragged": ms_ragged, "ms_plain": ms_plain, "speedup_vs_plain": ms_plain / ms_ragged,
            "input_elems": in_elems, "output_elems": out_elems,
            "output_elems_per_s": out_elems / (ms_ragged * 1e-3),
            "units": info["units"], "levels": info["levels"], "workgroups": info["workgroups"],
            "bit_identical_to_plain": same if same is None else bool(same),
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_call": algo}}


def bench_serve(count, rate, budget_bytes, seed=5):
    """Serving loop at a fixed offered load: `count` requests with Poisson
    arrivals at `rate` per second, each 1-512 columns of 2^log_n seeded
    values, log_n ~ U[12,24], blowup ~ U{1,2,4} (1: forward NTT, else coset
    LDE), admitted under an HBM budget of `budget_bytes`.  Latency = submit to
    completion (host clock); throughput = output elements over the makespan."""
    import numpy as np
    from boojum_cuda_amd import serve, workloads
    specs = workloads.serve_specs(count, rate, seed)
    srv = serve.Server(budget_bytes)
    try:
        warm = [srv.submit(serve.OP_LDE, 12, 4, lb, seed=1) for lb in (1, 2)] + [srv.submit(serve.OP_NTT, 16, 2)]
        srv.drain()
        for rid in warm:
            srv.release(rid)
        ids, out_elems, algo = [], 0, 0.0
        t0 = time.perf_counter()
        for k, (at, log_n, cols, lb) in enumerate(specs):
            while True:
                dt = t0 + at - time.perf_counter()
                if dt <= 0:
                    break
                time.sleep(min(dt, 0.0005))
            op = serve.OP_LDE if lb else serve.OP_NTT
            ids.append(srv.submit(op, log_n, cols, lb, seed=1000 + k))
            n = 1 << log_n
            out_elems += (n << lb) * cols
            algo += 8.0 * (n + (n << lb)) * cols
        srv.drain()
        infos = [srv.info(rid) for rid in ids]
        st = srv.stats()
        for rid in ids:
            srv.release(rid)
    finally:
        srv.destroy()
    assert all(i["status"] == 1 for i in infos), "a request failed"
    lat_ms = np.array([(i["t_done_ns"] - i["t_submit_ns"]) * 1e-6 for i in infos])
    span_s = (max(i["t_done_ns"] for i in infos) - min(i["t_submit_ns"] for i in infos)) * 1e-9
    gbs = algo / span_s / 1e9
    return {"workload": f"serving: {count} requests, Poisson arrivals at {rate}/s, cols ~ U[1,512], "
                        f"log_n ~ U[12,24], blowup ~ U{{1,2,4}}, seeded inputs, {budget_bytes / 1e9:.0f} GB budget",
            "offered_load_req_per_s": rate, "makespan_s": span_s,
            "p50_latency_ms": float(np.percentile(lat_ms, 50)), "p99_latency_ms": float(np.percentile(lat_ms, 99)),
            "mean_latency_ms": float(lat_ms.mean()), "output_elems_per_s": out_elems / span_s,
            "requests_per_s": count / span_s, "parts": st["parts_done"], "batches": st["batches_launched"],
            "max_batches_in_flight_seen": st["max_batches_in_flight_seen"],
            "arena_high_water_gb": st["arena_high_water_bytes"] / 1e9,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS}}


def bench_serve_sweep(rates, count, budget_bytes):
    """The serving leg at several offered loads (the same request
    distribution): served requests per second and latency per load, the
    capacity (most requests per second served) and the knee, the highest
    offered load still served at >= 80% of its rate (with `count` requests per
    point the last requests' own service time is part of the makespan, so an
    unsaturated point serves about 90% of its offered rate)."""
    points = []
    for rate in rates:
        r = bench_serve(count, rate, budget_bytes)
        points.append({"offered_req_per_s": rate, "served_req_per_s": r["requests_per_s"],
                       "p50_latency_ms": r["p50_latency_ms"], "p99_latency_ms": r["p99_latency_ms"],
                       "output_elems_per_s": r["output_elems_per_s"], "makespan_s": r["makespan_s"]})
        progress(f"serve sweep {rate}/s: served {r['requests_per_s']:.1f}/s, p50 {r['p50_latency_ms']:.1f} ms, "
                 f"p99 {r['p99_latency_ms']:.1f} ms")
    kept = [p["offered_req_per_s"] for p in points if p["served_req_per_s"] >= 0.8 * p["offered_req_per_s"]]
    return {"requests_per_point": count, "points": points, "knee_req_per_s": max(kept) if kept else None,
            "capacity_req_per_s": max(p["served_req_per_s"] for p in points)}


def bench_ops(log_n, cols, reps):
    """Order conversion and domain powers around the NTT on the batched
    shape: bit_reverse (out of place and in place, 16 B per element moved),
    omega_shift (16 B per element) and get_powers_of_w (8 B per element
    written).  HBM-bound kernels: GB/s against the 8 TB/s peak."""
    import torch
    from boojum_cuda_amd import ops
    n = 1 << log_n
    g = torch.Generator(device="cuda")
    g.manual_seed(6)
    x = uniform_field(cols * n, g)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        for r in range(reps):
            ev[2 * r].record(stream)
            fn()
            ev[2 * r + 1].record(stream)
        torch.cuda.synchronize()
        return sum(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps)) / reps

    total = cols * n
    out = {}
    for name, fn, bytes_per in [
            ("bit_reverse", lambda: ops.bit_reverse(x, y, n, cols), 16),
            ("bit_reverse_in_place", lambda: ops.bit_reverse_in_place(x, n, cols), 16),
            ("omega_shift", lambda: ops.omega_shift(x, log_n + 8, 0, False, 3, y), 16),
            ("get_powers_of_w", lambda: ops.get_powers_of_w(log_n + 8, 0, False, False, y), 8)]:
        ms = timed(fn)
        gbs = bytes_per * total / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
    del x, y
    return {"workload": f"{cols} x 2^{log_n} u64 ({total} elements)", **out}


def bench_barycentric(log_n, cols, reps):
    """Barycentric evaluation on the batched shape: Lagrange coefficients at a
    base / extension point (8 / 16 B written per element) and batch_eval of
    `cols` base columns (BaseAtBase, BaseAtExt) or cols / 2 extension columns
    (ExtAtExt) of 2^log_n rows: algorithmic bytes = the columns read once plus
    the coefficients (8 or 16 B per row).  GB/s against the 8 TB/s peak."""
    import torch
    from boojum_cuda_amd import barycentric as bary
    n = 1 << log_n
    g = torch.Generator(device="cuda")
    g.manual_seed(8)
    ys = uniform_field(cols * n, g)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        for r in range(reps):
            ev[2 * r].record(stream)
            fn()
            ev[2 * r + 1].record(stream)
        torch.cuda.synchronize()
        return sum(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps)) / reps

    out = {}
    x = uniform_field(2, g)
    common = torch.empty(2, dtype=torch.int64, device="cuda")
    coeffs = {False: torch.empty(n, dtype=torch.int64, device="cuda"),
              True: torch.empty(2 * n, dtype=torch.int64, device="cuda")}
    for ext in (False, True):
        ms = timed(lambda: bary.precompute_lagrange_coeffs(ext, x, common, 7, coeffs[ext], True))
        gbs = (16 if ext else 8) * n / (ms * 1e-3) / 1e9
        out["precompute_at_ext" if ext else "precompute_at_base"] = {"ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
    for name, kind, polys in [("base_at_base", bary.EVAL_BASE_AT_BASE, cols),
                              ("base_at_ext", bary.EVAL_BASE_AT_EXT, cols),
                              ("ext_at_ext", bary.EVAL_EXT_AT_EXT, cols // 2)]:
        w = 1 if kind == bary.EVAL_BASE_AT_BASE else 2
        partial, _ = bary.get_batch_eval_temp_storage_sizes(kind, n, polys)
        temp = torch.empty(partial * w, dtype=torch.int64, device="cuda")
        evals = torch.empty(polys * w, dtype=torch.int64, device="cuda")
        L = coeffs[kind != bary.EVAL_BASE_AT_BASE]
        ms = timed(lambda: bary.batch_eval(kind, ys, n, polys, L, temp, None, evals))
        algo = 8 * cols * n + 8 * w * n
        gbs = algo / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS, "polys": polys}
    del ys
    return {"workload": f"{cols} base columns (or {cols // 2} extension columns) x 2^{log_n}", **out}


def bench_poseidon(log_n, cols, reps):
    """Poseidon and Poseidon2 Merkle trees over one LDE coset: `cols` columns x
    2^log_n rows,
    one leaf per row (cols / 8 permutations per leaf), then the 2^log_n - 1
    node hashes.  The permutation is VALU-bound; GB/s is the leaf input read
    once.  Round constants are synthetic (the time does not depend on their
    values)."""
    import numpy as np
    import torch
    from boojum_cuda_amd import poseidon
    n = 1 << log_n
    poseidon.set_round_constants(np.random.default_rng(0).integers(0, 0xFFFFFFFF00000001, 360, dtype=np.uint64))
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    vals = uniform_field(cols * n, g)
    tree = torch.empty(8 * n, dtype=torch.int64, device="cuda")
    leaves, nodes = tree[:4 * n], tree[4 * n:]
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        for r in range(reps):
            ev[2 * r].record(stream)
            fn()
            ev[2 * r + 1].record(stream)
        torch.cuda.synchronize()
        return sum(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps)) / reps

    perms = n * ((cols + 7) // 8)
    out = {"workload": f"Merkle tree over {cols} columns x 2^{log_n} rows (one leaf per row)",
           "round_constants": "synthetic"}
    for p in (poseidon.POSEIDON, poseidon.POSEIDON2):
        ms_leaves = timed(lambda: poseidon.build_merkle_tree_leaves(vals, leaves, 0, p=p))
        ms_nodes = timed(lambda: poseidon.build_merkle_tree_nodes(leaves, nodes, log_n, p=p))
        out[p] = {"ms_leaves": ms_leaves, "ms_nodes": ms_nodes, "ms_tree": ms_leaves + ms_nodes,
                  "leaf_permutations_per_s": perms / (ms_leaves * 1e-3),
                  "node_permutations_per_s": (n - 1) / (ms_nodes * 1e-3),
                  "leaf_input_GB/s": 8 * cols * n / (ms_leaves * 1e-3) / 1e9}
    del vals
    return out


def bench_latency(reps=200):
    """Config 1: one column of 2^16, forward natural -> bit-reversed, out of
    place.  Reports
    - gpu_us_per_call: GPU time per call back to back, from HIP events around
      the replay of a graph holding `reps` calls (so the host's enqueue cost
      is not in it); every replay's output is checked against the eager one;
    - eager_enqueue_us_per_call: the same calls enqueued one by one from
      Python (bounded by the host's per-call cost when that exceeds the GPU's);
    - the host-observed latency of a call plus a stream synchronize (median,
      p99), and whether the inverse round trip returns the input."""
    import torch
    from boojum_cuda_amd import ntt
    log_n = 16
    n = 1 << log_n
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    x = uniform_field(n, g)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    for _ in range(10):
        ntt.batch_ntt_out_of_place(x, y, log_n, 1, 0, 0, n, n, False, False, 0, 0)
    torch.cuda.synchronize()
    want = y.clone()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(reps):
        ntt.batch_ntt_out_of_place(x, y, log_n, 1, 0, 0, n, n, False, False, 0, 0)
    ev[1].record(stream)
    torch.cuda.synchronize()
    eager_us = ev[0].elapsed_time(ev[1]) * 1e3 / reps

    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            for _ in range(reps):
                ntt.batch_ntt_out_of_place(x, y, log_n, 1, 0, 0, n, n, False, False, 0, 0)
    stream.wait_stream(side)
    torch.cuda.synchronize()
    replays_exact = True
    times = []
    for _ in range(5):
        y.zero_()
        torch.cuda.synchronize()
        ev[0].record(stream)
        graph.replay()
        ev[1].record(stream)
        torch.cuda.synchronize()
        times.append(ev[0].elapsed_time(ev[1]) * 1e3 / reps)
        replays_exact = replays_exact and bool(torch.equal(y, want))
    times.sort()
    del graph

    lat = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ntt.batch_ntt_out_of_place(x, y, log_n, 1, 0, 0, n, n, False, False, 0, 0)
        stream.synchronize()
        lat.append((time.perf_counter() - t0) * 1e6)
    lat.sort()
    ntt.batch_ntt_out_of_place(y, z, log_n, 1, 0, 0, n, n, True, True, 0, 0)  # b2n inverse
    torch.cuda.synchronize()
    return {"workload": "1 x 2^16 forward n2b, out of place",
            "gpu_us_per_call": times[len(times) // 2],
            "gpu_time_method": f"HIP events around a graph of {reps} back-to-back calls, median of 5 replays",
            "graph_replays_equal_eager": replays_exact,
            "eager_enqueue_us_per_call": eager_us,
            "host_latency_us_median": lat[len(lat) // 2], "host_latency_us_p99": lat[int(0.99 * (len(lat) - 1))],
            "inverse_round_trip_exact": bool(torch.equal(z, x))}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(log_n, cols, threads, warmup=3, runs=10):
    """The oracle's multithreaded NTT (threads over columns) on the full
    matrix: `warmup` untimed calls, then the median of `runs` timed calls."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    n = 1 << log_n
    inp = oracle.matrix_uniform(cols, n, 2)
    out = np.empty_like(inp)
    times = []
    for k in range(warmup + runs):
        t0 = time.perf_counter()
        oracle.batch_ntt(inp, log_n, cols, n, n, False, False, out=out, threads=threads)
        if k >= warmup:
            times.append(time.perf_counter() - t0)
    times.sort()
    dt = times[len(times) // 2]
    # 1 x 2^16 forward n2b on one thread, 3 warm-up + 15 timed runs, median
    one = oracle.matrix_uniform(1, 1 << 16, 1)
    single = []
    for k in range(18):
        t1 = time.perf_counter()
        oracle.batch_ntt(one, 16, 1, 1 << 16, 1 << 16, False, False)
        if k >= 3:
            single.append(time.perf_counter() - t1)
    single.sort()
    return {"value": cols * n / dt, "unit": "elem/s", "cores": threads, "kind": "port",
            "sample": f"{cols} columns x 2^{log_n}, forward n2b, {warmup} warm-up + median of {runs} calls "
                      f"({dt:.3f} s per call, min {times[0]:.3f}, max {times[-1]:.3f})",
            "nproc": os.cpu_count(), "cpus_allowed": len(os.sched_getaffinity(0)), "cpu_model": cpu_model(),
            "config1_single_thread_ms_median": single[len(single) // 2] * 1e3}


def measured_traffic(log_n, cols):
    """HBM bytes per call from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (copy-calibrated), or None for other shapes."""
    path = os.path.join(ROOT, "profiles", f"ntt_{cols}x2^{log_n}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("log_n") != log_n or rec.get("columns") != cols:
        return None
    return rec["hbm_bytes_per_call"]


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def other_legs(args):
    """Every measurement leg after the batched NTT, as one dict."""
    import torch
    log_n, cols = args.log_n, args.cols
    extras = {}
    if not args.no_latency:
        extras["latency"] = bench_latency()
        progress(f"latency done: {extras['latency']['gpu_us_per_call']:.1f} us per 2^16 call")
    if not args.no_lde:
        extras["lde"] = bench_lde(20, 128, 2, max(3, args.steps // 2), 2)
        torch.cuda.empty_cache()
        progress(f"coset LDE done: {extras['lde']['ms_per_call']:.3f} ms per call")
    if not args.no_ragged:
        extras["ragged"] = bench_ragged(args.ragged_jobs, 10, 23, 5, 1, plain=not args.ragged_no_plain)
        progress(f"ragged done: {extras['ragged']['ms_ragged']:.2f} ms vs plain {extras['ragged']['ms_plain']:.2f} ms")
    if not args.no_ops:
        torch.cuda.empty_cache()
        extras["ops"] = bench_ops(log_n, cols, 5)
        progress(f"ops done: bit_reverse {extras['ops']['bit_reverse']['GB/s']:.0f} GB/s")
    if not args.no_bary:
        torch.cuda.empty_cache()
        extras["barycentric"] = bench_barycentric(log_n, cols, 5)
        progress(f"barycentric done: base_at_ext {extras['barycentric']['base_at_ext']['GB/s']:.0f} GB/s")
    if not args.no_poseidon:
        torch.cuda.empty_cache()
        extras["poseidon"] = bench_poseidon(log_n, 128, 3)
        progress(f"poseidon done: tree {extras['poseidon']['poseidon']['ms_tree']:.2f} ms, "
                 f"poseidon2 {extras['poseidon']['poseidon2']['ms_tree']:.2f} ms")
    if not args.no_serve:
        torch.cuda.empty_cache()
        extras["serve"] = bench_serve(args.serve_requests, args.serve_rate, int(args.serve_budget_gb * 1e9))
        progress(f"serving done: p50 {extras['serve']['p50_latency_ms']:.1f} ms, "
                 f"p99 {extras['serve']['p99_latency_ms']:.1f} ms")
        if not args.no_serve_sweep:
            torch.cuda.empty_cache()
            extras["serve"]["sweep"] = bench_serve_sweep((10, 25, 50, 100, 200), 60, int(args.serve_budget_gb * 1e9))
    return extras


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--cols", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cols", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (default: every core this process may run on, at most "
                         "the 16-core host share of one GPU)")
    ap.add_argument("--no-lde", action="store_true", help="skip the coset-LDE measurement")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged-batch measurement")
    ap.add_argument("--ragged-jobs", type=int, default=1000)
    ap.add_argument("--ragged-no-plain", action="store_true",
                    help="skip the per-job comparison path of the ragged leg (profiling runs)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-ops", action="store_true", help="skip the bit-reverse / powers / omega-shift measurement")
    ap.add_argument("--no-bary", action="store_true", help="skip the barycentric-evaluation measurement")
    ap.add_argument("--no-latency", action="store_true", help="skip the single-column 2^16 latency measurement")
    ap.add_argument("--no-poseidon", action="store_true", help="skip the Poseidon Merkle-tree measurement")
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--serve-rate", type=float, default=25.0, help="offered load, requests per second")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-serve-sweep", action="store_true", help="skip the serving offered-load sweep")
    ap.add_argument("--legs-only", action="store_true",
                    help="skip the batched-NTT measurement and print only the other legs (for per-leg profiling)")
    ap.add_argument("--mode", choices=["ntt", "copy"], default="ntt",
                    help="copy: time a plain device copy of the same matrix (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (columns shard by range "
              f"with no communication, so one GPU's share is the measurement)", file=sys.stderr)
        sys.exit(2)

    import torch
    from boojum_cuda_amd import ntt
    from boojum_cuda_amd.context import Context

    torch.cuda.set_device(0)
    ctx = Context.create(12, 12)
    log_n, cols = args.log_n, args.cols
    n = 1 << log_n
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    x = uniform_field(cols * n, g)
    y = torch.empty_like(x)
    if args.mode == "copy":
        for _ in range(args.warmup):
            y.copy_(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            y.copy_(x)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({"mode": "copy", "bytes_per_copy": 16 * cols * n, "ms": dt * 1e3,
                          "GBs": 16 * cols * n / dt / 1e9}), flush=True)
        ctx.destroy()
        return
    variants = [(False, False), (True, False), (False, True), (True, True)]  # (bitrev_inputs, inverse)
    if args.legs_only:
        del x, y
        torch.cuda.empty_cache()
        print(json.dumps(other_legs(args)), flush=True)
        ctx.destroy()
        return

    def step():
        for br, inv in variants:
            ntt.batch_ntt_out_of_place(x, y, log_n, cols, 0, 0, n, n, br, inv, 0, 0)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps * len(variants))]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k = 0
    for _ in range(args.steps):
        for br, inv in variants:
            ev[k].record(stream)
            ntt.batch_ntt_out_of_place(x, y, log_n, cols, 0, 0, n, n, br, inv, 0, 0)
            ev[k + 1].record(stream)
            k += 2
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    call_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(0, k, 2)]
    avg_call_ms = sum(call_ms) / len(call_ms)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    elems_per_step = len(variants) * cols * n
    value = elems_per_step / (ms_per_step * 1e-3)
    algo_bytes = 16.0 * cols * n
    achieved = algo_bytes / (avg_call_ms * 1e-3) / 1e9
    # measured copy peak: a device-to-device copy of the same 2^28-element
    # matrix (16 B moved per element), timed like the transform calls
    y.copy_(x)
    cev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    cev[0].record(stream)
    for _ in range(5):
        y.copy_(x)
    cev[1].record(stream)
    torch.cuda.synchronize()
    copy_gbs = algo_bytes / (cev[0].elapsed_time(cev[1]) / 5 * 1e-3) / 1e9
    del x, y
    progress(f"batched NTT done: {avg_call_ms:.3f} ms per call")
    extras = other_legs(args)
    ctx.destroy()
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"batched NTT {cols} x 2^{log_n} Goldilocks, fwd/inv x n2b/b2n, out-of-place",
                   "columns": cols, "log_n": log_n, "calls_per_step": len(variants)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(log_n, cols),
                     "kernel": "one batch_ntt call (all its pass kernels)", "avg_call_ms": avg_call_ms,
                     "algorithmic_bytes_per_call": algo_bytes, "measured_copy_peak": copy_gbs,
                     "frac_of_copy_peak": achieved / copy_gbs},
    }
    rec.update(extras)
    if not args.no_cpu:
        # every host core this process may run on, at most the 16 a one-GPU
        # share of the host provides (on a one-GPU box that is all of them);
        # --cpu-threads overrides it on a larger host
        threads = args.cpu_threads or min(CPU_SHARE, len(os.sched_getaffinity(0)))
        progress(f"cpu baseline on {threads} threads")
        rec["cpu_baseline"] = cpu_baseline(log_n, args.cpu_cols, threads)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
