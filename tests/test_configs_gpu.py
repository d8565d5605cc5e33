"""GPU parity at the full shapes of the measured configurations, with the
same seeded jobs the benchmark runs (boojum_cuda_amd.workloads):

- coset LDE: 128 columns x 2^20, blowup 4, replayed from a captured hipGraph;
  coefficients and all four cosets of sampled columns against the oracle,
  plus Horner evaluation of the coefficients at points of g H_{2^22};
- ragged batch: the 1000-job batch (log_n uniform in [10, 23]); every job
  against the per-job path, every job with log_n >= 22 and 20 others against
  the oracle;
- serving: requests drawn from the serving distribution (1-512 columns,
  log_n uniform in [12, 24], blowup 1, 2 or 4) plus a log_n 24 blowup-4
  request split into parts, under a 200 GB budget; sampled columns of the
  kept outputs against the oracle, and every kept output equal to the same
  request run alone.

The oracle runs in a thread pool (its C calls release the GIL)."""
import ctypes
import os
import random
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, os.cpu_count() or 1))


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def lde_expected(evals, log_n, k):
    """Bit-reversed evaluations -> natural coefficients -> 2^k cosets, coset c
    in block brev(c), bit-reversed (the reference's call sequence)."""
    n = 1 << log_n
    coeffs = oracle.ntt_column(evals, log_n, True, True)
    out = np.empty(n << k, dtype=np.uint64)
    for c in range(1 << k):
        blk = oracle.bitrev(c, k)
        out[blk * n:(blk + 1) * n] = oracle.ntt_column(coeffs, log_n, False, False, k, c)
    return coeffs, out


def test_config_lde_128x2e20_blowup4_graph(gpu_ctx):
    import torch
    from boojum_cuda_amd import lde, workloads
    shape = workloads.LDE_SHAPE
    log_n, cols, k = shape["log_n"], shape["cols"], shape["log_blowup"]
    n = 1 << log_n
    g = torch.Generator(device="cuda")
    g.manual_seed(shape["seed"])
    evals = workloads.uniform_field(cols * n, g)
    coeffs = torch.empty_like(evals)
    out = torch.empty(cols * (n << k), dtype=torch.int64, device="cuda")
    graph = lde.LdeGraph(evals, coeffs, out, log_n, cols, n, n, n << k, k, True)
    try:
        graph.launch()
        first = out[: n << k].clone()
        out.zero_()
        graph.launch()  # a replay gives the same bits
        torch.cuda.synchronize()
        assert torch.equal(first, out[: n << k])
    finally:
        graph.destroy()
    sample = [0, 63, cols - 1]
    jobs = [(c, to_host(evals[c * n:(c + 1) * n]).copy(), to_host(coeffs[c * n:(c + 1) * n]).copy(),
             to_host(out[c * (n << k):(c + 1) * (n << k)]).copy()) for c in sample]

    def check(job):
        c, ev, co, got = job
        want_co, want = lde_expected(ev, log_n, k)
        assert np.array_equal(co, want_co), f"coefficients of column {c}"
        for cs in range(1 << k):
            blk = oracle.bitrev(cs, k)
            assert np.array_equal(got[blk * n:(blk + 1) * n], want[blk * n:(blk + 1) * n]), f"column {c} coset {cs}"
        # the column is the bit-reversed evaluation of the coefficients on g H_{n 2^k}
        natural = oracle.bitreverse(got)
        w = oracle.domain_generator(log_n + k)
        for m in [0, 1, 5, n - 1, n, 3 * n + 17, (n << k) - 1]:
            assert int(natural[m]) == oracle.horner(co, 7 * pow(w, m, oracle.P) % oracle.P), (c, m)
        return c

    with ThreadPoolExecutor(WORKERS) as ex:
        assert sorted(ex.map(check, jobs)) == sample
    del evals, coeffs, out
    torch.cuda.empty_cache()


def ragged_expected(op, log_n, lb, x):
    from boojum_cuda_amd import ragged
    if op == ragged.OP_NTT:
        return oracle.ntt_column(x, log_n, False, False)
    if op == ragged.OP_INTT:
        return oracle.ntt_column(x, log_n, True, True)
    return lde_expected(x, log_n, lb)[1]


@pytest.mark.timeout(900)
def test_config_ragged_1000_jobs(gpu_ctx):
    import torch
    from boojum_cuda_amd import ragged, workloads
    specs = workloads.ragged_specs()
    jobs = workloads.ragged_jobs(specs)
    assert len(jobs) == 1000 and max(s[0] for s in specs) == 23 and min(s[0] for s in specs) == 10
    plan = ragged.RaggedPlan(jobs)
    try:
        plan.run()  # one persistent launch for the whole batch
        torch.cuda.synchronize()
        plan.wait_status()
        mixed = [j.outputs.clone() for j in jobs]
    finally:
        plan.destroy()
    for j in jobs:
        j.outputs.zero_()
    ragged.run_plain(jobs)
    torch.cuda.synchronize()
    for k, (j, m) in enumerate(zip(jobs, mixed)):
        assert torch.equal(j.outputs, m), f"job {k} {specs[k]}: ragged differs from the per-job path"
    del mixed
    rng = random.Random(44)
    big = [k for k, s in enumerate(specs) if s[0] >= 22]
    rest = rng.sample([k for k, s in enumerate(specs) if s[0] < 22], 20)
    chosen = big + rest
    assert len(big) > 50

    def check(k):
        log_n, op, lb = specs[k]
        x = to_host(jobs[k].inputs).copy()
        got = to_host(jobs[k].outputs).copy()
        assert np.array_equal(got, ragged_expected(op, log_n, lb, x)), f"job {k} {specs[k]}"
        return k

    with ThreadPoolExecutor(WORKERS) as ex:
        assert sorted(ex.map(check, chosen)) == sorted(chosen)
    del jobs
    torch.cuda.empty_cache()


_hip = None


def hip_memcpy(dst, src, nbytes, kind):
    """hipMemcpy through the HIP runtime torch has loaded (kind 2: D2H, 3: D2D)."""
    global _hip
    if _hip is None:
        import torch
        torch.cuda.init()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        assert path, "libamdhip64 is not loaded"
        _hip = ctypes.CDLL(path)
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    rc = _hip.hipMemcpy(dst, src, nbytes, kind)
    assert rc == 0, rc


def kept_outputs(srv, rid, log_n, cols, lb):
    """A completed keep_outputs request's columns, copied into one device tensor."""
    import torch
    width = (1 << log_n) << lb
    dev = torch.empty(cols * width, dtype=torch.int64, device="cuda")
    seen = 0
    for p in srv.parts(rid):
        assert p["stride_out"] == width
        hip_memcpy(dev.data_ptr() + p["first_col"] * width * 8, p["outputs"], p["num_cols"] * width * 8, 3)
        seen += p["num_cols"]
    assert seen == cols
    return dev


@pytest.mark.timeout(900)
def test_config_serving_mix(gpu_ctx):
    import time
    import torch
    from boojum_cuda_amd import serve, workloads
    free, _ = torch.cuda.mem_get_info()
    budget = int(min(200e9, free - (72 << 30)))
    assert budget > 100e9, f"only {free / 2**30:.0f} GiB free on the device"
    # requests of the serving distribution (the benchmark's first twelve), kept
    # when their outputs are at most 4 GiB (up to 32 GiB in all), streamed
    # otherwise; and one log_n 24 blowup-4 request of 24 columns (12 GiB of
    # kept outputs, split into parts)
    reqs, kept_total = [], 0
    for k, (at, log_n, cols, lb) in enumerate(workloads.serve_specs(12, 25.0)):
        out_bytes = cols * (8 << (log_n + lb))
        keep = out_bytes <= 4 << 30 and kept_total + out_bytes <= 32 << 30
        kept_total += out_bytes if keep else 0
        reqs.append(dict(at=at, log_n=log_n, cols=cols, lb=lb, seed=1000 + k, keep=keep))
    reqs.insert(3, dict(at=reqs[3]["at"], log_n=24, cols=24, lb=2, seed=77, keep=True))
    srv = serve.Server(budget)
    try:
        t0 = time.perf_counter()
        for r in reqs:
            while time.perf_counter() < t0 + r["at"]:
                time.sleep(0.0005)
            op = serve.OP_LDE if r["lb"] else serve.OP_NTT
            r["id"] = srv.submit(op, r["log_n"], r["cols"], r["lb"], seed=r["seed"], keep_outputs=r["keep"])
        srv.drain()
        st = srv.stats()
        assert st["requests_done"] == len(reqs) and st["pending_parts"] == 0
        assert st["arena_high_water_bytes"] <= budget
        big = reqs[3]
        assert srv.info(big["id"])["parts"] >= 4, "the log_n 24 blowup-4 request was not split"
        kept = [r for r in reqs if r["keep"]]
        assert len(kept) >= 4
        mixed = {}
        for r in reqs:
            assert srv.info(r["id"])["status"] == 1
            if r["keep"]:
                mixed[r["id"]] = kept_outputs(srv, r["id"], r["log_n"], r["cols"], r["lb"])
            srv.release(r["id"])
        assert srv.stats()["arena_bytes_in_use"] == 0
        # each kept request alone on the idle server gives the same bits
        for r in kept:
            op = serve.OP_LDE if r["lb"] else serve.OP_NTT
            rid = srv.submit(op, r["log_n"], r["cols"], r["lb"], seed=r["seed"], keep_outputs=True)
            srv.wait(rid)
            alone = kept_outputs(srv, rid, r["log_n"], r["cols"], r["lb"])
            srv.release(rid)
            assert torch.equal(alone, mixed[r["id"]]), (r["log_n"], r["cols"], r["lb"])
            del alone
    finally:
        srv.destroy()
    # sampled columns against the oracle: the first and last column of the big
    # request, the last column of every other kept request
    work = []
    for r in kept:
        width = (1 << r["log_n"]) << r["lb"]
        cols = [0, r["cols"] - 1] if r is big else [r["cols"] - 1]
        for c in cols:
            work.append((r, c, to_host(mixed[r["id"]][c * width:(c + 1) * width]).copy()))
    del mixed
    torch.cuda.empty_cache()

    def check(item):
        r, c, got = item
        x = oracle.fill_raw(1 << r["log_n"], r["seed"], c)
        if r["lb"]:
            want = lde_expected(x, r["log_n"], r["lb"])[1]
        else:
            want = oracle.ntt_column(x, r["log_n"], False, False)
        assert np.array_equal(got, want), (r["log_n"], r["cols"], r["lb"], c)
        return c

    with ThreadPoolExecutor(WORKERS) as ex:
        assert len(list(ex.map(check, work))) == len(work)
