"""GPU parity of the ragged persistent kernel: a mixed batch of NTT, iNTT and
coset-LDE jobs of different sizes gives, job for job, the oracle's result
and the same bits as running each job alone."""
import random

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to("cuda")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def make_jobs(specs, seed):
    import torch
    from boojum_cuda_amd import ragged
    jobs, hosts = [], []
    for k, (op, log_n, cols, blow) in enumerate(specs):
        n = 1 << log_n
        stride_in = n + 16 * (k % 3)
        x = np.zeros(cols * stride_in, dtype=np.uint64)
        for c in range(cols):
            x[c * stride_in: c * stride_in + n] = oracle.fill_raw(n, seed + k, c)
        out_len = n << blow
        stride_out = out_len + 8 * (k % 2)
        j = ragged.Job(op=op, log_n=log_n, num_cols=cols, inputs=to_dev(x),
                       outputs=torch.zeros(cols * stride_out, dtype=torch.int64, device="cuda"),
                       stride_in=stride_in, stride_out=stride_out, log_blowup=blow)
        if op == ragged.OP_LDE:
            j.coeffs = torch.zeros(cols * n, dtype=torch.int64, device="cuda")
            j.stride_coeffs = n
        jobs.append(j)
        hosts.append(x)
    return jobs, hosts


def expected(job, x):
    from boojum_cuda_amd import ragged
    n = 1 << job.log_n
    outs = []
    for c in range(job.num_cols):
        col = x[c * job.stride_in: c * job.stride_in + n]
        if job.op == ragged.OP_NTT:
            outs.append(oracle.ntt_column(col, job.log_n, False, False))
        elif job.op == ragged.OP_INTT:
            outs.append(oracle.ntt_column(col, job.log_n, True, True))
        else:
            coeffs = oracle.ntt_column(col, job.log_n, True, True)
            full = np.zeros(n << job.log_blowup, dtype=np.uint64)
            for cs in range(1 << job.log_blowup):
                blk = oracle.bitrev(cs, job.log_blowup)
                full[blk * n:(blk + 1) * n] = oracle.ntt_column(coeffs, job.log_n, False, False, job.log_blowup, cs)
            outs.append(full)
    return outs


def check(jobs, hosts):
    for j, x in zip(jobs, hosts):
        got = to_host(j.outputs)
        want = expected(j, x)
        for c, w in enumerate(want):
            g = got[c * j.stride_out: c * j.stride_out + w.size]
            assert np.array_equal(g, w), (j.op, j.log_n, j.num_cols, j.log_blowup, c)


def specs_mixed(seed, count, log_lo, log_hi):
    from boojum_cuda_amd import ragged
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        kind = rng.choice(["ntt", "intt", "lde2", "lde4"])
        log_n = rng.randint(log_lo, log_hi)
        cols = rng.randint(1, 3)
        if kind == "ntt":
            out.append((ragged.OP_NTT, log_n, cols, 0))
        elif kind == "intt":
            out.append((ragged.OP_INTT, log_n, cols, 0))
        else:
            out.append((ragged.OP_LDE, log_n, cols, 1 if kind == "lde2" else 2))
    return out


def test_ragged_mixed_batch_matches_oracle(gpu_ctx):
    from boojum_cuda_amd import ragged
    specs = specs_mixed(1, 24, 8, 15) + [(ragged.OP_NTT, 20, 2, 0), (ragged.OP_INTT, 21, 1, 0),
                                          (ragged.OP_LDE, 17, 2, 2), (ragged.OP_LDE, 13, 1, 1)]
    jobs, hosts = make_jobs(specs, 100)
    plan = ragged.RaggedPlan(jobs)
    try:
        info = plan.info()
        assert info["levels"] >= 3 and info["units"] > 0
        plan.run()
        check(jobs, hosts)
        plan.wait_status()
        # a second run of the same plan (counters reset on the device)
        for j in jobs:
            j.outputs.zero_()
        plan.run()
        check(jobs, hosts)
        plan.wait_status()
    finally:
        plan.destroy()


def test_ragged_wait_timeout_is_reported(gpu_ctx):
    """The single launch's waits between levels are bounded.  A wait limit of
    one poll (set from the host) makes the early workgroups give up: the run
    still ends, its status reports the timeout, and with the default limit
    the next run of the same plan is exact and reports success."""
    import torch
    from boojum_cuda_amd import ragged
    from boojum_cuda_amd._lib import HipError
    specs = [(ragged.OP_LDE, 16, 2, 2), (ragged.OP_NTT, 20, 1, 0), (ragged.OP_INTT, 18, 3, 0)]
    jobs, hosts = make_jobs(specs, 900)
    plan = ragged.RaggedPlan(jobs)
    try:
        assert plan.info()["levels"] >= 3
        ragged.set_wait_limit(1)
        try:
            plan.run()
            torch.cuda.synchronize()
        finally:
            ragged.set_wait_limit(0)
        with pytest.raises(HipError) as e:
            plan.wait_status()
        assert e.value.code == 702
        for j in jobs:
            j.outputs.zero_()
        plan.run()
        torch.cuda.synchronize()
        plan.wait_status()
        check(jobs, hosts)
    finally:
        plan.destroy()


def test_ragged_run_in_a_graph_replays_exact(gpu_ctx):
    """bc_ragged_run on a stream under graph capture launches once per level
    (the single cooperative launch is not captured): every replay, with eager
    work of other shapes between replays, gives the eager run's bits."""
    import torch
    from boojum_cuda_amd import ragged
    specs = specs_mixed(3, 10, 9, 16) + [(ragged.OP_LDE, 15, 2, 2)]
    jobs, hosts = make_jobs(specs, 300)
    plan = ragged.RaggedPlan(jobs)
    other_jobs, _ = make_jobs(specs_mixed(4, 6, 8, 13), 400)
    other = ragged.RaggedPlan(other_jobs)
    try:
        plan.run()
        torch.cuda.synchronize()
        plan.wait_status()
        want = [to_host(j.outputs).copy() for j in jobs]
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                plan.run(stream=side)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        for k in range(3):
            for j in jobs:
                j.outputs.zero_()
            graph.replay()
            torch.cuda.synchronize()
            for j, w in zip(jobs, want):
                assert np.array_equal(to_host(j.outputs), w), k
            other.run()  # eager work between replays
            torch.cuda.synchronize()
        check(jobs, hosts)
    finally:
        other.destroy()
        plan.destroy()


def test_ragged_equals_plain_and_alone(gpu_ctx):
    from boojum_cuda_amd import ragged
    specs = specs_mixed(7, 40, 10, 18)
    jobs, hosts = make_jobs(specs, 500)
    plan = ragged.RaggedPlan(jobs)
    plan.run()
    mixed = [to_host(j.outputs).copy() for j in jobs]
    plan.destroy()
    for j in jobs:
        j.outputs.zero_()
    ragged.run_plain(jobs)
    plain = [to_host(j.outputs).copy() for j in jobs]
    for a, b in zip(mixed, plain):
        assert np.array_equal(a, b)
    # each job alone through its own one-job plan
    for k in [0, 5, 17, 39]:
        jobs[k].outputs.zero_()
        p1 = ragged.RaggedPlan([jobs[k]])
        p1.run()
        assert np.array_equal(to_host(jobs[k].outputs), mixed[k])
        p1.destroy()


def test_ragged_rejects_bad_jobs(gpu_ctx):
    import torch
    from boojum_cuda_amd import ragged
    x = torch.zeros(1 << 10, dtype=torch.int64, device="cuda")
    with pytest.raises(AssertionError):
        ragged.RaggedPlan([ragged.Job(op=ragged.OP_NTT, log_n=7, inputs=x, outputs=x)])  # below 2^8
    with pytest.raises(AssertionError):
        ragged.RaggedPlan([ragged.Job(op=ragged.OP_LDE, log_n=9, inputs=x, outputs=x, log_blowup=1)])  # no coeffs
