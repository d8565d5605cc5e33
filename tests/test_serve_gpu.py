"""GPU tests of the serving loop (bc_server_*): requests that arrive while
others run are admitted under an HBM budget and batched together; every
result must equal the oracle's and the result of the same request run alone,
bit for bit, however it was split and batched."""
import ctypes
import random
import time

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

_hip = None


def hip_runtime():
    """The HIP runtime already loaded in this process (by torch and the library)."""
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
    return _hip


def d2h(ptr, count):
    out = np.empty(count, dtype=np.uint64)
    rc = hip_runtime().hipMemcpy(out.ctypes.data, ptr, count * 8, 2)  # hipMemcpyDeviceToHost
    assert rc == 0, rc
    return out


def expected_column(op, log_n, lb, x):
    from boojum_cuda_amd import serve
    n = 1 << log_n
    if op == serve.OP_NTT:
        return oracle.ntt_column(x, log_n, False, False)
    if op == serve.OP_INTT:
        return oracle.ntt_column(x, log_n, True, True)
    coeffs = oracle.ntt_column(x, log_n, True, True)
    full = np.zeros(n << lb, dtype=np.uint64)
    for c in range(1 << lb):
        blk = oracle.bitrev(c, lb)
        full[blk * n:(blk + 1) * n] = oracle.ntt_column(coeffs, log_n, False, False, lb, c)
    return full


def gather(server, rid, log_n, cols, lb):
    """Columns of a completed keep_outputs request, from all of its parts."""
    width = (1 << log_n) << lb
    got = [None] * cols
    for p in server.parts(rid):
        assert p["stride_out"] == width
        block = d2h(p["outputs"], p["num_cols"] * width)
        for i in range(p["num_cols"]):
            got[p["first_col"] + i] = block[i * width:(i + 1) * width]
    assert all(g is not None for g in got)
    return got


def mixed_specs(seed, count):
    from boojum_cuda_amd import serve
    rng = random.Random(seed)
    specs = []
    for k in range(count):
        kind = rng.choice(["ntt", "intt", "lde1", "lde2"])
        op = {"ntt": serve.OP_NTT, "intt": serve.OP_INTT}.get(kind, serve.OP_LDE)
        lb = {"lde1": 1, "lde2": 2}.get(kind, 0)
        specs.append((op, rng.randint(8, 13), rng.randint(1, 5), lb, 1000 + k))
    return specs


def test_fill_seeded_matches_oracle(gpu_ctx):
    import torch
    from boojum_cuda_amd import serve
    log_n, cols, stride, seed, first = 10, 3, (1 << 10) + 8, 11, 2
    buf = torch.zeros(cols * stride, dtype=torch.int64, device="cuda")
    serve.fill_seeded(buf, log_n, cols, stride, seed, first)
    torch.cuda.synchronize()
    got = buf.cpu().numpy().view(np.uint64)
    for c in range(cols):
        assert np.array_equal(got[c * stride:c * stride + (1 << log_n)], oracle.fill_raw(1 << log_n, seed, first + c))


def test_serve_mix_matches_oracle_and_alone(gpu_ctx):
    from boojum_cuda_amd import serve
    specs = mixed_specs(5, 16)
    # a small budget and small parts: requests are split, queue for memory and share batches
    cfg = dict(budget_bytes=48 << 20, max_part_elems=1 << 15, max_batches_in_flight=3)
    srv = serve.Server(**cfg)
    try:
        rng = random.Random(9)
        ids = []
        for op, log_n, cols, lb, seed in specs:
            ids.append(srv.submit(op, log_n, cols, lb, seed=seed, keep_outputs=True))
            time.sleep(rng.random() * 0.002)
        mixed = []
        for rid, (op, log_n, cols, lb, seed) in zip(ids, specs):
            srv.wait(rid)
            info = srv.info(rid)
            assert info["status"] == 1 and info["parts_done"] == info["parts"]
            assert info["t_submit_ns"] <= info["t_first_launch_ns"] <= info["t_done_ns"]
            got = gather(srv, rid, log_n, cols, lb)
            for c in range(cols):
                want = expected_column(op, log_n, lb, oracle.fill_raw(1 << log_n, seed, c))
                assert np.array_equal(got[c], want), (op, log_n, cols, lb, c)
            mixed.append(got)
        st = srv.stats()
        assert st["requests_done"] == len(specs) and st["pending_parts"] == 0
        assert st["arena_high_water_bytes"] <= cfg["budget_bytes"]
        assert st["parts_done"] > len(specs)  # requests were split
        for rid in ids:
            srv.release(rid)
        assert srv.stats()["arena_bytes_in_use"] == 0
        # the same requests, each alone on an idle server
        for (op, log_n, cols, lb, seed), got_mixed in zip(specs, mixed):
            rid = srv.submit(op, log_n, cols, lb, seed=seed, keep_outputs=True)
            srv.wait(rid)
            alone = gather(srv, rid, log_n, cols, lb)
            srv.release(rid)
            for a, b in zip(alone, got_mixed):
                assert np.array_equal(a, b)
    finally:
        srv.destroy()


def test_serve_budget_admission(gpu_ctx):
    from boojum_cuda_amd import serve
    srv = serve.Server(budget_bytes=8 << 20, max_part_elems=1 << 20)
    try:
        # kept outputs (8 columns x 2 MiB) exceed the 8 MiB budget: refused
        with pytest.raises(serve.OutOfBudget):
            srv.submit(serve.OP_LDE, 16, 8, 2, seed=3, keep_outputs=True)
        # one column that cannot fit at all: refused
        with pytest.raises(serve.OutOfBudget):
            srv.submit(serve.OP_LDE, 20, 1, 2, seed=3)
        with pytest.raises(AssertionError):
            srv.submit(serve.OP_NTT, 7, 1)
        with pytest.raises(AssertionError):
            srv.submit(5, 10, 1)
        # streamed outputs: split into parts that run one after another within the budget
        rid = srv.submit(serve.OP_LDE, 16, 8, 2, seed=3)
        srv.wait(rid)
        info = srv.info(rid)
        assert info["status"] == 1 and info["parts"] >= 4
        st = srv.stats()
        assert st["arena_high_water_bytes"] <= 8 << 20 and st["arena_bytes_in_use"] == 0
        srv.release(rid)
    finally:
        srv.destroy()


def test_serve_device_inputs(gpu_ctx):
    import torch
    from boojum_cuda_amd import serve
    log_n, cols = 12, 3
    n = 1 << log_n
    stride = n + 16
    host = np.zeros(cols * stride, dtype=np.uint64)
    for c in range(cols):
        host[c * stride:c * stride + n] = oracle.fill_raw(n, 21, c)
    x = torch.from_numpy(host.view(np.int64)).to("cuda")
    torch.cuda.synchronize()
    srv = serve.Server(budget_bytes=16 << 20)
    try:
        rid = srv.submit(serve.OP_INTT, log_n, cols, inputs=x, stride_in=stride, keep_outputs=True)
        srv.wait(rid)
        got = gather(srv, rid, log_n, cols, 0)
        for c in range(cols):
            assert np.array_equal(got[c], oracle.ntt_column(host[c * stride:c * stride + n], log_n, True, True))
        srv.release(rid)
    finally:
        srv.destroy()
