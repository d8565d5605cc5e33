"""GPU check of bc_ragged_run_no_alloc, the allocation-free ragged entry point
(plan table and queue state in the library's static device array): two calls
in a row, the second on a smaller batch, each bit-identical to the same jobs
through bc_ragged_plan_create + bc_ragged_run, and the first also against the
oracle.  The C ABI is called directly through ctypes."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _jobs(specs, seed):
    import test_ragged_gpu as tr
    return tr.make_jobs(specs, seed)


def _run_plan(jobs):
    from boojum_cuda_amd import ragged
    plan = ragged.RaggedPlan(jobs)
    try:
        plan.run()
    finally:
        import torch
        torch.cuda.synchronize()
        plan.destroy()


def _outputs(jobs):
    import test_ragged_gpu as tr
    return [tr.to_host(j.outputs).copy() for j in jobs]


def _run_no_alloc_abi(jobs):
    import torch
    from boojum_cuda_amd import ragged
    from boojum_cuda_amd._lib import lib
    arr = ragged._array(jobs)
    rc = lib().bc_ragged_run_no_alloc(arr, len(jobs), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, rc
    torch.cuda.synchronize()


def test_no_alloc_matches_planned_run_twice(gpu_ctx):
    import test_ragged_gpu as tr
    from boojum_cuda_amd import ragged
    big = tr.specs_mixed(11, 30, 8, 16) + [(ragged.OP_NTT, 20, 2, 0), (ragged.OP_LDE, 17, 1, 2),
                                           (ragged.OP_INTT, 19, 1, 0)]
    small = tr.specs_mixed(12, 5, 8, 12)
    for specs, seed, with_oracle in ((big, 900, True), (small, 950, False)):
        jobs, hosts = _jobs(specs, seed)
        _run_plan(jobs)
        planned = _outputs(jobs)
        for j in jobs:
            j.outputs.fill_(-1)  # stale values the call must overwrite
            if j.coeffs is not None:
                j.coeffs.fill_(-1)
        _run_no_alloc_abi(jobs)
        got = _outputs(jobs)
        for k, (j, a, b) in enumerate(zip(jobs, got, planned)):
            width = (1 << j.log_n) << j.log_blowup
            for c in range(j.num_cols):  # column windows (the padding between columns is not written)
                w = slice(c * j.stride_out, c * j.stride_out + width)
                assert np.array_equal(a[w], b[w]), (seed, k, c, specs[k])
        if with_oracle:
            tr.check(jobs, hosts)


def test_no_alloc_empty_and_bad_jobs(gpu_ctx):
    import torch
    from boojum_cuda_amd import ragged
    from boojum_cuda_amd._lib import lib
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().bc_ragged_run_no_alloc(None, 0, s) == 0
    assert lib().bc_ragged_run_no_alloc(None, 3, s) == 1
    x = torch.zeros(1 << 10, dtype=torch.int64, device="cuda")
    bad = ragged.Job(op=7, log_n=10, inputs=x, outputs=x)
    assert lib().bc_ragged_run_no_alloc(ragged._array([bad]), 1, s) == 1
    torch.cuda.synchronize()
    assert oracle is not None
