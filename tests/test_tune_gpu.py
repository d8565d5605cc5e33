"""GPU: every tuned plan (pass split, strided pairs per workgroup, grid cap)
computes the same bits as the built-in plan, and the oracle's result."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log_n,cols", [(13, 3), (16, 2), (16, 8), (20, 2), (22, 1)])
def test_tuned_plans_match_builtin(gpu_ctx, tmp_path, log_n, cols):
    import torch
    from boojum_cuda_amd import autotune, ntt
    n = 1 << log_n
    host = np.concatenate([oracle.fill_raw(n, 77, c) for c in range(cols)])
    x = torch.from_numpy(host.view(np.int64)).to("cuda")
    y, z = torch.empty_like(x), torch.empty_like(x)
    lc = cols.bit_length() - 1

    def run():
        ntt.batch_ntt_out_of_place(x, y, log_n, cols, 0, 0, n, n, False, False, 0, 0)
        ntt.batch_ntt_out_of_place(y, z, log_n, cols, 0, 0, n, n, True, True, 0, 0)
        torch.cuda.synchronize()
        return y.clone(), z.clone()

    saved = str(tmp_path / "table.txt")
    autotune.save(saved)
    try:
        autotune.clear()
        ref_y, ref_z = run()
        for c in range(cols):
            col = host[c * n:(c + 1) * n]
            assert np.array_equal(ref_y[c * n:(c + 1) * n].cpu().numpy().view(np.uint64),
                                  oracle.ntt_column(col, log_n, False, False))
        assert np.array_equal(ref_z.cpu().numpy().view(np.uint64), oracle.canon(host))  # round trip
        variants = [autotune.TuneParams(c, k, 0, 0, 0, 0) for c, k in autotune.splits(log_n)]
        if log_n == 16:  # workgroup tile: one-wave 256-point passes and 4096-point pass kernels
            variants += [autotune.TuneParams(0, 0, 0, 0, 0, 0, t) for t in (8, 12)]
        variants += [autotune.TuneParams(0, 0, bpw, grid, 0, 0) for bpw in (1, 3, 32) for grid in (1, 700)]
        for p in variants:
            autotune.set_entry(log_n, lc, p)
            got_y, got_z = run()
            assert torch.equal(got_y, ref_y) and torch.equal(got_z, ref_z), p.key()
    finally:
        autotune.clear()
        autotune.load(saved)
