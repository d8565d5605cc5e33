"""GPU parity of Poseidon and Poseidon2 Merkle-tree hashing
(src/poseidon.rs:129-328),
shaped like the reference's own tests (src/poseidon.rs:495-700: leaves with
and without checkpoints, nodes, whole trees), bit-exact against the oracle
(oracle/poseidon_oracle.c) with the same 360 round constants.

The constants are the Grain-LFSR set of the Poseidon reference parameter
procedure, not boojum's (which the reference does not hold): the structure is
pinned (fused matrices equal the reference's tables, fused form equals the
textbook permutation), the digests under boojum's constants are not."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

P = oracle.P


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).to("cuda")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


@pytest.fixture(scope="module")
def consts(gpu_ctx):
    from boojum_cuda_amd import poseidon
    rc = oracle.pos_grain_constants()
    poseidon.set_round_constants(rc)
    return rc, oracle.pos_derive(rc)["words"]


HASHES = ["poseidon", "poseidon2"]


def oracle_leaves(vals, rows, cols, count, consts, p="poseidon", **kw):
    if p == "poseidon2":
        return oracle.pos2_leaves(vals, rows, cols, count, consts[0], **kw)
    return oracle.pos_leaves(vals, rows, cols, count, consts[0], consts[1], **kw)


def oracle_nodes(vals, count, consts, p="poseidon"):
    if p == "poseidon2":
        return oracle.pos2_nodes(vals, count, consts[0])
    return oracle.pos_nodes(vals, count, consts[0], consts[1])


@pytest.mark.parametrize("p", HASHES)
@pytest.mark.parametrize("log_rows", [0, 1, 2, 3])
@pytest.mark.parametrize("cols", [1, 3, 8, 13])
def test_leaves_match_oracle(consts, p, log_rows, cols):
    """src/poseidon.rs:495-550 (test_leaves): random values, several row
    groupings and column counts (13 pads the last chunk)."""
    import torch
    from boojum_cuda_amd import poseidon
    rows, count = 1 << log_rows, 300
    vals = np.random.default_rng(log_rows * 31 + cols).integers(0, 2**64, rows * cols * count, dtype=np.uint64)
    res = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    poseidon.launch_single_thread_leaves_kernel(to_dev(vals), res, log_rows, p=p)
    assert np.array_equal(to_host(res), oracle_leaves(vals, rows, cols, count, consts, p))


@pytest.mark.parametrize("p", HASHES)
@pytest.mark.parametrize("rows,c1,c2", [(1, 8, 5), (2, 4, 3), (8, 1, 2), (1, 16, 8)])
def test_leaves_checkpointed(consts, p, rows, c1, c2):
    """src/poseidon.rs:568-581 (checkpointed leaves): the columns in two calls,
    the first storing the capacity and the second loading it, equal one call."""
    import torch
    from boojum_cuda_amd import poseidon
    count = 257
    log_rows = rows.bit_length() - 1
    vals = np.random.default_rng(rows * 7 + c1).integers(0, P, rows * (c1 + c2) * count, dtype=np.uint64)
    split = rows * c1 * count
    res = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    poseidon.launch_single_thread_leaves_kernel(to_dev(vals[:split]), res, log_rows, store_intermediate=True, p=p)
    mid = to_host(res).copy()
    assert np.array_equal(mid, oracle_leaves(vals[:split], rows, c1, count, consts, p, store_intermediate=True))
    poseidon.launch_single_thread_leaves_kernel(to_dev(vals[split:]), res, log_rows, load_intermediate=True, p=p)
    assert np.array_equal(to_host(res), oracle_leaves(vals, rows, c1 + c2, count, consts, p))


@pytest.mark.parametrize("p", HASHES)
@pytest.mark.parametrize("count", [1, 2, 255, 256, 1000])
def test_nodes_match_oracle(consts, p, count):
    """src/poseidon.rs:583-626 (test_nodes)."""
    import torch
    from boojum_cuda_amd import poseidon
    vals = np.random.default_rng(count).integers(0, 2**64, 8 * count, dtype=np.uint64)
    res = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    poseidon.launch_single_thread_nodes_kernel(to_dev(vals), res, p=p)
    assert np.array_equal(to_host(res), oracle_nodes(vals, count, consts, p))


@pytest.mark.parametrize("p", HASHES)
@pytest.mark.parametrize("count", [16383, 16384, 16385, 20000])
def test_nodes_both_kernels(consts, p, count):
    """Layers of at most 2^14 nodes run the cooperative kernel (16 lanes per
    state), larger ones one lane per state: both against the oracle, on either
    side of the switch."""
    import torch
    from boojum_cuda_amd import poseidon
    vals = np.random.default_rng(count + 1).integers(0, 2**64, 8 * count, dtype=np.uint64)
    res = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    poseidon.launch_single_thread_nodes_kernel(to_dev(vals), res, p=p)
    assert np.array_equal(to_host(res), oracle_nodes(vals, count, consts, p))


@pytest.mark.parametrize("p", HASHES)
@pytest.mark.parametrize("log_rows", [0, 1, 2, 3])
@pytest.mark.parametrize("cols", [1, 3, 8, 13])
def test_cooperative_leaves_match_oracle(consts, p, log_rows, cols):
    """src/poseidon.rs:574-577 (poseidon2_cooperative_leaves, the test_leaves
    shapes) and 696-749 (cooperative equals single-thread): the cooperative
    kernel against the oracle and the single-thread kernel.  count 301 leaves
    a part-filled last workgroup."""
    import torch
    from boojum_cuda_amd import poseidon
    rows, count = 1 << log_rows, 301
    vals = np.random.default_rng(log_rows * 37 + cols).integers(0, 2**64, rows * cols * count, dtype=np.uint64)
    dv = to_dev(vals)
    coop = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    single = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    poseidon.launch_cooperative_leaves_kernel(dv, coop, log_rows, p=p)
    poseidon.launch_single_thread_leaves_kernel(dv, single, log_rows, p=p)
    got = to_host(coop)
    assert np.array_equal(got, oracle_leaves(vals, rows, cols, count, consts, p))
    assert np.array_equal(got, to_host(single))


@pytest.mark.parametrize("p", HASHES)
@pytest.mark.parametrize("rows,c1,c2", [(1, 8, 5), (2, 4, 3), (8, 1, 2)])
def test_cooperative_leaves_checkpointed(consts, p, rows, c1, c2):
    """src/poseidon.rs:578-581 (poseidon2_cooperative_leaves_checkpointed):
    capacity stored by one call and loaded by the next, mixing the cooperative
    and single-thread kernels across the checkpoint."""
    import torch
    from boojum_cuda_amd import poseidon
    count = 130
    log_rows = rows.bit_length() - 1
    vals = np.random.default_rng(rows * 11 + c1).integers(0, P, rows * (c1 + c2) * count, dtype=np.uint64)
    split = rows * c1 * count
    want_mid = oracle_leaves(vals[:split], rows, c1, count, consts, p, store_intermediate=True)
    want = oracle_leaves(vals, rows, c1 + c2, count, consts, p)
    for first, second in ((poseidon.launch_cooperative_leaves_kernel, poseidon.launch_cooperative_leaves_kernel),
                          (poseidon.launch_cooperative_leaves_kernel, poseidon.launch_single_thread_leaves_kernel),
                          (poseidon.launch_single_thread_leaves_kernel, poseidon.launch_cooperative_leaves_kernel)):
        res = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
        first(to_dev(vals[:split]), res, log_rows, store_intermediate=True, p=p)
        assert np.array_equal(to_host(res), want_mid)
        second(to_dev(vals[split:]), res, log_rows, load_intermediate=True, p=p)
        assert np.array_equal(to_host(res), want)


@pytest.mark.parametrize("p", HASHES)
@pytest.mark.parametrize("count", [1, 2, 255, 1000, 20000])
def test_cooperative_nodes_match_oracle(consts, p, count):
    """src/poseidon.rs:624-626 (poseidon2_cooperative_nodes) and 761-804
    (cooperative equals single-thread), including a layer above 2^14 nodes,
    where the single-thread entry point runs one lane per node."""
    import torch
    from boojum_cuda_amd import poseidon
    vals = np.random.default_rng(count + 7).integers(0, 2**64, 8 * count, dtype=np.uint64)
    dv = to_dev(vals)
    coop = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    single = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    poseidon.launch_cooperative_nodes_kernel(dv, coop, p=p)
    poseidon.launch_single_thread_nodes_kernel(dv, single, p=p)
    got = to_host(coop)
    assert np.array_equal(got, oracle_nodes(vals, count, consts, p))
    assert np.array_equal(got, to_host(single))


def test_cooperative_bad_arguments_raise(consts):
    """The cooperative entry points check their arguments as the single-thread
    ones do (src/poseidon.rs:139-155, 248-255)."""
    import torch
    from boojum_cuda_amd import poseidon
    vals = torch.zeros(8 * 3, dtype=torch.int64, device="cuda")
    with pytest.raises(AssertionError):  # results not whole digests
        poseidon.launch_cooperative_leaves_kernel(vals, torch.zeros(5, dtype=torch.int64, device="cuda"), 0)
    with pytest.raises(AssertionError):  # 6 values per leaf cannot store a capacity checkpoint
        poseidon.launch_cooperative_leaves_kernel(vals, torch.zeros(4 * 4, dtype=torch.int64, device="cuda"), 0,
                                                  store_intermediate=True, p="poseidon")
    with pytest.raises(AssertionError):  # values must be twice the results
        poseidon.launch_cooperative_nodes_kernel(vals, torch.zeros(4 * 5, dtype=torch.int64, device="cuda"))


@pytest.mark.parametrize("p", HASHES)
@pytest.mark.parametrize("log_leaves,log_rows,cols", [(0, 0, 3), (4, 0, 8), (10, 1, 5), (12, 0, 16)])
def test_merkle_tree_matches_oracle(consts, p, log_leaves, log_rows, cols):
    """src/poseidon.rs:628-694 (merkle_tree): every layer of the tree."""
    import torch
    from boojum_cuda_amd import poseidon
    leaves, rows = 1 << log_leaves, 1 << log_rows
    layers = log_leaves + 1
    vals = np.random.default_rng(log_leaves).integers(0, P, leaves * rows * cols, dtype=np.uint64)
    res = torch.zeros(8 * leaves, dtype=torch.int64, device="cuda")
    poseidon.build_merkle_tree(to_dev(vals), res, log_rows, layers_count=layers, p=p)
    got = to_host(res)
    layer = oracle_leaves(vals, rows, cols, leaves, consts, p)
    assert np.array_equal(got[:4 * leaves], layer)
    off, n = 4 * leaves, leaves
    for _ in range(layers - 1):
        n //= 2
        layer = oracle_nodes(layer, n, consts, p)
        assert np.array_equal(got[off:off + 4 * n], layer)
        off += 4 * n


@pytest.mark.parametrize("p", HASHES)
def test_merkle_tree_large_sampled(consts, p):
    """src/poseidon.rs:690-694 (merkle_tree_poseidon2_large: 2^(14 + 3)
    leaves, so layers use both node kernels): 64 sampled nodes of every layer
    against the oracle applied to their children, and sampled leaves."""
    import torch
    from boojum_cuda_amd import poseidon
    log_leaves = 17
    leaves = 1 << log_leaves
    g = torch.Generator(device="cuda")
    g.manual_seed(17)
    vals = torch.randint(-2**63, 2**63 - 1, (leaves,), dtype=torch.int64, device="cuda", generator=g)
    res = torch.zeros(8 * leaves, dtype=torch.int64, device="cuda")
    poseidon.build_merkle_tree(vals, res, 0, layers_count=log_leaves + 1, p=p)
    got = to_host(res)
    host = to_host(vals)
    rng = np.random.default_rng(3)
    pick = rng.choice(leaves, 64, replace=False)
    want = oracle_leaves(host[pick], 1, 1, 64, consts, p)
    for i in range(4):
        assert np.array_equal(got[i * leaves + pick], want[i * 64:(i + 1) * 64])
    off, n = 0, leaves
    while n > 1:
        below, n_below = got[off:off + 4 * n], n
        off += 4 * n
        n //= 2
        layer = got[off:off + 4 * n]
        pick = np.sort(rng.choice(n, min(n, 64), replace=False))
        k = pick.size
        # the picked nodes' children as a k-node layer: values[2 (g + i k) + j]
        children = np.empty(8 * k, dtype=np.uint64)
        for i in range(4):
            for j in range(2):
                children[2 * (np.arange(k) + i * k) + j] = below[i * n_below + 2 * pick + j]
        want = oracle_nodes(children, k, consts, p)
        for i in range(4):
            assert np.array_equal(layer[i * n + pick], want[i * k:(i + 1) * k]), (n, i)


@pytest.mark.parametrize("p", HASHES)
def test_leaves_large_sampled(consts, p):
    """2^20 leaves x 8 columns (one 8-element chunk each), 64 sampled leaves
    against the oracle, plus a second call giving the same digests."""
    import torch
    from boojum_cuda_amd import poseidon
    count, cols = 1 << 20, 8
    g = torch.Generator(device="cuda")
    g.manual_seed(20)
    vals = torch.randint(-2**63, 2**63 - 1, (cols * count,), dtype=torch.int64, device="cuda", generator=g)
    res = torch.zeros(4 * count, dtype=torch.int64, device="cuda")
    poseidon.launch_single_thread_leaves_kernel(vals, res, 0, p=p)
    got = to_host(res)
    host = to_host(vals)
    pick = np.random.default_rng(1).choice(count, 64, replace=False)
    sub = np.concatenate([host[c * count + pick] for c in range(cols)])  # the picked leaves as a 64-leaf matrix
    want = oracle_leaves(sub, 1, cols, 64, consts, p)
    for i in range(4):
        assert np.array_equal(got[i * count + pick], want[i * 64:(i + 1) * 64])
    res2 = torch.zeros_like(res)
    poseidon.launch_single_thread_leaves_kernel(vals, res2, 0, p=p)
    assert torch.equal(res, res2)


def test_empty_is_noop(consts):
    import torch
    from boojum_cuda_amd import poseidon
    e = torch.zeros(0, dtype=torch.int64, device="cuda")
    for p in HASHES:
        poseidon.launch_single_thread_leaves_kernel(e, e, 0, p=p)
        poseidon.launch_single_thread_nodes_kernel(e, e, p=p)


def test_bad_arguments_raise(consts):
    import torch
    from boojum_cuda_amd import poseidon
    v = torch.zeros(24, dtype=torch.int64, device="cuda")
    with pytest.raises(AssertionError):
        poseidon.launch_single_thread_leaves_kernel(v, torch.zeros(6, dtype=torch.int64, device="cuda"), 0)
    with pytest.raises(AssertionError):
        poseidon.launch_single_thread_nodes_kernel(v, torch.zeros(8, dtype=torch.int64, device="cuda"))
    with pytest.raises(AssertionError):
        poseidon.launch_single_thread_leaves_kernel(v, torch.zeros(6, dtype=torch.int64, device="cuda"), 0,
                                                    p="poseidon2")


def test_hashes_differ(consts):
    import torch
    from boojum_cuda_amd import poseidon
    vals = to_dev(np.arange(8 * 4, dtype=np.uint64))
    a = torch.zeros(16, dtype=torch.int64, device="cuda")
    b = torch.zeros(16, dtype=torch.int64, device="cuda")
    poseidon.launch_single_thread_leaves_kernel(vals, a, 0, p="poseidon")
    poseidon.launch_single_thread_leaves_kernel(vals, b, 0, p="poseidon2")
    assert not torch.equal(a, b)


# ---- Merkle openings: gather_rows / gather_merkle_paths -------------------------

def idx_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to("cuda")


@pytest.mark.parametrize("log_rpi,index_hi", [(1, 42), (1, 2048), (0, 4096), (3, 512)])
def test_gather_rows(gpu_ctx, log_rpi, index_hi):
    """src/poseidon.rs:817-862 (gather_rows): 16 columns of 2^12 rows, 42
    indexes, each index's 2^log_rpi rows of every column."""
    import torch
    from boojum_cuda_amd import poseidon
    src_rows, cols, count = 1 << 12, 16, 42
    rng = np.random.default_rng(log_rpi * 1000 + index_hi)
    idx = rng.integers(0, index_hi, count).astype(np.uint32)
    vals = rng.integers(0, 2**64, src_rows * cols, dtype=np.uint64)
    dst_rows = count << log_rpi
    res = torch.zeros(dst_rows * cols, dtype=torch.int64, device="cuda")
    poseidon.gather_rows(idx_dev(idx), log_rpi, to_dev(vals), src_rows, res, dst_rows)
    got = to_host(res)
    for i, index in enumerate(idx):
        for j in range(1 << log_rpi):
            for k in range(cols):
                assert got[k * dst_rows + (i << log_rpi) + j] == vals[(k << 12) + (int(index) << log_rpi) + j]


def test_gather_rows_skips_out_of_range(gpu_ctx):
    import torch
    from boojum_cuda_amd import poseidon
    vals = np.arange(64 * 2, dtype=np.uint64) + 100
    idx = np.array([3, 64, 10, 1 << 31], dtype=np.uint32)  # rows 64 and 2^31 do not exist
    res = torch.full((4 * 2,), 7, dtype=torch.int64, device="cuda")
    poseidon.gather_rows(idx_dev(idx), 0, to_dev(vals), 64, res, 4)
    got = to_host(res)
    assert got.tolist() == [103, 7, 110, 7, 167, 7, 174, 7]


@pytest.mark.parametrize("log_leaves,layers", [(12, 8), (5, 4), (2, 1)])
def test_gather_merkle_paths(gpu_ctx, log_leaves, layers):
    """src/poseidon.rs:864-916 (gather_merkle_paths): the sibling of each
    index's node on every layer, checked by the reference's recursion."""
    import torch
    from boojum_cuda_amd import poseidon
    count = 42
    rng = np.random.default_rng(log_leaves)
    idx = rng.integers(0, 1 << log_leaves, count).astype(np.uint32)
    vals = rng.integers(0, 2**64, 4 << (log_leaves + 1), dtype=np.uint64)
    res = torch.zeros(4 * count * layers, dtype=torch.int64, device="cuda")
    poseidon.gather_merkle_paths(idx_dev(idx), to_dev(vals), res, layers)
    got = to_host(res)
    values, results, ix = vals, got, idx.astype(np.int64)
    for _ in range(layers):
        half = values.size // 2
        layer, values = values[:half], values[half:]
        cur, results = results[:count * 4], results[count * 4:]
        n = layer.size // 4
        for c in range(4):
            assert np.array_equal(cur[c * count:(c + 1) * count], layer[c * n + (ix ^ 1)])
        ix = ix >> 1


def test_gather_bad_arguments_raise(gpu_ctx):
    import torch
    from boojum_cuda_amd import poseidon
    idx = idx_dev(np.zeros(4, dtype=np.uint32))
    v = torch.zeros(64, dtype=torch.int64, device="cuda")
    with pytest.raises(AssertionError):  # result rows != indexes << log_rpi
        poseidon.gather_rows(idx, 1, v, 16, torch.zeros(12, dtype=torch.int64, device="cuda"), 3)
    with pytest.raises(AssertionError):  # layers_count must be below log2(leaves)
        poseidon.gather_merkle_paths(idx, v, torch.zeros(4 * 4 * 3, dtype=torch.int64, device="cuda"), 3)
