"""CPU tests of the tuning database (bc_tune_*): host-only entry points, so
the library is loaded and called without a GPU."""
import os

import pytest


@pytest.fixture
def tune():
    from boojum_cuda_amd import autotune
    autotune.clear()
    yield autotune
    autotune.clear()


def test_splits_cover_valid_plans():
    from boojum_cuda_amd import autotune
    assert autotune.splits(12) == [(12, 0)]
    assert autotune.splits(13) == [(9, 1)]
    assert (12, 1) in autotune.splits(20) and (10, 2) in autotune.splits(20) and (12, 2) in autotune.splits(20)
    for log_n in range(13, 33):
        for c, k in autotune.splits(log_n):
            assert 9 <= c <= 12 and 4 * k <= log_n - c <= 8 * k


def test_set_get_nearest_bucket(tune):
    P = tune.TuneParams
    assert tune.get_entry(20, 256).key() == (0,) * 7
    tune.set_entry(20, 8, P(10, 2, 4, 1024, 2, 4))
    assert tune.get_entry(20, 256).key() == (10, 2, 4, 1024, 2, 4, 0)
    assert tune.get_entry(20, 300).key() == (10, 2, 4, 1024, 2, 4, 0)   # bucket 8
    assert tune.get_entry(20, 1).key() == (10, 2, 4, 1024, 2, 4, 0)     # nearest bucket of log_n 20
    assert tune.get_entry(21, 256).key() == (0,) * 7                  # other log_n: built-in
    tune.set_entry(20, 0, P(12, 1, 0, 0, 0, 0))
    assert tune.get_entry(20, 1).key()[:2] == (12, 1)
    assert tune.get_entry(20, 16).key()[:2] == (12, 1)                # bucket 4: 0 and 8 tie, smaller wins
    tune.set_entry(20, 0, None)
    assert tune.get_entry(20, 1).key()[:2] == (10, 2)


@pytest.mark.parametrize("log_n,params", [
    (20, (8, 2, 0, 0, 0, 0)),     # contiguous pass below 9
    (20, (12, 3, 0, 0, 0, 0)),    # 8 bits cannot make 3 strided passes of >= 4
    (20, (12, 0, 0, 0, 0, 0)),    # strided bits left over
    (13, (12, 1, 0, 0, 0, 0)),    # 1 strided bit
    (10, (9, 0, 0, 0, 0, 0)),     # log_n <= 12 is one contiguous pass
    (20, (0, 1, 0, 0, 0, 0)),     # strided count without a split
    (20, (0, 0, 65, 0, 0, 0)),    # pairs per workgroup too large
    (20, (0, 0, 0, 0, 65, 0)),    # ragged unit too large
    (20, (0, 0, 0, 0, 0, 0, 8)),  # one-wave 256-point passes only for 2^16
    (16, (0, 0, 0, 0, 0, 0, 10)), # tile must be 8 or 12
])
def test_invalid_entries_rejected(tune, log_n, params):
    with pytest.raises(RuntimeError):
        tune.set_entry(log_n, 4, tune.TuneParams(*params))


def test_save_load_round_trip(tune, tmp_path):
    P = tune.TuneParams
    entries = {(20, 8): (10, 2, 4, 1024, 2, 4, 0), (16, 0): (10, 1, 0, 0, 1, 1, 0), (12, 5): (12, 0, 0, 4096, 4, 0, 0),
               (16, 5): (0, 0, 0, 0, 0, 0, 8), (16, 1): (0, 0, 0, 0, 0, 0, 12)}
    for (log_n, lc), v in entries.items():
        tune.set_entry(log_n, lc, P(*v))
    path = str(tmp_path / "db.txt")
    tune.save(path)
    tune.clear()
    assert tune.get_entry(20, 256).key() == (0,) * 7
    tune.load(path)
    for (log_n, lc), v in entries.items():
        assert tune.get_entry(log_n, 1 << lc).key() == v
    text = open(path).read()
    assert text.startswith("#") and len(text.strip().splitlines()) == 1 + len(entries)


def test_malformed_database_rejected(tune, tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("20 8 10 2 4 1024\n")          # too few fields
    with pytest.raises(RuntimeError):
        tune.load(str(bad))
    bad.write_text("20 8 8 2 0 0 0 0\n")          # invalid split
    with pytest.raises(RuntimeError):
        tune.load(str(bad))
    assert tune.get_entry(20, 256).key() == (0,) * 7
    with pytest.raises(RuntimeError):
        tune.load(str(tmp_path / "missing.txt"))


def test_shipped_database_is_valid(tune):
    """The committed database parses and holds only valid entries."""
    from boojum_cuda_amd._lib import LIB_PATH
    path = os.path.join(os.path.dirname(LIB_PATH), "tuning", "mi355x.txt")
    if not os.path.exists(path):
        pytest.skip("no tuning database shipped")
    tune.load(path)


def test_builtin_split_is_a_valid_split_and_matches_the_planner_rule():
    """The tuner skips the library's own split (make_fast_plan): 12 contiguous
    bits and the fewest strided passes of 4..8 bits, the contiguous pass
    shrunk when fewer than 4 bits remain."""
    from boojum_cuda_amd import autotune
    want = {12: (12, 0), 13: (9, 1), 15: (11, 1), 16: (12, 1), 20: (12, 1), 21: (12, 2), 24: (12, 2),
            28: (12, 2), 29: (12, 3)}
    for log_n, split in want.items():
        assert autotune.builtin_split(log_n) == split, log_n
    for log_n in range(13, 33):
        assert autotune.builtin_split(log_n) in autotune.splits(log_n), log_n


def test_merge_keeps_every_bucket_not_tuned(tmp_path):
    """--merge-from carries over every entry outside the tuned
    (log_n, log_cols) buckets, including other column buckets of a tuned
    log_n (host-only: reads the file, no device)."""
    from boojum_cuda_amd import autotune
    db = tmp_path / "db.txt"
    db.write_text("# header\n"
                  "16 0 0 0 0 0 4 4 8\n"
                  "16 3 0 0 2 0 4 4 0\n"
                  "20 0 0 0 1 0 0 0 0\n"
                  "20 8 9 2 0 0 0 0 0\n")
    kept = autotune.kept_entries(str(db), {(20, 0), (16, 3)})
    assert [(n, c, p.key()) for n, c, p in kept] == [(16, 0, (0, 0, 0, 0, 4, 4, 8)), (20, 8, (9, 2, 0, 0, 0, 0, 0))]
