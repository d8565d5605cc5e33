"""Autotuner for the launch parameters of the batched NTT (bc_tune_* C ABI).

For each (log_n, column-count bucket) it times candidate plans on the GPU:
the workgroup tile (for 2^16 columns: one-wave 256-point passes or the
4096-point pass kernels), the split of log_n into a contiguous pass and
strided passes, the strided
(column, block) pairs per workgroup and the contiguous grid cap, by
coordinate search from the built-in plan, then the ragged work-unit sizes.
A candidate is timed as one forward natural->bit-reversed call plus one
inverse bit-reversed->natural call.  Before an entry is kept its outputs are
compared with the built-in plan's, bit for bit.  The result is a text
database (bc_tune_save) that bc_context_create loads.

    python -m boojum_cuda_amd.autotune --log-n 12-24 --log-cols 0-9 \\
        --out era-boojum-cuda_amd/tuning/mi355x.txt

No reference counterpart: the reference's plans are the static PLANS table
(src/ntt.rs:74-169).
"""
import argparse
import ctypes
import os
import sys
import time

from ._lib import lib

MAX_TUNE_ELEMS = 1 << 28  # larger shapes reuse the nearest tuned bucket
MIN_GAIN = 0.02           # a candidate must beat the incumbent by 2% (timing noise)
SMALL_LOG_N, SMALL_MAX_COLS = 16, 4  # the built-in latency-path rule (csrc/ntt_internal.hpp)


class TuneParams(ctypes.Structure):
    _fields_ = [("contig_log", ctypes.c_uint32), ("num_strided", ctypes.c_uint32), ("strided_bpw", ctypes.c_uint32),
                ("contig_grid", ctypes.c_uint32), ("ragged_unit_contig", ctypes.c_uint32),
                ("ragged_unit_strided", ctypes.c_uint32), ("tile_log", ctypes.c_uint32)]

    def key(self):
        return tuple(getattr(self, f) for f, _ in self._fields_)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


def set_entry(log_n, log_cols, params):
    _check(lib().bc_tune_set(log_n, log_cols, ctypes.byref(params) if params is not None else None), "tune_set")


def get_entry(log_n, num_cols):
    p = TuneParams()
    _check(lib().bc_tune_get(log_n, num_cols, ctypes.byref(p)), "tune_get")
    return p


def load(path):
    _check(lib().bc_tune_load(os.fsencode(path)), "tune_load")


def save(path):
    _check(lib().bc_tune_save(os.fsencode(path)), "tune_save")


def clear():
    _check(lib().bc_tune_clear(), "tune_clear")


def splits(log_n):
    """Valid (contig_log, num_strided) pairs for log_n (contiguous 9..12,
    strided passes of 4..8 bits each)."""
    if log_n <= 12:
        return [(log_n, 0)]
    out = []
    for c in range(9, 13):
        s = log_n - c
        for k in range(1, 8):
            if 4 * k <= s <= 8 * k:
                out.append((c, k))
    return out


def builtin_split(log_n):
    """The split the library uses without a tuned entry (make_fast_plan in
    csrc/ntt_api.hip): a 12-bit contiguous pass and the fewest strided passes
    of at most 8 bits, the contiguous pass shrunk when the rest is under 4."""
    if log_n <= 12:
        return (log_n, 0)
    c, s = 12, log_n - 12
    k = (s + 7) // 8
    if s < 4 * k:
        c -= 4 * k - s
    return (c, k)


class _Bench:
    """Buffers and timing of one (log_n, num_cols) shape."""

    def __init__(self, log_n, num_cols, reps):
        import torch
        from . import ntt
        self.torch, self.ntt = torch, ntt
        self.log_n, self.cols, self.reps = log_n, num_cols, reps
        n = 1 << log_n
        g = torch.Generator(device="cuda")
        g.manual_seed(log_n * 1000 + num_cols)
        self.x = torch.randint(-2**63, 2**63 - 1, (num_cols * n,), dtype=torch.int64, device="cuda", generator=g)
        self.y = torch.empty_like(self.x)
        self.z = torch.empty_like(self.x)

    def run(self):
        n = 1 << self.log_n
        self.ntt.batch_ntt_out_of_place(self.x, self.y, self.log_n, self.cols, 0, 0, n, n, False, False, 0, 0)
        self.ntt.batch_ntt_out_of_place(self.y, self.z, self.log_n, self.cols, 0, 0, n, n, True, True, 0, 0)

    def time_ms(self):
        torch = self.torch
        self.run()
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        best = float("inf")
        for _ in range(self.reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            self.run()
            b.record(stream)
            b.synchronize()
            best = min(best, a.elapsed_time(b))
        return best

    def outputs(self):
        self.run()
        self.torch.cuda.synchronize()
        return self.y.clone(), self.z.clone()

    def warm(self, seconds=0.15):
        """Runs the shape until `seconds` have passed, so the first timed
        configuration does not pay for clock ramp-up."""
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            self.run()
            self.torch.cuda.synchronize()


def _time_ragged(log_n, num_cols, reps):
    """One ragged launch of an NTT job and an iNTT job of this shape."""
    import torch
    from . import ragged
    n = 1 << log_n
    x = torch.randint(-2**63, 2**63 - 1, (num_cols * n,), dtype=torch.int64, device="cuda")
    y, z = torch.empty_like(x), torch.empty_like(x)
    plan = ragged.RaggedPlan([ragged.Job(op=ragged.OP_NTT, log_n=log_n, inputs=x, outputs=y, num_cols=num_cols),
                              ragged.Job(op=ragged.OP_INTT, log_n=log_n, inputs=x, outputs=z, num_cols=num_cols)])
    try:
        plan.run()
        torch.cuda.synchronize()
        best = float("inf")
        stream = torch.cuda.current_stream()
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            plan.run()
            b.record(stream)
            b.synchronize()
            best = min(best, a.elapsed_time(b))
        return best
    finally:
        plan.destroy()


def tune_shape(log_n, log_cols, reps=5, log=print):
    """Tunes one bucket (timed at num_cols = 2^log_cols, capped at
    MAX_TUNE_ELEMS elements) and installs the best entry."""
    num_cols = 1 << log_cols
    if (num_cols << log_n) > MAX_TUNE_ELEMS:
        return None
    set_entry(log_n, log_cols, None)
    bench = _Bench(log_n, num_cols, reps)
    ref = bench.outputs()
    bench.warm()
    base_ms = bench.time_ms()
    best = TuneParams()
    best_ms = base_ms

    def trial(p):
        set_entry(log_n, log_cols, p)
        return bench.time_ms()

    def better(ms, incumbent):
        return ms < incumbent * (1.0 - MIN_GAIN)

    # 0. workgroup tile: for 2^16 columns, the one-wave 256-point passes
    # against the 4096-point pass kernels (whichever the built-in rule does
    # not pick for this column count); the rest of the search applies to the
    # 4096-point kernels only
    if log_n == SMALL_LOG_N:
        alt = 12 if num_cols <= SMALL_MAX_COLS else 8
        p = TuneParams()
        p.tile_log = alt
        ms = trial(p)
        if better(ms, best_ms):
            best, best_ms = p, ms
    small = best.tile_log == 8 or (log_n == SMALL_LOG_N and best.tile_log == 0 and num_cols <= SMALL_MAX_COLS)
    # 1. pass split (the built-in split itself is not a candidate: an entry
    # that repeats it could only record timing noise)
    for c, k in ([] if small else splits(log_n)):
        if (c, k) == builtin_split(log_n):
            continue
        p = TuneParams(c, k, 0, 0, 0, 0, 0)
        ms = trial(p)
        if better(ms, best_ms):
            best, best_ms = p, ms
    # 2. strided pairs per workgroup
    if log_n > 12 and not small:
        for bpw in (1, 2, 4, 8, 16, 32):
            p = TuneParams(*best.key())
            p.strided_bpw = bpw
            ms = trial(p)
            if better(ms, best_ms):
                best, best_ms = p, ms
    # 3. contiguous grid cap (the built-in default is two rounds of resident
    # workgroups; 768 and 2304 are one and three rounds on 256 CUs x 3)
    for grid in (() if small else (512, 768, 1024, 2304, 4096, 8192)):
        p = TuneParams(*best.key())
        p.contig_grid = grid
        ms = trial(p)
        if better(ms, best_ms):
            best, best_ms = p, ms
    # confirmation: the built-in plan and the winner timed alternately, so a
    # gain that came from timing noise (or from a plan equal to the built-in
    # one) is not recorded
    if best.key() != TuneParams().key():
        b_ms, w_ms = [], []
        for _ in range(3):
            set_entry(log_n, log_cols, None)
            b_ms.append(bench.time_ms())
            set_entry(log_n, log_cols, best)
            w_ms.append(bench.time_ms())
        if not better(min(w_ms), min(b_ms)):
            best, best_ms = TuneParams(), min(b_ms)
    # 4. ragged work units (timed on the persistent path)
    set_entry(log_n, log_cols, best)
    _time_ragged(log_n, num_cols, 1)  # warm the persistent path
    r_best = _time_ragged(log_n, num_cols, reps)
    for uc in (4, 8, 16, 32, 64):
        for us in (4, 8, 16, 32, 64):
            p = TuneParams(*best.key())
            p.ragged_unit_contig, p.ragged_unit_strided = uc, us
            set_entry(log_n, log_cols, p)
            ms = _time_ragged(log_n, num_cols, reps)
            if better(ms, r_best):
                r_best = ms
                best.ragged_unit_contig, best.ragged_unit_strided = uc, us
    if (best.ragged_unit_contig, best.ragged_unit_strided) != (0, 0):  # confirm as above
        plain = TuneParams(*best.key())
        plain.ragged_unit_contig = plain.ragged_unit_strided = 0
        d_ms, u_ms = [], []
        for _ in range(3):
            set_entry(log_n, log_cols, plain)
            d_ms.append(_time_ragged(log_n, num_cols, reps))
            set_entry(log_n, log_cols, best)
            u_ms.append(_time_ragged(log_n, num_cols, reps))
        if not better(min(u_ms), min(d_ms)):
            best = plain
    # keep the entry only if it differs from the built-in plan and gives the same bits
    if best.key() == TuneParams().key():
        set_entry(log_n, log_cols, None)
    else:
        set_entry(log_n, log_cols, best)
        got = bench.outputs()
        if not (bench.torch.equal(got[0], ref[0]) and bench.torch.equal(got[1], ref[1])):
            set_entry(log_n, log_cols, None)
            raise RuntimeError(f"tuned plan {best.key()} for log_n={log_n} cols={num_cols} changed the result")
    log(f"log_n={log_n:2d} cols={num_cols:4d}: built-in {base_ms:.3f} ms -> {best_ms:.3f} ms "
        f"{best.key()} ragged {r_best:.3f} ms")
    return best, base_ms, best_ms


def _range(text):
    lo, _, hi = text.partition("-")
    return range(int(lo), int(hi or lo) + 1)


def kept_entries(path, tuned):
    """Entries of the database at `path` outside the (log_n, log_cols)
    buckets in `tuned`, as (log_n, log_cols, TuneParams): what --merge-from
    carries over (comment lines skipped)."""
    out = []
    for line in open(path):
        f = line.split()
        if not f or f[0].startswith("#") or (int(f[0]), int(f[1])) in tuned:
            continue
        out.append((int(f[0]), int(f[1]), TuneParams(*[int(v) for v in f[2:]])))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--log-n", default="12-24")
    ap.add_argument("--log-cols", default="0-9")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", required=True)
    ap.add_argument("--merge-from", default=None,
                    help="database whose entries for the (log_n, log_cols) buckets not tuned here are kept in --out")
    args = ap.parse_args(argv)
    import torch
    from .context import Context
    torch.cuda.set_device(0)
    ctx = Context.create(12, 12)
    clear()  # tune from the built-in plans, not from a loaded database
    t0 = time.time()
    for log_n in _range(args.log_n):
        for log_cols in _range(args.log_cols):
            tune_shape(log_n, log_cols, args.reps, log=lambda m: print(m, flush=True))
        torch.cuda.empty_cache()
    if args.merge_from:
        tuned = {(n, c) for n in _range(args.log_n) for c in _range(args.log_cols)}
        for log_n, log_cols, params in kept_entries(args.merge_from, tuned):
            set_entry(log_n, log_cols, params)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    save(args.out)
    print(f"saved {args.out} in {time.time() - t0:.0f} s", flush=True)
    ctx.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
