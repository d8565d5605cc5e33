"""Seeded workload shapes of the measured configurations, shared by the
benchmark and the GPU parity tests so that both run exactly the same jobs.

- coset LDE: 128 columns of 2^20 bit-reversed evaluations, blowup 4, seed 3;
- ragged batch: 1000 single-column jobs, log_n uniform in [10, 23], op
  uniform over {NTT, iNTT, LDE blowup 2, LDE blowup 4}, seed 4;
- serving: Poisson arrivals, 1-512 columns, log_n uniform in [12, 24],
  blowup uniform over {1, 2, 4} (1: forward NTT), seed 5.
"""
import random

from .ragged import OP_INTT, OP_LDE, OP_NTT

LDE_SHAPE = {"log_n": 20, "cols": 128, "log_blowup": 2, "seed": 3}
RAGGED_SHAPE = {"count": 1000, "log_lo": 10, "log_hi": 23, "seed": 4}
SERVE_SHAPE = {"log_lo": 12, "log_hi": 24, "cols_hi": 512, "seed": 5}


def uniform_field(count, gen):
    """Uniform draws in [0, p) on the device as int64 bit patterns: full-range
    u64 draws, the 2^32 - 1 values >= p folded down by p (probability 2^-32
    each)."""
    import torch
    x = torch.randint(-2**63, 2**63 - 1, (count,), dtype=torch.int64, device="cuda", generator=gen)
    big = (x < 0) & (x >= -(2**32 - 1))  # u64 value >= p
    return torch.where(big, x + (2**32 - 1), x)


def ragged_specs(count=RAGGED_SHAPE["count"], log_lo=RAGGED_SHAPE["log_lo"], log_hi=RAGGED_SHAPE["log_hi"],
                 seed=RAGGED_SHAPE["seed"]):
    """(log_n, op, log_blowup) of single-column ragged jobs."""
    rng = random.Random(seed)
    kinds = [(OP_NTT, 0), (OP_INTT, 0), (OP_LDE, 1), (OP_LDE, 2)]
    return [(rng.randint(log_lo, log_hi),) + kinds[rng.randrange(4)] for _ in range(count)]


def ragged_jobs(specs, seed=RAGGED_SHAPE["seed"]):
    """Device buffers for ragged specs: uniform inputs drawn on the device in
    spec order from one generator seeded with `seed`, uninitialised outputs
    and coefficient buffers."""
    import torch
    from .ragged import Job
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    jobs = []
    for log_n, op, lb in specs:
        n = 1 << log_n
        x = uniform_field(n, g)
        out = torch.empty(n << lb, dtype=torch.int64, device="cuda")
        j = Job(op=op, log_n=log_n, inputs=x, outputs=out, log_blowup=lb)
        if op == OP_LDE:
            j.coeffs = torch.empty(n, dtype=torch.int64, device="cuda")
        jobs.append(j)
    return jobs


def serve_specs(count, rate, seed=SERVE_SHAPE["seed"], log_lo=SERVE_SHAPE["log_lo"], log_hi=SERVE_SHAPE["log_hi"],
                cols_hi=SERVE_SHAPE["cols_hi"]):
    """(arrival_s, log_n, cols, log_blowup) of Poisson arrivals at `rate` per
    second; log_blowup 0 is a forward NTT, 1 and 2 are coset LDEs."""
    rng = random.Random(seed)
    specs, t = [], 0.0
    for _ in range(count):
        t += rng.expovariate(rate)
        lb = rng.choice([0, 1, 2])
        specs.append((t, rng.randint(log_lo, log_hi), rng.randint(1, cols_hi), lb))
    return specs
