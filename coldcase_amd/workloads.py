"""workloads -- the benchmark configurations' instances, shared by bench.py
and the tests so that both run exactly the same cells.

  C2  single cell K6(9,4): q=6 n=9 R=4, M=660, random init, tenure 10
  C3  single cell K7(9,4): M=1480, a uniform random code (numpy seed 3)
  C4  64 independent instances, (q, n, R) drawn uniformly from
      {5,6,7,8} x {7,8,9} x {2,3,4} with numpy.random.default_rng(2026),
      M = ceil(1.3 q^n / |ball|); a draw with M > 65535 (uint16 counters,
      cov/engine/covsearch2e.c:1117) is redrawn and reported; instance i
      searches from seed 1000003 + 7919 i (cov/engine/NOTES.md:196-203)
  C5  Poisson arrivals (seed 7, 20 req/s) of verify requests (30%) and
      search bursts over six cells up to K6(10,4) M=2784; record codes as
      verify inputs where the caller provides them
"""
import math

C2 = (6, 9, 4, 660)
C3 = (7, 9, 4, 1480)
C5_SHAPES = [(5, 7, 2, 400), (6, 6, 3, 41), (6, 8, 4, 169), (6, 9, 4, 660), (7, 9, 4, 1480), (6, 10, 4, 2784)]
TENURE = 10


def ball(q, n, R):
    return sum(math.comb(n, i) * (q - 1) ** i for i in range(min(R, n) + 1))


def c4_specs(seed=2026, count=64):
    """The batched configuration's (q, n, R, M) instances, and the draws that
    were redrawn because their M exceeds the uint16 counters."""
    import numpy as np
    rng = np.random.default_rng(seed)
    specs, redrawn = [], []
    while len(specs) < count:
        q = int(rng.choice([5, 6, 7, 8]))
        n = int(rng.choice([7, 8, 9]))
        R = int(rng.choice([2, 3, 4]))
        M = math.ceil(1.3 * q ** n / ball(q, n, R))
        if M > 65535:
            redrawn.append((q, n, R, M))
            continue
        specs.append((q, n, R, M))
    return specs, redrawn


def c4_seed(i, s=1):
    """Search seed of instance i (the engine's portfolio convention)."""
    return s * 1000003 + i * 7919


def c5_trace(n_req=60, moves=100000, codes=None, seed=7, rate=20.0, verify_frac=0.3):
    """The serving configuration's request trace (coldcase_amd.serve.Request list)."""
    from .serve import poisson_workload
    return poisson_workload(seed, n_req, rate, C5_SHAPES, verify_frac=verify_frac, moves=moves, codes=codes)
