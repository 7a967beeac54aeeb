"""The serving loop on the GPU: every served result equals the same request
run alone (tolerance 0), verify answers match the oracle, admission keeps
resident memory within the budget, and oversized requests are refused."""
import os

import numpy as np
import pytest

from oracle.verify_oracle import mindist
from tests._util import DATA, cli_code_sha, golden_runs, read_code

pytestmark = pytest.mark.gpu

SHAPES = [(5, 7, 2, 400), (6, 6, 3, 41), (4, 8, 3, 60), (3, 8, 2, 120), (6, 9, 4, 660)]


def _trace(n_req=14, rate=2000.0, moves=2000, seed=5):
    # arrivals 0.5 ms apart on average and searches of tens of ms: they overlap
    from coldcase_amd.serve import poisson_workload
    return poisson_workload(seed, n_req, rate, SHAPES, verify_frac=0.35, moves=moves)


def _alone(r):
    from coldcase_amd.serve import Server
    return Server(tick_steps=10 ** 9).run([r], realtime=False)[0]


@pytest.mark.parametrize('tick,cap', [(64, None), (1000, None), (64, 2)])
def test_served_equals_alone(tick, cap):
    from coldcase_amd.serve import Server, summarize
    print('case serve mix tick=%d lane_cap=%s' % (tick, cap), flush=True)
    reqs = _trace()
    assert {r.kind for r in reqs} == {'verify', 'search'}
    srv = Server(tick_steps=tick, lane_cap=cap)
    res = srv.run(reqs, realtime=True)
    s = summarize(res)
    assert s['errors'] == 0 and s['completed'] == len(reqs)
    assert srv.peak_cells >= 2                      # searches really overlapped
    for r, got in zip(reqs, res):
        assert got.rid == r.rid
        assert got.answer() == _alone(r).answer(), r.rid
        assert got.finished >= got.admitted >= got.arrival - 1e-6


def test_verify_answers_match_oracle():
    from coldcase_amd.serve import Request, Server
    rng = np.random.default_rng(9)
    reqs = []
    for rid, (q, n, R, M) in enumerate([(5, 7, 2, 400), (4, 8, 3, 60), (3, 8, 2, 30)]):
        code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
        reqs.append(Request(rid, 'verify', q, n, R, M, code=code))
    res = Server().run(reqs, realtime=False)
    for r, got in zip(reqs, res):
        d = mindist(r.code, r.n, r.q)
        far = np.flatnonzero(d > r.R)
        assert got.uncovered == far.size
        assert got.radius == int(d.max())
        assert got.hist == np.bincount(d, minlength=r.n + 1)[:r.n + 1].tolist()
        want = [tuple(int(x) for x in np.base_repr(int(w), r.q).zfill(r.n)) for w in far[:5]]
        assert [tuple(w) for w in got.witnesses] == want


def test_budget_admission_and_refusal():
    from coldcase_amd.engine import mem_bytes
    from coldcase_amd.serve import Request, Server
    print('case serve budget', flush=True)
    small = [Request(i, 'search', 5, 7, 2, 400, seed=100 + i, moves=200) for i in range(5)]
    one = mem_bytes(5, 7, 2, 400)
    big = Request(9, 'search', 6, 10, 4, 2784, moves=10)
    srv = Server(hbm_budget=2 * one + one // 2)      # room for two at a time
    res = srv.run(small + [big], realtime=False)
    assert srv.peak_bytes <= 2 * one + one // 2 and srv.peak_cells == 2
    for r, got in zip(small, res[:5]):
        assert not got.error and got.answer() == _alone(r).answer()
    assert res[5].error


def test_serve_k6_10_4_search_and_record_verify():
    """The serving configuration's largest cell, K6(10,4) M=2784 (121 MB of
    counts): a search from the fixture's seed and a verify of the record
    code, in a mix with smaller requests.  The search equals its run alone
    and the reference's fixture (best uncovered, iterations, kicks, the
    published code); the verify finds the record code covering."""
    from coldcase_amd.serve import Request, Server
    print('case serve K6(10,4) search + record verify', flush=True)
    gold = [r for r in golden_runs() if r['tag'] == 'C5' and r['iters'] == 30][0]
    rec = read_code(os.path.join(DATA, 'K6_10_4_M2784.txt'), 10)
    reqs = [Request(0, 'search', 6, 10, 4, 2784, arrival=0.0, seed=gold['seed'], moves=30),
            Request(1, 'verify', 6, 10, 4, 2784, arrival=0.001, code=rec),
            Request(2, 'search', 6, 9, 4, 660, arrival=0.002, seed=11, moves=500),
            Request(3, 'search', 5, 7, 2, 400, arrival=0.003, seed=12, moves=2000),
            Request(4, 'verify', 6, 9, 4, 660, arrival=0.004,
                    code=read_code(os.path.join(DATA, 'K6_9_4_M660.txt'), 9)),
            Request(5, 'search', 6, 10, 4, 2784, arrival=0.005, seed=gold['seed'] + 1, moves=20)]
    srv = Server(tick_steps=8)
    res = srv.run(reqs, realtime=True)
    assert all(not r.error for r in res)
    assert srv.peak_cells >= 2
    for r, got in zip(reqs, res):
        assert got.answer() == _alone(r).answer(), r.rid
    s0 = res[0]
    assert (s0.best_uncovered, s0.iters, s0.kicks) == (gold['uncovered'], gold['iters_done'], gold['kicks'])
    assert cli_code_sha(s0.best_code, 6, 2784) == gold['out_sha256']
    assert res[1].uncovered == 0 and res[1].radius == 4 and res[1].witnesses == []
    assert res[4].uncovered == 0 and res[4].radius == 4
