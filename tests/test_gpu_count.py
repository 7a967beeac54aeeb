"""The distance-count transform (N_d layers, gain and loss maps) and the map
calls on the GPU, through the C ABI, against brute force on the grids of
cov/gemm2/test_dct.py:44-45 (with repeated codewords), under small LDS
budgets that force many passes and partial tiles, and at K7(9,4) size
through invariants: sum_d N_d = cnt, min{d : N_d > 0} = the distance
transform, the maps equal the exact ball gathers."""
import numpy as np
import pytest

from oracle.verify_oracle import count_layers, mindist

pytestmark = pytest.mark.gpu

GRIDS = [(3, 4, 2), (3, 5, 3), (4, 4, 3), (2, 6, 2), (5, 4, 2), (6, 3, 2), (4, 6, 4), (3, 6, 5)]


def _cell(q, n, R, code):
    from coldcase_amd import Cell
    c = Cell(q, n, R, len(code))
    c.set_code(code)
    return c


@pytest.mark.parametrize('lds', [0, 1024, 4096])
@pytest.mark.parametrize('q,n,R', GRIDS)
def test_layers_and_maps_match_brute_force(q, n, R, lds):
    from coldcase_amd.engine import CT_CODE, CT_PRIVATE, CT_UNCOVERED
    print('case count K%d(%d,%d) lds=%d' % (q, n, R, lds), flush=True)
    rng = np.random.default_rng(q * 100 + n * 10 + R)
    N = q ** n
    M = int(rng.integers(2, max(3, N // 7)))
    code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
    code[-2:] = code[:2]                                   # repeated codewords
    idx = code.astype(np.int64) @ np.array([q ** (n - 1 - p) for p in range(n)], dtype=np.int64)
    c = _cell(q, n, R, code)
    if lds:
        c.set_count_tile(lds)
    want = count_layers(idx, n, q, R)
    assert np.array_equal(c.count_transform(CT_CODE, layers=True), want)
    cnt = c.cnt().astype(np.int64)
    assert np.array_equal(want.sum(axis=0), cnt)
    unc = np.flatnonzero(cnt == 0)
    gain = c.count_transform(CT_UNCOVERED, want_map=True)
    assert np.array_equal(gain, count_layers(unc, n, q, R).sum(axis=0))
    assert np.array_equal(c.count_transform(CT_UNCOVERED, layers=True), count_layers(unc, n, q, R))
    loss = c.count_transform(CT_PRIVATE, want_map=True)
    assert np.array_equal(loss, count_layers(np.flatnonzero(cnt == 1), n, q, R).sum(axis=0))
    # the map calls on the loss map now on the device
    mx = c.map_max()
    assert mx == loss.max()
    if mx > 0:
        words, vals, tot = c.map_extract(max(1, mx // 2), 1 << 20)
        sel = np.flatnonzero(loss >= max(1, mx // 2))
        assert tot == sel.size and np.array_equal(words, sel) and np.array_equal(vals, loss[sel])
        h = c.map_hist(16, mx)
        pos = loss[loss > 0].astype(np.int64)
        assert np.array_equal(h, np.bincount(np.minimum((pos - 1) * 16 // mx, 15), minlength=16))
    probe = rng.choice(N, size=min(40, N), replace=False)
    assert np.array_equal(c.map_read(probe), loss[probe])
    assert np.array_equal(c.ball_gather(probe, 1), loss[probe])
    assert np.array_equal(c.ball_gather(probe, 0), gain[probe])
    c.close()


def test_k7_9_4_invariants():
    """Full size: the code's layers sum to the recount, their first non-zero
    layer is the distance transform (and all are zero exactly where it
    exceeds R), and the gain and loss maps equal exact ball gathers."""
    from coldcase_amd.engine import CT_CODE, CT_PRIVATE, CT_UNCOVERED
    print('case count K7(9,4) M=1480', flush=True)
    q, n, R, M = 7, 9, 4, 1480
    code = np.random.default_rng(3).integers(0, q, size=(M, n)).astype(np.uint8)
    c = _cell(q, n, R, code)
    L = c.count_transform(CT_CODE, layers=True)
    cnt = c.cnt()
    assert np.array_equal(L.sum(axis=0), cnt)
    dist, _, _ = c.distance_transform(want_array=True)
    nz = L > 0
    first = np.where(nz.any(axis=0), nz.argmax(axis=0), 127)
    assert np.array_equal(np.minimum(first, 127), np.where(dist <= R, dist, 127))
    del L, nz, first
    rng = np.random.default_rng(5)
    probe = rng.choice(q ** n, size=300, replace=False)
    gain = c.count_transform(CT_UNCOVERED, want_map=True)
    assert np.array_equal(gain[probe], c.ball_gather(probe, 0))
    loss = c.count_transform(CT_PRIVATE, want_map=True)
    pw = np.array([q ** (n - 1 - p) for p in range(n)], dtype=np.int64)
    cw = code.astype(np.int64) @ pw
    assert np.array_equal(loss[cw], c.ball_gather(cw, 1))
    # every uncovered word contributes its ball to the gain map
    from coldcase_amd.verify_cov import ball_volume
    assert int(gain.sum()) == int((cnt == 0).sum()) * ball_volume(q, n, R)
    assert int(loss.sum()) == int((cnt == 1).sum()) * ball_volume(q, n, R)
    c.close()


def test_load_words_equals_recount():
    from coldcase_amd import Cell
    print('case load_words K6(9,4)', flush=True)
    q, n, R, M = 6, 9, 4, 660
    code = np.random.default_rng(8).integers(0, q, size=(M, n)).astype(np.uint8)
    pw = np.array([q ** (n - 1 - p) for p in range(n)], dtype=np.int64)
    a = _cell(q, n, R, code)
    with Cell(q, n, R, M) as b:
        b.load_words(code[:500].astype(np.int64) @ pw)
        b.load_words(code.astype(np.int64) @ pw)        # a second load starts from zero
        assert np.array_equal(a.cnt(), b.cnt()) and a.uncovered() == b.uncovered()
    a.close()
