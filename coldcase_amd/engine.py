"""ctypes binding of libcovhip.so and the `Cell` handle.

Follows the reference's ctypes convention (cov/gemm2/dctlib.py:19-39,
cov/gemm2/chain.py:11-57): plain pointers and sizes, int status codes, the
library owning device memory.  Unlike dct.cu's process-global state
(cov/gemm2/dct.cu:35-49) every cell is its own handle with its own HIP stream,
so several cells can be resident and running at once.

There is no CPU fallback: if the library is missing or cannot load, every
entry point raises.
"""
import ctypes
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'lib', 'libcovhip.so')

COV_OK = 0
COV_TR_W = 16
COV_DSEL_RANDOM = 1 << 60
MODE_PLAIN, MODE_GRAPH = 0, 1
GRAPH_STEPS = 40          # COV_GRAPH_STEPS: steps per captured graph
LEAVE_AUTO, LEAVE_WALK, LEAVE_KEPT = 0, 1, 2
CT_CODE, CT_UNCOVERED, CT_PRIVATE = 0, 1, 2
TR_FIELDS = ['it', 'u', 'ncand', 'ncw', 'bi', 'bp', 'bv', 'oldv', 'dsel', 'dunc',
             'unc', 'best', 'kick', 'draws', 'since', 'nties']

# the exported symbols, as declared in include/covhip.h
EXPORTS = ['cov_version', 'cov_strerror', 'cov_mem_bytes', 'cov_create', 'cov_destroy',
           'cov_last_error', 'cov_sync', 'cov_set_code', 'cov_set_code_flags', 'cov_get_code', 'cov_get_best',
           'cov_recount', 'cov_ball_update', 'cov_count_uncovered', 'cov_gather_uncovered',
           'cov_first_uncovered_from', 'cov_gather_uncovered_noalloc', 'cov_read_cnt', 'cov_eval_moves',
           'cov_eval_codewords', 'cov_commit',
           'cov_distance_transform', 'cov_distance_stats', 'cov_gather_far', 'cov_search_init', 'cov_search_params',
           'cov_search_wide', 'cov_search_run', 'cov_search_launch', 'cov_search_state', 'cov_query', 'cov_search_limits',
           'cov_search_ttt', 'cov_set_launch',
           'cov_get_launch', 'cov_set_leave_mode', 'cov_set_transform_tile', 'cov_timing', 'cov_timing_read', 'cov_search_phases',
           'cov_search_bytes', 'cov_count_transform', 'cov_set_count_tile', 'cov_map_max', 'cov_map_hist',
           'cov_map_extract', 'cov_map_read', 'cov_ball_gather', 'cov_load_words', 'cov_batch_create', 'cov_batch_destroy', 'cov_batch_last_error',
           'cov_batch_size', 'cov_batch_launch', 'cov_batch_sync', 'cov_batch_query', 'cov_batch_run']


class CovError(RuntimeError):
    pass


class State(ctypes.Structure):
    _fields_ = [('it', ctypes.c_int64), ('uncovered', ctypes.c_int64),
                ('best_uncovered', ctypes.c_int64), ('kicks', ctypes.c_int64),
                ('since_improve', ctypes.c_int64), ('draws', ctypes.c_int64),
                ('list_n', ctypes.c_int64), ('done', ctypes.c_int32),
                ('list_valid', ctypes.c_int32), ('fault', ctypes.c_int32),
                ('reserved', ctypes.c_int32)]


_lib = None


def lib_path():
    return LIB


def lib():
    """Load libcovhip.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise CovError('libcovhip.so is not built (run __graft_entry__.build())')
    L = ctypes.CDLL(LIB)
    vp, i32, i64, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
    sig = {
        'cov_version': (i32, []),
        'cov_strerror': (ctypes.c_char_p, [i32]),
        'cov_mem_bytes': (i64, [i32, i32, i32, i32]),
        'cov_create': (i32, [i32, i32, i32, i32, ctypes.POINTER(vp)]),
        'cov_destroy': (i32, [vp]),
        'cov_last_error': (ctypes.c_char_p, [vp]),
        'cov_sync': (i32, [vp]),
        'cov_set_code': (i32, [vp, vp]),
        'cov_set_code_flags': (i32, [vp, vp, i32]),
        'cov_get_code': (i32, [vp, vp]),
        'cov_get_best': (i32, [vp, vp, vp]),
        'cov_recount': (i32, [vp]),
        'cov_ball_update': (i32, [vp, vp, i32, i32]),
        'cov_count_uncovered': (i64, [vp]),
        'cov_gather_uncovered': (i64, [vp, vp, i64]),
        'cov_first_uncovered_from': (i64, [vp, i64]),
        'cov_gather_uncovered_noalloc': (i64, [vp, vp, i64]),
        'cov_read_cnt': (i32, [vp, i64, i64, vp]),
        'cov_eval_moves': (i32, [vp, vp, i32, vp, vp]),
        'cov_eval_codewords': (i32, [vp, i32, vp, vp]),
        'cov_commit': (i32, [vp, i32, i32, i32, vp]),
        'cov_distance_transform': (i32, [vp, vp, vp, vp]),
        'cov_gather_far': (i64, [vp, i32, vp, i64]),
        'cov_distance_stats': (i32, [vp, i32, vp, vp]),
        'cov_search_init': (i32, [vp, u64, vp, i32, i32]),
        'cov_search_params': (i32, [vp, i32, i64, i32, i32]),
        'cov_search_wide': (i32, [vp, i32]),
        'cov_search_run': (i64, [vp, i64, i32, vp]),
        'cov_search_launch': (i32, [vp, i64, i32]),
        'cov_search_state': (i32, [vp, ctypes.POINTER(State)]),
        'cov_query': (i32, [vp]),
        'cov_search_limits': (i32, [vp, i64, ctypes.c_double]),
        'cov_search_ttt': (i32, [vp, vp, vp]),
        'cov_set_launch': (i32, [vp, i32, i32]),
        'cov_set_leave_mode': (i32, [vp, i32]),
        'cov_get_launch': (i32, [vp, vp, vp]),
        'cov_set_transform_tile': (i32, [vp, i32, i32]),
        'cov_timing': (i32, [vp, i32]),
        'cov_timing_read': (i32, [vp, vp, vp, i32]),
        'cov_search_phases': (i32, [vp, vp, i32]),
        'cov_search_bytes': (i32, [vp, vp, vp, vp, vp]),
        'cov_count_transform': (i32, [vp, i32, vp, vp]),
        'cov_set_count_tile': (i32, [vp, i32]),
        'cov_map_max': (i32, [vp, vp]),
        'cov_map_hist': (i32, [vp, i32, i32, vp]),
        'cov_map_extract': (i64, [vp, i32, vp, vp, i64]),
        'cov_map_read': (i32, [vp, vp, i32, vp]),
        'cov_ball_gather': (i32, [vp, vp, i32, i32, vp]),
        'cov_load_words': (i32, [vp, vp, i32]),
        'cov_batch_create': (i32, [ctypes.POINTER(vp), i32, ctypes.POINTER(vp)]),
        'cov_batch_destroy': (i32, [vp]),
        'cov_batch_last_error': (ctypes.c_char_p, [vp]),
        'cov_batch_size': (i32, [vp]),
        'cov_batch_launch': (i32, [vp, i64, i32]),
        'cov_batch_sync': (i32, [vp]),
        'cov_batch_query': (i32, [vp]),
        'cov_batch_run': (i32, [vp, i64, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def mem_bytes(q, n, R, M):
    """Device bytes a cell of this shape needs (the serving loop's admission)."""
    r = lib().cov_mem_bytes(q, n, R, M)
    if r < 0:
        raise CovError('cov_mem_bytes: %s' % lib().cov_strerror(int(r)).decode())
    return r


class Cell:
    """One (q, n, R, M) cell resident on the GPU: coverage counts, the code,
    and the device-resident tabu search."""

    def __init__(self, q, n, R, M):
        self.q, self.n, self.R, self.M = q, n, R, M
        self.N = q ** n
        h = ctypes.c_void_p()
        rc = lib().cov_create(q, n, R, M, ctypes.byref(h))
        if rc != COV_OK:
            raise CovError('cov_create(q=%d n=%d R=%d M=%d): %s'
                           % (q, n, R, M, lib().cov_strerror(rc).decode()))
        self.h = h
        self.pw = np.array([q ** (n - 1 - p) for p in range(n)], dtype=np.int64)

    # ------------------------------------------------------------ plumbing
    def _check(self, rc, what):
        if rc < 0:
            msg = lib().cov_last_error(self.h)
            raise CovError('%s: %s (%s)' % (what, lib().cov_strerror(int(rc)).decode(),
                                             msg.decode() if msg else ''))
        return rc

    def close(self):
        if getattr(self, 'h', None):
            lib().cov_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        self._check(lib().cov_sync(self.h), 'cov_sync')

    # ------------------------------------------------------- code, coverage
    def word_index(self, words):
        return np.asarray(words, dtype=np.int64).reshape(-1, self.n) @ self.pw

    def set_code(self, code, recount=True):
        """Upload M x n digits; rebuild the coverage counts unless recount=False
        (enough for distance_transform / gather_far)."""
        c = np.ascontiguousarray(code, dtype=np.uint8).reshape(self.M, self.n)
        self._check(lib().cov_set_code_flags(self.h, _p(c), 0 if recount else 1), 'cov_set_code')

    def code(self):
        c = np.zeros((self.M, self.n), np.uint8)
        self._check(lib().cov_get_code(self.h, _p(c)), 'cov_get_code')
        return c

    def best(self):
        c = np.zeros((self.M, self.n), np.uint8)
        b = np.zeros(1, np.int64)
        self._check(lib().cov_get_best(self.h, _p(c), _p(b)), 'cov_get_best')
        return c, int(b[0])

    def recount(self):
        self._check(lib().cov_recount(self.h), 'cov_recount')

    def ball_update(self, centers, delta):
        c = np.ascontiguousarray(centers, dtype=np.int64).ravel()
        self._check(lib().cov_ball_update(self.h, _p(c), len(c), int(delta)), 'cov_ball_update')

    def uncovered(self):
        return self._check(lib().cov_count_uncovered(self.h), 'cov_count_uncovered')

    def gather_uncovered(self, cap):
        out = np.zeros(max(int(cap), 1), np.int64)
        tot = self._check(lib().cov_gather_uncovered(self.h, _p(out), int(cap)), 'cov_gather_uncovered')
        return out[:min(int(cap), tot)], tot

    def gather_uncovered_noalloc(self, cap):
        """gather_uncovered without device allocation (a static per-GPU
        scratch array; q^n <= 2^26, cap <= 65536; calls on one GPU must not
        overlap)."""
        out = np.zeros(max(int(cap), 1), np.int64)
        tot = self._check(lib().cov_gather_uncovered_noalloc(self.h, _p(out), int(cap)),
                          'cov_gather_uncovered_noalloc')
        return out[:min(int(cap), tot)], tot

    def first_uncovered_from(self, start):
        return lib().cov_first_uncovered_from(self.h, int(start))

    def cnt(self, off=0, length=None):
        length = self.N - off if length is None else length
        out = np.zeros(length, np.uint16)
        self._check(lib().cov_read_cnt(self.h, int(off), int(length), _p(out)), 'cov_read_cnt')
        return out

    # ------------------------------------------------------------- evaluator
    def eval_moves(self, u, codewords):
        u = np.ascontiguousarray(u, dtype=np.uint8)
        cw = np.ascontiguousarray(codewords, dtype=np.int32).ravel()
        d = np.zeros((len(cw), self.n), np.int64)
        self._check(lib().cov_eval_moves(self.h, _p(u), len(cw), _p(cw), _p(d)), 'cov_eval_moves')
        return d

    def eval_codewords(self, codewords):
        """(k, n, q) deltas of every move of the given codewords."""
        cw = np.ascontiguousarray(codewords, dtype=np.int32).ravel()
        d = np.zeros((len(cw), self.n, self.q), np.int64)
        self._check(lib().cov_eval_codewords(self.h, len(cw), _p(cw), _p(d)), 'cov_eval_codewords')
        return d

    def commit(self, i, p, v):
        d = np.zeros(1, np.int64)
        self._check(lib().cov_commit(self.h, int(i), int(p), int(v), _p(d)), 'cov_commit')
        return int(d[0])

    # -------------------------------------------------- distance transform
    def distance_transform(self, want_array=False):
        dist = np.zeros(self.N, np.uint8) if want_array else None
        hist = np.zeros(self.n + 1, np.int64)
        rad = ctypes.c_int(0)
        self._check(lib().cov_distance_transform(self.h, _p(dist), _p(hist), ctypes.byref(rad)),
                    'cov_distance_transform')
        return dist, hist, rad.value

    def distance_stats(self, r):
        """The transform without its histogram: (#{dist <= r}, radius)."""
        w = np.zeros(1, np.int64)
        rad = ctypes.c_int(0)
        self._check(lib().cov_distance_stats(self.h, int(r), _p(w), ctypes.byref(rad)), 'cov_distance_stats')
        return int(w[0]), rad.value

    def gather_far(self, r, cap):
        """After distance_transform(): (first `cap` words with dist > r, total)."""
        out = np.zeros(max(int(cap), 1), np.int64)
        tot = self._check(lib().cov_gather_far(self.h, int(r), _p(out), int(cap)), 'cov_gather_far')
        return out[:min(int(cap), tot)], tot

    # ---------------------------------------------------------------- search
    def search_init(self, seed, seed_code=None, greedy=False):
        if seed_code is None:
            sc, m = None, 0
        else:
            sc = np.ascontiguousarray(seed_code, dtype=np.uint8).reshape(-1, self.n)
            m = sc.shape[0]
        self._check(lib().cov_search_init(self.h, ctypes.c_uint64(int(seed)), _p(sc), m, int(bool(greedy))),
                    'cov_search_init')

    def search_params(self, tenure=0, stall=5000, kick=2, cand=100000):
        self._check(lib().cov_search_params(self.h, int(tenure), int(stall), int(kick), int(cand)),
                    'cov_search_params')

    def search_wide(self, on=True):
        """--wide: every move of a candidate codeword is a candidate while the
        uncovered list is valid and short (set before search_init)."""
        self._check(lib().cov_search_wide(self.h, int(bool(on))), 'cov_search_wide')

    def run(self, nsteps, mode=MODE_PLAIN, trace=False):
        tr = np.zeros((max(int(nsteps), 1), COV_TR_W), np.int64) if trace else None
        ran = self._check(lib().cov_search_run(self.h, int(nsteps), int(mode), _p(tr)), 'cov_search_run')
        return tr[:ran] if trace else ran

    def launch(self, nsteps, mode=MODE_GRAPH):
        self._check(lib().cov_search_launch(self.h, int(nsteps), int(mode)), 'cov_search_launch')

    def query(self):
        """True when every launch on the cell's stream has finished (no wait)."""
        return self._check(lib().cov_query(self.h), 'cov_query') == 1

    def limits(self, target=-1, seconds=0.0):
        """Stop before a step once best_uncovered <= target (target < 0: none)
        or once `seconds` from now have passed on the device clock (0: none)."""
        self._check(lib().cov_search_limits(self.h, int(target), float(seconds)), 'cov_search_limits')

    def ttt(self):
        """(iteration that first reached the target or -1, seconds since then or -1)."""
        it = np.zeros(1, np.int64)
        age = np.zeros(1, np.float64)
        self._check(lib().cov_search_ttt(self.h, _p(it), _p(age)), 'cov_search_ttt')
        return int(it[0]), float(age[0])

    def state(self):
        s = State()
        self._check(lib().cov_search_state(self.h, ctypes.byref(s)), 'cov_search_state')
        return {k: getattr(s, k) for k, _ in State._fields_ if k != 'reserved'}

    def check_fault(self, st=None):
        """Raise CovError if the search stopped on a device-detected error."""
        st = self.state() if st is None else st
        if st['fault']:
            raise CovError({1: 'inconsistent counters: no candidate move',
                            2: 'candidate sample larger than the rng ring allows'}.get(st['fault'], 'fault %d'
                                                                                         % st['fault']))
        return st

    def set_launch(self, eval_groups, commit_groups):
        self._check(lib().cov_set_launch(self.h, int(eval_groups), int(commit_groups)), 'cov_set_launch')

    def set_leave_mode(self, mode):
        """LEAVE_WALK, LEAVE_KEPT or LEAVE_AUTO: how steps find the leaving
        side of the deltas (the moves are the same)."""
        self._check(lib().cov_set_leave_mode(self.h, int(mode)), 'cov_set_leave_mode')

    def get_launch(self):
        e = np.zeros(1, np.int32)
        c = np.zeros(1, np.int32)
        self._check(lib().cov_get_launch(self.h, _p(e), _p(c)), 'cov_get_launch')
        return int(e[0]), int(c[0])

    def set_transform_tile(self, lds_bytes, threads=0):
        self._check(lib().cov_set_transform_tile(self.h, int(lds_bytes), int(threads)), 'cov_set_transform_tile')

    def timing(self, on):
        self._check(lib().cov_timing(self.h, int(bool(on))), 'cov_timing')

    def timing_read(self):
        """(ms, launches) per kind: 0 evaluator, 1 commit, 2 transform pass,
        3 recount, 4 selection, 5 post, 6 count-transform pass."""
        ms = np.zeros(7, np.float64)
        n = np.zeros(7, np.int64)
        self._check(lib().cov_timing_read(self.h, _p(ms), _p(n), 7), 'cov_timing_read')
        return ms, n

    # ------------------------------------------------ count transform, maps
    def count_transform(self, source, layers=False, want_map=False):
        """N_d for d = 0..R of the source set (CT_CODE, CT_UNCOVERED,
        CT_PRIVATE): the (R+1, q^n) layers if `layers`, else the map sum_d
        N_d is kept on the device (and returned if `want_map`)."""
        L = np.zeros((self.R + 1, self.N), np.int32) if layers else None
        m = np.zeros(self.N, np.int32) if (want_map and not layers) else None
        self._check(lib().cov_count_transform(self.h, int(source), _p(L), _p(m)), 'cov_count_transform')
        return L if layers else m

    def set_count_tile(self, lds_bytes):
        self._check(lib().cov_set_count_tile(self.h, int(lds_bytes)), 'cov_set_count_tile')

    def map_max(self):
        v = np.zeros(1, np.int32)
        self._check(lib().cov_map_max(self.h, _p(v)), 'cov_map_max')
        return int(v[0])

    def map_hist(self, nbins, vmax):
        h = np.zeros(int(nbins), np.int64)
        self._check(lib().cov_map_hist(self.h, int(nbins), int(vmax), _p(h)), 'cov_map_hist')
        return h

    def map_extract(self, thr, cap):
        """(words with map value >= thr in ascending order, their values, total)."""
        idx = np.zeros(max(int(cap), 1), np.int64)
        val = np.zeros(max(int(cap), 1), np.int32)
        tot = self._check(lib().cov_map_extract(self.h, int(thr), _p(idx), _p(val), int(cap)), 'cov_map_extract')
        m = min(int(cap), tot)
        return idx[:m], val[:m], tot

    def map_read(self, words):
        w = np.ascontiguousarray(words, dtype=np.int64).ravel()
        out = np.zeros(max(len(w), 1), np.int32)
        self._check(lib().cov_map_read(self.h, _p(w), len(w), _p(out)), 'cov_map_read')
        return out[:len(w)]

    def ball_gather(self, centers, target):
        c = np.ascontiguousarray(centers, dtype=np.int64).ravel()
        out = np.zeros(max(len(c), 1), np.int32)
        self._check(lib().cov_ball_gather(self.h, _p(c), len(c), int(target), _p(out)), 'cov_ball_gather')
        return out[:len(c)]

    def load_words(self, words):
        """Coverage counts of a word list of any length (the LNS engine's code)."""
        w = np.ascontiguousarray(words, dtype=np.int64).ravel()
        self._check(lib().cov_load_words(self.h, _p(w), len(w)), 'cov_load_words')

    def phases(self):
        """Microseconds in the narrow phases since search_init."""
        us = np.zeros(12, np.float64)
        self._check(lib().cov_search_phases(self.h, _p(us), 12), 'cov_search_phases')
        return dict(zip(['select', 'post', 'begin', 'pick', 'list', 'shuffle', 'rng', 'kick',
                         'eval_span', 'commit_span', 'gap_c2e', 'gap_e2c'], us.tolist()))

    def search_bytes(self):
        """(evaluator bytes, commit bytes, reference move bytes B_move, steps counted)."""
        v = np.zeros(4, np.int64)
        self._check(lib().cov_search_bytes(self.h, _p(v[0:]), _p(v[1:]), _p(v[2:]), _p(v[3:])),
                    'cov_search_bytes')
        return tuple(int(x) for x in v)


class Batch:
    """Cells of any shapes advanced by the same launches (one evaluator and
    one commit launch per batched step).  Each cell keeps the trajectory it
    has alone.  The cells stay owned by the caller and must outlive the batch;
    while a batched run is in flight they are not to be used."""

    def __init__(self, cells):
        self.cells = list(cells)
        arr = (ctypes.c_void_p * len(self.cells))(*[c.h.value for c in self.cells])
        h = ctypes.c_void_p()
        rc = lib().cov_batch_create(arr, len(self.cells), ctypes.byref(h))
        if rc != COV_OK:
            raise CovError('cov_batch_create(%d cells): %s' % (len(self.cells), lib().cov_strerror(rc).decode()))
        self.h = h

    def _check(self, rc, what):
        if rc < 0:
            msg = lib().cov_batch_last_error(self.h)
            raise CovError('%s: %s (%s)' % (what, lib().cov_strerror(int(rc)).decode(),
                                             msg.decode() if msg else ''))
        return rc

    def launch(self, nsteps, mode=MODE_GRAPH):
        self._check(lib().cov_batch_launch(self.h, int(nsteps), int(mode)), 'cov_batch_launch')

    def sync(self):
        self._check(lib().cov_batch_sync(self.h), 'cov_batch_sync')

    def query(self):
        """True when the launched steps have finished (does not wait)."""
        return self._check(lib().cov_batch_query(self.h), 'cov_batch_query') == 1

    def run(self, nsteps, mode=MODE_GRAPH):
        self._check(lib().cov_batch_run(self.h, int(nsteps), int(mode)), 'cov_batch_run')

    def close(self):
        if getattr(self, 'h', None):
            lib().cov_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Fleet:
    """Groups of cells, each group a Batch on its own stream, advanced
    concurrently in chunks of steps: a group of cheap cells runs at its own
    pace instead of the pace of the most expensive cell of one big batch,
    and its short dependent phases overlap the long gathers of the others.
    Every cell keeps the trajectory it has alone."""

    def __init__(self, groups):
        self.batches = []
        try:
            for g in groups:
                if g:
                    self.batches.append(Batch(g))
        except Exception:
            self.close()
            raise

    def run(self, nsteps, chunk=2048, mode=MODE_GRAPH, poll=0.0005):
        """Advance every group by nsteps; returns when all are done."""
        left = [int(nsteps)] * len(self.batches)
        busy = [False] * len(self.batches)
        while True:
            idle = True
            for i, b in enumerate(self.batches):
                if busy[i] and b.query():
                    busy[i] = False
                if not busy[i] and left[i] > 0:
                    k = min(int(chunk), left[i])
                    b.launch(k, mode=mode)
                    left[i] -= k
                    busy[i] = True
                    idle = False
            if not any(busy) and not any(left):
                for b in self.batches:
                    b.sync()               # raises on a cell's fault
                return
            if idle:
                time.sleep(poll)

    def close(self):
        for b in self.batches:
            b.close()
        self.batches = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
