"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, as the checker; the product
(coldcase_amd) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')
CLI_PATH = os.path.join(HERE, 'cso')
REF_PATH = os.path.join(HERE, '_ref', 'covsearch2e')

TR_FIELDS = ['it', 'u', 'ncand', 'ncw', 'bi', 'bp', 'bv', 'oldv', 'dsel', 'dunc',
             'unc', 'best', 'kick', 'draws', 'since', 'nties']
TR_W = len(TR_FIELDS)
DSEL_RANDOM = 1 << 60

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('oracle not built: make -C oracle')
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64
        L.cso_create.restype = vp
        L.cso_create.argtypes = [i32, i32, i32, i32, u64]
        L.cso_destroy.argtypes = [vp]
        L.cso_init.argtypes = [vp, vp, i32, i32]
        L.cso_run.restype = i64
        L.cso_run.argtypes = [vp, i64, i32, i64, i32, i32, vp]
        L.cso_set_kick_order.argtypes = [vp, i32]
        L.cso_set_threads.argtypes = [vp, i32]
        L.cso_set_target.argtypes = [vp, i64]
        L.cso_set_wide.restype = i32
        L.cso_set_wide.argtypes = [vp, i32]
        L.cso_eval_codeword_wide.argtypes = [vp, i32, vp]
        L.cso_ttt_it.restype = i64
        L.cso_ttt_it.argtypes = [vp]
        for f in ('cso_uncovered', 'cso_best_uncovered', 'cso_iters', 'cso_kicks', 'cso_space'):
            getattr(L, f).restype = i64
            getattr(L, f).argtypes = [vp]
        for f in ('cso_code', 'cso_best', 'cso_cnt'):
            getattr(L, f).restype = vp
            getattr(L, f).argtypes = [vp]
        L.cso_rng_state.argtypes = [vp, vp]
        L.cso_set_rng_state.argtypes = [vp, vp]
        L.cso_set_code.argtypes = [vp, vp]
        L.cso_draws.restype = i64
        L.cso_draws.argtypes = [vp]
        L.cso_move_delta.restype = i64
        L.cso_move_delta.argtypes = [vp, i32, i32, i32]
        L.cso_eval_codeword.argtypes = [vp, i32, vp, vp, i32]
        L.cso_commit.restype = i64
        L.cso_commit.argtypes = [vp, i32, i32, i32]
        L.cso_ball.argtypes = [vp, i32, i32]
        L.cso_pick.restype = i64
        L.cso_pick.argtypes = [vp]
        L.cso_unique_cover.restype = i64
        L.cso_unique_cover.argtypes = [vp, i32]
        L.cso_distinct_fill.argtypes = [vp, vp, i32, vp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One search state on the CPU: q, n, R, M and the rng seed."""

    def __init__(self, q, n, R, M, seed, threads=1):
        self.q, self.n, self.R, self.M = q, n, R, M
        self.h = lib().cso_create(q, n, R, M, seed)
        if not self.h:
            raise ValueError('bad parameters')
        lib().cso_set_threads(self.h, threads)
        self.N = lib().cso_space(self.h)

    def close(self):
        if self.h:
            lib().cso_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self, seed_code=None, greedy=False):
        if seed_code is None:
            sc = np.zeros((0, self.n), np.uint8)
        else:
            sc = np.ascontiguousarray(seed_code, dtype=np.uint8).reshape(-1, self.n)
        rc = lib().cso_init(self.h, _ptr(sc), sc.shape[0], int(bool(greedy)))
        if rc != 0:
            raise ValueError('bad seed code')
        self._keep = sc
        return self

    def run(self, iters, tabu=10, stall=5000, kick=2, cand=100000, trace=False):
        tr = np.zeros((max(iters, 1), TR_W), np.int64) if trace else None
        done = lib().cso_run(self.h, iters, tabu, stall, kick, cand,
                             _ptr(tr) if trace else None)
        return tr[:done] if trace else done

    def set_wide(self, on=True):
        if lib().cso_set_wide(self.h, int(bool(on))) != 0:
            raise MemoryError('cso_set_wide')
        return self

    def eval_codeword_wide(self, i):
        d = np.zeros(self.n * self.q, np.int64)
        lib().cso_eval_codeword_wide(self.h, int(i), _ptr(d))
        return d.reshape(self.n, self.q)

    def set_target(self, target):
        lib().cso_set_target(self.h, int(target))

    @property
    def ttt_it(self):
        return lib().cso_ttt_it(self.h)

    def set_code(self, code):
        c = np.ascontiguousarray(code, dtype=np.uint8).reshape(self.M, self.n)
        if lib().cso_set_code(self.h, _ptr(c)) != 0:
            raise ValueError('digit out of range')

    @property
    def uncovered(self):
        return lib().cso_uncovered(self.h)

    @property
    def best_uncovered(self):
        return lib().cso_best_uncovered(self.h)

    @property
    def iters(self):
        return lib().cso_iters(self.h)

    @property
    def draws(self):
        return lib().cso_draws(self.h)

    @property
    def kicks(self):
        return lib().cso_kicks(self.h)

    def code(self):
        p = lib().cso_code(self.h)
        return np.ctypeslib.as_array((ctypes.c_uint8 * (self.M * self.n)).from_address(p)).reshape(self.M, self.n).copy()

    def best(self):
        p = lib().cso_best(self.h)
        return np.ctypeslib.as_array((ctypes.c_uint8 * (self.M * self.n)).from_address(p)).reshape(self.M, self.n).copy()

    def cnt(self):
        p = lib().cso_cnt(self.h)
        return np.ctypeslib.as_array((ctypes.c_uint16 * self.N).from_address(p)).copy()

    def rng_state(self):
        s = np.zeros(4, np.uint64)
        lib().cso_rng_state(self.h, _ptr(s))
        return s

    def set_rng_state(self, s):
        s = np.ascontiguousarray(s, dtype=np.uint64)
        lib().cso_set_rng_state(self.h, _ptr(s))

    def move_delta(self, i, p, v):
        return lib().cso_move_delta(self.h, i, p, v)

    def eval_codeword(self, i, u, use_list=True):
        u = np.ascontiguousarray(u, dtype=np.uint8)
        d = np.full(self.n, DSEL_RANDOM, np.int64)
        lib().cso_eval_codeword(self.h, i, _ptr(u), _ptr(d), int(use_list))
        return d

    def commit(self, i, p, v):
        return lib().cso_commit(self.h, i, p, v)

    def ball(self, i, delta):
        lib().cso_ball(self.h, i, delta)

    def pick(self):
        return lib().cso_pick(self.h)

    def unique_cover(self, i):
        return lib().cso_unique_cover(self.h, i)

    def distinct_fill(self, words):
        w = np.ascontiguousarray(words, dtype=np.uint8).reshape(-1, self.n)
        out = np.zeros((self.M, self.n), np.uint8)
        lib().cso_distinct_fill(self.h, _ptr(w), w.shape[0], _ptr(out))
        return out


def format_code(code, q):
    """Code file text as the reference solver writes it (:836-846)."""
    digits = '0123456789abcdefghijklmnopqrstuvwxyz'
    lines = []
    for w in code:
        if q <= 36:
            lines.append(''.join(digits[int(v)] for v in w))
        else:
            lines.append(' '.join(str(int(v)) for v in w))
    return ''.join(l + '\n' for l in lines)
