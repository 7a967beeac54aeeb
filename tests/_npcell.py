"""A CPU stand-in for the device cell, for small spaces: the calls the LNS
engine makes (coverage counts of a word list, ball updates and gathers, the
count-transform maps and the map calls), computed by brute force over a
pairwise distance matrix.  Test infrastructure: every answer is exact, so a
run of the engine on it must take the same decisions as on the GPU."""
import numpy as np

from oracle.verify_oracle import digits_table


class NumpyCell:
    def __init__(self, q, n, R, M=1):
        self.q, self.n, self.R, self.M = q, n, R, M
        self.N = q ** n
        d = digits_table(q, n).astype(np.int8)
        D = np.zeros((self.N, self.N), np.int8)
        for p in range(n):
            D += d[:, None, p] != d[None, :, p]
        self.B = (D <= R).astype(np.int32)
        self.cnt_ = np.zeros(self.N, np.int64)
        self.map = None

    def close(self):
        pass

    def load_words(self, words):
        w = np.asarray(words, dtype=np.int64)
        self.cnt_ = self.B[:, w].sum(axis=1).astype(np.int64) if len(w) else np.zeros(self.N, np.int64)

    def ball_update(self, centers, delta):
        for c in np.asarray(centers, dtype=np.int64).ravel():
            self.cnt_ += delta * self.B[:, c]

    def uncovered(self):
        return int((self.cnt_ == 0).sum())

    def cnt(self):
        return self.cnt_.astype(np.uint16)

    def count_transform(self, source, layers=False, want_map=False):
        S = (self.cnt_ == (0 if source == 1 else 1)).astype(np.int32)
        self.map = self.B @ S
        return self.map.copy() if want_map else None

    def map_max(self):
        return int(self.map.max())

    def map_hist(self, nbins, vmax):
        v = self.map[self.map > 0].astype(np.int64)
        return np.bincount(np.minimum((v - 1) * nbins // vmax, nbins - 1), minlength=nbins).astype(np.int64)

    def map_extract(self, thr, cap):
        sel = np.flatnonzero(self.map >= thr)
        return sel[:cap], self.map[sel[:cap]].astype(np.int32), int(sel.size)

    def map_read(self, words):
        return self.map[np.asarray(words, dtype=np.int64)].astype(np.int32)

    def ball_gather(self, centers, target):
        c = np.asarray(centers, dtype=np.int64)
        return (self.B[:, c] * (self.cnt_ == target)[:, None]).sum(axis=0).astype(np.int32)
