"""CPU oracle of cghip_matrix_stats (TEST INFRASTRUCTURE ONLY).

Brute-force numpy restatement of the 36 matrix statistics: per row,
q0 = log a_ii, q1 = sum_{j!=i} |a_ij| / a_ii, q2/q3 = log of the largest /
smallest nonzero off-diagonal |a_ij| / a_ii, q4 = q2 - q3 (q2..q4 over rows
that have nonzero off-diagonal entries); mean, population std, min, max of
each; fractions of rows of length 1..15 and >= 16.  No reference
counterpart (the reference computes no matrix statistics).
"""
from __future__ import annotations

import numpy as np


def matrix_stats(ptr, index, value) -> np.ndarray:
    ptr = np.asarray(ptr, np.int64)
    index = np.asarray(index, np.int64)
    value = np.asarray(value, np.float64)
    n = ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(ptr))
    isdiag = index == rows
    d = np.zeros(n)
    d[rows[isdiag]] = value[isdiag]
    if np.any(d <= 0):
        raise ValueError("row without a positive diagonal")
    off = ~isdiag & (value != 0)
    a = np.abs(value)
    s = np.bincount(rows[off], weights=a[off], minlength=n)
    smax = np.zeros(n)
    np.maximum.at(smax, rows[off], a[off])
    smin = np.full(n, np.inf)
    np.minimum.at(smin, rows[off], a[off])
    has = smax > 0
    q = [np.log(d), s / d]
    q2 = np.log(smax[has] / d[has])
    q3 = np.log(smin[has] / d[has])
    q += [q2, q3, q2 - q3]
    out = []
    for v in q:
        if v.size:
            out += [v.mean(), v.std(), v.min(), v.max()]
        else:
            out += [0.0, 0.0, 0.0, 0.0]
    lens = np.diff(ptr)
    bins = np.clip(lens, 1, 16) - 1
    out += list(np.bincount(bins, minlength=16)[:16] / n)
    return np.array(out)
