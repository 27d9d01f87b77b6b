"""Spike-list comparison (Kilosort's ground-truth matcher, kilosort/bench.py:150-169).

``f1_same_template`` matches two spike lists template by template with the greedy two-pointer
rule of ``nmatch`` (a spike of one list matches the nearest-following unmatched-order spike of
the other within +-dt samples) and returns 2*matches / (n1 + n2).
"""
from __future__ import annotations

import numpy as np


def nmatch(ss0, ss, dt=6):
    """Greedy match of sorted times ss against sorted ss0 within dt (kilosort/bench.py:150-169)."""
    i = j = n0 = 0
    ntmax = len(ss0)
    while i < len(ss):
        while j + 1 < ntmax and ss0[j + 1] <= ss[i] + dt:
            j += 1
        if ntmax and abs(int(ss0[j]) - int(ss[i])) <= dt:
            n0 += 1
        i += 1
    return n0


def f1_same_template(t_ref, y_ref, t_new, y_new, dt=1):
    t_ref, y_ref = np.asarray(t_ref), np.asarray(y_ref)
    t_new, y_new = np.asarray(t_new), np.asarray(y_new)
    if len(t_ref) == 0 and len(t_new) == 0:
        return 1.0
    n = 0
    for u in np.union1d(np.unique(y_ref), np.unique(y_new)):
        a = np.sort(t_ref[y_ref == u])
        b = np.sort(t_new[y_new == u])
        if len(a) and len(b):
            n += nmatch(a, b, dt)
    return 2.0 * n / (len(t_ref) + len(t_new))
