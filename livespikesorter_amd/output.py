"""Output step of the sorter (host side, through liblss): nearest-centroid labels, duplicate
removal with state carried across windows, and the reference's spikeOutput.txt writer
(src/SpikeSorter/OnlineSpikesV2.cu:1028-1043, 1117-1164; kilosort/postprocessing.py:10-30)."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import check, lib


def label_spikes(x, y, cx, cy) -> np.ndarray:
    """Index of the nearest cluster centroid for each spike position."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    cx = np.ascontiguousarray(cx, np.float32)
    cy = np.ascontiguousarray(cy, np.float32)
    out = np.zeros(len(x), np.int32)
    check(lib().lss_label_spikes(x.ctypes.data, y.ctypes.data, len(x), cx.ctypes.data, cy.ctypes.data,
                                 len(cx), out.ctypes.data), "lss_label_spikes")
    return out


class Deduplicator:
    """Drops same-label spikes closer than dt samples to the previous kept one; the state
    persists across windows, so overlapping windows do not double-count."""

    def __init__(self, n_labels: int, dt: int = 7):
        self.dt = int(dt)
        self.last = np.full(int(n_labels), np.iinfo(np.int64).min // 2, np.int64)

    def __call__(self, sample, label, amplitude=None):
        o = np.argsort(np.asarray(sample), kind="stable")
        s = np.ascontiguousarray(np.asarray(sample)[o], np.int64)
        l = np.ascontiguousarray(np.asarray(label)[o], np.int32)
        a = None if amplitude is None else np.ascontiguousarray(np.asarray(amplitude)[o], np.float32)
        n_out = ct.c_int64()
        check(lib().lss_dedup_spikes(s.ctypes.data, l.ctypes.data, None if a is None else a.ctypes.data,
                                     len(s), self.dt, self.last.ctypes.data, len(self.last), ct.byref(n_out)),
              "lss_dedup_spikes")
        n = n_out.value
        return s[:n], l[:n], (None if a is None else a[:n])


def write_spikes_txt(path, sample, label, amplitude, y, recording_offset: int = 0, append: bool = True):
    s = np.ascontiguousarray(sample, np.int64)
    l = np.ascontiguousarray(label, np.int32)
    a = np.ascontiguousarray(amplitude, np.float32)
    yy = np.ascontiguousarray(y, np.float32)
    check(lib().lss_write_spikes_txt(str(path).encode(), s.ctypes.data, l.ctypes.data, a.ctypes.data,
                                     yy.ctypes.data, len(s), int(recording_offset), int(append)),
          "lss_write_spikes_txt")
