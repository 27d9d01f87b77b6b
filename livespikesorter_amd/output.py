"""Output step of the sorter (through liblss): nearest-centroid labels, duplicate removal with
state carried across windows, and the reference's spikeOutput.txt writer
(src/SpikeSorter/OnlineSpikesV2.cu:1028-1043, 1117-1164; kilosort/postprocessing.py:10-30); and
the feature-space cluster assignment of the reference's Python online sorter on the device
(src/Python/OnlineKilosort.py:211-284)."""
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


def load_cluster_centroids(directory):
    """(cx, cy) float32 arrays from an oss_input directory's cluster_centroids.npy
    ([x1, y1, x2, y2, ...], src/SpikeSorter/OnlineSpikesV2.cu:612-628)."""
    d = str(directory).encode()
    n = ct.c_int32()
    check(lib().lss_load_cluster_centroids(d, None, None, 0, ct.byref(n)), "lss_load_cluster_centroids")
    cx = np.zeros(n.value, np.float32)
    cy = np.zeros(n.value, np.float32)
    check(lib().lss_load_cluster_centroids(d, cx.ctypes.data, cy.ctypes.data, n.value, ct.byref(n)),
          "lss_load_cluster_centroids")
    return cx, cy


def load_channel_map(directory) -> np.ndarray:
    """int32 channel list from an oss_input directory's channelMap.npy (the channels OnlineSpikesV2
    fetches, src/SpikeSorter/OnlineSpikesV2.cu:474, 657-661)."""
    d = str(directory).encode()
    n = ct.c_int32()
    check(lib().lss_load_channel_map(d, None, 0, ct.byref(n)), "lss_load_channel_map")
    out = np.zeros(n.value, np.int32)
    check(lib().lss_load_channel_map(d, out.ctypes.data, n.value, ct.byref(n)), "lss_load_channel_map")
    return out


def cluster_centroids_pca(sorter):
    """[T][n_nearest * n_pcs] float32 device tensor: each template's PC features on its nearest
    channels, the cluster centroids of cluster_centroids_pca_compute
    (src/Python/OnlineKilosort.py:211-231) with pc_feature_ind[t] = the template's nearest
    channels (Kilosort's iCC[:, iU[t]])."""
    import torch
    out = torch.empty((sorter.T, sorter.nC * sorter.K), dtype=torch.float32, device="cuda")
    check(lib().lss_cluster_centroids_pca(sorter._h, out.data_ptr()), "lss_cluster_centroids_pca")
    return out


def assign_to_clusters(tF, centroids_pca, spike_xs, spike_ys, cluster_xy, R, device="cuda"):
    """Cluster of each spike in PC-feature space (src/Python/OnlineKilosort.py:233-284), same
    arguments: tF [N][nC][D] features, centroids_pca [K][nC*D], spike positions, cluster_xy
    {cluster index: (x, y)} (clusters absent from it never qualify), R the distance limit in um.
    The best of 2 <tF, mu_k> - |mu_k|^2 over clusters within R of the spike, first on ties; -1
    where none is within R.  Runs on the device (lss_assign_clusters); returns int64 labels."""
    import torch
    if len(tF) == 0:
        return np.zeros(0, np.int64)
    dev = torch.device(device)
    X = torch.as_tensor(np.asarray(tF, np.float32) if not torch.is_tensor(tF) else tF).to(dev, torch.float32)
    X = X.reshape(X.shape[0], -1).contiguous()
    mu = torch.as_tensor(np.asarray(centroids_pca, np.float32) if not torch.is_tensor(centroids_pca)
                         else centroids_pca).to(dev, torch.float32).contiguous()
    n, D = X.shape
    K = mu.shape[0]
    if mu.ndim != 2 or mu.shape[1] != D:
        raise ValueError("centroids_pca must be [K][nC*D] with the features' width")
    cx = torch.full((K,), float("nan"), dtype=torch.float32)
    cy = torch.full((K,), float("nan"), dtype=torch.float32)
    for idx, (x, y) in cluster_xy.items():
        cx[idx] = float(x)
        cy[idx] = float(y)
    cx, cy = cx.to(dev), cy.to(dev)
    xs = torch.as_tensor(np.asarray(spike_xs, np.float32)).to(dev).contiguous()
    ys = torch.as_tensor(np.asarray(spike_ys, np.float32)).to(dev).contiguous()
    if xs.numel() != n or ys.numel() != n:
        raise ValueError("one position per spike")
    label = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream    # after the copies above, then waited for
    check(lib().lss_assign_clusters(X.data_ptr(), n, D, mu.data_ptr(), K, xs.data_ptr(), ys.data_ptr(),
                                    cx.data_ptr(), cy.data_ptr(), float(R), label.data_ptr(), stream),
          "lss_assign_clusters")
    return label.cpu().numpy().astype(np.int64)
