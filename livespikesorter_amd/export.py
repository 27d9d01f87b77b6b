"""Writes an ``oss_input`` artifact directory in the layout LiveSpikeSorter's exporter produces
(src/Python/run_online_spikes.py:545-576) from an artifacts dict (U, wPCA, Wrot, drift, hp_filter,
xc, yc and the scalars), so that ``lss_create_from_dir`` / ``Sorter.from_dir`` can load it.

Arrays derived from the templates (iCC, iU, Ucc, template waveforms, the permuted copies) are
written for compatibility with the reference's loader; liblss rebuilds them itself.  As in the
reference exporter, ``drift_matrix.npy`` holds the transpose of the drift matrix.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def write_oss_input(art: dict, out_dir, with_ctc: bool = False, cluster_xy=None, channel_map=None) -> Path:
    d = Path(out_dir)
    d.mkdir(parents=True, exist_ok=True)
    U = np.asarray(art["U"], np.float32)                 # [T][K][C]
    W = np.asarray(art["wPCA"], np.float32)              # [K][nt]
    T, K, C = U.shape
    nt = W.shape[1]
    xc = np.asarray(art["xc"], np.float64)
    yc = np.asarray(art["yc"], np.float64)
    nC = int(art.get("nC", 10))
    np.save(d / "Wall3.npy", U)
    np.save(d / "wPCA.npy", W)
    np.save(d / "wPCA_permuted.npy", np.ascontiguousarray(W.T))
    np.save(d / "whiteningMat.npy", np.asarray(art["Wrot"], np.float32))
    if art.get("drift") is not None:
        np.save(d / "drift_matrix.npy", np.asarray(art["drift"], np.float32).T)
    np.save(d / "hp_filter.npy", np.asarray(art["hp_filter"], np.float32))
    # acquisition channels, template and channel masks (crop_methods.py:171-175); channel_map
    # defaults to the channels 0..C-1
    cmap = np.arange(C) if channel_map is None else np.asarray(channel_map)
    if cmap.shape != (C,):
        raise ValueError(f"channel_map must hold {C} entries")
    np.save(d / "channelMap.npy", cmap.astype(np.int32))
    np.save(d / "channelMask.npy", np.arange(C, dtype=np.int64))
    np.save(d / "templateMap.npy", np.arange(T, dtype=np.int64))
    np.save(d / "xc.npy", xc)
    np.save(d / "yc.npy", yc)
    # prepare_extract (kilosort/template_matching.py:20-26), ties broken by channel index
    ds = (xc - xc[:, None]) ** 2 + (yc - yc[:, None]) ** 2
    iCC = np.argsort(ds, 0, kind="stable")[:nC].astype(np.int64)
    iU = np.argmax((U.astype(np.float64) ** 2).sum(1), -1).astype(np.int64)
    Ucc = U[np.arange(T)[None, :], :, iCC[:, iU]]        # [nC][T][K]
    np.save(d / "iCC.npy", iCC)
    np.save(d / "iU.npy", iU)
    np.save(d / "Ucc.npy", np.ascontiguousarray(Ucc, np.float32))
    # template waveforms [T][nt][C] (run_online_spikes.py:539)
    pre_wf = np.einsum("tkc,km->tmc", U.astype(np.float64), W.astype(np.float64)).astype(np.float32)
    np.save(d / "preclustered_template_waveforms.npy", pre_wf)
    if with_ctc:
        Wd = W.astype(np.float64)
        wtw = np.zeros((K, K, 2 * nt + 1))
        for l in range(2 * nt + 1):
            for tau in range(nt):
                a = nt - l + tau
                if 0 <= a < nt:
                    wtw[:, :, l] += np.outer(Wd[:, a], Wd[:, tau])
        utu = np.einsum("ikc,jmc->ijkm", U.astype(np.float64), U.astype(np.float64))
        ctc = np.einsum("ijkm,kml->ijl", utu, wtw).astype(np.float32)
        np.save(d / "ctc.npy", ctc)
        np.save(d / "ctc_permuted.npy", np.ascontiguousarray(ctc.transpose(1, 0, 2)))
    if cluster_xy is not None:
        # label centroids, flattened [x1, y1, x2, y2, ...] as OnlineSpikesV2 reads them
        np.save(d / "cluster_centroids.npy", np.asarray(cluster_xy, np.float32).reshape(-1, 2))
    with open(d / "misc.txt", "w") as f:
        f.write(f"nt0min:{int(art.get('nt0min', 20))}\n")
        f.write(f"numNearestChans:{nC}\n")
        f.write(f"Th_learned:{int(round(float(art.get('Th', 8))))}\n")
        f.write(f"duplicate_spike_bins:{int(art.get('dup_bins', 7))}\n")
    return d
