"""Training-side Kilosort4 artifacts built in FP64 on the GPU through liblss.

Host-side mirror of the Python the artifacts come from, with the same names and argument meaning:

* ``get_whitening_matrix`` / ``whitening_local`` / ``whitening_from_covariance``
  (kilosort/preprocessing.py:8-35, 96-113): covariance of the CAR + highpass filtered batches
  (every ``nskip``-th, the last batch excluded), local ZCA rows over the ``nrange`` nearest sites.
* ``drift_inverse`` (iKxx, kilosort/datashift.py:205-219) and ``get_drift_matrix``
  (kilosort/preprocessing.py:44-69); ``Sorter.set_drift`` applies M to the whitening operator.
* ``extract_snippets`` (kilosort/spikedetect.py:32-47) and ``extract_wPCA``
  (extract_wPCA_wTEMP's wPCA, kilosort/spikedetect.py:50-72: TruncatedSVD of the normalised
  snippets of every ``nskip``-th whitened batch, capped at 500000 snippets batch by batch).

Tensors are torch device tensors of the current HIP device; the work runs on torch's current
stream and every call returns once it is done.  There is no CPU path: without liblss.so the calls
raise.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from ._lib import check, lib
from .sorter import Sorter, pad_batch

__all__ = ["covariance_add", "whitening_from_covariance", "whitening_local", "FilterStage",
           "get_whitening_matrix", "drift_inverse", "get_drift_matrix", "extract_snippets",
           "wpca", "extract_wPCA"]


def _stream():
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def _f64host(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _rows(x: torch.Tensor, dtype) -> None:
    if x.dtype != dtype or x.dim() != 2 or x.stride(1) != 1 or not x.is_cuda:
        raise ValueError(f"expected a 2-d {dtype} device tensor with unit column stride")


def covariance_add(x_cw: torch.Tensor, cc: torch.Tensor) -> torch.Tensor:
    """cc += x x^T / n for the [C][n] float32 rows x (row stride may exceed n); FP64."""
    _rows(x_cw, torch.float32)
    _rows(cc, torch.float64)
    C, n = x_cw.shape
    check(lib().lss_covariance_add(x_cw.data_ptr(), x_cw.stride(0), C, n, cc.data_ptr(), cc.stride(0),
                                   _stream()), "lss_covariance_add")
    return cc


def whitening_from_covariance(CC: torch.Tensor, eps: float = 1e-6) -> torch.Tensor:
    """ZCA whitening matrix of an n x n covariance (n <= 96), FP64."""
    CC = CC.to(torch.float64).contiguous()
    n = CC.shape[0]
    W = torch.empty_like(CC)
    check(lib().lss_whitening_from_covariance(CC.data_ptr(), n, n, float(eps), W.data_ptr(), n, _stream()),
          "lss_whitening_from_covariance")
    return W


def whitening_local(CC: torch.Tensor, xc, yc, nrange: int = 32, eps: float = 1e-6) -> torch.Tensor:
    """Per-channel whitening rows over the nrange nearest sites (ties by channel index), FP64."""
    CC = CC.to(torch.float64).contiguous()
    C = CC.shape[0]
    xc, yc = _f64host(xc), _f64host(yc)
    W = torch.empty_like(CC)
    check(lib().lss_whitening_local(CC.data_ptr(), C, C, xc.ctypes.data, yc.ctypes.data, int(nrange), float(eps),
                                    W.data_ptr(), C, _stream()), "lss_whitening_local")
    return W


class FilterStage:
    """CAR + highpass of padded int16 batches (the filter of BinaryFiltered without whitening,
    kilosort/io.py:757-775), on a sorter handle whose template tables are placeholders."""

    def __init__(self, n_chan: int, hp_filter, xc, yc, max_window: int, nt: int = 61, do_car: bool = True,
                 fs: float = 30000.0):
        C = int(n_chan)
        art = dict(U=np.ones((1, 1, C), np.float32) / np.sqrt(C), wPCA=np.eye(1, nt, nt // 2, dtype=np.float32),
                   Wrot=np.eye(C, dtype=np.float32), drift=None, hp_filter=np.asarray(hp_filter, np.float32),
                   xc=np.asarray(xc, np.float32), yc=np.asarray(yc, np.float32), nt0min=nt // 2, nC=1, fs=fs)
        self.sorter = Sorter(art, max_window=max_window, max_spikes=1, do_car=do_car)

    def __call__(self, window_dev: torch.Tensor) -> torch.Tensor:
        return self.sorter.stage_filter(window_dev)

    def close(self):
        self.sorter.close()


def get_whitening_matrix(rec: np.ndarray, xc, yc, hp_filter, NT: int = 60000, nt: int = 61, nskip: int = 25,
                         nrange: int = 32, do_car: bool = True, fs: float = 30000.0, device="cuda:0"):
    """Whitening matrix of an int16 [n_samples][C] recording: CC averaged over batches
    0, nskip, 2 nskip, ... < n_batches - 1 of the filtered padded batches' NT central samples,
    then whitening_local.  Returns (Wrot, CC), FP64 device tensors."""
    n_samples, C = rec.shape
    n_batches = int(np.ceil(n_samples / NT))
    filt = FilterStage(C, hp_filter, xc, yc, NT + 2 * nt, nt=nt, do_car=do_car, fs=fs)
    try:
        CC = torch.zeros((C, C), dtype=torch.float64, device=device)
        k = 0
        for j in range(0, n_batches - 1, nskip):
            win, _ = pad_batch(rec, j, NT, nt)
            X = filt(torch.from_numpy(win).to(device))
            covariance_add(X[:, nt:-nt], CC)
            k += 1
        if k == 0:
            raise ValueError("the recording has fewer than two batches")
        CC /= k
        return whitening_local(CC, xc, yc, nrange=nrange), CC
    finally:
        filt.close()


def drift_inverse(xc, yc, sig_interp: float = 20.0, lam: float = 0.01, device="cuda:0"):
    """inv(Kxx + lam I) over the probe sites (FP64 device tensor) and the Newton-Schulz steps."""
    xc, yc = _f64host(xc), _f64host(yc)
    C = len(xc)
    out = torch.empty((C, C), dtype=torch.float64, device=device)
    it = ct.c_int32()
    check(lib().lss_drift_inverse(xc.ctypes.data, yc.ctypes.data, C, float(sig_interp), float(lam), out.data_ptr(), C,
                                  ct.byref(it), _stream()), "lss_drift_inverse")
    return out, it.value


def get_drift_matrix(xc, yc, dshift, iKxx: torch.Tensor, sig_interp: float = 20.0, yblk=None) -> torch.Tensor:
    """M = Kyx @ iKxx for the batch's shifts dshift (one per block; yblk the block centres when
    there are several), FP64 device tensor [C][C]."""
    xc, yc = _f64host(xc), _f64host(yc)
    ds = _f64host(np.atleast_1d(dshift))
    nb = len(ds)
    yb = _f64host(yblk) if yblk is not None else None
    if nb > 1 and (yb is None or len(yb) != nb):
        raise ValueError("yblk must give one centre per block")
    C = len(xc)
    iK = iKxx.to(torch.float64).contiguous()
    M = torch.empty((C, C), dtype=torch.float64, device=iK.device)
    check(lib().lss_drift_matrix(xc.ctypes.data, yc.ctypes.data, C, None if yb is None else yb.ctypes.data,
                                 ds.ctypes.data, nb, float(sig_interp), iK.data_ptr(), C, M.data_ptr(), C, _stream()),
          "lss_drift_matrix")
    return M


def extract_snippets(X_wc: torch.Tensor, nt: int = 61, twav_min: int = 20, Th_single_ch: float = 6.0,
                     loc_range=(4, 5), long_range=(6, 30), max_clips: int | None = None, no_alloc: bool = False):
    """Snippets at the isolated single-channel peaks of a whitened [W][C] window, in channel then
    sample order.  Returns (clips [n][nt] float32, chan [n], time [n], n_found).  no_alloc runs
    lss_extract_snippets_noalloc, whose scratch is a fixed device array (C * ceil(W / 64) up to
    about 520000)."""
    _rows(X_wc, torch.float32)
    W, C = X_wc.shape
    fn = "lss_extract_snippets_noalloc" if no_alloc else "lss_extract_snippets"
    call = getattr(lib(), fn)
    lr = (ct.c_int32 * 2)(*[int(v) for v in loc_range])
    gr = (ct.c_int32 * 2)(*[int(v) for v in long_range])
    n = ct.c_int64()
    # count first, then write exactly the needed rows
    check(call(X_wc.data_ptr(), X_wc.stride(0), C, W, nt, twav_min, float(Th_single_ch), lr, gr, None, 0, None, None,
               ct.byref(n), _stream()), fn)
    cap = n.value if max_clips is None else min(n.value, int(max_clips))
    clips = torch.empty((max(cap, 1), nt), dtype=torch.float32, device=X_wc.device)
    chan = torch.empty(max(cap, 1), dtype=torch.int32, device=X_wc.device)
    tim = torch.empty(max(cap, 1), dtype=torch.int32, device=X_wc.device)
    if cap > 0:
        check(call(X_wc.data_ptr(), X_wc.stride(0), C, W, nt, twav_min, float(Th_single_ch), lr, gr, clips.data_ptr(),
                   cap, chan.data_ptr(), tim.data_ptr(), ct.byref(n), _stream()), fn)
    return clips[:cap], chan[:cap], tim[:cap], n.value


def wpca(clips: torch.Tensor, n_pcs: int = 6):
    """(wPCA [n_pcs][nt], singular values [n_pcs]) of the row-normalised snippets, FP64."""
    clips = clips.to(torch.float32).contiguous()
    n, nt = clips.shape
    out = torch.empty((n_pcs, nt), dtype=torch.float64, device=clips.device)
    sv = torch.empty(n_pcs, dtype=torch.float64, device=clips.device)
    check(lib().lss_wpca(clips.data_ptr(), n, nt, int(n_pcs), out.data_ptr(), sv.data_ptr(), _stream()), "lss_wpca")
    return out, sv


def extract_wPCA(sorter: Sorter, rec: np.ndarray, NT: int = 60000, nskip: int = 25, twav_min: int = 20,
                 Th_single_ch: float = 6.0, n_pcs: int = 6, max_clips: int = 500000, device="cuda:0"):
    """wPCA from the whitened batches 0, nskip, ... of an int16 recording through the sorter's
    filter and whitening (its M @ Wrot), stopping before the batch that would overflow max_clips.
    Returns (wPCA, singular values, number of snippets)."""
    nt = sorter.nt
    n_batches = int(np.ceil(rec.shape[0] / NT))
    parts, total = [], 0
    for j in range(0, n_batches, nskip):
        win, _ = pad_batch(rec, j, NT, nt)
        x = sorter.stage_whiten(sorter.stage_filter(torch.from_numpy(win).to(device)))
        clips, _, _, n = extract_snippets(x, nt=nt, twav_min=twav_min, Th_single_ch=Th_single_ch)
        if total + n > max_clips:
            break
        parts.append(clips)
        total += n
    if total == 0:
        raise ValueError("no isolated peaks above Th_single_ch")
    w, sv = wpca(torch.cat(parts), n_pcs)
    return w, sv, total
