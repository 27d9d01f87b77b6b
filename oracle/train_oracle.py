"""CPU oracle of the training-side Kilosort4 artifacts (test infrastructure only).

This module is the checker, never the product: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it.  It restates in numpy FP64 the builders of
the Kilosort 4.0.13 copy vendored by LiveSpikeSorter at
``src/Python/run_online_spikes/Lib/site-packages/kilosort/`` (``ks/`` below): whitening, drift
interpolation and the wPCA temporal basis.  Each function cites the file:line it follows.  The
reference runs these in torch FP32 on the GPU (``CC``, ``torch.linalg.svd``) and FP32 numpy +
scikit-learn (snippets, ``TruncatedSVD``); the restatement is FP64 throughout, and where the
reference's library routine fixes a convention (SVD singular values, interp1d extrapolation,
TruncatedSVD's sign rule, max_pool2d/avg_pool2d padding) the pin below checks it against that
library routine.

Pinning (tests/test_train_oracle.py):

* ``covariance``                - a constructed X = A Z with orthonormal Z rows gives CC = A A^T;
                                  brute-force loops on a tiny case.
* ``whitening_from_covariance`` - eigh-based here; checked against the reference's routine
                                  (torch.linalg.svd in FP64, E / sqrt(D + eps) @ E.T), against the
                                  diagonal closed form, and W CC W = I - eps (CC + eps I)^-1.
* ``neighbours`` / ``whitening_local`` - nrange = C reproduces the full ZCA rows; a block-diagonal
                                  covariance gives the per-block ZCA; the neighbour sets equal
                                  numpy's default argsort sets at nrange = 32 on linear and
                                  staggered probe layouts (ties broken by channel index here).
* ``kernel2d`` / ``drift_inverse`` - brute-force loops; (Kxx + lam I) iKxx = I; eigen closed form.
* ``drift_shifts``              - scipy ``interp1d(kind='linear', fill_value='extrapolate')`` at
                                  points inside and outside the block centres.
* ``drift_matrix``              - zero shift gives the closed form Kxx (Kxx + lam I)^-1.
* ``isolated_peaks``            - torch ``max_pool2d`` / ``avg_pool2d`` (count_include_pad) as
                                  ks/spikedetect.py:17-47 call them; brute-force loops on a tiny
                                  window.
* ``wpca``                      - scikit-learn ``TruncatedSVD`` (arpack and randomized) on the same
                                  normalised snippets, components and singular values.
"""
from __future__ import annotations

import numpy as np


# ----------------------------------------------------------------------------------------------
# whitening (ks/preprocessing.py:8-35, 96-113)
# ----------------------------------------------------------------------------------------------
def covariance(batches) -> np.ndarray:
    """ks/preprocessing.py:101-110: CC = mean over batches of X X^T / n, X = [C][n] filtered."""
    CC = None
    k = 0
    for X in batches:
        X = np.asarray(X, dtype=np.float64)
        P = X @ X.T / X.shape[1]
        CC = P if CC is None else CC + P
        k += 1
    return CC / k


def whitening_from_covariance(CC: np.ndarray, eps: float = 1e-6) -> np.ndarray:
    """ks/preprocessing.py:8-17: W = E diag(1/sqrt(D + eps)) E^T with E, D the SVD of the symmetric
    PSD CC, i.e. D = |eigenvalues| (computed here with eigh)."""
    d, E = np.linalg.eigh(np.asarray(CC, dtype=np.float64))
    return (E / np.sqrt(np.abs(d) + eps)) @ E.T


def neighbours(xc, yc, j: int, nrange: int) -> np.ndarray:
    """ks/preprocessing.py:27-29: the nrange sites nearest to site j by squared distance; ties by
    channel index (a stable sort)."""
    xc = np.asarray(xc, np.float64)
    yc = np.asarray(yc, np.float64)
    ds = (xc[j] - xc) ** 2 + (yc[j] - yc) ** 2
    return np.argsort(ds, kind="stable")[:nrange]


def whitening_local(CC: np.ndarray, xc, yc, nrange: int = 32, eps: float = 1e-6) -> np.ndarray:
    """ks/preprocessing.py:19-35: row j = first row of the ZCA of CC over j's neighbours."""
    C = CC.shape[0]
    W = np.zeros((C, C))
    for j in range(C):
        ix = neighbours(xc, yc, j, nrange)
        W[j, ix] = whitening_from_covariance(CC[np.ix_(ix, ix)], eps)[0]
    return W


# ----------------------------------------------------------------------------------------------
# drift interpolation (ks/preprocessing.py:44-69, ks/datashift.py:176-184, 205-219)
# ----------------------------------------------------------------------------------------------
def kernel2d(x: np.ndarray, y: np.ndarray, sig: float) -> np.ndarray:
    """ks/datashift.py:181-184: exp(-|x_i - y_j|^2 / (2 sig^2)) for [n][2] site arrays."""
    ds = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    return np.exp(-ds / (2 * sig ** 2))


def sites(xc, yc) -> np.ndarray:
    return np.stack((np.asarray(xc, np.float64), np.asarray(yc, np.float64)), 1)


def drift_inverse(xc, yc, sig_interp: float = 20.0, lam: float = 0.01) -> np.ndarray:
    """ks/datashift.py:215-219: iKxx = inv(Kxx + lam I)."""
    xp = sites(xc, yc)
    Kxx = kernel2d(xp, xp, sig_interp)
    return np.linalg.inv(Kxx + lam * np.eye(len(xp)))


def drift_shifts(yc, dshift, yblk=None) -> np.ndarray:
    """ks/preprocessing.py:48-53: one block shifts every site by dshift; several blocks
    interpolate linearly over the block centres yblk, extrapolating the end segments."""
    yc = np.asarray(yc, np.float64)
    ds = np.atleast_1d(np.asarray(dshift, np.float64))
    if len(ds) == 1:
        return np.full(len(yc), ds[0])
    yb = np.asarray(yblk, np.float64)
    o = np.argsort(yb, kind="stable")
    yb, ds = yb[o], ds[o]
    hi = np.clip(np.searchsorted(yb, yc, side="left"), 1, len(yb) - 1)
    lo = hi - 1
    return (ds[hi] - ds[lo]) / (yb[hi] - yb[lo]) * (yc - yb[lo]) + ds[lo]


def drift_matrix(xc, yc, dshift, iKxx: np.ndarray, sig_interp: float = 20.0, yblk=None) -> np.ndarray:
    """ks/preprocessing.py:44-69: M = Kyx @ iKxx with the sites shifted down by the shifts."""
    xp = sites(xc, yc)
    yp = xp.copy()
    yp[:, 1] -= drift_shifts(yc, dshift, yblk)
    return kernel2d(yp, xp, sig_interp) @ iKxx


# ----------------------------------------------------------------------------------------------
# wPCA (ks/spikedetect.py:17-79)
# ----------------------------------------------------------------------------------------------
def _box_max(a: np.ndarray, rc: int, rt: int) -> np.ndarray:
    """Maximum over [c-rc, c+rc] x [t-rt, t+rt], out-of-range cells excluded."""
    C, W = a.shape
    p = np.full((C + 2 * rc, W + 2 * rt), -np.inf)
    p[rc:rc + C, rt:rt + W] = a
    out = np.full((C, W), -np.inf)
    for dc in range(2 * rc + 1):
        for dt in range(2 * rt + 1):
            np.maximum(out, p[dc:dc + C, dt:dt + W], out=out)
    return out


def _box_count(a: np.ndarray, rc: int, rt: int) -> np.ndarray:
    """Integer count of a 0/1 array over [c-rc, c+rc] x [t-rt, t+rt] (zero outside)."""
    C, W = a.shape
    s = np.zeros((C + 1, W + 1), np.int64)
    s[1:, 1:] = np.cumsum(np.cumsum(a.astype(np.int64), 0), 1)
    c0 = np.clip(np.arange(C) - rc, 0, C)
    c1 = np.clip(np.arange(C) + rc + 1, 0, C)
    t0 = np.clip(np.arange(W) - rt, 0, W)
    t1 = np.clip(np.arange(W) + rt + 1, 0, W)
    return (s[c1][:, t1] - s[c0][:, t1] - s[c1][:, t0] + s[c0][:, t0])


def isolated_peaks(X: np.ndarray, nt: int, Th_single_ch: float, loc_range=(4, 5), long_range=(6, 30)):
    """ks/spikedetect.py:32-43 on a whitened [C][W] window: (channel, sample) of the isolated
    peaks in row-major (nonzero) order."""
    Xabs = np.abs(np.asarray(X, np.float32))
    Xmax = _box_max(Xabs, loc_range[0], loc_range[1])
    ispeak = (Xmax == Xabs) & (Xabs > np.float32(Th_single_ch))
    n = _box_count(ispeak, long_range[0], long_range[1])
    iso = (n == 1) & ispeak
    iso[:, :nt] = False
    iso[:, -nt:] = False
    c, t = np.nonzero(iso)
    return c, t


def extract_snippets(X: np.ndarray, nt: int = 61, twav_min: int = 20, Th_single_ch: float = 6.0,
                     loc_range=(4, 5), long_range=(6, 30)):
    """ks/spikedetect.py:32-47: clips[i] = X[c_i, t_i - twav_min + arange(nt)]."""
    X = np.asarray(X, np.float32)
    c, t = isolated_peaks(X, nt, Th_single_ch, loc_range, long_range)
    idx = t[:, None] - twav_min + np.arange(nt)[None, :]
    return X[c[:, None], idx], c, t


def wpca(clips: np.ndarray, n_pcs: int = 6):
    """ks/spikedetect.py:69-72: TruncatedSVD(n_pcs) of the row-normalised clips (uncentred): the
    leading eigenvectors of the Gram matrix, each signed so its largest-|.| entry (first on ties)
    is positive; returns (components [n_pcs][nt], singular values)."""
    Q = np.asarray(clips, np.float64)
    Q = Q / np.sqrt((Q ** 2).sum(1, keepdims=True))
    d, V = np.linalg.eigh(Q.T @ Q)
    o = np.argsort(-d, kind="stable")[:n_pcs]
    comp = V[:, o].T.copy()
    for r in range(n_pcs):
        if comp[r, np.argmax(np.abs(comp[r]))] < 0:
            comp[r] = -comp[r]
    return comp, np.sqrt(np.maximum(d[o], 0.0))
