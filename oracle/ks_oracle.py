"""CPU oracle for the Kilosort4 online template-matching path (test infrastructure only).

This module is the checker, never the product: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it.  It restates, in plain CPU torch (FP32 or
FP64, chosen per call), the algorithm of the Kilosort 4.0.13 copy vendored by LiveSpikeSorter at
``src/Python/run_online_spikes/Lib/site-packages/kilosort/`` (abbreviated ``ks/`` below).  Every
function cites the file:line it follows.  Nothing here imports or runs the reference.

Pinning (what checks each function independently of the restatement itself; the tests named
are in ``tests/test_oracle.py``):

* ``padded_batch``        - brute-force edge replication on tiny recordings (first/middle/last
                            batch, single batch).
* ``filter_batch``        - mean: closed form; median: numpy lower median; highpass: scipy
                            ``filtfilt`` of the batch (circular, on a periodic extension) and a
                            brute-force circular correlation; whitening: direct matmul.
* ``highpass_taps``       - scipy butter/filtfilt invariants: zero-phase symmetry, DC gain ~0,
                            unit gain far above the cutoff.
* ``prepare_matching``    - brute-force quadruple loop on tiny U/wPCA; symmetry
                            ctc[i,j,l] = ctc[j,i,2nt-l]; ctc[i,i,nt] = nm[i] for orthonormal wPCA.
* ``prepare_extract``     - brute force on a 4-channel probe.
* ``run_matching``        - the reference's own integer KAT (``src/Python/conv1d_test.py``:14-43)
                            for projection/contraction/score/max-pool; brute-force single-spike
                            recovery (exact amplitude, exact residual); explicit tie cases; the
                            multi-spike update against a pure-Python spike-by-spike pursuit that
                            rescores the projection of the residual every pass (a tiny integer
                            case with an exactly tied plateau and two peaks of one pass
                            nt < d <= 2nt apart: spike list, amplitudes, residual and final
                            scores); the linearity invariant B = project(Xres) after the last
                            pass of a crowded FP64 window; ``_sub_windows`` against torch's own
                            in-place indexed subtract with repeated columns.  Both pursuit checks
                            fail when the even/odd split is dropped (``UPDATE_GROUPS = 1``).
* ``extract_features``    - brute force on tiny inputs.
* ``spike_positions``     - closed form (all mass on one channel gives that channel's position).
* ``nmatch``              - hand-checked lists.
* ``cluster_centroids_pca`` - the reference layout (Wall [T][C][P], pc_feature_ind [T][nC]) gives
                            the same rows as Ucc of ``prepare_extract``.
* ``assign_to_clusters``  - 2<x,mu> - |mu|^2 = |x|^2 - |x-mu|^2: the label is the nearest centroid
                            in feature space among clusters within R (brute force with
                            ``scipy.spatial.distance.cdist``); no cluster in range gives -1.
* ``residual_energy``     - equals |Xres|^2 of ``run_matching`` for the pursuit's own spikes and
                            amplitudes; its autograd gradient against FP64 central differences
                            (the loss is quadratic in U and in wPCA, so central differences are
                            exact up to rounding); brute-force loops on a tiny case.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch
import torch.nn.functional as F


# ----------------------------------------------------------------------------------------------
# batching (ks/io.py:584-637)
# ----------------------------------------------------------------------------------------------
def n_batches(n_samples: int, NT: int) -> int:
    """ks/io.py:465: number of NT batches covering the recording."""
    return int(math.ceil(n_samples / NT))


def padded_batch(rec: np.ndarray, ibatch: int, NT: int, nt: int, dtype=torch.float32) -> torch.Tensor:
    """Padded batch [C][NT+2nt] from an int16 recording rec [n_samples][C].

    Follows ``BinaryRWFile.padded_batch_to_torch`` (ks/io.py:584-637): batch 0 reads NT+nt samples
    into columns nt.. and replicates column nt into the first nt columns; the last batch reads what
    is left from ibatch*NT-nt and replicates its last column to the end; middle batches read
    NT+2nt samples.  A recording with one batch takes the batch-0 branch only, so its tail stays
    zero (ks/io.py:622-631: the ``elif`` is not reached).
    """
    n_samples, C = rec.shape
    nb = n_batches(n_samples, NT)
    if ibatch == 0:
        bstart, bend = 0, min(n_samples, NT + nt)
    else:
        bstart = ibatch * NT - nt
        bend = min(n_samples, bstart + NT + 2 * nt)
    data = torch.from_numpy(np.ascontiguousarray(rec[bstart:bend].T)).to(dtype)
    nsamp = data.shape[-1]
    X = torch.zeros((C, NT + 2 * nt), dtype=dtype)
    if ibatch == 0:
        X[:, nt:nt + nsamp] = data
        X[:, :nt] = X[:, nt:nt + 1]
    elif ibatch == nb - 1:
        X[:, :nsamp] = data
        X[:, nsamp:] = X[:, nsamp - 1:nsamp]
    else:
        X[:] = data
    return X


def window_start(ibatch: int, NT: int, nt: int) -> int:
    """Recording sample of padded column 0 (ks/io.py:597, 625: bstart - nt for batch 0)."""
    return ibatch * NT - nt


# ----------------------------------------------------------------------------------------------
# preprocessing (ks/preprocessing.py:121-154, ks/io.py:757-790)
# ----------------------------------------------------------------------------------------------
def highpass_taps(fs: float = 30000.0, cutoff: float = 300.0) -> np.ndarray:
    """ks/preprocessing.py:121-136: filtfilt impulse response of a 3rd-order Butterworth
    highpass, 30122 taps with the impulse at 15061; returned in float64."""
    from scipy.signal import butter, filtfilt
    NT = 30122
    b, a = butter(3, cutoff, fs=fs, btype="high")
    x = np.zeros(NT)
    x[NT // 2] = 1
    return filtfilt(b, a, x).copy()


def fft_highpass(hp_filter: torch.Tensor, NT: int) -> torch.Tensor:
    """ks/preprocessing.py:138-154: pad (centred) or crop the taps to NT and FFT them."""
    ft = hp_filter.shape[0]
    if ft < NT:
        pad = (NT - ft) // 2
        z = torch.zeros
        return torch.fft.fft(torch.cat((z(pad, dtype=hp_filter.dtype), hp_filter,
                                        z(pad + (NT - pad * 2 - ft), dtype=hp_filter.dtype))))
    if ft > NT:
        crop = (ft - NT) // 2
        return torch.fft.fft(hp_filter[crop:crop + NT])
    return torch.fft.fft(hp_filter)


def filter_batch(X: torch.Tensor, hp_filter: torch.Tensor | None, Wmat: torch.Tensor | None,
                 do_car: bool = True) -> torch.Tensor:
    """``BinaryFiltered.filter`` (ks/io.py:757-790) without channel map / sign / artifact steps.

    X [C][W'] -> mean over time removed per channel (io.py:765), lower median over channels
    removed per sample (io.py:768), circular FFT highpass then fftshift (io.py:771-774),
    then ``Wmat @ X`` where Wmat = M @ Wrot (io.py:781-789).
    """
    X = X - X.mean(1).unsqueeze(1)
    if do_car:
        X = X - torch.median(X, 0)[0]
    if hp_filter is not None:
        fwav = fft_highpass(hp_filter.to(X.dtype), NT=X.shape[1])
        X = torch.real(torch.fft.ifft(torch.fft.fft(X) * torch.conj(fwav)))
        X = torch.fft.fftshift(X, dim=-1)
    if Wmat is not None:
        X = Wmat.to(X.dtype) @ X
    return X


# ----------------------------------------------------------------------------------------------
# matching setup (ks/template_matching.py:20-26, 110-121)
# ----------------------------------------------------------------------------------------------
def prepare_matching(U: torch.Tensor, wPCA: torch.Tensor) -> torch.Tensor:
    """ks/template_matching.py:110-121: ctc[i,j,l] = sum_{k,m} (U_i U_j^T)[k,m] WtW[k,m,l] with
    WtW = flip(conv1d(wPCA, wPCA, padding=nt)).  U [T][K][C], wPCA [K][nt] -> [T][T][2nt+1]."""
    nt = wPCA.shape[1]
    W = wPCA.contiguous()
    WtW = F.conv1d(W.reshape(-1, 1, nt), W.reshape(-1, 1, nt), padding=nt)
    WtW = torch.flip(WtW, [2])
    UtU = torch.einsum("ikl, jml -> ijkm", U, U)
    return torch.einsum("ijkm, kml -> ijl", UtU, WtW)


def prepare_extract(xc: np.ndarray, yc: np.ndarray, U: torch.Tensor, nC: int):
    """ks/template_matching.py:20-26.  iCC [nC][C] nearest channels of each channel, iU [T] the
    channel of largest template energy, Ucc [nC][T][K].  Ties in distance are broken by channel
    index (stable sort); the reference uses numpy's default quicksort there, whose tie order is
    unspecified."""
    ds = (xc - xc[:, None]) ** 2 + (yc - yc[:, None]) ** 2
    iCC = torch.from_numpy(np.argsort(ds, 0, kind="stable")[:nC].astype(np.int64))
    iU = torch.argmax((U ** 2).sum(1), -1)
    Ucc = U[torch.arange(U.shape[0]), :, iCC[:, iU]]
    return iCC, iU, Ucc


def template_norms(U: torch.Tensor) -> torch.Tensor:
    """ks/template_matching.py:141: nm[t] = sum_{k,c} U[t,k,c]^2."""
    return (U ** 2).sum(-1).sum(-1)


def template_waves(U: torch.Tensor, wPCA: torch.Tensor) -> torch.Tensor:
    """Template waveforms [T][nt][C] = sum_k U[t,k,c] wPCA[k,m]; the tensor the residual update
    subtracts (ks/template_matching.py:197), stored like the exporter's
    ``preclustered_template_waveforms`` (run_online_spikes.py:539)."""
    return torch.einsum("tkc, km -> tmc", U, wPCA)


# ----------------------------------------------------------------------------------------------
# matching pursuit (ks/template_matching.py:128-202)
# ----------------------------------------------------------------------------------------------
def project(X: torch.Tensor, U: torch.Tensor, wPCA: torch.Tensor) -> torch.Tensor:
    """ks/template_matching.py:145-146: B[t,w] = sum_{k,c} U[t,k,c] sum_m X[c, w+m-nt//2] wPCA[k,m]
    ('same' zero padding).  Returns B [T][W']."""
    nt = wPCA.shape[1]
    Bp = F.conv1d(X.unsqueeze(1), wPCA.unsqueeze(1), padding=nt // 2)
    return torch.einsum("ijk, kjl -> il", U, Bp)


def score_argmax(B: torch.Tensor, nm: torch.Tensor, nt: int):
    """ks/template_matching.py:159, 165-167: Cf = relu(B)^2/nm with nt edge columns zeroed,
    then max / first argmax over templates."""
    Cf = torch.relu(B) ** 2 / nm.unsqueeze(-1)
    Cf[:, :nt] = 0
    Cf[:, -nt:] = 0
    return torch.max(Cf, 0)


def find_peaks(Cfmax: torch.Tensor, nt: int, Th: float) -> torch.Tensor:
    """ks/template_matching.py:168-174: max-pool (2nt+1, -inf padding), keep columns where the
    pooled value exceeds Th^2 and equals the column's own max (|diff| < 1e-9).  Ascending."""
    Cmax = F.max_pool1d(Cfmax.unsqueeze(0).unsqueeze(0), (2 * nt + 1), stride=1, padding=nt)
    cnd1 = Cmax[0, 0] > Th ** 2
    cnd2 = torch.abs(Cmax[0, 0] - Cfmax) < 1e-9
    return torch.nonzero(cnd1 * cnd2)[:, 0]


def _sub_windows(A: torch.Tensor, cols: torch.Tensor, vals: torch.Tensor):
    """A[:, cols] -= vals with cols [n][L], vals [R][n][L] (ks/template_matching.py:197-198).

    When windows of one group overlap (only possible for exactly tied peaks) the reference's
    ``index_put_`` has no defined order; here the later spike's write wins, as a sequential
    gather-subtract-scatter gives."""
    flat = cols.reshape(-1)
    if torch.unique(flat).numel() == flat.numel():
        A[:, cols] -= vals
        return
    snap = A[:, cols].clone()
    for s in range(cols.shape[0]):
        A[:, cols[s]] = snap[:, s] - vals[:, s]


# Peaks of one pass are subtracted in this many interleaved groups (ks/template_matching.py:194:
# n = 2, even-indexed peaks first).  Consecutive peaks of a pass can sit nt < d <= 2nt apart, so
# their +-nt score windows overlap; members of one group are two peaks apart (> 2nt) and never
# overlap unless peaks tie exactly.
UPDATE_GROUPS = 2


@dataclass
class MatchResult:
    st: torch.Tensor        # [nsp][2] int64: (padded column iX, template)
    amps: torch.Tensor      # [nsp][1]
    Xres: torch.Tensor      # [C][W'] residual
    passes: int             # passes that found at least one spike
    per_pass: list          # spikes found in each pass
    B: torch.Tensor = None  # [T][W'] scores after the last update


def run_matching(X: torch.Tensor, U: torch.Tensor, wPCA: torch.Tensor, ctc: torch.Tensor,
                 Th: float = 8.0, max_passes: int = 100, nm: torch.Tensor | None = None,
                 waves: torch.Tensor | None = None, B: torch.Tensor | None = None) -> MatchResult:
    """ks/template_matching.py:128-202 (run_matching), with the pass cap as a parameter
    (100 in the reference Python, 50 in the CUDA driver OnlineSpikesV2.cu:903).

    X [C][W'] preprocessed batch, U [T][K][C], wPCA [K][nt], ctc [T][T][2nt+1].
    ``nm``/``waves``/``B`` may be supplied to share exactly the same FP32 tensors with the device
    path (stage-level parity); otherwise they are computed as the reference does.
    """
    nt = wPCA.shape[1]
    if nm is None:
        nm = template_norms(U)
    if waves is None:
        waves = template_waves(U, wPCA)           # [T][nt][C]
    if B is None:
        B = project(X, U, wPCA)
    else:
        B = B.clone()
    trange = torch.arange(-nt, nt + 1)
    tiwave = torch.arange(-(nt // 2), nt // 2 + 1)
    Xres = X.clone()
    st_l, amp_l, per_pass = [], [], []
    for _ in range(max_passes):
        Cfmax, imax = score_argmax(B, nm, nt)
        iX = find_peaks(Cfmax, nt, Th)
        if iX.numel() == 0:
            break
        iY = imax[iX]
        amp = B[iY, iX] / nm[iY]
        st_l.append(torch.stack((iX, iY), 1))
        amp_l.append(amp)
        per_pass.append(int(iX.numel()))
        n = UPDATE_GROUPS
        for j in range(n):
            ix, iy, a = iX[j::n], iY[j::n], amp[j::n]
            if ix.numel() == 0:
                continue
            # Xres[:, iX+tiwave] -= amp * einsum('ijk, jl -> kil', U[iY], W)   (:197)
            wv = waves[iy].permute(2, 0, 1)               # [C][n][nt]
            _sub_windows(Xres, ix[:, None] + tiwave, a[None, :, None] * wv)
            # B[:, iX+trange] -= amp * ctc[:, iY, :]                            (:198)
            _sub_windows(B, ix[:, None] + trange, a[None, :, None] * ctc[:, iy, :])
    if st_l:
        st = torch.cat(st_l, 0)
        amps = torch.cat(amp_l, 0).unsqueeze(1)
    else:
        st = torch.zeros((0, 2), dtype=torch.int64)
        amps = torch.zeros((0, 1), dtype=X.dtype)
    return MatchResult(st, amps, Xres, len(per_pass), per_pass, B)


# ----------------------------------------------------------------------------------------------
# template refinement: residual energy of a fixed spike set (no reference counterpart)
# ----------------------------------------------------------------------------------------------
def residual_energy(X: torch.Tensor, col: torch.Tensor, tmp: torch.Tensor, amp: torch.Tensor,
                    U: torch.Tensor, wPCA: torch.Tensor, chunk: int = 512):
    """Loss of template refinement for a fixed spike set, differentiable in U and wPCA.

    R = X - sum_s amp_s * wave_{tmp_s}[:, m] placed at column col_s - nt//2 + m, with
    wave = template_waves(U, wPCA) (the update of ks/template_matching.py:197); contributions of
    different spikes add (the reference's successive passes subtract them one after another) and
    columns outside [0, W') are dropped.  Returns (sum R^2, R).  X [C][W'], U [T][K][C],
    wPCA [K][nt], col/tmp int64 [n], amp [n].  For the pursuit's own spikes and amplitudes, R is
    the pursuit's residual Xres.
    """
    C, W = X.shape
    nt = wPCA.shape[1]
    tiw = torch.arange(-(nt // 2), nt // 2 + 1)
    waves = template_waves(U, wPCA)                       # [T][nt][C]
    R = X
    for i0 in range(0, col.numel(), chunk):
        c, t, a = col[i0:i0 + chunk], tmp[i0:i0 + chunk], amp[i0:i0 + chunk]
        idx = c[:, None] + tiw                            # [n][nt]
        v = a[:, None, None] * waves[t]                   # [n][nt][C]
        ok = (idx >= 0) & (idx < W)
        R = R.index_add(1, idx[ok], -v[ok].T)
    return (R * R).sum(), R


# ----------------------------------------------------------------------------------------------
# features, positions, times (ks/template_matching.py:50-51, 53-59, 68; ks/postprocessing.py)
# ----------------------------------------------------------------------------------------------
def extract_features(Xres, st, amps, iCC, iU, Ucc, wPCA):
    """ks/template_matching.py:50-51: tF [nsp][nC][K] =
    Xres[iCC[:, iU[t]], iX + tiwave] @ wPCA^T + amp * Ucc[:, t]."""
    nt = wPCA.shape[1]
    tiwave = torch.arange(-(nt // 2), nt // 2 + 1)
    xfeat = Xres[iCC[:, iU[st[:, 1:2]]], st[:, :1] + tiwave] @ wPCA.T
    xfeat = xfeat + amps * Ucc[:, st[:, 1]]
    return xfeat.transpose(0, 1).contiguous()


def spike_positions(tF, st, iCC, iU, xc, yc):
    """ks/postprocessing.py:33-46: mass-weighted channel positions over the nC nearest channels."""
    tmass = (tF ** 2).sum(-1)
    tmass = tmass / tmass.sum(1, keepdim=True)
    xc = torch.as_tensor(xc, dtype=tF.dtype)
    yc = torch.as_tensor(yc, dtype=tF.dtype)
    chs = iCC[:, iU[st[:, 1]]]
    xs = (xc[chs.T] * tmass).sum(1)
    ys = (yc[chs.T] * tmass).sum(1)
    return xs, ys


def spike_samples(st, wstart: int, nt: int, nt0min: int):
    """ks/template_matching.py:68 with wstart = ibatch*NT - nt: recording sample of each spike."""
    return st[:, 0].to(torch.int64) + wstart - nt // 2 + nt0min


def remove_duplicates(spike_times, spike_clusters, dt=15):
    """ks/postprocessing.py:10-30: drop same-cluster spikes within dt samples (time-sorted)."""
    keep = np.zeros(len(spike_times), bool)
    t0s = {}
    for i in range(len(spike_times)):
        t, c = int(spike_times[i]), int(spike_clusters[i])
        t0 = t0s.get(c, t - dt)
        if t >= t0 + dt:
            t0s[c] = t
            keep[i] = True
    return spike_times[keep], spike_clusters[keep], keep


# ----------------------------------------------------------------------------------------------
# spike-list comparison (ks/bench.py:150-169)
# ----------------------------------------------------------------------------------------------
def nmatch(ss0, ss, dt=6):
    """ks/bench.py:150-169: greedy two-pointer match of sorted times ss against sorted ss0."""
    i = j = n0 = 0
    ntmax = len(ss0)
    is_matched = np.zeros(len(ss), bool)
    is_matched0 = np.zeros(len(ss0), bool)
    while i < len(ss):
        while j + 1 < ntmax and ss0[j + 1] <= ss[i] + dt:
            j += 1
        if ntmax and abs(ss0[j] - ss[i]) <= dt:
            n0 += 1
            is_matched[i] = 1
            is_matched0[j] = 1
        i += 1
    return n0, is_matched, is_matched0


def spike_f1(t_ref, y_ref, t_new, y_new, dt=1):
    """F1 between two spike lists, matching per template with ``nmatch`` (dt samples)."""
    t_ref, y_ref = np.asarray(t_ref), np.asarray(y_ref)
    t_new, y_new = np.asarray(t_new), np.asarray(y_new)
    if len(t_ref) == 0 and len(t_new) == 0:
        return 1.0
    n_match = 0
    for u in np.union1d(np.unique(y_ref), np.unique(y_new)):
        a = np.sort(t_ref[y_ref == u])
        b = np.sort(t_new[y_new == u])
        if len(a) and len(b):
            n_match += nmatch(a, b, dt)[0]
    return 2.0 * n_match / (len(t_ref) + len(t_new))


# ----------------------------------------------------------------------------------------------
# whole-window pipeline
# ----------------------------------------------------------------------------------------------
@dataclass
class WindowResult:
    samples: np.ndarray     # [nsp] int64 recording samples
    templates: np.ndarray   # [nsp] int64
    amps: np.ndarray        # [nsp]
    tF: np.ndarray          # [nsp][nC][K]
    xs: np.ndarray
    ys: np.ndarray
    X: torch.Tensor         # preprocessed batch [C][W']
    match: MatchResult


def sort_window(Xraw: torch.Tensor, art: dict, wstart: int, dtype=torch.float64,
                max_passes: int = 100, drop_negative: bool = True, ctc: torch.Tensor | None = None) -> WindowResult:
    """One padded window through the whole path in ``dtype``:
    filter (ks/io.py:757-790) -> run_matching (template_matching.py:128-202) -> features
    (:50-51) -> negative-time drop (:53-59) -> positions (postprocessing.py:33-46) -> samples (:68).

    ``ctc`` (from ``prepare_matching``) may be passed to keep the load-time table out of a timing.
    ``art`` holds numpy arrays: U [T][K][C], wPCA [K][nt], Wrot [C][C], drift [C][C] (or None),
    hp_filter, xc, yc and scalars Th, nt0min, nC.
    """
    cast = lambda a: torch.as_tensor(np.asarray(a)).to(dtype)
    U, wPCA = cast(art["U"]), cast(art["wPCA"])
    Wrot = cast(art["Wrot"])
    Wm = Wrot if art.get("drift") is None else cast(art["drift"]) @ Wrot
    hp = cast(art["hp_filter"])
    nt = wPCA.shape[1]
    X = filter_batch(Xraw.to(dtype), hp, Wm)
    if ctc is None:                       # load-time table (template_matching.py:38); may be given
        ctc = prepare_matching(U, wPCA)
    ctc = ctc.to(dtype)
    m = run_matching(X, U, wPCA, ctc, Th=float(art["Th"]), max_passes=max_passes)
    iCC, iU, Ucc = prepare_extract(np.asarray(art["xc"], np.float64), np.asarray(art["yc"], np.float64),
                                   U, int(art["nC"]))
    tF = extract_features(m.Xres, m.st, m.amps, iCC, iU, Ucc, wPCA)
    samples = spike_samples(m.st, wstart, nt, int(art["nt0min"]))
    keep = samples >= 0 if drop_negative else torch.ones_like(samples, dtype=torch.bool)
    st, amps, tF, samples = m.st[keep], m.amps[keep], tF[keep], samples[keep]
    xs, ys = spike_positions(tF, st, iCC, iU, art["xc"], art["yc"]) if len(st) else (
        torch.zeros(0), torch.zeros(0))
    return WindowResult(samples.numpy(), st[:, 1].numpy(), amps[:, 0].numpy(), tF.numpy(),
                        xs.numpy(), ys.numpy(), X, m)


# ----------------------------------------------------------------------------------------------
# feature-space cluster assignment of LSS's Python online sorter
# (src/Python/OnlineKilosort.py:211-284)
# ----------------------------------------------------------------------------------------------
def cluster_centroids_pca(Wall: np.ndarray, pc_feature_ind: np.ndarray) -> np.ndarray:
    """centroids[t] = Wall[t, pc_feature_ind[t], :] flattened (OnlineKilosort.py:211-231);
    Wall [T][C][P], pc_feature_ind [T][nC] -> [T][nC*P]."""
    T = Wall.shape[0]
    return np.stack([Wall[t, pc_feature_ind[t], :].reshape(-1) for t in range(T)])


def assign_to_clusters(tF: np.ndarray, centroids_pca: np.ndarray, xs, ys, cluster_xy: dict, R: float,
                       dtype=np.float64) -> np.ndarray:
    """OnlineKilosort.py:233-284 in ``dtype``: scores 2 X mu^T - |mu|^2 (X = tF flattened per
    spike), clusters farther than R from the spike (or without a position) masked to -inf, first
    argmax, -1 where no cluster is in range."""
    X = np.asarray(tF, dtype).reshape(len(tF), -1)
    mu = np.asarray(centroids_pca, dtype)
    scores = 2 * (X @ mu.T) - (mu ** 2).sum(1)
    K = mu.shape[0]
    cx = np.full(K, np.nan, dtype)
    cy = np.full(K, np.nan, dtype)
    for k, (x, y) in cluster_xy.items():
        cx[k], cy[k] = x, y
    xs = np.asarray(xs, dtype)
    ys = np.asarray(ys, dtype)
    with np.errstate(invalid="ignore"):
        mask = (xs[:, None] - cx[None, :]) ** 2 + (ys[:, None] - cy[None, :]) ** 2 <= dtype(R) ** 2
    scores = np.where(mask, scores, -np.inf)
    best = np.argmax(scores, axis=1)
    best[~mask.any(axis=1)] = -1
    return best.astype(np.int64)
