"""Seeded synthetic probes, recordings and Kilosort4-style artifacts.

A probe here is a two-column layout (32 um horizontal, 20 um vertical pitch).  Each unit has a
biphasic sum-of-Gaussians temporal shape with its trough at sample ``nt0min`` of an ``nt``-sample
window, a spatial footprint exp(-d/30 um), lognormal amplitudes and a Poisson spike train with a
refractory period.  The artifacts the sorter loads are then built the way Kilosort4 builds them at
training time, in float64 on the host:

* ``hp_filter``: filtfilt impulse response of the 300 Hz 3rd-order Butterworth highpass
  (ks/preprocessing.py:121-136);
* ``Wrot``: ZCA whitening of each channel's ``nrange`` nearest neighbours from the covariance of
  the CAR + highpassed first batch (ks/preprocessing.py:8-35, 96-119);
* ``drift``: the drift interpolation matrix for a vertical shift (ks/preprocessing.py:44-69 with
  iKxx from ks/datashift.py:215-219); zero shift gives Kxx (Kxx + 0.01 I)^-1;
* ``wPCA``: the leading ``n_pcs`` right singular vectors of unit-norm waveform snippets (the
  TruncatedSVD of ks/spikedetect.py:74-75, done here as an FP64 eigen-solve);
* ``U``: whitened unit templates projected on wPCA, layout [T][K][C] (Kilosort's Wall3).

The recording is int16 [n_samples][C] (interleaved, as an acquisition system delivers it).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class SynthConfig:
    n_chan: int = 64
    n_units: int = 32
    n_samples: int = 120000
    NT: int = 60000
    fs: float = 30000.0
    nt: int = 61
    nt0min: int = 20
    n_pcs: int = 6
    nC: int = 10
    noise_sd: float = 10.0
    amp_mu: float = float(np.log(80.0))
    amp_sigma: float = 0.3
    rate_lo: float = 1.0
    rate_hi: float = 20.0
    refractory: int = 60
    footprint_um: float = 30.0
    Th: float = 8.0
    hp_cutoff: float = 300.0
    nrange: int = 32
    sig_interp: float = 20.0
    drift_um: float = 0.0          # vertical probe shift applied to the drift matrix (um)
    drift_rate_um_per_s: float = 0.0  # drift baked into the recording (um per second)
    t_offset_s: float = 0.0        # recording time of sample 0 (the drift clock)
    dup_bins: int = 7              # int(0.25 ms * fs / 1000), ks/run_kilosort.py:306
    seed: int = 0


# Workloads of the benchmark suite (sizes in samples at 30 kHz).
CONFIGS = {
    1: dict(n_chan=64, n_units=32, n_samples=120000, NT=60000),
    2: dict(n_chan=384, n_units=600, n_samples=300000, NT=60000),
    3: dict(n_chan=384, n_units=1000, n_samples=1800000, NT=60000),
    # 2 s at 58-60 s of a probe that drifts linearly 0-20 um over 60 s; the templates are those
    # of drift 0, so the window's units sit ~19-20 um away from them
    4: dict(n_chan=384, n_units=600, n_samples=60000, NT=60000, drift_rate_um_per_s=20.0 / 60.0,
            t_offset_s=58.0),
}


def probe_geometry(n_chan: int):
    """Two columns (x in {0, 32} um), rows every 20 um."""
    idx = np.arange(n_chan)
    xc = 32.0 * (idx % 2)
    yc = 20.0 * (idx // 2)
    return xc.astype(np.float64), yc.astype(np.float64)


def highpass_taps(fs: float = 30000.0, cutoff: float = 300.0) -> np.ndarray:
    """30122-tap filtfilt impulse response, impulse at 15061 (ks/preprocessing.py:121-136)."""
    from scipy.signal import butter, filtfilt
    n = 30122
    b, a = butter(3, cutoff, fs=fs, btype="high")
    x = np.zeros(n)
    x[n // 2] = 1.0
    return filtfilt(b, a, x).copy()


def _temporal_shapes(rng, n_units, nt, nt0min):
    m = np.arange(nt, dtype=np.float64)
    s1 = rng.uniform(1.2, 2.5, n_units)
    s2 = rng.uniform(3.0, 7.0, n_units)
    dl = rng.uniform(5.0, 12.0, n_units)
    r = rng.uniform(0.15, 0.45, n_units)
    sh = (-np.exp(-(m - nt0min) ** 2 / (2 * s1[:, None] ** 2))
          + r[:, None] * np.exp(-(m - nt0min - dl[:, None]) ** 2 / (2 * s2[:, None] ** 2)))
    return sh / np.abs(sh[:, nt0min:nt0min + 1])


def whitening_from_covariance(CC: np.ndarray) -> np.ndarray:
    """ks/preprocessing.py:8-17 (ZCA, eps=1e-6), FP64."""
    E, D, _ = np.linalg.svd(CC)
    return (E / (D + 1e-6) ** 0.5) @ E.T


def whitening_local(CC, xc, yc, nrange=32):
    """ks/preprocessing.py:19-35: per-channel ZCA row over its nrange nearest channels (FP64,
    distance ties broken by channel index)."""
    C = CC.shape[0]
    Wrot = np.zeros((C, C))
    for j in range(C):
        ds = (xc[j] - xc) ** 2 + (yc[j] - yc) ** 2
        ix = np.argsort(ds, kind="stable")[:nrange]
        Wrot[j, ix] = whitening_from_covariance(CC[np.ix_(ix, ix)])[0]
    return Wrot


def drift_matrix(xc, yc, dshift_um: float, sig_interp: float = 20.0):
    """ks/preprocessing.py:44-69 (one block) with iKxx = inv(Kxx + 0.01 I) (ks/datashift.py:215-219)."""
    xp = np.stack((xc, yc), 1)
    yp = xp.copy()
    yp[:, 1] -= dshift_um
    k2 = lambda a, b: np.exp(-((a[:, None] - b) ** 2).sum(-1) / (2 * sig_interp ** 2))
    Kxx = k2(xp, xp)
    iKxx = np.linalg.inv(Kxx + 0.01 * np.eye(len(xc)))
    return k2(yp, xp) @ iKxx


def _car_highpass(xraw: np.ndarray, taps: np.ndarray) -> np.ndarray:
    """CAR + zero-phase highpass of a [C][n] float64 block, for covariance estimation only."""
    from scipy.signal import fftconvolve
    x = xraw - xraw.mean(1, keepdims=True)
    x = x - np.median(x, 0, keepdims=True)
    c = len(taps) // 2
    g = taps[c - 800:c + 801]
    return fftconvolve(x, g[None, ::-1], mode="same", axes=1)


@dataclass
class Synth:
    cfg: SynthConfig
    rec: np.ndarray                 # int16 [n_samples][C]
    gt_times: np.ndarray            # trough sample of every spike
    gt_units: np.ndarray
    gt_amps: np.ndarray
    art: dict = field(default_factory=dict)


def generate(cfg: SynthConfig) -> Synth:
    rng = np.random.default_rng(cfg.seed)
    C, T, nt, n = cfg.n_chan, cfg.n_units, cfg.nt, cfg.n_samples
    xc, yc = probe_geometry(C)
    ux = rng.uniform(-8.0, 40.0, T)
    uy = rng.uniform(0.0, yc.max(), T)
    shapes = _temporal_shapes(rng, T, nt, cfg.nt0min)                 # [T][nt]
    dist = np.sqrt((ux[:, None] - xc) ** 2 + (uy[:, None] - yc) ** 2)  # [T][C]
    foot = np.exp(-dist / cfg.footprint_um)
    foot[foot < 1e-3] = 0.0

    # spike trains: refractory + exponential ISIs at rate U(lo, hi)
    times, units = [], []
    for u in range(T):
        rate = rng.uniform(cfg.rate_lo, cfg.rate_hi)
        mean_isi = cfg.fs / rate
        k = int(n / mean_isi * 1.5) + 10
        isi = cfg.refractory + rng.exponential(max(mean_isi - cfg.refractory, 1.0), k)
        t = np.cumsum(isi).astype(np.int64) + cfg.nt0min
        t = t[(t >= cfg.nt0min) & (t < n - nt)]
        times.append(t)
        units.append(np.full(len(t), u, np.int64))
    times = np.concatenate(times)
    units = np.concatenate(units)
    amps = rng.lognormal(cfg.amp_mu, cfg.amp_sigma, len(times))
    order = np.argsort(times, kind="stable")
    times, units, amps = times[order], units[order], amps[order]

    data = rng.normal(0.0, cfg.noise_sd, (n, C)).astype(np.float32)
    # drift baked into the recording moves each unit vertically over time
    for u in range(T):
        sel = units == u
        tu, au = times[sel], amps[sel]
        if cfg.drift_rate_um_per_s != 0.0:
            for s, a in zip(tu, au):
                y = uy[u] + cfg.drift_rate_um_per_s * (cfg.t_offset_s + s / cfg.fs)
                f = np.exp(-np.sqrt((ux[u] - xc) ** 2 + (y - yc) ** 2) / cfg.footprint_um)
                f[f < 1e-3] = 0.0
                ch = np.nonzero(f)[0]
                data[s - cfg.nt0min:s - cfg.nt0min + nt, ch] += (a * shapes[u][:, None] * f[ch]).astype(np.float32)
            continue
        ch = np.nonzero(foot[u])[0]
        wv = (shapes[u][:, None] * foot[u][ch][None, :]).astype(np.float32)   # [nt][nch]
        base = tu - cfg.nt0min
        for m in range(nt):
            data[np.ix_(base + m, ch)] += au[:, None].astype(np.float32) * wv[m][None, :]
    rec = np.clip(np.rint(data), -32768, 32767).astype(np.int16)
    del data

    art = build_artifacts(cfg, rec, shapes, foot, xc, yc)
    return Synth(cfg, rec, times, units, amps, art)


def build_artifacts(cfg: SynthConfig, rec, shapes, foot, xc, yc) -> dict:
    """Training-side artifacts in FP64, stored as the float32 / int arrays the sorter loads."""
    nt, K = cfg.nt, cfg.n_pcs
    taps = highpass_taps(cfg.fs, cfg.hp_cutoff)
    nb = min(cfg.NT, rec.shape[0])
    xb = _car_highpass(rec[:nb].T.astype(np.float64), taps)
    CC = xb @ xb.T / xb.shape[1]
    Wrot = whitening_local(CC, xc, yc, cfg.nrange)
    M = drift_matrix(xc, yc, cfg.drift_um, cfg.sig_interp)
    Wm = M @ Wrot

    # temporal basis from highpassed, jittered unit shapes (uncentred SVD as TruncatedSVD)
    c = len(taps) // 2
    g = taps[c - 800:c + 801]
    snips = []
    for s in shapes:
        for sh in (-2, -1, 0, 1, 2):
            z = np.zeros(nt + 1600)
            z[800:800 + nt] = np.roll(s, sh)
            f = np.convolve(z, g[::-1], mode="same")[800:800 + nt]
            snips.append(f / np.linalg.norm(f))
    S = np.asarray(snips)
    evals, evecs = np.linalg.eigh(S.T @ S)
    wPCA = evecs[:, ::-1][:, :K].T.copy()                        # [K][nt]
    for k in range(K):
        if wPCA[k, np.argmax(np.abs(wPCA[k]))] < 0:
            wPCA[k] = -wPCA[k]

    # whitened templates at the median amplitude, projected on wPCA -> U [T][K][C]
    # the raw template amp0 * shape[m] * foot[c] is separable, so its highpassed, whitened PC
    # projection is amp0 * (wPCA @ hp(shape))[k] * (Wm @ foot)[c]
    amp0 = np.exp(cfg.amp_mu)
    T, C = shapes.shape[0], len(xc)
    U = np.zeros((T, K, C))
    for u in range(T):
        z = np.zeros(nt + 1600)
        z[800:800 + nt] = shapes[u]
        hp = np.convolve(z, g[::-1], mode="same")[800:800 + nt]     # [nt]
        U[u] = amp0 * np.outer(wPCA @ hp, Wm @ foot[u])              # [K][C]

    return dict(
        U=U.astype(np.float32), wPCA=wPCA.astype(np.float32), Wrot=Wrot.astype(np.float32),
        drift=M.astype(np.float32), hp_filter=taps.astype(np.float32),
        xc=xc.astype(np.float32), yc=yc.astype(np.float32),
        Th=float(cfg.Th), nt0min=int(cfg.nt0min), nC=int(cfg.nC), dup_bins=int(cfg.dup_bins),
        fs=float(cfg.fs), NT=int(cfg.NT), nt=int(nt))


def make(config_id: int = 1, seed: int = 0, **over) -> Synth:
    kw = dict(CONFIGS[config_id])
    kw.update(over)
    return generate(SynthConfig(seed=seed, **kw))
